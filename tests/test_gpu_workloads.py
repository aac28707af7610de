"""The workloads of the measurement at their own sizes and horizons, where the
oracle is far too slow to run: size-independent properties of the HIP path.

* weak plume, 1M real particles, FP64 form, 60 steps (through the first
  adaptive-h update at step 50): particle and mass bookkeeping, keys unique
  and in THashTable order, device reductions equal exactly rounded host sums;
  the FP32-pair form evaluates the same pairs every step and stays within
  1e-5 relative L1 of the FP64 form on every field;
* weak plume, 4M real particles, 500 steps: per-field relative L1 of the
  FP32 forms against the FP64 form after 100 and 500 steps, next to the
  chaos floors (the FP64 form against itself with every initial coordinate
  moved by one FP64 ulp, and by one FP32 ulp);
* adjoint, weak plume at 512k real particles, 200 steps, a checkpoint every
  20: the replay reproduces the recorded H, one segment (no recompute) gives
  the same gradients, and backward / forward <= 3;
* strong plume, 16M real particles, FP32-pair: 100 forward steps keep the
  bookkeeping invariants, and the adjoint over the last 50 steps gives the
  same gradients at checkpoint intervals 10 and 6.
"""
import math
import sys
import time

import numpy as np
import pytest

import oracle
import psph
from helpers import weak_plume_config

pytestmark = pytest.mark.gpu

SEED = 20261204
FIELDS = ("x", "y", "z", "vx", "vy", "vz", "rho", "e", "p", "c", "T", "xi", "h")


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def sync():
    import torch
    torch.cuda.synchronize()


def strong_plume_config(n_real):
    cfg = psph.default_config(3)
    lx, ly, lz = 40000.0, 40000.0, 30000.0
    cfg.h_air = (lx * ly * lz / n_real) ** (1.0 / 3.0)
    for k, (lo, hi) in enumerate([(-lx / 2, lx / 2), (-ly / 2, ly / 2), (0.0, lz)]):
        cfg.dom_lo[k] = lo
        cfg.dom_hi[k] = hi
    cfg.vent_z = 0.0
    cfg.vent_r, cfg.vent_v, cfg.vent_T, cfg.ng0 = 700.0, 150.0, 1273.0, 0.05
    cfg.vent_p = 101325.0
    psph.finalize_config(cfg)
    return cfg


def rel_l1_by_key(A, B):
    """Per-field relative L1 of A against B over the particles with the same
    key triple in both (the inlet schedule is the same; the runs may differ in
    the few particles that escaped)."""
    na, nb = len(A["key0"]), len(B["key0"])
    k = [np.concatenate([A[f], B[f]]) for f in ("key0", "key1", "key2")]
    src = np.concatenate([np.zeros(na, bool), np.ones(nb, bool)])
    idx = np.concatenate([np.arange(na), np.arange(nb)])
    o = np.lexsort((src, k[2], k[1], k[0]))
    k, src, idx = [x[o] for x in k], src[o], idx[o]
    same = (k[0][1:] == k[0][:-1]) & (k[1][1:] == k[1][:-1]) & (k[2][1:] == k[2][:-1]) & ~src[:-1] & src[1:]
    ia, ib = idx[:-1][same], idx[1:][same]
    assert len(ia) >= 0.999 * max(na, nb), (len(ia), na, nb)
    out = {}
    for f in FIELDS:
        a, b = A[f][ia], B[f][ib]
        den = np.abs(b).sum()
        out[f] = float(np.abs(a - b).sum() / den) if den > 0 else 0.0
    return out


def check_invariants(cfg, P, Q, outs, cons, H):
    inj = sum(o.n_injected for o in outs)
    rem = sum(o.n_removed for o in outs)
    assert inj > 0
    assert len(Q) == len(P) + inj - rem
    assert outs[-1].n_total == len(Q)
    assert outs[-1].n_real == np.isin(Q["bc"], (100, 99)).sum()
    m_want = math.fsum(P["m"]) + inj * cfg.vent_mass
    if rem == 0:
        assert abs(math.fsum(Q["m"]) - m_want) <= 1e-13 * m_want
    k = (Q["key0"].astype(np.uint64) << np.uint64(32)) | Q["key1"].astype(np.uint64)
    kk = np.stack([k, Q["key2"].astype(np.uint64)], 1)
    assert len(np.unique(kk, axis=0)) == len(Q)
    rng = np.random.default_rng(3)
    for i in rng.integers(0, len(Q) - 1, 5000):
        a = (oracle.slot(int(Q["key0"][i]), int(Q["key1"][i]), cfg.p_table_size),
             int(Q["key0"][i]), int(Q["key1"][i]), int(Q["key2"][i]))
        b = (oracle.slot(int(Q["key0"][i + 1]), int(Q["key1"][i + 1]), cfg.p_table_size),
             int(Q["key0"][i + 1]), int(Q["key1"][i + 1]), int(Q["key2"][i + 1]))
        assert a < b, (i, a, b)
    assert set(np.unique(Q["bc"])) <= {0, 1, 2, 99, 100}
    assert set(np.unique(Q["phase"])) <= {1, 2}
    r = np.isin(Q["bc"], (100, 99))
    m, v = Q["m"][r], [Q["vx"][r], Q["vy"][r], Q["vz"][r]]
    host = [math.fsum(m), math.fsum(m * Q["e"][r]),
            math.fsum(0.5 * m * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)),
            math.fsum(m * v[0]), math.fsum(m * v[1]), math.fsum(m * v[2]), math.fsum(m * Q["xi"][r])]
    scale = [abs(x) for x in host]
    scale[3:6] = [math.fsum(m * np.abs(c)) for c in v]
    for j, (g, h_, s_) in enumerate(zip(cons, host, scale)):
        assert abs(g - h_) <= 1e-12 * s_, (j, g, h_)
    mxi = m * Q["xi"][r]
    H_host = math.fsum(mxi * Q["z"][r]) / math.fsum(mxi)
    assert abs(H - H_host) <= 1e-12 * abs(H_host)
    return inj, rem


def run_form(cfg, P, precision, steps, cap_extra=400000):
    c = cfg.copy()
    c.precision = precision
    with psph.Solver(c, len(P) + cap_extra) as G:
        G.load(P)
        outs = [G.step() for _ in range(steps)]
        Q = G.read()
        cons = G.conservation()
        H = G.diag_height()
    return Q, outs, cons, H


def test_weak_plume_1M_fp64_and_fp32_pair():
    cfg = weak_plume_config(1.0e6)
    P = psph.gen_plume(cfg, psph.meteo_standard(), SEED, 0.01)
    say(f"weak plume 1M: {len(P)} particles, spacing {cfg.h_air:.2f} m; FP64 60 steps")
    Q64, o64, cons, H = run_form(cfg, P, psph.PSPH_FP64, 60)
    inj, rem = check_invariants(cfg, P, Q64, o64, cons, H)
    say(f"FP64: injected {inj}, removed {rem}, H = {H:.6g} m, pairs/step {o64[-1].n_pairs}; FP32-pair 60 steps")
    Q32, o32, _, _ = run_form(cfg, P, psph.PSPH_FP32_PAIR, 60)
    # the same pairs while the two forms' positions agree to the last bits
    assert [o.n_pairs for o in o32[:2]] == [o.n_pairs for o in o64[:2]]
    err = rel_l1_by_key(Q32, Q64)
    say("FP32-pair vs FP64 after 60 steps: " + " ".join(f"{k}={v:.1e}" for k, v in err.items()))
    assert max(err.values()) <= 1e-5, err


def test_weak_plume_4M_500_steps_against_chaos_floor():
    """Per-field relative L1 against the FP64 form after 100 and 500 steps,
    next to two floors: the FP64 form started with every initial coordinate
    moved by one FP64 ulp, and by one FP32 ulp (the size of the rounding an
    FP32 pair term carries).  The plume is chaotic: any difference grows
    until it saturates, so at 500 steps the FP32 forms are judged against the
    FP64 run's own sensitivity to a perturbation of FP32 size.

    Measured (MI355X, vx): after 100 steps FP32-pair 3.6e-7, FP32-exp 8.6e-5,
    FP64 + FP32 ulp 2.6e-4, FP64 + FP64 ulp 4.3e-13; after 500 steps 0.13,
    0.18, 0.26 and 2.9e-6.  The FP32-pair form holds the 1e-5 bar at 100
    steps; the FP32-exp form (FP32 exponentials) holds it only over shorter
    windows.  Both stay within the FP32-size floor at both horizons; xi,
    saturated at 500 steps, is the closest (0.57 and 0.43 against 0.53: the
    floor is one realisation of a chaotic quantity, hence the 25 % margin)."""
    cfg = weak_plume_config(4.0e6)
    P = psph.gen_plume(cfg, psph.meteo_standard(), SEED, 0.01)

    def moved(ulp_of):
        Q = P.copy()
        for f in ("x", "y", "z"):
            a = P[f]
            if ulp_of == np.float64:
                Q[f][:] = np.nextafter(a, np.inf)
            else:
                Q[f][:] = a + np.abs(np.spacing(a.astype(np.float32)).astype(np.float64))
        return Q

    runs = (("fp64", psph.PSPH_FP64, P), ("fp64_ulp64", psph.PSPH_FP64, moved(np.float64)),
            ("fp64_ulp32", psph.PSPH_FP64, moved(np.float32)), ("fp32_exp", psph.PSPH_FP32_EXP, P),
            ("fp32_pair", psph.PSPH_FP32_PAIR, P))
    res = {100: {}, 500: {}}
    for name, prec, inp in runs:
        say(f"weak plume 4M, {name}: 500 steps")
        t0 = time.perf_counter()
        c = cfg.copy()
        c.precision = prec
        with psph.Solver(c, len(inp) + 600000) as G:
            G.load(inp)
            G.run(100)
            res[100][name] = G.read()
            G.run(400)
            res[500][name] = G.read()
        say(f"{name}: {time.perf_counter() - t0:.1f} s")
    err = {h: {k: rel_l1_by_key(res[h][k], res[h]["fp64"]) for k in res[h] if k != "fp64"} for h in res}
    for h in (100, 500):
        say(f"after {h} steps: field " + " ".join(f"{k:>11s}" for k in err[h]))
        for f in FIELDS:
            say(f"after {h} steps: {f:5s} " + " ".join(f"{err[h][k][f]:11.3e}" for k in err[h]))
    for f in FIELDS:
        # 100 steps: the FP32-pair form within the 1e-5 bar
        assert err[100]["fp32_pair"][f] <= 1e-5, (100, "fp32_pair", f, err[100]["fp32_pair"][f])
        # both forms no further from the FP64 run than the FP64 run itself
        # moves for a one-time FP32-size change of its input (25% margin)
        for h in (100, 500):
            for k in ("fp32_exp", "fp32_pair"):
                assert err[h][k][f] <= 1.25 * err[h]["fp64_ulp32"][f], (h, k, f, err[h][k][f])


def test_adjoint_512k_200_steps_checkpoint_20():
    """The adjoint configuration: 512k real particles, 200 steps, a checkpoint
    every 20; backward at most 3x the plain forward.

    Central differences through the frozen replay: over the first 20 steps
    the 16 seeded directions at eps 1e-6 agree to 1e-4 (measured 1.9e-6 and
    7.7e-6 in two builds).  Over the full 200 steps the map is chaotic
    (dH/d ln v_vent ~ 3e6 m): the two-point quotient leaves the linear range
    at eps 1e-6 (3e-2 on the same directions) and FP64 rounding of H,
    amplified like any perturbation, takes over below eps ~1e-9 (0.11 at
    1e-10, 0.35 at 1e-11).  At the bottom of that curve a four-point stencil
    at eps 3e-9 agrees to 1e-2 (measured 8.4e-4 and 1.8e-3: the difference
    quotient's own error there is one realisation of the amplified
    rounding, and moves with any change to the rounding of the step), and
    the vent velocity alone, with the largest signal, at eps 1e-10 to 1e-4
    (measured 1.3e-5 and 6.1e-6)."""
    from psph.torch_adjoint import build_state
    cfg = weak_plume_config(5.12e5)
    met = psph.meteo_standard()
    theta0 = (cfg.vent_v, cfg.vent_T, 1.0 - cfg.ng0)
    theta = np.array(theta0)
    _, P = build_state(cfg, met, SEED, 0.01, *theta0, None)
    rng = np.random.default_rng(SEED + 1000)
    dirs = []
    for d in range(16):
        v = rng.standard_normal(3 + 3 * len(P))
        dirs.append(v / np.linalg.norm(v))
    N = 200
    with psph.Solver(cfg, len(P) + 400 * N + 50000) as S:

        def H(n, eps, v):
            dth, dv = v[:3], v[3:].reshape(len(P), 3)
            c, Pp = build_state(cfg, met, SEED, 0.01, *(theta * (1 + eps * dth)), eps * theta[0] * dv)
            S.set_vent(c.vent_v, c.vent_T, c.ng0)
            S.load(Pp)
            S.replay(n)
            return S.diag_height()

        def ad_dir(r, v):
            g = np.array([r["dH_dvent_v"], r["dH_dvent_T"], r["dH_dmass_frac"]])
            return float(np.sum(v[:3] * theta * g) + theta[0] * np.sum(v[3:].reshape(len(P), 3) * r["grad_v"]))

        def reset():
            # the config's own vent state (1 - (1 - ng0) is not ng0 to the last bit)
            S.set_vent(cfg.vent_v, cfg.vent_T, cfg.ng0)
            S.load(P)

        S.load(P)
        S.step()                      # first-touch warm-up
        S.load(P)
        sync()
        t0 = time.perf_counter()
        S.run(N)
        sync()
        fwd_ms = 1e3 * (time.perf_counter() - t0)
        H_plain = S.diag_height()
        say(f"adjoint 512k: {len(P)} particles; plain forward {N} steps {fwd_ms:.0f} ms; adjoint K=20")
        # 20 steps: the 16 directions at eps 1e-6
        reset()
        r20 = S.adjoint_run(20, 20, grads=True)
        worst20 = 0.0
        for d, v in enumerate(dirs):
            fdv = (H(20, 1e-6, v) - H(20, -1e-6, v)) / 2e-6
            worst20 = max(worst20, abs(ad_dir(r20, v) - fdv) / abs(fdv))
        say(f"20 steps, 16 directions at eps 1e-6: worst relative difference {worst20:.2e}")
        assert worst20 <= 1e-4, worst20
        # 200 steps
        reset()
        r = S.adjoint_run(N, 20, grads=True)
        say(f"K=20: H={r['H']:.9f} dH/dv={r['dH_dvent_v']:.6e} dH/dT={r['dH_dvent_T']:.6e} "
            f"dH/dmf={r['dH_dmass_frac']:.6e} backward {r['backward_ms']:.0f} ms")
        assert r["H"] == H_plain
        reset()
        S.replay(N)
        assert abs(S.diag_height() - r["H"]) <= 1e-9 * abs(r["H"])
        e_v = np.zeros(3 + 3 * len(P))
        e_v[0] = 1.0
        fdv = (H(N, 1e-10, e_v) - H(N, -1e-10, e_v)) / 2e-10
        rel_v = abs(ad_dir(r, e_v) - fdv) / abs(fdv)
        say(f"200 steps, vent velocity at eps 1e-10: AD {ad_dir(r, e_v):.9e} FD {fdv:.9e} rel {rel_v:.2e}")
        eps, worst = 3e-9, 0.0
        for d, v in enumerate(dirs):
            fd4 = (8 * (H(N, eps, v) - H(N, -eps, v)) - (H(N, 2 * eps, v) - H(N, -2 * eps, v))) / (12 * eps)
            rel = abs(ad_dir(r, v) - fd4) / abs(fd4)
            worst = max(worst, rel)
            say(f"direction {d}: AD {ad_dir(r, v):.9e} FD(4-point, eps 3e-9) {fd4:.9e} rel {rel:.2e}")
        say(f"200 steps, 16 directions, four-point stencil at eps 3e-9: worst relative difference {worst:.2e}")
        assert rel_v <= 1e-4, rel_v
        assert worst <= 1e-2, worst
        say("adjoint K=200 (one segment, no recompute)")
        reset()
        r1 = S.adjoint_run(N, N, grads=True)
    for k in ("H", "dH_dvent_v", "dH_dvent_T", "dH_dmass_frac"):
        assert abs(r1[k] - r[k]) <= 1e-12 * abs(r[k]), (k, r1[k], r[k])
    g, g0 = r["grad_v"], r1["grad_v"]
    assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max()
    ratio = r["backward_ms"] / fwd_ms
    say(f"backward / forward = {ratio:.2f}")
    assert ratio <= 3.0, ratio


def test_strong_plume_16M_forward_and_adjoint():
    cfg = strong_plume_config(16.0e6)
    cfg.precision = psph.PSPH_FP32_PAIR
    say("strong plume 16M: generating on the host")
    P = psph.gen_plume(cfg, psph.meteo_standard(), 20261206, 0.01)
    say(f"{len(P)} particles, spacing {cfg.h_air:.1f} m; FP32-pair 100 steps")
    cap = int(len(P) * 1.2) + 200000
    with psph.Solver(cfg, cap) as S:
        S.load(P)
        t0 = time.perf_counter()
        outs = [S.step() for _ in range(50)]
        t50, step50 = S.time()
        Q50 = S.read()
        S.load(Q50, t50, step50)
        outs += [S.step() for _ in range(50)]
        sync()
        say(f"100 steps in {time.perf_counter() - t0:.1f} s, pairs/step {outs[-1].n_pairs}")
        Q = S.read()
        cons = S.conservation()
        H = S.diag_height()
        inj, rem = check_invariants(cfg, P, Q, outs, cons, H)
        del P, Q
        say(f"injected {inj}, removed {rem}, H = {H:.6g} m")
    # the adjoint over steps 51-100, from the state the forward run held; each
    # checkpoint interval in a context of its own (a context keeps its
    # adjoint's records and checkpoints for the next run)
    res = {}
    for K in (10, 6):
        with psph.Solver(cfg, cap) as S:
            S.load(Q50, t50, step50)
            t0 = time.perf_counter()
            res[K] = S.adjoint_run(50, K, grads=False)
            say(f"adjoint steps {step50 + 1}-{step50 + 50}, K={K}: H={res[K]['H']:.9f} "
                f"dH/dv={res[K]['dH_dvent_v']:.9e} dH/dT={res[K]['dH_dvent_T']:.9e} "
                f"dH/dmf={res[K]['dH_dmass_frac']:.9e} ({time.perf_counter() - t0:.1f} s)")
    assert res[10]["H"] == H
    for k in ("H", "dH_dvent_v", "dH_dvent_T", "dH_dmass_frac"):
        assert abs(res[6][k] - res[10][k]) <= 1e-12 * abs(res[10][k]), (k, res[6][k], res[10][k])
