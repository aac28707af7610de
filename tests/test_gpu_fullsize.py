"""The weak plume at its full size (4M real particles, 6.8M with ghosts and
the vent column), where the oracle is too slow to run: size-independent
properties of the HIP path.

* load/read round trip: the generated state comes back bit for bit, in the
  THashTable order;
* after three steps: the particle count and total mass follow the inlet's and
  the escape's counts exactly; the keys are unique and in THashTable order
  (slot, then key words; slot from the oracle on a sample of neighbours);
  bc/phase values stay in their domains; the device reductions
  (psph_conservation, psph_diag_height) equal exactly-rounded host sums of the
  read-back state to 1e-12;
* the FP64 and FP32-pair forms evaluate the same kernel-support pairs, step by
  step, and stay within 1e-6 relative L1 of each other;
* the measurement's window (steps 56-110 after a 55-step FP64 pre-roll, wall
  refresh on): the FP32-exp form evaluates the same pairs as the FP64 form in
  its first step (within 1e-7 later, as positions drift) and stays within 1e-5
  relative L1 of it on every field.
"""
import math
import sys

import numpy as np
import pytest

import oracle
import psph
from helpers import weak_plume_config

pytestmark = pytest.mark.gpu

N_REAL = 4.0e6
SEED = 20261204


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def plume():
    cfg = weak_plume_config(N_REAL)
    say("full-size weak plume: generating on the host")
    P = psph.gen_plume(cfg, psph.meteo_standard(), SEED, 0.01)
    say(f"generated {len(P)} particles")
    return cfg, P


def run(cfg, P, precision, steps):
    c = cfg.copy()
    c.precision = precision
    outs = []
    with psph.Solver(c, len(P) + 200000) as G:
        G.load(P)
        Q0 = G.read()
        for s in range(steps):
            outs.append(G.step())
        Q = G.read()
        cons = G.conservation()
        H = G.diag_height()
    return Q0, Q, outs, cons, H


def test_fullsize_invariants(plume):
    cfg, P = plume
    say("FP64, 3 steps")
    Q0, Q, outs, cons, H = run(cfg, P, psph.PSPH_FP64, 3)
    for f in psph.ALL_FIELDS:
        assert np.array_equal(Q0[f], P[f]), f
    inj = sum(o.n_injected for o in outs)
    rem = sum(o.n_removed for o in outs)
    assert inj > 0
    assert len(Q) == len(P) + inj - rem
    assert outs[-1].n_total == len(Q)
    assert outs[-1].n_real == np.isin(Q["bc"], (100, 99)).sum()
    if rem == 0:
        m_want = math.fsum(P["m"]) + inj * cfg.vent_mass
        assert abs(math.fsum(Q["m"]) - m_want) <= 1e-13 * m_want
    # keys unique and in THashTable order
    k = (Q["key0"].astype(np.uint64) << np.uint64(32)) | Q["key1"].astype(np.uint64)
    kk = np.stack([k, Q["key2"].astype(np.uint64)], 1)
    assert len(np.unique(kk, axis=0)) == len(Q)
    rng = np.random.default_rng(3)
    idx = rng.integers(0, len(Q) - 1, 20000)
    for i in idx:
        a = (oracle.slot(int(Q["key0"][i]), int(Q["key1"][i]), cfg.p_table_size),
             int(Q["key0"][i]), int(Q["key1"][i]), int(Q["key2"][i]))
        b = (oracle.slot(int(Q["key0"][i + 1]), int(Q["key1"][i + 1]), cfg.p_table_size),
             int(Q["key0"][i + 1]), int(Q["key1"][i + 1]), int(Q["key2"][i + 1]))
        assert a < b, (i, a, b)
    assert set(np.unique(Q["bc"])) <= {0, 1, 2, 99, 100}
    assert set(np.unique(Q["phase"])) <= {1, 2}
    # device reductions against exactly rounded host sums
    r = np.isin(Q["bc"], (100, 99))
    m, v = Q["m"][r], [Q["vx"][r], Q["vy"][r], Q["vz"][r]]
    host = [math.fsum(m), math.fsum(m * Q["e"][r]),
            math.fsum(0.5 * m * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)),
            math.fsum(m * v[0]), math.fsum(m * v[1]), math.fsum(m * v[2]), math.fsum(m * Q["xi"][r])]
    scale = [abs(x) for x in host]
    scale[3:6] = [math.fsum(m * np.abs(c)) for c in v]          # momentum sums cancel
    for k_, (g, h_, s_) in enumerate(zip(cons, host, scale)):
        assert abs(g - h_) <= 1e-12 * s_, (k_, g, h_)
    mxi = m * Q["xi"][r]
    H_host = math.fsum(mxi * Q["z"][r]) / math.fsum(mxi)
    assert abs(H - H_host) <= 1e-12 * abs(H_host)
    say(f"injected {inj}, removed {rem}, H = {H:.6g} m")


def test_fullsize_forms_same_pairs(plume):
    cfg, P = plume
    say("FP64 and FP32-pair, 2 steps each")
    _, Q64, o64, _, _ = run(cfg, P, psph.PSPH_FP64, 2)
    _, Q32, o32, _, _ = run(cfg, P, psph.PSPH_FP32_PAIR, 2)
    assert [o.n_pairs for o in o64] == [o.n_pairs for o in o32]
    assert all(o.n_pairs > 100 * len(P) for o in o64)
    assert np.array_equal(Q64["key0"], Q32["key0"]) and np.array_equal(Q64["bc"], Q32["bc"])
    worst = 0.0
    for f in ("x", "y", "z", "vx", "vy", "vz", "rho", "e", "p", "xi", "h"):
        a, b = Q32[f], Q64[f]
        d = float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))
        worst = max(worst, d)
        assert d <= 1e-6, (f, d)
    say(f"pairs per step {[o.n_pairs for o in o64]}, worst rel L1 {worst:.2e}")


def _rel_l1_by_key(A, B, fields):
    """Per-field relative L1 of A against B over the particles present in both
    (matched by their Hilbert key words; shared keys dropped)."""
    def key(P):
        k = P["key0"].astype(np.uint64) << np.uint64(32) | P["key1"].astype(np.uint64)
        u, idx, cnt = np.unique(k, return_index=True, return_counts=True)
        keep = idx[cnt == 1]
        return k[keep], keep
    ka, sa = key(A)
    kb, sb = key(B)
    common, ia, ib = np.intersect1d(ka, kb, assume_unique=True, return_indices=True)
    ia, ib = sa[ia], sb[ib]
    out = {}
    for f in fields:
        a, b = A[f][ia], B[f][ib]
        out[f] = float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))
    return len(common), out


def test_fullsize_fp32_exp_bench_window(plume):
    """The measurement's window and its accuracy bar: 55 FP64 steps with the
    wall-ghost refresh (through the step-50 h adaptation), then steps 56-110 of
    the FP64 and of the FP32-exp form from that state (one more adaptation at
    step 100): the same pairs in the first step and within 1e-7 after, and
    every field within 1e-5 relative L1 of the FP64 form (measured: 3.9e-7 on
    vx, 2.3e-8 on rho)."""
    cfg, P = plume
    base = cfg.copy()
    base.wall_bc = 1
    cap = len(P) + 400 * 110 + 100000
    c = base.copy()
    c.precision = psph.PSPH_FP64
    say("bench window: 55 FP64 steps")
    with psph.Solver(c, cap) as G:
        G.load(P)
        G.apply_bcond()
        G.run(55)
        Q = G.read()
        t0, s0 = G.time()
    res = {}
    for prec in (psph.PSPH_FP64, psph.PSPH_FP32_EXP):
        c = base.copy()
        c.precision = prec
        say(f"steps 56-110, precision {prec}")
        with psph.Solver(c, cap) as G:
            G.load(Q, t0, s0)
            outs = [G.step() for _ in range(55)]
            res[prec] = (G.read(), outs)
    (A, oa), (B, ob) = res[psph.PSPH_FP32_EXP], res[psph.PSPH_FP64]
    # the same pair predicate on positions that drift apart by ~1e-10: a pair
    # on a kernel-box face can flip after some steps (2 of 2.9e9 at step 66)
    dp = max(abs(a.n_pairs - b.n_pairs) / b.n_pairs for a, b in zip(oa, ob))
    say(f"pairs per step: largest relative difference {dp:.1e}")
    assert oa[0].n_pairs == ob[0].n_pairs and dp <= 1e-7
    fields = ("x", "y", "z", "vx", "vy", "vz", "rho", "e", "p", "c", "T", "xi", "h")
    n, err = _rel_l1_by_key(A, B, fields)
    say(f"compared {n} of {len(B)} particles; " + " ".join(f"{f}={v:.1e}" for f, v in err.items()))
    assert n >= 0.999 * len(B)
    for f, v in err.items():
        assert v <= 1e-5, (f, v)
