#!/usr/bin/env python3
"""Throughput of the MI355X SPH particle step on the weak dusty-gas plume.

One "step" is one pass of the reference's per-step hot path (timestep ->
neighbour search -> [adaptive h every 50 steps] -> momentum/energy RHS ->
density/mass fraction -> velocity filter -> position update -> inlet) over
the whole single-GPU domain.  Inputs are synthetic (seeded lattice + 1%
jitter, standard atmosphere, vent at the origin) and resident in HBM before
the timed region.  Prints one JSON line (see the repository's driver contract).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--particles N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "plume-sph_amd"))

import numpy as np  # noqa: E402

import psph  # noqa: E402

METRIC = ("particle-updates/s FP32-pair vs FP64 GPU at 1e-5 rel L1; adjoint/forward time")
HBM_PEAK_GBS = 8000.0            # MI355X HBM3E
FP64_VECTOR_PEAK_TFS = 78.6      # MI355X FP64 vector
FP32_VECTOR_PEAK_TFS = 157.3     # MI355X FP32 vector
# Algorithmic HBM bytes per particle-update of the RHS pass: read x,sv,v (72),
# e,rho,p,c,m,h,T,xi (64), flags (4); write v,e,p,c,T (56) -> 196 B.
RHS_BYTES_PER_UPDATE = 196.0
# Algorithmic flops per kernel-support pair in the RHS pass (mul/add/div = 1,
# FMA = 2, exp = 1).
RHS_FLOPS_PER_PAIR = 116.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def plume_config(n_real: float, dims, vent):
    """Plume box dims = (lx, ly, lz) m with N real particles (spacing
    (V/N)^(1/3)); vent = (radius m, velocity m/s, T K, gas fraction), p = p(0)."""
    lx, ly, lz = dims
    cfg = psph.default_config(3)
    cfg.h_air = (lx * ly * lz / n_real) ** (1.0 / 3.0)
    for k, (lo, hi) in enumerate([(-lx / 2, lx / 2), (-ly / 2, ly / 2), (0.0, lz)]):
        cfg.dom_lo[k] = lo
        cfg.dom_hi[k] = hi
    cfg.vent_z = 0.0
    cfg.vent_r, cfg.vent_v, cfg.vent_T, cfg.ng0 = vent
    cfg.vent_p = 101325.0
    psph.finalize_config(cfg)
    return cfg


def weak_plume_config(n_real: float, lx=8000.0, ly=8000.0, lz=12000.0):
    """Weak plume: 8 x 8 x 12 km, vent r = 54 m, 135 m/s, 1053 K, erupted-material
    gas fraction 0.03."""
    return plume_config(n_real, (lx, ly, lz), (54.0, 135.0, 1053.0, 0.03))


def strong_plume_config(n_real: float):
    """Strong plume: 40 x 40 x 30 km, vent r = 700 m, 150 m/s, 1273 K, erupted-
    material gas fraction 0.05."""
    return plume_config(n_real, (40000.0, 40000.0, 30000.0), (700.0, 150.0, 1273.0, 0.05))


def strong_leg(args):
    """Strong plume at scale: FP32-pair forward for 100 steps, then the adjoint
    of the plume height over the last 50 with the checkpoint interval chosen
    so that one segment of step records fills at most 30 % of the free HBM."""
    sys.path.insert(0, os.path.join(ROOT, "plume-sph_amd"))
    import torch
    cfg = strong_plume_config(args.strong_particles)
    cfg.precision = psph.PSPH_FP32_PAIR
    met = psph.meteo_standard()
    t0 = time.perf_counter()
    P = psph.gen_plume(cfg, met, 20261206, 0.01)
    n0 = len(P)
    log(f"[bench] strong: {n0} particles, spacing {cfg.h_air:.1f} m, generated in {time.perf_counter() - t0:.1f}s")
    S = psph.Solver(cfg, int(n0 * 1.2) + 200000)
    S.load(P)
    del P
    first = args.strong_steps - args.strong_adjoint
    sync()
    t0 = time.perf_counter()
    S.run(first)
    sync()
    w_first = time.perf_counter() - t0
    t_sim, step = S.time()
    Q = S.read()
    _, n_real_mid = S.count()
    S.load(Q, t_sim, step)
    sync()
    t0 = time.perf_counter()
    S.run(args.strong_adjoint)
    sync()
    w_last = time.perf_counter() - t0
    n_total, n_real = S.count()
    log(f"[bench] strong: steps 1-{first} in {w_first:.1f}s, steps {first + 1}-{args.strong_steps} in "
        f"{w_last:.1f}s ({1e3 * w_last / args.strong_adjoint:.1f} ms/step, {n_real} real of {n_total})")
    free, _ = torch.cuda.mem_get_info()
    # a step record holds 30 doubles and 6 words per particle of capacity; the
    # adjoint's own arrays, visit lists and checkpoints take much of the rest
    per_step = 264.0 * S.capacity
    # the segment of step records takes at most 30 % of the free HBM: the
    # reverse sweep's two-role visit masks (512 B per visit, ~100 GB at 16M)
    # and the checkpoints need the rest
    K = int(max(1, min(args.strong_adjoint, 0.30 * free / per_step)))
    if args.strong_ckpt > 0:
        K = min(K, args.strong_ckpt)
    log(f"[bench] strong adjoint: {free / 1e9:.0f} GB free, {per_step / 1e9:.1f} GB per step record, "
        f"checkpoint every {K}")
    S.load(Q, t_sim, step)
    del Q
    r = S.adjoint_run(args.strong_adjoint, K, grads=False)
    S.close()
    log(f"[bench] strong adjoint: {args.strong_adjoint} steps, checkpoint every {K}: recorded forward "
        f"{r['forward_ms']:.0f} ms, backward {r['backward_ms']:.0f} ms")
    return {"workload": f"strong plume, 40x40x30 km, {args.strong_particles / 1e6:g}M real particles, FP32-pair "
                        f"forward {args.strong_steps} steps, adjoint over the last {args.strong_adjoint}",
            "particles_total": n_total, "particles_real": n_real, "spacing_m": cfg.h_air,
            "forward_ms_per_step_last": 1e3 * w_last / args.strong_adjoint,
            "forward_ms_per_step_first": 1e3 * w_first / first,
            "particle_updates_per_s_last": 0.5 * (n_real_mid + n_real) * args.strong_adjoint / w_last,
            "adjoint_checkpoint_every": K, "hbm_free_gb_at_adjoint": free / 1e9,
            "H": r["H"], "dH_dvent_v": r["dH_dvent_v"], "dH_dvent_T": r["dH_dvent_T"],
            "dH_dmass_frac": r["dH_dmass_frac"], "adjoint_recorded_forward_ms": r["forward_ms"],
            "adjoint_backward_ms": r["backward_ms"],
            "backward_over_forward": r["backward_ms"] / (1e3 * w_last)}


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(kernel, {}).get("hbm_bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(n_sample: float, steps: int, spacing: float, n_preroll: int, wall_bc: int):
    """The FP64 CPU restatement (oracle) on a bounded sample of the same
    workload: same lattice spacing, vent and atmosphere, smaller domain, timed
    over the same steps as the GPU window (the sample's first n_preroll steps,
    through the first h adaptation, run in the GPU's FP64 form, which equals
    the restatement to 1e-10)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    side = (n_sample * spacing ** 3 / 1.5) ** (1.0 / 3.0)
    cfg = psph.default_config(3)
    cfg.h_air = spacing
    for k, (lo, hi) in enumerate([(-side / 2, side / 2), (-side / 2, side / 2), (0.0, 1.5 * side)]):
        cfg.dom_lo[k] = lo
        cfg.dom_hi[k] = hi
    cfg.vent_z = 0.0
    cfg.vent_r, cfg.vent_v, cfg.vent_T, cfg.ng0, cfg.vent_p = 54.0, 135.0, 1053.0, 0.03, 101325.0
    cfg.wall_bc = wall_bc
    psph.finalize_config(cfg)
    P = psph.gen_plume(cfg, psph.meteo_standard(), 20261204, 0.01)
    P, t_sim, step = preroll(cfg, P, n_preroll)
    n_real = int(np.count_nonzero(P["bc"] == 100))
    threads = oracle.lib().orc_set_threads(0)
    S = oracle.State(cfg, P, t_sim)
    S.st.step = step
    S.st.ioutput, S.st.next_output = 0, cfg.t_output      # the output clock as psph_load sets it
    while S.st.next_output <= t_sim and cfg.t_output > 0:
        S.st.ioutput += 1
        S.st.next_output = (S.st.ioutput + 1) * cfg.t_output
    t0 = time.perf_counter()
    for _ in range(steps):
        S.step()
    wall = time.perf_counter() - t0
    return {"value": n_real * steps / wall, "unit": "particle-updates/s", "cores": int(threads),
            "kind": "port",
            "sample": f"oracle (FP64 C restatement, OpenMP) on a {side / 1000:.2f}x{side / 1000:.2f}x"
                      f"{1.5 * side / 1000:.2f} km weak-plume box at the same {spacing:.1f} m spacing, "
                      f"{n_real} real particles, steps {step + 1}-{step + steps} ({wall:.1f} s; steps "
                      f"1-{step} run in the GPU FP64 form)"}


def preroll(cfg, P, nsteps):
    """Runs the FP64 form (the parity form) nsteps steps from P and returns the
    state and clock after them: the input of every timed form, so that the
    timed window lies after the first adaptive-h update (step 50), in the
    regime that 451 of the workload's 500 steps run in."""
    c = cfg.copy()
    c.precision = psph.PSPH_FP64
    G = psph.Solver(c, len(P) + 400 * nsteps + 100000)
    G.load(P)
    if c.wall_bc:
        G.apply_bcond()
    G.run(nsteps)
    Q = G.read()
    t_sim, step = G.time()
    G.close()
    return Q, t_sim, step


def run_form(cfg, P, args, form, t0_sim=0.0, step0=0):
    cap = len(P) + 400 * (args.warmup + args.steps) + 100000
    G = psph.Solver(cfg, cap)
    G.load(P, t0_sim, step0)
    if cfg.wall_bc and step0 == 0:
        G.apply_bcond()
    t0 = time.perf_counter()
    for s in range(args.warmup):
        G.step()
    sync()
    log(f"[bench] {form}: {args.warmup} warm-up steps in {time.perf_counter() - t0:.1f}s")
    _, n_real_start = G.count()
    G.profile(True)
    pairs = 0
    sync()
    t0 = time.perf_counter()
    for s in range(args.steps):
        o = G.step()
        pairs += o.n_pairs
    sync()
    wall = time.perf_counter() - t0
    prof = G.profile_read()
    n_total_end, n_real_end = G.count()
    t_sim, step = G.time()
    n_real = 0.5 * (n_real_start + n_real_end)
    log(f"[bench] {form}: {args.steps} steps in {wall:.3f}s ({1e3 * wall / args.steps:.2f} ms/step), "
        f"sim time {t_sim:.3f}s after step {step}")
    for k, (ms, n) in prof.items():
        log(f"[bench]   {k:11s} {ms:9.2f} ms over {n} launches")
    final = None if args.no_compare else G.read()
    G.close()
    return {"value": n_real * args.steps / wall, "wall": wall, "prof": prof, "pairs": pairs,
            "n_real": n_real, "n_total": n_total_end, "t_sim": t_sim, "step": step, "final": final,
            "ms_per_step": 1e3 * wall / args.steps}


def adjoint_leg(args):
    """Adjoint of the plume height: weak plume at the given size, N steps with a
    checkpoint every K; backward time over the plain FP64 forward of the same
    steps, and central differences through the frozen replay on seeded
    directions over (vent parameters, every initial velocity): two-point
    over a short horizon, four-point over the full one, plus an eps sweep of
    the vent velocity at the full horizon."""
    sys.path.insert(0, os.path.join(ROOT, "plume-sph_amd"))
    from psph.torch_adjoint import build_state
    cfg = weak_plume_config(args.adjoint_particles)
    met = psph.meteo_standard()
    theta0 = np.array([cfg.vent_v, cfg.vent_T, 1.0 - cfg.ng0])
    _, P = build_state(cfg, met, args.seed, 0.01, *theta0, None)
    N, K = args.adjoint_steps, args.adjoint_ckpt
    cap = len(P) + 400 * N + 50000
    S = psph.Solver(cfg, cap)
    S.load(P)
    S.step()                              # first-touch / allocation warm-up
    S.load(P)
    sync()
    t0 = time.perf_counter()
    S.run(N)
    sync()
    fwd_ms = 1e3 * (time.perf_counter() - t0)
    H_plain = S.diag_height()
    V0 = theta0[0]

    def H_at(n, eps, dth, dv):
        """H after n steps of the frozen replay from theta0 (1 + eps dth), v0 + eps V0 dv"""
        c, Pp = build_state(cfg, met, args.seed, 0.01, *(theta0 * (1 + eps * dth)),
                            None if dv is None else eps * V0 * dv)
        S.set_vent(c.vent_v, c.vent_T, c.ng0)
        S.load(Pp)
        S.replay(n)
        return S.diag_height()

    def fd2(n, eps, dth, dv):
        return (H_at(n, eps, dth, dv) - H_at(n, -eps, dth, dv)) / (2 * eps)

    def fd4(n, eps, dth, dv):
        return (8 * (H_at(n, eps, dth, dv) - H_at(n, -eps, dth, dv)) -
                (H_at(n, 2 * eps, dth, dv) - H_at(n, -2 * eps, dth, dv))) / (12 * eps)

    rng = np.random.default_rng(args.seed + 1000)
    units = []
    for d in range(args.adjoint_fd):
        v = rng.standard_normal(3 + 3 * len(P))
        units.append(v / np.linalg.norm(v))

    def check(n, r, fd, eps, tag):
        g = np.array([r["dH_dvent_v"], r["dH_dvent_T"], r["dH_dmass_frac"]])
        out = []
        for d, v in enumerate(units):
            dth, dv = v[:3], v[3:].reshape(len(P), 3)
            ad = float(np.sum(dth * theta0 * g) + V0 * np.sum(dv * r["grad_v"]))
            fdv = fd(n, eps, dth, dv)
            out.append({"direction": d, "adjoint": ad, "fd": fdv, "rel_err": abs(ad - fdv) / abs(fdv)})
            log(f"[bench] adjoint FD {tag} direction {d} (eps {eps:g}, {n} steps): AD {ad:.9e} FD {fdv:.9e} "
                f"rel {out[-1]['rel_err']:.2e}")
        return out

    # the short horizon, where the map is not yet chaotic at FP64 resolution
    n_short = min(20, N)
    S.load(P)
    r_short = S.adjoint_run(n_short, n_short, grads=True)
    dirs_short = check(n_short, r_short, fd2, args.adjoint_eps, "two-point")
    S.set_vent(cfg.vent_v, cfg.vent_T, cfg.ng0)
    S.load(P)
    r = S.adjoint_run(N, K, grads=True)
    log(f"[bench] adjoint: {len(P)} particles, {N} steps, ckpt {K}: forward {fwd_ms:.0f} ms, "
        f"recorded forward {r['forward_ms']:.0f} ms, backward {r['backward_ms']:.0f} ms")
    dirs = check(N, r, fd4, args.adjoint_eps4, "four-point")
    sweep = []
    if args.adjoint_sweep:
        dth = np.array([1.0, 0.0, 0.0])
        ad = float(theta0[0] * r["dH_dvent_v"])
        for eps in (1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10):
            fdv = fd2(N, eps, dth, None)
            sweep.append({"eps": eps, "adjoint": ad, "fd": fdv, "rel_err": abs(ad - fdv) / abs(fdv)})
            log(f"[bench] adjoint FD sweep vent_v eps {eps:g}: AD {ad:.9e} FD {fdv:.9e} rel {sweep[-1]['rel_err']:.2e}")
    S.set_vent(cfg.vent_v, cfg.vent_T, cfg.ng0)
    S.close()
    return {"workload": f"weak plume, {args.adjoint_particles / 1e3:g}k real particles, {N} steps, "
                        f"checkpoint every {K}",
            "particles_total": len(P), "H": r["H"], "H_plain_forward": H_plain,
            "dH_dvent_v": r["dH_dvent_v"], "dH_dvent_T": r["dH_dvent_T"], "dH_dmass_frac": r["dH_dmass_frac"],
            "forward_ms": fwd_ms, "recorded_forward_ms": r["forward_ms"], "backward_ms": r["backward_ms"],
            "backward_over_forward": r["backward_ms"] / fwd_ms,
            "fd_short": {"steps": n_short, "eps": args.adjoint_eps, "stencil": "two-point",
                         "directions": dirs_short,
                         "max_rel_err": max((x["rel_err"] for x in dirs_short), default=None)},
            "fd_full": {"steps": N, "eps": args.adjoint_eps4, "stencil": "four-point",
                        "directions": dirs, "max_rel_err": max((x["rel_err"] for x in dirs), default=None)},
            "fd_sweep_vent_v": sweep,
            "note": "central differences of H through the frozen replay (psph_replay_run: the recorded run's "
                    "dt, h updates, inlet counts, ghost->real and escape events, cell order and pair set), "
                    "on seeded unit directions over (vent v, T, erupted mass fraction, every particle's initial "
                    "velocity); AD and FD are d H / d ln theta along the direction.  Over the full horizon the "
                    "map is chaotic: the two-point quotient leaves the linear range at eps 1e-6 and FP64 "
                    "rounding of H takes over below eps 1e-9 (the vent-velocity sweep shows both ends), so the "
                    "full-horizon directions use a four-point stencil at the bottom of that curve"}


def rel_l1_by_key(A, B):
    """Per-field relative L1 of A against B over the particles present in both."""
    def key(P):
        k = P["key0"].astype(np.uint64) << np.uint64(32) | P["key1"].astype(np.uint64)
        u, idx, cnt = np.unique(k, return_index=True, return_counts=True)
        keep = idx[cnt == 1]                      # drop the (rare) shared Hilbert keys
        return k[keep], keep
    ka, sa = key(A)
    kb, sb = key(B)
    common, ia, ib = np.intersect1d(ka, kb, assume_unique=True, return_indices=True)
    ia, ib = sa[ia], sb[ib]
    out = {"particles_compared": int(len(common)), "particles_a": int(len(ka)), "particles_b": int(len(kb))}
    for f in ("x", "y", "z", "vx", "vy", "vz", "rho", "e", "p", "c", "T", "xi", "h"):
        a, b = A[f][ia], B[f][ib]
        den = np.sum(np.abs(b))
        out[f] = float(np.sum(np.abs(a - b)) / den) if den > 0 else 0.0
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50,
                    help="timed steps (default: one adaptive-h period, steps 61-110 with the update at 100)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps of each form before its window")
    ap.add_argument("--preroll", type=int, default=-1,
                    help="steps run in the FP64 form to make the input of every form (default: the "
                         "adaptive-h cadence + 5 = 55, so that every timed step follows the first h update)")
    ap.add_argument("--particles", type=float, default=4.0e6, help="real particles")
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--precision", choices=["fp64", "fp32-exp", "fp32-pair", "fp32-pure"], default="fp32-exp",
                    help="form whose throughput is the reported value")
    ap.add_argument("--no-compare", action="store_true", help="skip the run of the other form")
    ap.add_argument("--cpu-sample", type=float, default=6.0e4, help="real particles in the CPU sample")
    ap.add_argument("--cpu-steps", type=int, default=3, help="steps of the CPU sample (~20 s on 16 cores)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--adjoint-particles", type=float, default=5.12e5)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-ckpt", type=int, default=20)
    ap.add_argument("--adjoint-fd", type=int, default=16, help="seeded directions checked by central differences")
    ap.add_argument("--adjoint-eps", type=float, default=1e-6,
                    help="relative step of the two-point central differences over the short horizon")
    ap.add_argument("--adjoint-eps4", type=float, default=3e-9,
                    help="relative step of the four-point differences over the full horizon")
    ap.add_argument("--no-adjoint-sweep", dest="adjoint_sweep", action="store_false",
                    help="skip the eps sweep of the vent-velocity difference quotient")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-wall-bc", dest="wall_bc", action="store_false",
                    help="forward forms without the wall-ghost refresh that the reference's loop makes after "
                         "every step (main.cc:468)")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint leg (profiling)")
    ap.add_argument("--strong", action="store_true", help="also run the 16M strong-plume leg (minutes)")
    ap.add_argument("--strong-only", action="store_true", help="run only the strong-plume leg")
    ap.add_argument("--strong-particles", type=float, default=16.0e6)
    ap.add_argument("--strong-steps", type=int, default=100)
    ap.add_argument("--strong-adjoint", type=int, default=50, help="adjoint over the last N steps")
    ap.add_argument("--strong-ckpt", type=int, default=0,
                    help="checkpoint interval of the strong adjoint (default: sized to the free HBM)")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="GiB: run only a plain device copy of this size (PMC calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs one GPU's whole domain; multi-GPU sharding is a "
                                   "design note only (run with --gpus 1)", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.copy_calibration > 0:
        import torch
        n = int(args.copy_calibration * (1 << 30) // 8)
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}))
        return

    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_leg(args)}))
        return
    if args.strong_only:
        print(json.dumps({"strong": strong_leg(args)}))
        return

    base = weak_plume_config(args.particles)
    base.wall_bc = 1 if args.wall_bc else 0
    met = psph.meteo_standard()
    t0 = time.perf_counter()
    P = psph.gen_plume(base, met, args.seed, 0.01)
    n_total0 = len(P)
    n_real0 = int(np.count_nonzero(P["bc"] == psph.BC_REAL))
    log(f"[bench] generated {n_total0} particles ({n_real0} real, spacing {base.h_air:.2f} m) "
        f"in {time.perf_counter() - t0:.1f}s")
    n_pre = args.preroll if args.preroll >= 0 else base.sml_update_int + 5
    t0 = time.perf_counter()
    P, t_pre, step_pre = preroll(base, P, n_pre)
    log(f"[bench] pre-roll: {n_pre} FP64 steps in {time.perf_counter() - t0:.1f}s, t = {t_pre:.3f}s, "
        f"{len(P)} particles")

    forms = [args.precision] + ([] if args.no_compare else
                                [p for p in ("fp32-exp", "fp32-pair", "fp64", "fp32-pure") if p != args.precision])
    codes = {"fp32-exp": psph.PSPH_FP32_EXP, "fp32-pair": psph.PSPH_FP32_PAIR, "fp32-pure": psph.PSPH_FP32_PURE,
             "fp64": psph.PSPH_FP64}
    runs = {}
    for form in forms:
        cfg = base.copy()
        cfg.precision = codes[form]
        runs[form] = run_form(cfg, P, args, form, t_pre, step_pre)
    del P
    main_run = runs[args.precision]
    value, wall, prof, pairs, n_real = (main_run[k] for k in ("value", "wall", "prof", "pairs", "n_real"))
    t_sim, step, n_total_end = main_run["t_sim"], main_run["step"], main_run["n_total"]
    cfg = base

    compare = None
    if "fp64" in runs and len(runs) > 1:
        B = runs["fp64"]["final"]
        compare = {"fp64_value": runs["fp64"]["value"],
                   "note": "all forms from the same input and step count; the FP64 form is the parity-checked "
                           "one (1e-10 vs the CPU oracle on small cases). fp32-exp (the reported form) runs the "
                           "viscosity, heat and energy pair terms in FP32 and evaluates every pair weight's "
                           "exponential with the FP32 hardware instruction on an FP64 argument rounded once to FP32, "
                           "h, 1/h, norms, pressure terms and sums FP64, except the density weights, FP32 "
                           "throughout and summed in FP64; fp32-pair evaluates the weights' exponential in FP64 "
                           "(slower, ~100x closer to FP64); fp32-pure runs the RHS and density pair arithmetic "
                           "all in FP32 (per-particle factors too)"}
        for form in ("fp32-exp", "fp32-pair", "fp32-pure"):
            if form not in runs:
                continue
            key = form.replace("-", "_")
            compare[f"{key}_value"] = runs[form]["value"]
            compare[f"speedup_{key}_over_fp64"] = runs[form]["value"] / runs["fp64"]["value"]
            compare[f"rel_l1_{key}_vs_fp64"] = rel_l1_by_key(runs[form]["final"], B)
            log(f"[bench] {form} / fp64 = {compare[f'speedup_{key}_over_fp64']:.3f}; rel L1 "
                + " ".join(f"{k}={v:.1e}" for k, v in compare[f"rel_l1_{key}_vs_fp64"].items()))
    for r in runs.values():
        r.pop("final", None)

    rhs_ms, rhs_n = prof["rhs"]
    rhs_s = rhs_ms / max(rhs_n, 1) / 1e3
    bytes_per_launch = RHS_BYTES_PER_UPDATE * n_real
    achieved = bytes_per_launch / rhs_s / 1e9
    traffic = traffic_from_profiles("rhs_" + args.precision)
    kname = "k_pairs_v<3, Bound<3, RhsPass<%s>>> (momentum/energy RHS)" % {
        "fp32-exp": "float, true, true", "fp32-pair": "float, true", "fp32-pure": "float, false",
        "fp64": "double, false"}[args.precision]
    adapt_n = sum(1 for s in range(step - args.steps + 1, step + 1)
                  if base.sml_update_int > 0 and s % base.sml_update_int == 0)
    roofline = {"kernel": kname, "bound": "hbm", "limiter": "VALU issue (roofline_valu)",
                "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "avg_launch_ms": rhs_s * 1e3, "algorithmic_bytes_per_launch": bytes_per_launch}
    flops = RHS_FLOPS_PER_PAIR * (pairs / max(args.steps, 1)) / rhs_s / 1e12
    vpeak = FP64_VECTOR_PEAK_TFS if args.precision == "fp64" else FP32_VECTOR_PEAK_TFS
    roofline_valu = {"bound": "valu_fp64" if args.precision == "fp64" else "valu_fp32",
                     "achieved": flops, "peak": vpeak, "unit": "TFLOP/s", "frac": flops / vpeak,
                     "pairs_per_particle": pairs / max(args.steps, 1) / n_real}

    adjoint = None
    if not args.no_adjoint:
        try:
            adjoint = adjoint_leg(args)
        except Exception as e:   # reported as missing, the forward figure stands
            log(f"[bench] adjoint leg failed: {e}")
            adjoint = {"error": str(e)}

    strong = None
    if args.strong:
        try:
            strong = strong_leg(args)
        except Exception as e:   # reported as missing, the headline stands
            log(f"[bench] strong leg failed: {e}")
            strong = {"error": str(e)}

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args.cpu_sample, args.cpu_steps, cfg.h_air, n_pre, base.wall_bc)
        except Exception as e:   # the GPU figure stands without the baseline
            log(f"[bench] cpu baseline failed: {e}")
            cpu = None

    out = {
        "metric": METRIC, "value": value, "unit": "particle-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32-exp": "f32-pair+f64", "fp32-pair": "f32-pair+f64",
                  "fp32-pure": "f32-pair"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"weak dusty-gas plume, 8x8x12 km, {args.particles / 1e6:g}M real particles, "
                               f"{ {'fp64': 'FP64 form', 'fp32-exp': 'FP32-pair form with FP32 exponentials', 'fp32-pair': 'FP32-pair form', 'fp32-pure': 'FP32-pure form'}[args.precision]}",
                   "particles_real": n_real, "particles_total": n_total_end,
                   "spacing_m": cfg.h_air, "vent": "r=54 m, 135 m/s, 1053 K, gas fraction 0.03",
                   "kernel": "gaussian-box", "wall_ghost_refresh": bool(args.wall_bc),
                   "preroll_steps": n_pre, "timed_steps": [step - args.steps + 1, step],
                   "adaptations_in_window": adapt_n, "sim_time_s": t_sim},
        "roofline": roofline, "roofline_valu": roofline_valu, "cpu_baseline": cpu,
        "forms": compare,
        "adjoint": adjoint,
        "strong": strong,
        "stage_ms_per_step": {k: v[0] / args.steps for k, v in prof.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
