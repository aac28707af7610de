"""Shared builders for the parity tests: small plume and shock-tube inputs."""
from __future__ import annotations

import numpy as np

import psph

# Parity bars of the FP64 form: neighbour sets and ordering bit-exact;
# density, velocity, energy (and the EOS outputs) within 1e-10 relative.
REL_TOL_FP64 = 1e-10
FLOAT_FIELDS_3D = ["x", "y", "z", "vx", "vy", "vz", "svx", "svy", "svz", "rho", "e", "m", "h",
                   "p", "c", "T", "gam", "cp", "xi"]
FLOAT_FIELDS_1D = ["x", "vx", "rho", "e", "m", "h", "p", "c", "gam", "xi"]
EXACT_FIELDS = ["key0", "key1", "key2", "bc", "phase", "involved"]


def small_plume_config(h_air=100.0, half=600.0, top=1200.0, vent_r=54.0, vent_v=135.0,
                       vent_T=1053.0, ng0=0.03, sml_update_int=50):
    cfg = psph.default_config(3)
    cfg.h_air = h_air
    for k, (lo, hi) in enumerate([(-half, half), (-half, half), (0.0, top)]):
        cfg.dom_lo[k] = lo
        cfg.dom_hi[k] = hi
    cfg.vent_z = 0.0
    cfg.vent_r, cfg.vent_v, cfg.vent_T, cfg.ng0 = vent_r, vent_v, vent_T, ng0
    cfg.vent_p = 101325.0
    cfg.sml_update_int = sml_update_int
    psph.finalize_config(cfg)
    return cfg


def small_plume(seed=20261202, jitter=0.01, **kw):
    cfg = small_plume_config(**kw)
    met = psph.meteo_standard()
    P = psph.gen_plume(cfg, met, seed, jitter)
    return cfg, met, P


def weak_plume_config(n_real=4.0e6, lx=8000.0, ly=8000.0, lz=12000.0):
    """The weak dusty-gas plume of the throughput measurement: 8 x 8 x 12 km
    with n_real real particles (spacing (V/N)^(1/3)), vent r = 54 m, 135 m/s,
    1053 K, erupted-material gas fraction 0.03, p(0) = 101325 Pa."""
    cfg = psph.default_config(3)
    cfg.h_air = (lx * ly * lz / n_real) ** (1.0 / 3.0)
    for k, (lo, hi) in enumerate([(-lx / 2, lx / 2), (-ly / 2, ly / 2), (0.0, lz)]):
        cfg.dom_lo[k] = lo
        cfg.dom_hi[k] = hi
    cfg.vent_z = 0.0
    cfg.vent_r, cfg.vent_v, cfg.vent_T, cfg.ng0 = 54.0, 135.0, 1053.0, 0.03
    cfg.vent_p = 101325.0
    psph.finalize_config(cfg)
    return cfg


def sod_config(n_left=800, n_right=100, eta=1.3, t_end=0.2):
    """Config 1: Sod tube on [-0.5, 0.5], equal-mass particles, plain SPH."""
    cfg = psph.default_config(1)
    cfg.dom_lo[0], cfg.dom_hi[0] = -0.5, 0.5
    cfg.h_air = 0.5 / n_left
    cfg.eta = eta
    cfg.t_end = t_end
    cfg.t_output = t_end
    psph.finalize_config(cfg)
    return cfg


def sod(n_left=800, n_right=100, **kw):
    cfg = sod_config(n_left, n_right, **kw)
    P = psph.gen_shock_tube(cfg, (1.0, 1.0, 0.0), (0.125, 0.1, 0.0), 0.0, 40, 20, 35, True)
    return cfg, P


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.shape != b.shape:
        return np.inf
    scale = np.max(np.abs(b)) if b.size else 0.0
    if scale == 0.0:
        return float(np.max(np.abs(a))) if a.size else 0.0
    return float(np.max(np.abs(a - b)) / scale)


def compare(P, Q, fields, tol, exact=EXACT_FIELDS):
    """Max-norm relative error per float field; exact fields must be equal."""
    assert len(P) == len(Q), f"particle count {len(P)} vs {len(Q)}"
    for k in exact:
        bad = np.nonzero(P[k] != Q[k])[0]
        assert bad.size == 0, f"{k} differs at {bad[:10]} ({P[k][bad[:5]]} vs {Q[k][bad[:5]]})"
    errs = {k: rel_err(P[k], Q[k]) for k in fields}
    worst = {k: v for k, v in errs.items() if not (v <= tol)}
    assert not worst, f"fields above {tol}: {worst}"
    return errs


def same_lists(off_a, idx_a, off_b, idx_b):
    """Neighbour sets equal per particle (order ignored)."""
    assert len(off_a) == len(off_b)
    n = len(off_a) - 1
    for i in range(n):
        a = np.sort(idx_a[off_a[i]:off_a[i + 1]])
        b = np.sort(idx_b[off_b[i]:off_b[i + 1]])
        if a.shape != b.shape or not np.array_equal(a, b):
            return i
    return -1


def rel_l1(a, b):
    """Relative L1 error sum|a-b| / sum|b| (the FP32-pair acceptance metric)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sum(np.abs(b))
    return float(np.sum(np.abs(a - b)) / den) if den > 0 else float(np.sum(np.abs(a)))
