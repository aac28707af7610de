"""Pins for the oracle's wall-ghost refresh (apply_bcond, bcond.cc:24-249),
on the CPU.

Closed forms that fix the restatement: a Shepard average of a uniform state
returns that state; the image velocity is the mirror image of the
neighbours' in the ground plane; the mass fraction sums only particles
outside the image's own bucket (bcond.cc:140-146), so it equals the uniform
value when that bucket holds no contributor and is zero when only that bucket
does; a ghost with no contributor in reach gets rho = 1 and zero velocity,
energy and mass fraction; only wall ghosts of the bucket layer that straddles
the ground change.
"""
import numpy as np
import pytest

import oracle
import psph
from helpers import small_plume

RHO0, V0, E0, XI0 = 1.1, (1.5, -2.0, 3.0), 2.2e5, 0.4


def uniform_plume():
    cfg, met, P = small_plume()
    contr = np.isin(P["bc"], (100, 99, 1))
    P["rho"][contr] = RHO0
    for k, f in enumerate(("vx", "vy", "vz")):
        P[f][contr] = V0[k]
    P["e"][contr] = E0
    P["xi"][contr] = XI0
    return cfg, P


def buckets(cfg, P):
    del_ = cfg.bucket_edge
    return np.stack([np.floor((P[f] - cfg.ext_lo[k]) / del_).astype(int) for k, f in enumerate(("x", "y", "z"))], 1)


def imaged(cfg, P):
    """wall ghosts of the bucket layer that contains the ground z = dom_lo[2]"""
    bzg = int(np.floor((cfg.dom_lo[2] - cfg.ext_lo[2]) / cfg.bucket_edge))
    return (P["bc"] == 2) & (buckets(cfg, P)[:, 2] == bzg)


def run(cfg, P):
    S = oracle.State(cfg, P)
    S.apply_bcond()
    return S.particles()


def by_key(P):
    k = (P["key0"].astype(np.uint64) << np.uint64(32)) | P["key1"].astype(np.uint64)
    return {int(a): i for i, a in enumerate(k)}


def test_uniform_state_is_reproduced_and_velocity_mirrored():
    cfg, P = uniform_plume()
    Q = run(cfg, P)
    m = imaged(cfg, Q)
    assert m.sum() > 100
    assert np.allclose(Q["rho"][m], RHO0, rtol=1e-13, atol=0)
    assert np.allclose(Q["e"][m], E0, rtol=1e-13, atol=0)
    assert np.allclose(Q["vx"][m], V0[0], rtol=1e-13)
    assert np.allclose(Q["vy"][m], V0[1], rtol=1e-13)
    assert np.allclose(Q["vz"][m], -V0[2], rtol=1e-13)          # reflected in the ground plane
    # only the neighbour buckets feed the mass fraction: between 0 (image
    # neighbourhood inside its own bucket) and the uniform value
    assert np.all((Q["xi"][m] >= 0) & (Q["xi"][m] <= XI0 * (1 + 1e-13)))
    assert np.any(Q["xi"][m] == 0) and np.any(Q["xi"][m] > 0.5 * XI0)
    # EOS of the new state
    for i in np.nonzero(m)[0][:20]:
        p, c, T, gam, cp = oracle.eos(cfg, Q["rho"][i], Q["e"][i], Q["xi"][i])
        assert Q["p"][i] == pytest.approx(p, rel=1e-14) and Q["c"][i] == pytest.approx(c, rel=1e-14)


def test_only_ground_layer_wall_ghosts_change():
    cfg, P = uniform_plume()
    Q = run(cfg, P)
    iP, iQ = by_key(P), by_key(Q)
    m = imaged(cfg, Q)
    keys_changed = {k for k, i in iQ.items() if m[i]}
    same = [k for k in iQ if k not in keys_changed]
    assert len(same) > len(keys_changed) > 0
    a = np.array([iP[k] for k in same])
    b = np.array([iQ[k] for k in same])
    for f in ("rho", "vx", "vy", "vz", "e", "xi", "p", "c"):
        assert np.array_equal(P[f][a], Q[f][b]), f
    # deeper wall ghosts exist and are among the unchanged ones
    deep = (Q["bc"] == 2) & ~m
    assert deep.sum() > 0


def test_mass_fraction_excludes_the_image_bucket():
    cfg, P = uniform_plume()
    B = buckets(cfg, P)
    m = imaged(cfg, P)
    g = np.nonzero(m)[0][len(np.nonzero(m)[0]) // 2]
    img = np.array([P["x"][g], P["y"][g], 2 * cfg.dom_lo[2] - P["z"][g]])
    bi = np.floor((img - np.array(cfg.ext_lo[:])) / cfg.bucket_edge).astype(int)
    own = np.all(B == bi, axis=1) & np.isin(P["bc"], (100, 99, 1))
    # (a) no contributor in the image's bucket: xi = the uniform value
    Pa = P.copy()
    Pa["bc"][own] = 2
    Qa = run(cfg, Pa)
    ia = by_key(Qa)[int((np.uint64(P["key0"][g]) << np.uint64(32)) | np.uint64(P["key1"][g]))]
    assert Qa["xi"][ia] == pytest.approx(XI0, rel=1e-13)
    # (b) only the image's bucket contributes: xi = 0, the rest still averages
    Pb = P.copy()
    others = np.isin(P["bc"], (100, 99, 1)) & ~own
    Pb["bc"][others] = 2
    Qb = run(cfg, Pb)
    ib = by_key(Qb)[int((np.uint64(P["key0"][g]) << np.uint64(32)) | np.uint64(P["key1"][g]))]
    assert Qb["xi"][ib] == 0.0
    assert Qb["rho"][ib] == pytest.approx(RHO0, rel=1e-13)


def test_no_contributor_in_reach():
    cfg, P = uniform_plume()
    contr = np.isin(P["bc"], (100, 99, 1))
    P["bc"][contr] = 0                      # erupt ghosts do not contribute
    Q = run(cfg, P)
    m = imaged(cfg, Q)
    assert np.all(Q["rho"][m] == 1.0)
    for f in ("vx", "vy", "vz", "e", "xi"):
        assert np.all(Q[f][m] == 0.0), f
