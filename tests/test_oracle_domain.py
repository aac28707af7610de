"""Domain adaptation (ADJUST_DOMAIN): the oracle's restatement pinned by the
geometry the reference's rules fix, and the product's initial state of an
adaptive run equal to the oracle's bit for bit (CPU; host code only).

Reference: bucket types preprocess.cc:380-440; initial flags dataread.cc:
52-75, 326-457; add_air.cc:491-813; scan_outside_layer.cc:44-121;
adapt_domain.cc:41-194; add_pressure_ghost.cc:37-316; add_wall_ghost.cc:
45-293; setup_erupt.cc:221-253; main.cc:128-132, 256-304, 433-436.
"""
import numpy as np
import pytest

import oracle
import psph
from helpers import small_plume_config

UNDERGROUND, MIXED, OVERGROUND, PRESS_BC = 1, 2, 3, 4
SEED, JITTER = 20261202, 0.01


def adaptive_config(idom_half=400.0, idom_top=800.0, half=1600.0, top=2400.0):
    cfg = small_plume_config(h_air=100.0, half=half, top=top, sml_update_int=10)
    cfg.adapt_domain = 1
    for k, (lo, hi) in enumerate([(-idom_half, idom_half), (-idom_half, idom_half), (0.0, idom_top)]):
        cfg.idom_lo[k] = lo
        cfg.idom_hi[k] = hi
    return cfg


def bucket_index(cfg, x, y, z):
    p = np.stack([x, y, z], 1)
    return np.floor((p - np.array(cfg.ext_lo[:])) / cfg.bucket_edge).astype(int)


def test_bucket_types_follow_the_domain_faces():
    cfg = adaptive_config()
    nb = [int(round((cfg.ext_hi[k] - cfg.ext_lo[k]) / cfg.bucket_edge)) for k in range(3)]
    for bz in range(nb[2]):
        zlo = cfg.ext_lo[2] + bz * cfg.bucket_edge
        zhi = zlo + cfg.bucket_edge
        for by in range(nb[1]):
            for bx in range(nb[0]):
                t, bt = oracle.bucket_type(cfg, bx, by, bz)
                xlo = cfg.ext_lo[0] + bx * cfg.bucket_edge
                ylo = cfg.ext_lo[1] + by * cfg.bucket_edge
                lo = (xlo, ylo, zlo)
                hi = (xlo + cfg.bucket_edge, ylo + cfg.bucket_edge, zhi)
                outside = any(hi[k] < cfg.dom_lo[k] or lo[k] > cfg.dom_hi[k] for k in range(3))
                straddle = any(lo[k] < cfg.dom_lo[k] <= hi[k] or lo[k] <= cfg.dom_hi[k] < hi[k]
                               for k in range(3))
                if outside:
                    assert t == (UNDERGROUND if zhi < cfg.dom_lo[2] else PRESS_BC), (bx, by, bz, t)
                elif straddle:
                    assert t == MIXED, (bx, by, bz)
                else:
                    assert t == OVERGROUND, (bx, by, bz)
                # the ground layer: the bucket's lower face below the ground
                assert (bt[4] == -1) == (zlo < cfg.dom_lo[2])


def test_initial_state_layers():
    """Real air exactly in the buckets that meet the initial domain (above the
    ground), pressure ghosts in the empty buckets next to them, wall ghosts
    below the ground layer, the vent's ghosts in the duct."""
    cfg = adaptive_config()
    met = psph.meteo_standard()
    S = oracle.domain_setup(cfg, met, SEED, JITTER)
    P = S.particles()
    b = bucket_index(cfg, P["x"], P["y"], P["z"])
    ground = cfg.dom_lo[2]
    lo = np.array(cfg.ext_lo[:]) + b * cfg.bucket_edge
    hi = lo + cfg.bucket_edge
    meets_idom = np.all((hi >= np.array(cfg.idom_lo[:])) & (lo <= np.array(cfg.idom_hi[:])), 1)
    real = P["bc"] == 100
    assert np.all(meets_idom[real]) and np.all(P["z"][real] > ground)
    assert np.all(P["involved"][real] == 1)
    pg = P["bc"] == 1
    assert pg.sum() > 0 and not np.any(meets_idom[pg]) and np.all(P["z"][pg] > ground)
    # every pressure-ghost bucket borders a bucket of the initial domain
    rb = {tuple(x) for x in b[real]}
    for t in {tuple(x) for x in b[pg]}:
        assert any((t[0] + dx, t[1] + dy, t[2] + dz) in rb
                   for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)), t
    wg = P["bc"] == 2
    assert np.all(P["z"][wg] < ground)
    assert np.all(P["bc"][P["phase"] == 2] == 0)
    # the potentially involved buckets are the initial domain's
    bv = oracle.bucket_values(S)
    assert set(np.unique(bv)) <= {0, 1}


def test_adaptation_grows_around_involved_material():
    """A state whose involved particle sits in the outer layer of the region:
    the scan reports it, the adaptation turns the bordering buckets' particles
    real (potentially involved) and wraps them in a new pressure-ghost layer;
    no particle is lost and the added ones are ghosts of this step."""
    cfg = adaptive_config()
    met = psph.meteo_standard()
    S = oracle.domain_setup(cfg, met, SEED, JITTER)
    S.st.step = 7
    P0 = S.particles()
    real = np.nonzero(P0["bc"] == 100)[0]
    # the real particle farthest from the vent along x
    i = real[np.argmax(P0["x"][real])]
    S.st.inv[i] = 2
    oracle.lib().orc_domain_scan(S.cfg, S.st)
    assert S.st.adapt_pending >= 1
    n0 = S.st.n
    oracle.lib().orc_domain_adapt(S.cfg, S.st)
    P1 = S.particles()
    assert len(P1) == n0 + S.st.last_adapted > n0
    new = P1["key2"] == 7
    assert new.sum() == S.st.last_adapted and np.all(np.isin(P1["bc"][new], (1, 2)))
    grown = (P1["bc"] == 100) & (P1["key2"] == 0) & (P1["phase"] == 1)
    assert grown.sum() > (P0["bc"] == 100).sum()
    bv = oracle.bucket_values(S)
    ib = bucket_index(cfg, P0["x"][i:i + 1], P0["y"][i:i + 1], P0["z"][i:i + 1])[0]
    assert bv[ib[2], ib[1], ib[0]] == 3
    # the buckets beyond the involved one along +x are potentially involved now
    assert bv[ib[2], ib[1], ib[0] + 1] == 1


def test_product_initial_state_equals_oracle():
    cfg = adaptive_config()
    met = psph.meteo_standard()
    P = psph.gen_plume(cfg, met, SEED, JITTER)
    Q = oracle.domain_setup(cfg, met, SEED, JITTER).particles()
    assert len(P) == len(Q)
    for k in psph.ALL_FIELDS:
        assert np.array_equal(P[k], Q[k]), k


def test_initial_domain_equal_to_the_domain_fills_it():
    """With idom = dom every bucket of the domain is potentially involved: the
    real air of the adaptive set-up is the real air of the full-domain
    generator inside dom."""
    cfg = adaptive_config(idom_half=1600.0, idom_top=2400.0)
    met = psph.meteo_standard()
    A = oracle.domain_setup(cfg, met, SEED, JITTER).particles()
    c2 = cfg.copy()
    c2.adapt_domain = 0
    F = psph.gen_plume(c2, met, SEED, JITTER)
    inside = (F["bc"] == 100)
    ra = A["bc"] == 100
    ka = set(zip(A["key0"][ra], A["key1"][ra]))
    kf = set(zip(F["key0"][inside], F["key1"][inside]))
    assert kf <= ka
