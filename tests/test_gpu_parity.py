"""HIP path vs the CPU oracle on the same seeded inputs (FP64 form).

Bars: neighbour sets, particle order, keys, bc/phase/involved bit-exact;
every floating-point field within 1e-10 relative (max-norm over the field).
"""
import sys

import numpy as np
import pytest

import oracle
import psph
from helpers import (FLOAT_FIELDS_1D, FLOAT_FIELDS_3D, REL_TOL_FP64, compare, same_lists,
                     small_plume, sod)

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def test_plume_stage_by_stage():
    """Every stage of one 3D step, with adaptive_sml forced on this step."""
    cfg, met, P = small_plume(sml_update_int=1)
    say(f"plume stage-by-stage: {len(P)} particles")
    S = oracle.State(cfg, P)
    with psph.Solver(cfg, len(P) + 20000) as G:
        G.load(P)
        dt_o, dt_g = S.timestep(), G.timestep()
        assert dt_g == dt_o, (dt_g, dt_o)
        dt_o, dt_g = S.incrtime(dt_o), G.incrtime(dt_g)
        assert dt_g == dt_o
        say("neighbours")
        off_o, idx_o = S.neighbors()
        G.search_neighs_consth()
        off_g, idx_g = G.neighbors()
        assert off_o[-1] == off_g[-1], (off_o[-1], off_g[-1])
        assert same_lists(off_g, idx_g, off_o, idx_o) == -1
        say("adaptive_sml")
        S.adaptive_sml()
        G.adaptive_sml()
        compare(G.read(), S.particles(), FLOAT_FIELDS_3D, REL_TOL_FP64)
        say("mom_engr_update")
        S.rhs(dt_o)
        G.mom_engr_update(dt_g)
        compare(G.read(), S.particles(), FLOAT_FIELDS_3D, REL_TOL_FP64)
        say("smooth_density")
        S.density()
        G.smooth_density()
        compare(G.read(), S.particles(), FLOAT_FIELDS_3D, REL_TOL_FP64)
        say("smooth_velocity")
        S.vsmooth()
        G.smooth_velocity()
        compare(G.read(), S.particles(), FLOAT_FIELDS_3D, REL_TOL_FP64)
        say("update_pos + inject")
        S.update_pos(dt_o)
        S.remove_escaped()
        n_o = S.inject(dt_o)
        S.sort()
        G.update_pos(dt_g)
        n_g = G.add_new_erupt(dt_g)
        assert n_g == n_o
        compare(G.read(), S.particles(), FLOAT_FIELDS_3D, REL_TOL_FP64)


def test_lists_wider_than_a_bucket():
    """Smoothing lengths of 2.2 h_air make the 4.25 h list box (9.35 h_air)
    wider than a reference bucket (8 h_air): the list then holds the own
    bucket and, from the 26 adjacent buckets only, the particles inside the
    box (neighs.cc:37-211); a particle two buckets away is outside even when
    inside the box.  Lists, one adaptive-h update and one RHS step against
    the oracle (whose bucket search equals its brute force)."""
    cfg, met, P = small_plume()
    P["h"][:] = 2.2 * cfg.h_air
    say(f"lists wider than a bucket: {len(P)} particles, h = {2.2 * cfg.h_air:.1f}")
    S = oracle.State(cfg, P)
    with psph.Solver(cfg, len(P) + 20000) as G:
        G.load(P)
        off_o, idx_o = S.neighbors()
        off_b, idx_b = S.neighbors(brute=True)
        assert same_lists(off_o, idx_o, off_b, idx_b) == -1
        G.search_neighs_consth()
        off_g, idx_g = G.neighbors()
        assert off_o[-1] == off_g[-1], (off_o[-1], off_g[-1])
        assert same_lists(off_g, idx_g, off_o, idx_o) == -1
        S.adaptive_sml()
        G.adaptive_sml()
        compare(G.read(), S.particles(), FLOAT_FIELDS_3D, REL_TOL_FP64)
        dt = S.timestep()
        S.rhs(dt)
        G.mom_engr_update(dt)
        errs = compare(G.read(), S.particles(), FLOAT_FIELDS_3D, REL_TOL_FP64)
        say(f"max rel err {max(errs.values()):.3e}")


@pytest.mark.parametrize("nsteps,adapt", [(6, 3), (4, 50)])
def test_plume_steps(nsteps, adapt):
    cfg, met, P = small_plume(sml_update_int=adapt)
    say(f"plume {nsteps} steps, adaptive every {adapt}: {len(P)} particles")
    S = oracle.State(cfg, P)
    with psph.Solver(cfg, len(P) + 20000) as G:
        G.load(P)
        for s in range(nsteps):
            S.step()
            o = G.step()
            assert o.step == S.st.step
            assert o.dt > 0
        tg, sg = G.time()
        assert sg == S.st.step and abs(tg - S.st.time) <= 1e-12 * abs(S.st.time)
        Q = G.read()
        assert o.n_real == np.isin(Q["bc"], (100, 99)).sum() and o.n_total == len(Q)
        errs = compare(Q, S.particles(), FLOAT_FIELDS_3D, REL_TOL_FP64)
        say(f"max rel err {max(errs.values()):.3e}")
        cons_g, cons_o = G.conservation(), S.conservation()
        assert np.allclose(cons_g, cons_o, rtol=1e-10, atol=1e-10 * np.abs(cons_o).max())
        assert abs(G.diag_height() - S.diag_height()) <= 1e-10 * max(1.0, abs(S.diag_height()))


def test_plume_steps_wall_bc():
    """The reference's loop with the wall-ghost refresh (apply_bcond before the
    loop and after every step, main.cc:196, 468), adaptive h at step 3."""
    cfg, met, P = small_plume(sml_update_int=3)
    cfg.wall_bc = 1
    say(f"plume with wall refresh, 6 steps: {len(P)} particles")
    S = oracle.State(cfg, P)
    S.apply_bcond()
    with psph.Solver(cfg, len(P) + 20000) as G:
        G.load(P)
        G.apply_bcond()
        A0, B0 = G.read(), S.particles()
        walls = B0["bc"] == 2
        assert not np.array_equal(B0["rho"][walls], P["rho"][P["bc"] == 2])   # the refresh changed them
        errs = compare(A0, B0, FLOAT_FIELDS_3D, REL_TOL_FP64)
        say(f"after the set-up refresh: max rel err {max(errs.values()):.3e}")
        for s in range(6):
            S.step()
            G.step()
        errs = compare(G.read(), S.particles(), FLOAT_FIELDS_3D, REL_TOL_FP64)
        say(f"after 6 steps: max rel err {max(errs.values()):.3e}")


def test_sod_steps():
    """Config 1 (1D Sod, plain SPH): adaptive h at set-up and every step."""
    cfg, P = sod()
    say(f"sod: {len(P)} particles")
    S = oracle.State(cfg, P)
    S.neighbors()
    S.adaptive_sml()
    with psph.Solver(cfg, len(P) + 16) as G:
        G.load(P)
        G.search_neighs_consth()
        G.adaptive_sml()
        compare(G.read(), S.particles(), FLOAT_FIELDS_1D, REL_TOL_FP64)
        for s in range(40):
            S.step()
            G.step()
        errs = compare(G.read(), S.particles(), FLOAT_FIELDS_1D, REL_TOL_FP64)
        say(f"sod max rel err {max(errs.values()):.3e}")


def test_inlet_past_capacity_is_an_error():
    """The inlet adds the vent's new particles each step; a context sized too
    small for them reports ECAPACITY before any device work of the inlet."""
    cfg, met, P = small_plume(sml_update_int=50)
    say(f"inlet past capacity: {len(P)} particles, capacity {len(P) + 8}")
    with psph.Solver(cfg, len(P) + 8) as G:
        G.load(P)
        with pytest.raises(psph.PsphError) as e:
            for s in range(4):
                G.step()
        assert e.value.code == -4 and "inject" in str(e.value)
