"""Domain adaptation on the GPU against the oracle (FP64 form): a small
initial domain around the vent, grown bucket by bucket as involved material
reaches its outer layer (main.cc:256-304, 433-436).

Bars: particle counts, keys, bc, phase and involved bit-exact and every
floating-point field within 1e-10 relative, at every step where the region
grows and at the end; the run must grow the region several times.
"""
import sys

import numpy as np
import pytest

import oracle
import psph
from helpers import FLOAT_FIELDS_3D, REL_TOL_FP64, compare
from test_oracle_domain import SEED, JITTER, adaptive_config

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("wall_bc,nsteps", [(0, 22), (1, 12)])
def test_domain_grows_like_the_oracle(wall_bc, nsteps):
    cfg = adaptive_config()
    cfg.wall_bc = wall_bc
    met = psph.meteo_standard()
    P = psph.gen_plume(cfg, met, SEED, JITTER)
    say(f"adaptive domain, wall_bc {wall_bc}: {len(P)} particles, {nsteps} steps")
    S = oracle.State(cfg, P)
    oracle.domain_init(S, met, SEED, JITTER)
    if wall_bc:
        S.apply_bcond()
    grown = 0
    with psph.Solver(cfg, 4 * len(P) + 40000) as G:
        G.load(P)
        G.domain_init(met, SEED, JITTER)
        if wall_bc:
            G.apply_bcond()
        for s in range(nsteps):
            S.step()
            o = G.step()
            assert o.n_domain == S.st.last_adapted, (s, o.n_domain, S.st.last_adapted)
            n_total, _ = G.count()
            assert n_total == len(S), (s, n_total, len(S))
            if o.n_domain:
                grown += 1
                errs = compare(G.read(), S.particles(), FLOAT_FIELDS_3D, REL_TOL_FP64)
                say(f"step {o.step}: region grew by {o.n_domain} ghosts, {n_total} particles, "
                    f"max rel err {max(errs.values()):.3e}")
        errs = compare(G.read(), S.particles(), FLOAT_FIELDS_3D, REL_TOL_FP64)
        say(f"after {nsteps} steps: max rel err {max(errs.values()):.3e}")
    assert grown >= 2
