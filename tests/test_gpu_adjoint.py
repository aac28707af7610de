"""Adjoint of the plume diagnostic against central finite differences.

H = sum m xi z / sum m xi (real particles) after N steps of the weak-plume
step map, differentiated with its discrete schedule frozen (time steps,
adapted smoothing lengths, inlet counts, erupt->real and escape events); the
finite differences run the same frozen map (psph_replay_run), on 16 seeded
random directions over (vent velocity, vent temperature, erupted mass
fraction, every particle's initial velocity), plus each vent parameter on its
own, plus the smoothed-velocity gradient from a state in motion (at rest, H
is not differentiable in sv: the heat term of a pair is proportional to
|v_ij| and the viscosity switches at v_ij . r_ij = 0).

The kernel is truncated at |s_k| = 3 with a jump of W(3) = exp(-9) N; the
replay evaluates every pair predicate on the recorded run's positions, so a
perturbation never moves a pair across a box face in the map the difference
quotients see (the adjoint treats the box indicators as constant too), and
eps = 1e-6, the measurement's own step, serves every horizon here.
"""
import numpy as np
import pytest
import torch

import psph
from psph.torch_adjoint import build_state, plume_height
from helpers import small_plume_config

pytestmark = pytest.mark.gpu

SEED, JITTER = 20261202, 0.01
FD_TOL = 1e-4


def plume_case(sml_update_int=10):
    cfg = small_plume_config(sml_update_int=sml_update_int)
    met = psph.meteo_standard()
    return cfg, met


class FD:
    def __init__(self, cfg, met, nsteps):
        self.cfg, self.met, self.n = cfg, met, nsteps
        self.theta0 = np.array([cfg.vent_v, cfg.vent_T, 1.0 - cfg.ng0])
        _, P = build_state(cfg, met, SEED, JITTER, *self.theta0, None)
        self.np = len(P)
        self.S = psph.Solver(cfg, len(P) + 20000)
        self.S.load(P)

    def H(self, theta, dv):
        c, P = build_state(self.cfg, self.met, SEED, JITTER, *theta, dv)
        self.S.set_vent(c.vent_v, c.vent_T, c.ng0)
        self.S.load(P)
        self.S.replay(self.n)
        return self.S.diag_height()

    def directional(self, dth, dv, eps):
        """central difference of H along theta0 (1 + eps dth), dv0 = eps V0 dv"""
        hp = self.H(self.theta0 * (1 + eps * dth), eps * self.theta0[0] * dv)
        hm = self.H(self.theta0 * (1 - eps * dth), -eps * self.theta0[0] * dv)
        return (hp - hm) / (2 * eps)


def ad_directional(r, theta0, dth, dv):
    g = np.array([r["dH_dvent_v"], r["dH_dvent_T"], r["dH_dmass_frac"]])
    return float(np.sum(dth * theta0 * g) + theta0[0] * np.sum(dv * r["grad_v"]))


def make_case(N, K):
    cfg, met = plume_case()
    fd = FD(cfg, met, N)
    r = fd.S.adjoint_run(N, K)
    print(f"N={N} H={r['H']:.6f} dH/dV={r['dH_dvent_v']:.6e} dH/dT={r['dH_dvent_T']:.6e} "
          f"dH/dmf={r['dH_dmass_frac']:.6e} n={r['n_particles']} fwd {r['forward_ms']:.0f} ms "
          f"bwd {r['backward_ms']:.0f} ms", flush=True)
    return fd, r


@pytest.fixture(scope="module")
def adjoint_case():
    return make_case(24, 7)           # segments 7, 7, 7, 3; h adapted at steps 10 and 20


@pytest.fixture(scope="module")
def short_case():
    return make_case(6, 4)            # segments 4, 2


def test_replay_reproduces_recorded_run(adjoint_case):
    fd, r = adjoint_case
    H0 = fd.H(fd.theta0, None)
    assert abs(H0 - r["H"]) <= 1e-9 * abs(r["H"]), (H0, r["H"])


@pytest.mark.parametrize("k", [0, 1, 2])
def test_vent_parameters_fd(adjoint_case, k):
    fd, r = adjoint_case
    dth = np.zeros(3)
    dth[k] = 1.0
    dv = np.zeros((fd.np, 3))
    ad = ad_directional(r, fd.theta0, dth, dv)
    fdv = fd.directional(dth, dv, 1e-6)
    print(f"param {k}: AD {ad:.9e} FD {fdv:.9e}", flush=True)
    assert abs(ad - fdv) <= FD_TOL * abs(fdv), (k, ad, fdv)


def random_directions(fd, r, eps):
    rng = np.random.default_rng(SEED + 1000)
    worst = 0.0
    for d in range(16):
        v = rng.standard_normal(3 + 3 * fd.np)
        v /= np.linalg.norm(v)
        dth, dv = v[:3], v[3:].reshape(fd.np, 3)
        ad = ad_directional(r, fd.theta0, dth, dv)
        fdv = fd.directional(dth, dv, eps)
        rel = abs(ad - fdv) / abs(fdv)
        worst = max(worst, rel)
        print(f"direction {d}: AD {ad:.9e} FD {fdv:.9e} rel {rel:.2e}", flush=True)
    return worst


def test_random_directions_fd_eps_1e6(short_case):
    fd, r = short_case
    assert random_directions(fd, r, 1e-6) <= FD_TOL


def test_random_directions_fd_24_steps(adjoint_case):
    fd, r = adjoint_case
    assert random_directions(fd, r, 1e-6) <= FD_TOL


@pytest.fixture(scope="module")
def wall_case():
    """The map with the wall-ghost refresh at the end of every step
    (bcond.cc:24-249 after the step, main.cc:468): 24 steps, segments 7, 7,
    7, 3.  The vent sits on the ground, so the erupted material meets the
    refreshed ghosts from the first step."""
    cfg, met = plume_case()
    cfg.wall_bc = 1
    fd = FD(cfg, met, 24)
    r = fd.S.adjoint_run(24, 7)
    print(f"wall refresh: H={r['H']:.6f} dH/dV={r['dH_dvent_v']:.6e} dH/dT={r['dH_dvent_T']:.6e} "
          f"dH/dmf={r['dH_dmass_frac']:.6e}", flush=True)
    return fd, r


def test_wall_refresh_random_directions_fd(wall_case):
    """The refresh reversed (Shepard sums at the images, the mirror, the
    EOS) against the frozen replay, whose refresh takes its contributor sets
    from the recorded run: the replay reproduces H, then the 16 directions."""
    fd, r = wall_case
    H0 = fd.H(fd.theta0, None)
    assert abs(H0 - r["H"]) <= 1e-9 * abs(r["H"]), (H0, r["H"])
    assert random_directions(fd, r, 1e-6) <= FD_TOL


@pytest.mark.parametrize("k", [0, 1, 2])
def test_wall_refresh_vent_parameters_fd(wall_case, k):
    fd, r = wall_case
    dth = np.zeros(3)
    dth[k] = 1.0
    dv = np.zeros((fd.np, 3))
    ad = ad_directional(r, fd.theta0, dth, dv)
    fdv = fd.directional(dth, dv, 1e-6)
    print(f"wall refresh, param {k}: AD {ad:.9e} FD {fdv:.9e}", flush=True)
    assert abs(ad - fdv) <= FD_TOL * abs(fdv), (k, ad, fdv)


def test_wall_refresh_checkpoint_interval_invariant(wall_case):
    """The refresh's records of a recomputed segment equal the recorded run's."""
    fd, r = wall_case
    _, P = build_state(fd.cfg, fd.met, SEED, JITTER, *fd.theta0, None)
    fd.S.set_vent(fd.cfg.vent_v, fd.cfg.vent_T, fd.cfg.ng0)
    fd.S.load(P)
    r2 = fd.S.adjoint_run(24, 24)
    for k in ("H", "dH_dvent_v", "dH_dvent_T", "dH_dmass_frac"):
        assert np.isclose(r2[k], r[k], rtol=1e-9, atol=0), (k, r2[k], r[k])
    assert np.allclose(r2["grad_v"], r["grad_v"], rtol=1e-8, atol=1e-14 * np.abs(r["grad_v"]).max())


def test_torch_function_routes_adjoint(adjoint_case):
    fd, r = adjoint_case
    th = [torch.tensor(float(x), dtype=torch.float64, requires_grad=True) for x in fd.theta0]
    dv0 = torch.zeros((fd.np, 3), dtype=torch.float64, requires_grad=True)
    H = plume_height(*th, dv0, cfg=fd.cfg, met=fd.met, seed=SEED, jitter=JITTER, nsteps=24, ckpt_every=7)
    H.backward()
    assert abs(H.item() - r["H"]) <= 1e-9 * abs(r["H"])
    assert np.isclose(th[0].grad.item(), r["dH_dvent_v"], rtol=1e-9)
    assert np.isclose(th[1].grad.item(), r["dH_dvent_T"], rtol=1e-9)
    assert np.isclose(th[2].grad.item(), r["dH_dmass_frac"], rtol=1e-9)
    assert np.allclose(dv0.grad.numpy(), r["grad_v"], rtol=1e-9, atol=0)


def test_checkpoint_interval_invariant(adjoint_case):
    fd, r = adjoint_case
    _, P = build_state(fd.cfg, fd.met, SEED, JITTER, *fd.theta0, None)
    fd.S.set_vent(fd.cfg.vent_v, fd.cfg.vent_T, fd.cfg.ng0)
    fd.S.load(P)
    r2 = fd.S.adjoint_run(24, 24)            # one segment: no recompute
    for k in ("H", "dH_dvent_v", "dH_dvent_T", "dH_dmass_frac"):
        assert np.isclose(r2[k], r[k], rtol=1e-9, atol=0), (k, r2[k], r[k])
    assert np.allclose(r2["grad_v"], r["grad_v"], rtol=1e-8, atol=1e-14 * np.abs(r["grad_v"]).max())


def test_inlet_ghost_mass_terms_every_interval():
    """A shallow vent column (150 m instead of 12 h_air): ghosts the inlet adds
    during the window reach the surface, meet real particles and carry a mass
    adjoint (their mass is the vent's, a function of T and the gas fraction).
    Every checkpoint interval must give the one-segment gradient, which runs
    no recompute: a recompute restores the id counter of its checkpoint, so
    the mass terms must take the id range of the whole run."""
    cfg = small_plume_config(sml_update_int=10)
    cfg.vent_bot = -150.0
    P = psph.gen_plume(cfg, psph.meteo_standard(), SEED, JITTER)
    res = {}
    for K in (20, 4, 7):
        with psph.Solver(cfg, len(P) + 40000) as S:
            S.load(P)
            res[K] = S.adjoint_run(20, K, grads=True)
        print(f"K={K}: dH/dT={res[K]['dH_dvent_T']:.9e} dH/dmf={res[K]['dH_dmass_frac']:.9e}", flush=True)
    # the recompute reproduces the recorded steps up to the last bit of a few
    # per-particle values; summing the mass terms of the first segment's inlet
    # ghosts only moves dH/dT and dH/dmass_frac by 15-50 % here
    for K in (4, 7):
        for k in ("H", "dH_dvent_v", "dH_dvent_T", "dH_dmass_frac"):
            assert abs(res[K][k] - res[20][k]) <= 1e-12 * abs(res[20][k]), (K, k, res[K][k], res[20][k])
        g, g0 = res[K]["grad_v"], res[20]["grad_v"]
        assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max(), (K, np.abs(g - g0).max())


def test_smoothed_velocity_gradient_in_motion():
    """dH/dsv, one particle component at a time, from a state 6 steps into the
    run.  (Random directions over every particle are not a fair check of sv:
    an erupt ghost that has just turned real still has sv = 0, as do the
    ghosts, so some pairs sit exactly at v_ij = 0, where the heat term has a
    |v_ij| kink.)"""
    cfg, met = plume_case()
    theta0 = np.array([cfg.vent_v, cfg.vent_T, 1.0 - cfg.ng0])
    _, P = build_state(cfg, met, SEED, JITTER, *theta0, None)
    with psph.Solver(cfg, len(P) + 20000) as S:
        S.load(P)
        S.run(6)
        t, step = S.time()
        Q = S.read()
        S.load(Q, t, step)
        r = S.adjoint_run(6, 3)
        assert r["n_particles"] == len(Q)
        g = r["grad_sv"]
        real = np.nonzero(Q["bc"] == psph.BC_REAL)[0]
        rng = np.random.default_rng(SEED + 2000)
        pick = list(real[np.argsort(-np.abs(g[real]).max(1))[:4]]) + list(rng.choice(real, 4, replace=False))
        # particles with dH/dsv ~ 1e-4 of the largest move H by ~1e-11 m at this
        # step: 1e-5 relative keeps the difference well above FP64 round-off of H
        eps = 1e-5 * cfg.vent_v
        worst = 0.0
        for i in pick:
            for k, f in enumerate(("svx", "svy", "svz")):
                hs = []
                for sgn in (1, -1):
                    Qp = Q.copy()
                    Qp[f][i] += sgn * eps
                    S.load(Qp, t, step)
                    S.replay(6)
                    hs.append(S.diag_height())
                fdv = (hs[0] - hs[1]) / (2 * eps)
                err = abs(g[i, k] - fdv) / max(abs(fdv), 1e-2 * np.abs(g).max())
                worst = max(worst, err)
                print(f"particle {i} {f}: AD {g[i, k]:.9e} FD {fdv:.9e} rel {err:.2e}", flush=True)
        assert worst <= FD_TOL, worst
