"""Closed-form pins of the oracle functions whose other checks a plausible
slip would pass: the THashTable slot, 3D adaptive smoothing length, the
turbulent heat-conduction term, the energy-smoothing ratio, and the
erupt-ghost -> real / escape bookkeeping of the position update.

Each expected value below is derived from the mathematics of the reference's
formulas (sums over an explicitly enumerated lattice, identities such as
grad W . dx = -2 r^2 W / h^2 for the Gaussian), not from the oracle's code.
"""
import math

import numpy as np
import pytest

import oracle
import psph
from helpers import small_plume_config

UMAX = 4294967295              # 2^32 - 1, the key scale of THashTable::hash
EPS_HALF, PRANDTL = 0.4, 0.85  # EPSILON_HALF, PRANDTL_NUM (constant.h)


# ---------------------------------------------------------------- slot
@pytest.mark.parametrize("nslots", [4000000, 4000007, 10])
def test_slot_special_values(nslots):
    """THashTable::hash (thashtab.cc:209-222): slot = floor(k0/umax*S01) +
    floor(k1/umax*S02), S02 = N/10 (integer division), S01 = N - S02.  The
    corners of the key square land on 0, S01, S02 and N exactly."""
    s02 = nslots // 10
    s01 = nslots - s02
    assert oracle.slot(0, 0, nslots) == 0
    assert oracle.slot(UMAX, 0, nslots) == s01
    assert oracle.slot(0, UMAX, nslots) == s02
    assert oracle.slot(UMAX, UMAX, nslots) == nslots


def test_slot_floor_boundaries_and_monotonicity():
    """With N = 10 (S01 = 9, S02 = 1) the first key that reaches slot 1 is the
    smallest k0 with 9 k0 >= umax, i.e. ceil(umax / 9) = 477218589; the key
    before it still floors to 0.  k1 moves the slot only at k1 = umax."""
    first = -(-UMAX // 9)
    assert first == 477218589
    assert oracle.slot(first - 1, 0, 10) == 0
    assert oracle.slot(first, 0, 10) == 1
    assert oracle.slot(0, UMAX - 1, 10) == 0
    assert oracle.slot(first, UMAX, 10) == 2
    rng = np.random.default_rng(11)
    k0 = np.sort(rng.integers(0, UMAX, 300, dtype=np.int64))
    for k1 in (0, 12345, UMAX // 3):
        s = [oracle.slot(int(a), k1, 4000000) for a in k0]
        assert all(b >= a for a, b in zip(s, s[1:]))
    k1 = np.sort(rng.integers(0, UMAX, 300, dtype=np.int64))
    s = [oracle.slot(7, int(b), 4000000) for b in k1]
    assert all(b >= a for a, b in zip(s, s[1:]))
    # the key[0] part spans 9/10 of the table, the key[1] part 1/10
    assert oracle.slot(UMAX // 2 + 1, 0, 4000000) == 1800000
    assert oracle.slot(0, UMAX // 2 + 1, 4000000) == 200000


# ---------------------------------------------------------------- adaptive h
def uniform_lattice_plume():
    """Perfect lattice (no jitter) with one mass everywhere and every lattice
    particle contributing to density (wall ghosts relabelled as pressure
    ghosts); h = spacing a."""
    cfg = small_plume_config()
    met = psph.meteo_standard()
    P = psph.gen_plume(cfg, met, 1, 0.0)
    P["m"][:] = 2.5e5
    P["bc"][P["bc"] == 2] = 1
    return cfg, P


def gauss3(h):
    return (1.0 / (math.sqrt(math.pi) * h)) ** 3


def adaptive_closed_form(cfg, L, m, h0, loops):
    """Iterates h <- eta cbrt(m / rho*) with rho* = sum m W(n a / H*, H*),
    H* = C_smooth h, over the stale list of the particle at lattice index L:
    offsets n with max|n_k| <= 4 (inside 4.25 h0) or in the particle's own
    bucket (8 x 8 x 8 lattice cells), truncated to the kernel box |n_k a| <=
    3 H*.  Stops after `loops` iterations or when |dh / h| < thresh."""
    a = cfg.h_air
    B = [L[k] // 8 for k in range(3)]
    R = 20
    offs = []
    for n0 in range(-R, R + 1):
        for n1 in range(-R, R + 1):
            for n2 in range(-R, R + 1):
                n = (n0, n1, n2)
                same = all((L[k] + n[k]) // 8 == B[k] for k in range(3))
                if same or max(abs(v) for v in n) <= 4:
                    offs.append(n)
    h = h0
    seq = []
    for _ in range(loops):
        Hs = cfg.c_smooth * h
        terms = [math.exp(-(n[0] ** 2 + n[1] ** 2 + n[2] ** 2) * a * a / (Hs * Hs))
                 for n in offs if all(abs(v) * a <= 3 * Hs for v in n)]
        rho = m * gauss3(Hs) * math.fsum(terms)
        hn = cfg.eta * (m / rho) ** (1.0 / 3.0)
        seq.append(hn)
        err = abs((hn - h) / h)
        h = hn
        if err < cfg.thresh:
            break
    return seq


@pytest.mark.parametrize("loops", [1, 2])
def test_adaptive_sml_lattice_closed_form(loops):
    """neighs.cc:479-693 (ADAPTIVE_SML 1, 3D): eta = 1.5, C_smooth = 2, the
    list built at h = a (own bucket + 4.25 h box) is reused unchanged."""
    cfg, P = uniform_lattice_plume()
    cfg.num_loop = loops
    S = oracle.State(cfg, P)
    S.neighbors()
    S.adaptive_sml()
    Q = S.particles()
    a = cfg.h_air
    L = np.round((np.stack([Q["x"], Q["y"], Q["z"]], 1) - np.array(cfg.ext_lo)) / a - 0.5).astype(int)
    picks = []
    # real particles above z = 600 m: the whole +-6 lattice box around them is
    # filled with contributing particles (air, or ghosts of the extended domain)
    for want in ((0, 0, 2), (4, 3, 7), (7, 7, 7), (2, 5, 3)):     # position inside the bucket
        sel = np.nonzero((Q["bc"] == 100) & np.all(L % 8 == want, axis=1) & (Q["z"] > 600))[0]
        assert sel.size, want
        picks.append(sel[0])
    seen = set()
    for i in picks:
        seq = adaptive_closed_form(cfg, tuple(L[i]), 2.5e5, a, loops)
        assert len(seq) == loops
        assert Q["h"][i] == pytest.approx(seq[-1], rel=1e-12)
        seen.add(round(seq[-1], 9))
    # the own-bucket extension depends on where the particle sits in its bucket
    assert len(seen) == len(picks)
    if loops == 2:
        # second pass: H* = 2 h1 reaches past the stale list, so truncation matters
        full = adaptive_closed_form(cfg, tuple(L[picks[1]]), 2.5e5, a, 1)[0]
        assert abs(Q["h"][picks[1]] - full) > 1e-3 * full


def test_adaptive_sml_stops_below_threshold():
    """With many iterations allowed the loop stops at the first iterate whose
    relative change is below thresh (1e-5)."""
    cfg, P = uniform_lattice_plume()
    cfg.num_loop = 60
    S = oracle.State(cfg, P)
    S.neighbors()
    S.adaptive_sml()
    Q = S.particles()
    a = cfg.h_air
    L = np.round((np.stack([Q["x"], Q["y"], Q["z"]], 1) - np.array(cfg.ext_lo)) / a - 0.5).astype(int)
    i = np.nonzero((Q["bc"] == 100) & np.all(L % 8 == (3, 3, 3), axis=1) & (Q["z"] > 600))[0][0]
    seq = adaptive_closed_form(cfg, tuple(L[i]), 2.5e5, a, 60)
    assert 2 < len(seq) < 60
    assert Q["h"][i] == pytest.approx(seq[-1], rel=1e-12)
    assert abs(seq[-1] - seq[-2]) / seq[-2] < cfg.thresh <= abs(seq[-2] - seq[-3]) / seq[-3]


# ---------------------------------------------------------------- heat conduction
def two_particles(Ti, Tj, lans=1, heat_scale=1.0):
    cfg = psph.default_config(3)
    cfg.h_air = 1.0
    for k in range(3):
        cfg.dom_lo[k], cfg.dom_hi[k] = -20.0, 20.0
    cfg.g = 0.0
    cfg.lans = lans
    cfg.heat_scale = heat_scale
    cfg.vent_z = -20.0
    psph.finalize_config(cfg)
    P = psph.Particles(2)
    P["x"][:] = [0.0, 0.9]
    P["y"][:] = [0.0, -0.4]
    P["z"][:] = [0.0, 0.7]
    P["vx"][:] = [1.0, -2.0]
    P["svx"][:] = [1.5, -2.5]
    P["svy"][:] = [0.3, 1.1]
    P["svz"][:] = [-0.7, 0.2]
    P["rho"][:] = [1.0, 1.3]
    P["p"][:] = [1.0e5, 1.2e5]
    P["c"][:] = [300.0, 360.0]
    P["m"][:] = [1.0, 1.7]
    P["h"][:] = [1.0, 1.2]
    P["cp"][:] = [1000.0, 1300.0]
    P["T"][:] = [Ti, Tj]
    P["e"][:] = [2.0e5, 2.1e5]
    P["bc"][:] = 100
    P["phase"][:] = 1
    P["key0"][:] = [0, 1]
    return cfg, P


def rhs_e(cfg, P, dt):
    S = oracle.State(cfg, P)
    S.neighbors()
    S.rhs(dt)
    Q = S.particles()
    return Q["e"] - P["e"]


@pytest.mark.parametrize("heat_scale", [1.0, 2.0])
def test_heat_conduction_two_particle_closed_form(heat_scale):
    """SPH_epsilon_heat_conductivity (sph_lib.cc:1399-1439) and compute_F
    (:257-281) inside mom_engr_update (LANS on).  Only the heat term reads T,
    so the change of de when T_j moves is exactly
        dt m_j/(rho_i rho_j) k_ij (T_i - T_j) F_ij,
    k_ij = 10 eps/2 Cp_ij r^2 |v_ij|^2 / (Pr sum_k |v_ij,k dx_k|) / (h_ij c_ij),
    and, for the Gaussian at scale lam (W_heat = W(lam s, h/lam)),
    F_ij = grad W_heat . dx / r^2 = -2 lam^2 W_heat / h^2."""
    dt = 1e-3
    cfg, P0 = two_particles(400.0, 400.0, heat_scale=heat_scale)
    _, P1 = two_particles(400.0, 250.0, heat_scale=heat_scale)
    _, P2 = two_particles(400.0, 100.0, heat_scale=heat_scale)
    d0, d1, d2 = rhs_e(cfg, P0, dt), rhs_e(cfg, P1, dt), rhs_e(cfg, P2, dt)
    heat1, heat2 = d1 - d0, d2 - d0

    dx = np.array([P0["x"][0] - P0["x"][1], P0["y"][0] - P0["y"][1], P0["z"][0] - P0["z"][1]])
    vij = np.array([P0["svx"][0] - P0["svx"][1], P0["svy"][0] - P0["svy"][1], P0["svz"][0] - P0["svz"][1]])
    h = 0.5 * (1.0 + 1.2)
    r2 = float(dx @ dx)
    lam = heat_scale
    s = dx / h
    Wh = (lam / (math.sqrt(math.pi) * h)) ** 3 * math.exp(-lam * lam * float(s @ s))
    assert np.all(np.abs(lam * s) <= 3)
    F = -2 * lam * lam * Wh / (h * h)
    cp, cbar = 0.5 * (1000.0 + 1300.0), 0.5 * (300.0 + 360.0)
    k = 10 * EPS_HALF * cp * r2 * float(vij @ vij) / (PRANDTL * float(np.sum(np.abs(vij * dx)))) / (h * cbar)
    want_i = dt * 1.7 / (1.0 * 1.3) * k * (400.0 - 250.0) * F
    assert want_i < 0                         # heat flows from the hot particle
    assert heat1[0] == pytest.approx(want_i, rel=1e-7)
    # particle j: m_i/(rho_i rho_j) k (T_j - T_i) F, F even in dx
    want_j = dt * 1.0 / (1.0 * 1.3) * k * (250.0 - 400.0) * F
    assert heat1[1] == pytest.approx(want_j, rel=1e-7)
    # linear in the temperature difference, and the pair exchanges heat only
    assert heat2[0] == pytest.approx(2 * heat1[0], rel=1e-7)
    assert 1.0 * heat1[0] + 1.7 * heat1[1] == pytest.approx(0.0, abs=1e-9 * abs(heat1[0]))


def test_heat_conduction_off_without_lans_and_for_parallel_motion():
    """LANS off: no heat term.  sum_k |v_ij,k dx_k| = 0 (relative velocity
    perpendicular to every axis component of dx): k_ij = 0."""
    dt = 1e-3
    cfg, P0 = two_particles(400.0, 400.0, lans=0)
    _, P1 = two_particles(400.0, 100.0, lans=0)
    assert np.array_equal(rhs_e(cfg, P0, dt), rhs_e(cfg, P1, dt))
    cfg, P0 = two_particles(400.0, 400.0)
    _, P1 = two_particles(400.0, 100.0)
    for Q in (P0, P1):
        Q["z"][1] = 0.0          # dx = (-0.9, 0.4, 0)
        Q["svx"][:] = 0.0        # v_ij = (0, 0, *): every v_ij,k dx_k is zero
        Q["svy"][:] = 0.0
    assert np.array_equal(rhs_e(cfg, P0, dt), rhs_e(cfg, P1, dt))


# ---------------------------------------------------------------- energy smoothing
def energy_lattice(h_center=2.0):
    """9^3 real particles at unit spacing, m = rho = 1, h = 2 (h_center at the
    middle particle), e = e0 except e0 + delta at the middle."""
    cfg = psph.default_config(3)
    cfg.h_air = 1.0
    for k in range(3):
        cfg.dom_lo[k], cfg.dom_hi[k] = -20.0, 20.0
    cfg.g = 0.0
    cfg.vent_z = -20.0
    psph.finalize_config(cfg)
    n = 9
    g = np.arange(n) - 4.0
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    N = n ** 3
    P = psph.Particles(N)
    P["x"][:], P["y"][:], P["z"][:] = X.ravel(), Y.ravel(), Z.ravel()
    P["m"][:] = 1.0
    P["rho"][:] = 1.0
    P["h"][:] = 2.0
    P["e"][:] = 3.0e5
    mid = (4 * n + 4) * n + 4
    P["h"][mid] = h_center
    P["e"][mid] += 1.0e3
    P["bc"][:] = 100
    P["phase"][:] = 1
    P["key0"][:] = np.arange(N)
    return cfg, P, mid


def w_e(nvec, h, ratio=4.0):
    """Energy-smoothing weight m/rho W(ratio dx/h, h/ratio) at unit spacing."""
    if any(abs(ratio * v / h) > 3 for v in nvec):
        return 0.0
    return (ratio / (math.sqrt(math.pi) * h)) ** 3 * math.exp(-ratio * ratio * sum(v * v for v in nvec) / (h * h))


@pytest.mark.parametrize("h_center", [2.0, 2.4])
def test_energy_smoothing_single_perturbation(h_center):
    """density_massfrac.cc:162-168 (HAVE_ENERGY_SMOOTH 1, E_SMOOTH_RATIO 4,
    h_ij = (h_i + h_j)/2): e_i <- sum w_ij e_j / sum w_ij.  A single
    perturbation delta at the middle particle c reaches a neighbour i as
    delta w_ic / sum_j w_ij; with h = 2 only |n_k| <= 1 lie in the box
    |4 n_k / h| <= 3."""
    cfg, P, mid = energy_lattice(h_center)
    S = oracle.State(cfg, P)
    S.neighbors()
    S.density()
    Q = S.particles()
    pos = np.stack([Q["x"], Q["y"], Q["z"]], 1)
    delta, e0 = 1.0e3, 3.0e5
    checked = 0
    for i in range(len(Q)):
        n = np.round(pos[i]).astype(int)
        if np.max(np.abs(n)) > 2:
            continue
        if np.max(np.abs(n)) == 2:
            assert Q["e"][i] == pytest.approx(e0, rel=1e-15), n      # outside the box of c
            continue
        num = den = 0.0
        for d0 in (-1, 0, 1):
            for d1 in (-1, 0, 1):
                for d2 in (-1, 0, 1):
                    j = n + np.array([d0, d1, d2])
                    is_c = not np.any(j)
                    hij = 0.5 * ((h_center if not np.any(n) else 2.0) + (h_center if is_c else 2.0))
                    w = w_e((d0, d1, d2), hij)
                    den += w
                    num += w * (e0 + (delta if is_c else 0.0))
        assert Q["e"][i] == pytest.approx(num / den, rel=1e-13), n
        checked += 1
    assert checked == 27
    # closed form of the uniform-h case for a face neighbour
    if h_center == 2.0:
        i = int(np.nonzero(np.all(np.round(pos).astype(int) == (1, 0, 0), axis=1))[0][0])
        assert Q["e"][i] - e0 == pytest.approx(delta * math.exp(-4.0) / (1 + 2 * math.exp(-4.0)) ** 3, rel=1e-9)


# ---------------------------------------------------------------- position update, escape
def update_case():
    cfg = small_plume_config()
    cfg.erupt_start, cfg.erupt_end = 0.0, 100.0
    z0 = cfg.dom_lo[2]
    hi0 = cfg.dom_hi[0]
    rows = [
        # bc, x, y, z, vz, svz, h
        (0, 10.0, 0.0, z0 - 0.5, 10.0, 0.0, 7.0),        # 0: ghost crosses the ground -> real
        (0, 10.0, 0.0, z0 - 2.0, 10.0, 0.0, 7.0),        # 1: ghost stays below
        (0, 10.0, 0.0, z0 - 1.0, 10.0, 0.0, 7.0),        # 2: ghost lands exactly on z_lo -> real
        (100, 0.0, 0.0, 500.0, 3.0, -4.0, 90.0),         # 3: real moves with sv
        (1, 0.0, 0.0, 500.0, 3.0, -4.0, 90.0),           # 4: pressure ghost does not move
        (100, hi0 + 0.9e-8, 0.0, 500.0, 0.0, 0.0, 90.0),  # 5: inside the 1e-8 tolerance
        (100, hi0 + 1.1e-8, 0.0, 500.0, 0.0, 0.0, 90.0),  # 6: escaped
        (100, 0.0, cfg.dom_lo[1] - 1.1e-8, 500.0, 0.0, 0.0, 90.0),  # 7: escaped at y_lo
        (100, 0.0, 0.0, cfg.dom_hi[2] + 0.5e-8, 0.0, 0.0, 90.0),    # 8: inside tolerance at z_hi
        (1, hi0 + 50.0, 0.0, 500.0, 0.0, 0.0, 90.0),     # 9: ghost outside: never removed
        (2, 0.0, 0.0, z0 - 30.0, 0.0, 0.0, 90.0),        # 10: wall ghost below ground stays
    ]
    P = psph.Particles(len(rows))
    for i, (bc, x, y, z, vz, svz, h) in enumerate(rows):
        P["bc"][i], P["x"][i], P["y"][i], P["z"][i] = bc, x, y, z
        P["vz"][i], P["svz"][i], P["h"][i] = vz, svz, h
        P["key0"][i] = i
        P["phase"][i] = 2 if bc == 0 else 1
        P["involved"][i] = 0
    return cfg, P


def test_update_pos_ghost_turns_real_and_escape():
    """update_pos.cc:34-168: real particles move with sv (LANS); erupt ghosts
    move with v while the eruption is on and turn real (bc 100, h = h_air,
    involved = 2) at z >= z_lo; other ghosts stay.  Escape
    (Bucket::determine_escape, bucket.cc:435-452): a real particle beyond a
    domain face by more than 1e-8 is removed."""
    cfg, P = update_case()
    dt = 0.1
    S = oracle.State(cfg, P)
    S.st.time = 1.0
    S.update_pos(dt)
    Q = S.particles()
    assert Q["bc"][0] == 100 and Q["h"][0] == cfg.h_air and Q["involved"][0] == 2
    assert Q["z"][0] == P["z"][0] + dt * 10.0
    assert Q["bc"][1] == 0 and Q["h"][1] == 7.0 and Q["involved"][1] == 0
    assert Q["z"][2] == cfg.dom_lo[2] and Q["bc"][2] == 100
    assert Q["z"][3] == 500.0 + dt * -4.0           # smoothed velocity, not v
    assert Q["z"][4] == 500.0 and Q["bc"][4] == 1
    S.remove_escaped()
    R = S.particles()
    assert S.st.last_removed == 2
    assert sorted(R["key0"].tolist()) == [0, 1, 2, 3, 4, 5, 8, 9, 10]


def test_update_pos_no_ghost_motion_outside_eruption():
    cfg, P = update_case()
    S = oracle.State(cfg, P)
    S.st.time = 150.0                       # after erupt_end
    S.update_pos(0.1)
    Q = S.particles()
    assert np.array_equal(Q["z"][:3], P["z"][:3])
    assert np.all(Q["bc"][:3] == 0)
    assert Q["z"][3] == 500.0 + 0.1 * -4.0


# ---------------------------------------------------------------- RHS: pressure, SPH-epsilon, gravity
def rhs_pair(lans, p, vel_j, dt=1e-4):
    """Two real particles, i at the origin, j at (1.1, 0.5, -0.3); smoothed
    velocities sv_i = 0, sv_j = vel_j (raw v = sv); returns (Q, inputs)."""
    cfg = psph.default_config(3)
    cfg.h_air = 1.0
    for k in range(3):
        cfg.dom_lo[k], cfg.dom_hi[k] = -20.0, 20.0
    cfg.g = 0.0
    cfg.lans = lans
    cfg.vent_z = -20.0
    psph.finalize_config(cfg)
    P = psph.Particles(2)
    P["x"][:] = [0.0, 1.1]
    P["y"][:] = [0.0, 0.5]
    P["z"][:] = [0.0, -0.3]
    for k, v in zip(("vx", "vy", "vz"), vel_j):
        P[k][:] = [0.0, v]
        P["s" + k][:] = [0.0, v]
    P["rho"][:] = [1.0, 1.25]
    P["p"][:] = p
    P["c"][:] = [300.0, 340.0]
    P["T"][:] = 300.0
    P["cp"][:] = 1000.0
    P["m"][:] = [1.0, 2.0]
    P["h"][:] = [1.0, 1.4]
    P["e"][:] = 1.0
    P["bc"][:] = 100
    P["phase"][:] = 1
    P["key0"][:] = [0, 1]
    S = oracle.State(cfg, P)
    S.neighbors()
    S.rhs(dt)
    return S.particles(), P


def pair_geometry():
    dx = np.array([-1.1, -0.5, 0.3])          # x_i - x_j
    h = 1.2                                   # (h_i + h_j) / 2
    s = dx / h
    W = (1 / (math.sqrt(math.pi) * h)) ** 3 * math.exp(-float(s @ s))
    gradW = -2 * s / h * W
    return dx, gradW


def test_rhs_pressure_pair_closed_form():
    """Symmetric p/rho^2 form (mom_engr_update.cc:96-330, LANS off, receding
    pair so Pi = 0): dv_i = -dt m_j (p_j/rho_j^2 + p_i/rho_i^2) grad_i W, which
    pushes i away from j; de_i = dt/2 m_j (...) grad_i W . v_ij."""
    dt = 1e-4
    vel_j = (1.0, 0.4, -0.2)                  # j moves away from i: dx . v_ij > 0
    Q, P = rhs_pair(0, [2.0e5, 3.0e5], vel_j, dt)
    dx, gradW = pair_geometry()
    vij = -np.array(vel_j)
    assert dx @ vij > 0
    M = 2.0 * (3.0e5 / 1.25 ** 2 + 2.0e5 / 1.0)
    dv = -dt * M * gradW
    assert np.allclose([Q["vx"][0], Q["vy"][0], Q["vz"][0]], dv, rtol=1e-12, atol=0)
    assert dv @ dx > 0                        # repulsion: along x_i - x_j
    assert Q["e"][0] - 1.0 == pytest.approx(dt * 0.5 * M * float(gradW @ vij), rel=1e-9)


def test_rhs_sph_epsilon_momentum_closed_form():
    """SPH_epsilon_mom (sph_lib.cc:1331-1338) with LANS: Pi -= eps/2 |v_ij|^2 /
    rho_j; with zero pressure and a receding pair it is the only term:
    dv_i = dt m_j 0.4 |v_ij|^2 / rho_j grad_i W (an attraction)."""
    dt = 1e-4
    vel_j = (1.0, 0.4, -0.2)
    Q, _ = rhs_pair(1, [0.0, 0.0], vel_j, dt)
    dx, gradW = pair_geometry()
    vij = -np.array(vel_j)
    dv = dt * 2.0 * EPS_HALF * float(vij @ vij) / 1.25 * gradW
    assert np.allclose([Q["vx"][0], Q["vy"][0], Q["vz"][0]], dv, rtol=1e-12, atol=0)
    assert dv @ dx < 0
    Q0, _ = rhs_pair(0, [0.0, 0.0], vel_j, dt)
    assert Q0["vx"][0] == 0.0 and Q0["vy"][0] == 0.0 and Q0["vz"][0] == 0.0


def test_rhs_gravity_single_particle():
    """An isolated particle (its list holds only itself): v_z += -g dt and
    e += dt v_z g_z with the smoothed velocity under LANS
    (mom_engr_update.cc:343-354)."""
    cfg = psph.default_config(3)
    cfg.h_air = 1.0
    for k in range(3):
        cfg.dom_lo[k], cfg.dom_hi[k] = -20.0, 20.0
    cfg.vent_z = -20.0
    psph.finalize_config(cfg)
    assert cfg.g == 9.80665
    P = psph.Particles(1)
    P["vx"][:], P["vy"][:], P["vz"][:] = 1.0, 2.0, 3.0
    P["svx"][:], P["svy"][:], P["svz"][:] = 1.0, 2.0, 5.0
    P["rho"][:], P["p"][:], P["c"][:], P["m"][:], P["h"][:] = 1.0, 1e5, 300.0, 1.0, 1.0
    P["e"][:] = 2.0e5
    P["bc"][:] = 100
    P["phase"][:] = 1
    dt = 0.01
    S = oracle.State(cfg, P)
    S.neighbors()
    S.rhs(dt)
    Q = S.particles()
    assert (Q["vx"][0], Q["vy"][0]) == (1.0, 2.0)
    assert Q["vz"][0] == pytest.approx(3.0 - 9.80665 * dt, rel=1e-15)
    assert Q["e"][0] == pytest.approx(2.0e5 + dt * 5.0 * -9.80665, rel=1e-15)
