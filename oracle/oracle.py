"""ctypes wrapper of the CPU FP64 restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY -- the parity checker.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only to check or time the reference algorithm; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class OrcMeteo(C.Structure):
    _fields_ = [("rows", C.c_int), ("nprop", C.c_int), ("z", C.POINTER(C.c_double)),
                ("data", C.POINTER(C.c_double))]


class OrcState(C.Structure):
    _fields_ = ([("n", C.c_size_t), ("cap", C.c_size_t)]
                + [("x", C.POINTER(C.c_double) * 3), ("v", C.POINTER(C.c_double) * 3),
                   ("sv", C.POINTER(C.c_double) * 3)]
                + [(f, C.POINTER(C.c_double)) for f in
                   ["rho", "e", "m", "h", "p", "c", "T", "gam", "cp", "xi"]]
                + [("key", C.POINTER(C.c_uint32) * 3)]
                + [(f, C.POINTER(C.c_int32)) for f in ["bc", "phase", "inv"]]
                + [("nb_off", C.POINTER(C.c_int64)), ("nb_idx", C.POINTER(C.c_uint32)),
                   ("nb_cap", C.c_size_t),
                   ("time", C.c_double), ("next_output", C.c_double), ("step", C.c_int64),
                   ("ioutput", C.c_int),
                   ("last_injected", C.c_int64), ("last_removed", C.c_int64),
                   ("last_pairs", C.c_int64),
                   ("bval", C.POINTER(C.c_uint8)), ("berupt", C.POINTER(C.c_uint8)),
                   ("nbk", C.c_int32 * 3), ("adapt_pending", C.c_int32), ("lat_seed", C.c_uint32),
                   ("lat_jitter", C.c_double), ("met", OrcMeteo), ("last_adapted", C.c_int64)])


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    import psph  # noqa: F401  (psph_config / psph_soa layouts)
    from psph import Config, SoA
    L = C.CDLL(LIB_PATH)
    S, CFG = C.POINTER(OrcState), C.POINTER(Config)
    sig = {
        "orc_weight": (C.c_double, [C.POINTER(C.c_double), C.c_double, C.c_int]),
        "orc_d_weight": (C.c_double, [C.POINTER(C.c_double), C.c_double, C.c_int, C.c_int]),
        "orc_eos": (None, [CFG, C.c_double, C.c_double, C.c_double] + [C.POINTER(C.c_double)] * 5),
        "orc_hsfc3d_bits": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "orc_thsfc3d": (None, [C.POINTER(C.c_double), C.c_uint32, C.POINTER(C.c_uint32)]),
        "orc_slot": (C.c_int64, [C.POINTER(C.c_uint32), C.c_int, C.c_int]),
        "orc_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "orc_meteo_from_arrays": (None, [C.POINTER(OrcMeteo), C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "orc_meteo_free": (None, [C.POINTER(OrcMeteo)]),
        "orc_meteo_interp": (C.c_double, [C.POINTER(OrcMeteo), C.c_double, C.c_int]),
        "orc_state_init": (C.c_int, [S, C.c_size_t]),
        "orc_state_free": (None, [S]),
        "orc_state_load": (C.c_int, [S, C.POINTER(SoA), C.c_size_t]),
        "orc_state_store": (C.c_int, [S, C.POINTER(SoA)]),
        "orc_sort_canonical": (None, [CFG, S]),
        "orc_air_lattice": (C.c_int, [CFG, C.POINTER(OrcMeteo), C.c_uint32, C.c_double, S]),
        "orc_domain_init": (C.c_int, [CFG, C.POINTER(OrcMeteo), C.c_uint32, C.c_double, S]),
        "orc_domain_setup": (C.c_int, [CFG, C.POINTER(OrcMeteo), C.c_uint32, C.c_double, S]),
        "orc_domain_scan": (C.c_int, [CFG, S]),
        "orc_domain_adapt": (C.c_int, [CFG, S]),
        "orc_bucket_type": (None, [CFG, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "orc_setup_erupt": (C.c_int, [CFG, S]),
        "orc_shock_tube": (C.c_int, [CFG] + [C.c_double] * 7 + [C.c_int] * 4 + [S]),
        "orc_timestep": (C.c_double, [CFG, S]),
        "orc_incrtime": (None, [CFG, S, C.POINTER(C.c_double)]),
        "orc_neighbors": (C.c_int, [CFG, S]),
        "orc_neighbors_bruteforce": (C.c_int, [CFG, S]),
        "orc_adaptive_sml": (C.c_int, [CFG, S]),
        "orc_rhs": (C.c_int, [CFG, S, C.c_double]),
        "orc_density": (C.c_int, [CFG, S]),
        "orc_vsmooth": (C.c_int, [CFG, S]),
        "orc_update_pos": (C.c_int, [CFG, S, C.c_double]),
        "orc_remove_escaped": (C.c_int, [CFG, S]),
        "orc_inject": (C.c_int, [CFG, S, C.c_double]),
        "orc_step": (C.c_int, [CFG, S]),
        "orc_apply_bcond": (C.c_int, [CFG, S]),
        "orc_iferupt": (C.c_int, [CFG, S]),
        "orc_conservation": (None, [S, C.POINTER(C.c_double)]),
        "orc_diag_height": (C.c_double, [S]),
        "orc_set_threads": (C.c_int, [C.c_int]),
        "orc_get_threads": (C.c_int, []),
        "orc_last_error": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    _lib = L
    return L


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(f"oracle error {rc}: {lib().orc_last_error().decode()}")


class State:
    """Oracle particle state (kept in the reference's iteration order)."""

    def __init__(self, cfg, P=None, t0=0.0):
        self.cfg = cfg.copy()
        self.st = OrcState()
        _check(lib().orc_state_init(C.byref(self.st), max(16, len(P) if P is not None else 16)))
        self.st.time = t0
        self.st.next_output = cfg.t_output
        if P is not None:
            s = P.soa()
            _check(lib().orc_state_load(C.byref(self.st), C.byref(s), len(P)))

    def __del__(self):
        try:
            lib().orc_state_free(C.byref(self.st))
        except Exception:
            pass

    def __len__(self):
        return self.st.n

    def particles(self):
        from psph import Particles
        P = Particles(self.st.n)
        s = P.soa()
        _check(lib().orc_state_store(C.byref(self.st), C.byref(s)))
        return P

    def sort(self):
        lib().orc_sort_canonical(C.byref(self.cfg), C.byref(self.st))

    # stages
    def timestep(self):
        return lib().orc_timestep(C.byref(self.cfg), C.byref(self.st))

    def incrtime(self, dt):
        d = C.c_double(dt)
        lib().orc_incrtime(C.byref(self.cfg), C.byref(self.st), C.byref(d))
        return d.value

    def neighbors(self, brute=False):
        f = lib().orc_neighbors_bruteforce if brute else lib().orc_neighbors
        _check(f(C.byref(self.cfg), C.byref(self.st)))
        n = self.st.n
        off = np.ctypeslib.as_array(self.st.nb_off, (n + 1,)).copy()
        tot = int(off[-1])
        idx = np.ctypeslib.as_array(self.st.nb_idx, (max(tot, 1),))[:tot].copy() if tot else np.zeros(0, np.uint32)
        return off, idx

    def adaptive_sml(self):
        _check(lib().orc_adaptive_sml(C.byref(self.cfg), C.byref(self.st)))

    def rhs(self, dt):
        _check(lib().orc_rhs(C.byref(self.cfg), C.byref(self.st), dt))

    def density(self):
        _check(lib().orc_density(C.byref(self.cfg), C.byref(self.st)))

    def vsmooth(self):
        _check(lib().orc_vsmooth(C.byref(self.cfg), C.byref(self.st)))

    def update_pos(self, dt):
        _check(lib().orc_update_pos(C.byref(self.cfg), C.byref(self.st), dt))

    def remove_escaped(self):
        _check(lib().orc_remove_escaped(C.byref(self.cfg), C.byref(self.st)))

    def inject(self, dt):
        _check(lib().orc_inject(C.byref(self.cfg), C.byref(self.st), dt))
        return self.st.last_injected

    def step(self):
        _check(lib().orc_step(C.byref(self.cfg), C.byref(self.st)))

    def apply_bcond(self):
        _check(lib().orc_apply_bcond(C.byref(self.cfg), C.byref(self.st)))

    def conservation(self):
        out = (C.c_double * 7)()
        lib().orc_conservation(C.byref(self.st), out)
        return np.array(out[:])

    def diag_height(self):
        return lib().orc_diag_height(C.byref(self.st))


def meteo(met):
    m = OrcMeteo()
    z = np.ascontiguousarray(met.z, np.float64)
    d = np.ascontiguousarray(met.data, np.float64)
    lib().orc_meteo_from_arrays(C.byref(m), len(z), d.shape[0], z.ctypes.data_as(C.POINTER(C.c_double)),
                                d.ctypes.data_as(C.POINTER(C.c_double)))
    return m


def air_lattice(cfg, met, seed, jitter):
    """Oracle restatement of the synthetic plume input (lattice + vent)."""
    S = State(cfg)
    m = meteo(met)
    _check(lib().orc_air_lattice(C.byref(S.cfg), C.byref(m), seed, jitter, C.byref(S.st)))
    _check(lib().orc_setup_erupt(C.byref(S.cfg), C.byref(S.st)))
    lib().orc_meteo_free(C.byref(m))
    return S


def domain_setup(cfg, met, seed, jitter):
    """Oracle restatement of the initial state of a run with domain adaptation
    (the initial domain's buckets, their pressure-ghost and wall-ghost layers,
    the vent) and its bucket flags."""
    S = State(cfg)
    m = meteo(met)
    _check(lib().orc_domain_setup(C.byref(S.cfg), C.byref(m), seed, jitter, C.byref(S.st)))
    lib().orc_meteo_free(C.byref(m))
    return S


def domain_init(S, met, seed, jitter):
    """Bucket flags of an oracle state for domain adaptation."""
    m = meteo(met)
    _check(lib().orc_domain_init(C.byref(S.cfg), C.byref(m), seed, jitter, C.byref(S.st)))
    lib().orc_meteo_free(C.byref(m))


def bucket_values(S):
    """has_involved of every bucket, shape (nbz, nby, nbx)."""
    nb = tuple(S.st.nbk)
    n = nb[0] * nb[1] * nb[2]
    return np.ctypeslib.as_array(S.st.bval, (n,)).copy().reshape(nb[2], nb[1], nb[0])


def bucket_type(cfg, bx, by, bz):
    bt = (C.c_int * 6)()
    t = C.c_int()
    lib().orc_bucket_type(C.byref(cfg), bx, by, bz, bt, C.byref(t))
    return t.value, tuple(bt)


def shock_tube(cfg, left, right, mid=0.0, n_nosave_l=40, n_nosave_r=20, n_ghost=35, equal_mass=True):
    S = State(cfg)
    _check(lib().orc_shock_tube(C.byref(S.cfg), *left, *right, mid, n_nosave_l, n_nosave_r, n_ghost,
                                int(equal_mass), C.byref(S.st)))
    return S


def weight(s, h, dim):
    a = (C.c_double * 3)(*(list(s) + [0] * (3 - len(s))))
    return lib().orc_weight(a, h, dim)


def d_weight(s, h, d, dim):
    a = (C.c_double * 3)(*(list(s) + [0] * (3 - len(s))))
    return lib().orc_d_weight(a, h, d, dim)


def eos(cfg, rho, e, xi):
    o = [C.c_double() for _ in range(5)]
    o[4].value = 0.0
    lib().orc_eos(C.byref(cfg), rho, e, xi, *[C.byref(v) for v in o])
    return tuple(v.value for v in o)   # p, c, T, gam, cp


def hsfc3d_bits(cx, cy, cz):
    c = (C.c_uint32 * 3)(cx, cy, cz)
    k = (C.c_uint32 * 2)()
    lib().orc_hsfc3d_bits(c, k)
    return k[0], k[1]


def thsfc3d(normc, add_step=0):
    a = (C.c_double * 3)(*normc)
    k = (C.c_uint32 * 3)()
    lib().orc_thsfc3d(a, add_step, k)
    return tuple(k)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return tuple(o)


def slot(k0, k1, table_size, dim=3):
    k = (C.c_uint32 * 3)(k0, k1, 0)
    return lib().orc_slot(k, table_size, dim)
