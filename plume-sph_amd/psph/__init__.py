"""psph -- Python mirror of the Plume-SPH solver-step interface over the
MI355X C ABI (include/psph.h), loaded with ctypes.

The reference drives its hot path through C++ free functions declared in
src/plume/particler.h (search_neighs_consth, adaptive_sml, mom_engr_update,
smooth_density, smooth_velocity, update_pos, add_new_erupt, timestep); the
same names are exposed here as methods of :class:`Solver`, each calling the
corresponding C entry point.  Errors raise :class:`PsphError` with the code
and message of the library (the reference asserts or exits instead).

The library must be built (``plume-sph_amd/build.sh``); there is no CPU
fallback: without the HIP library every compute call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PSPH_LIB names another build of the library (a variant under study)
LIB_PATH = os.environ.get("PSPH_LIB") or os.path.join(_HERE, "libpsph.so")

PSPH_FP64, PSPH_FP32_PAIR, PSPH_FP32_PURE, PSPH_FP32_EXP = 0, 1, 2, 3
PSPH_KERNEL_GAUSS_BOX = 0
BC_ERUPT_GHOST, BC_PRESSURE_GHOST, BC_WALL_GHOST, BC_NOSAVE_REAL, BC_REAL = 0, 1, 2, 99, 100

ERRORS = {0: "OK", -1: "EINVAL", -2: "ENOMEM", -3: "EINVARIANT", -4: "ECAPACITY",
          -5: "EHIP", -6: "EIO", -7: "ENOTSUP"}
(PSPH_OK, PSPH_EINVAL, PSPH_ENOMEM, PSPH_EINVARIANT, PSPH_ECAPACITY, PSPH_EHIP, PSPH_EIO,
 PSPH_ENOTSUP) = 0, -1, -2, -3, -4, -5, -6, -7


class PsphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"psph error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


_INT_FIELDS = ["dim", "precision", "kernel", "lans", "energy_smooth", "pghost_contrib", "list_all",
               "num_loop", "sml_update_int", "adapt_at_setup", "vent_profile", "max_steps",
               "num_erupt_particles", "p_table_size", "wall_bc"]
_DBL_FIELDS = ["cutoff", "cutoff2", "list_factor", "alpha", "beta", "ata", "eps_half", "prandtl",
               "e_smooth_ratio", "heat_scale", "g", "cfl", "thresh", "c_smooth", "eta",
               "ratio_sml_dx", "msfrc_thresh", "ng0", "cvs", "cvg", "cva", "rg", "ra", "gamma_p",
               "h_air", "bucket_edge"]
_DBL_TAIL = ["t_end", "t_output", "erupt_start", "erupt_end", "vent_r", "vent_v", "vent_T",
             "vent_p", "vent_x", "vent_y", "vent_z", "vent_rho", "vent_e", "vent_gamma", "vent_c",
             "vent_sml", "vent_mass", "vent_bot"]


class Config(C.Structure):
    """psph_config (include/psph.h)."""
    _fields_ = ([(f, C.c_int32) for f in _INT_FIELDS] + [("adapt_domain", C.c_int32)]
                + [(f, C.c_double) for f in _DBL_FIELDS]
                + [("dom_lo", C.c_double * 3), ("dom_hi", C.c_double * 3),
                   ("ext_lo", C.c_double * 3), ("ext_hi", C.c_double * 3)]
                + [(f, C.c_double) for f in _DBL_TAIL]
                + [("idom_lo", C.c_double * 3), ("idom_hi", C.c_double * 3), ("reserved_d", C.c_double * 2)])

    def copy(self) -> "Config":
        c = Config()
        C.memmove(C.byref(c), C.byref(self), C.sizeof(Config))
        return c

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


class SoA(C.Structure):
    """psph_soa (include/psph.h)."""
    _fields_ = ([(f, C.POINTER(C.c_double)) for f in
                 ["x", "y", "z", "vx", "vy", "vz", "svx", "svy", "svz", "rho", "e", "m", "h", "p",
                  "c", "T", "gam", "cp", "xi"]]
                + [(f, C.POINTER(C.c_uint32)) for f in ["key0", "key1", "key2"]]
                + [(f, C.POINTER(C.c_int32)) for f in ["bc", "phase", "involved"]])


class StepOut(C.Structure):
    _fields_ = [("dt", C.c_double), ("time", C.c_double), ("step", C.c_int64),
                ("n_total", C.c_int64), ("n_real", C.c_int64), ("n_injected", C.c_int64),
                ("n_removed", C.c_int64), ("n_pairs", C.c_int64),
                ("output", C.c_int64), ("n_domain", C.c_int64)]


class AdjointOut(C.Structure):
    _fields_ = [("H", C.c_double), ("dH_dvent_v", C.c_double), ("dH_dvent_T", C.c_double),
                ("dH_dng0", C.c_double), ("dH_dmass_frac", C.c_double), ("forward_ms", C.c_double),
                ("backward_ms", C.c_double), ("steps", C.c_int64), ("checkpoints", C.c_int64),
                ("n_particles", C.c_int64), ("reserved", C.c_double * 4)]


class Meteo(C.Structure):
    _fields_ = [("rows", C.c_int32), ("nprop", C.c_int32),
                ("z", C.POINTER(C.c_double)), ("data", C.POINTER(C.c_double))]


D_FIELDS = ["x", "y", "z", "vx", "vy", "vz", "svx", "svy", "svz", "rho", "e", "m", "h", "p", "c",
            "T", "gam", "cp", "xi"]
U_FIELDS = ["key0", "key1", "key2"]
I_FIELDS = ["bc", "phase", "involved"]
ALL_FIELDS = D_FIELDS + U_FIELDS + I_FIELDS


class Particles:
    """Host particle set: one numpy array per field, psph_soa view on demand."""

    def __init__(self, n: int = 0, arrays: dict | None = None):
        if arrays is not None:
            self.a = {k: np.ascontiguousarray(arrays[k]) for k in ALL_FIELDS}
        else:
            self.a = {}
            for k in D_FIELDS:
                self.a[k] = np.zeros(n, np.float64)
            for k in U_FIELDS:
                self.a[k] = np.zeros(n, np.uint32)
            for k in I_FIELDS:
                self.a[k] = np.zeros(n, np.int32)

    def __len__(self):
        return len(self.a["x"])

    def __getitem__(self, k):
        return self.a[k]

    def soa(self) -> SoA:
        s = SoA()
        for k in D_FIELDS:
            setattr(s, k, self.a[k].ctypes.data_as(C.POINTER(C.c_double)))
        for k in U_FIELDS:
            setattr(s, k, self.a[k].ctypes.data_as(C.POINTER(C.c_uint32)))
        for k in I_FIELDS:
            setattr(s, k, self.a[k].ctypes.data_as(C.POINTER(C.c_int32)))
        return s

    def take(self, n: int) -> "Particles":
        return Particles(arrays={k: v[:n].copy() for k, v in self.a.items()})

    def copy(self) -> "Particles":
        return Particles(arrays={k: v.copy() for k, v in self.a.items()})


_lib = None


def lib() -> C.CDLL:
    """Load libpsph.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PsphError(-7, f"{LIB_PATH} is not built: run plume-sph_amd/build.sh")
    # One HIP runtime per process: torch ships its own libamdhip64 (same
    # soname, another file).  Loaded first, it is the one libpsph binds to;
    # loaded after libpsph had pulled in the system's copy, a second runtime
    # comes up that finds no GPU.  So torch, where installed, is loaded first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P, CFG, SOA = C.c_void_p, C.POINTER(Config), C.POINTER(SoA)
    sig = {
        "psph_config_default": (C.c_int, [CFG, C.c_int]),
        "psph_config_vent": (C.c_int, [CFG]),
        "psph_config_domain": (C.c_int, [CFG]),
        "psph_read_deck": (C.c_int, [C.c_char_p, CFG, C.c_char_p, C.c_size_t, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "psph_meteo_read": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "psph_meteo_standard": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_int32)]),
        "psph_keys": (C.c_int, [CFG, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_size_t, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "psph_gen_plume": (C.c_int, [CFG, C.POINTER(Meteo), C.c_uint32, C.c_double, SOA, C.c_size_t, C.POINTER(C.c_size_t)]),
        "psph_gen_shock_tube": (C.c_int, [CFG, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double, C.c_int32, C.c_int32, C.c_int32, C.c_int32, SOA, C.c_size_t, C.POINTER(C.c_size_t)]),
        "psph_create": (C.c_int, [CFG, C.c_size_t, C.POINTER(P)]),
        "psph_destroy": (C.c_int, [P]),
        "psph_load": (C.c_int, [P, SOA, C.c_size_t, C.c_double, C.c_int64]),
        "psph_read": (C.c_int, [P, SOA, C.c_size_t, C.POINTER(C.c_size_t)]),
        "psph_count": (C.c_int, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
        "psph_time": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
        "psph_get_config": (C.c_int, [P, CFG]),
        "psph_device_view": (C.c_int, [P, SOA, C.POINTER(C.c_size_t), C.POINTER(C.c_void_p)]),
        "psph_timestep": (C.c_int, [P, C.POINTER(C.c_double)]),
        "psph_advance_time": (C.c_int, [P, C.POINTER(C.c_double)]),
        "psph_neighbors": (C.c_int, [P]),
        "psph_adapt_h": (C.c_int, [P]),
        "psph_rhs": (C.c_int, [P, C.c_double]),
        "psph_density": (C.c_int, [P]),
        "psph_vsmooth": (C.c_int, [P]),
        "psph_update_pos": (C.c_int, [P, C.c_double]),
        "psph_inject": (C.c_int, [P, C.c_double, C.POINTER(C.c_int64)]),
        "psph_apply_bcond": (C.c_int, [P]),
        "psph_step": (C.c_int, [P, C.POINTER(StepOut)]),
        "psph_domain_init": (C.c_int, [P, C.POINTER(Meteo), C.c_uint32, C.c_double]),
        "psph_run": (C.c_int, [P, C.c_int64, C.POINTER(StepOut)]),
        "psph_get_neighbors": (C.c_int, [P, C.POINTER(C.c_int64), C.c_size_t, C.POINTER(C.c_uint32), C.c_size_t]),
        "psph_conservation": (C.c_int, [P, C.POINTER(C.c_double)]),
        "psph_diag_height": (C.c_int, [P, C.POINTER(C.c_double)]),
        "psph_conservation_noalloc": (C.c_int, [SOA, C.c_size_t, C.c_int32, C.c_void_p, C.c_void_p]),
        "psph_last_error": (C.c_char_p, []),
        "psph_version": (C.c_char_p, []),
        "psph_abi_sizeof": (C.c_size_t, [C.c_int]),
        "psph_profile": (C.c_int, [P, C.c_int]),
        "psph_profile_read": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
        "psph_adjoint_run": (C.c_int, [P, C.c_int64, C.c_int64, C.POINTER(AdjointOut), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_size_t]),
        "psph_replay_run": (C.c_int, [P, C.c_int64]),
        "psph_set_vent": (C.c_int, [P, C.c_double, C.c_double, C.c_double]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    for which, st in enumerate((Config, SoA, StepOut, AdjointOut)):
        if L.psph_abi_sizeof(which) != C.sizeof(st):
            raise PsphError(-1, f"ABI mismatch for {st.__name__}: C {L.psph_abi_sizeof(which)} vs ctypes {C.sizeof(st)}")
    _lib = L
    return L


def check(rc: int):
    if rc != 0:
        raise PsphError(rc, lib().psph_last_error().decode(errors="replace"))


# ---------------------------------------------------------------------------
# configuration and inputs
# ---------------------------------------------------------------------------
def default_config(dim: int = 3) -> Config:
    c = Config()
    check(lib().psph_config_default(C.byref(c), dim))
    return c


def finalize_config(c: Config) -> Config:
    """Extended domain + vent-derived values (psph_config_domain, psph_config_vent)."""
    check(lib().psph_config_domain(C.byref(c)))
    check(lib().psph_config_vent(C.byref(c)))
    return c


def read_deck(path: str, cfg: Config) -> tuple[str, int, int]:
    buf = C.create_string_buffer(4096)
    rows, cols = C.c_int32(), C.c_int32()
    check(lib().psph_read_deck(path.encode(), C.byref(cfg), buf, 4096, C.byref(rows), C.byref(cols)))
    return buf.value.decode(), rows.value, cols.value


@dataclass
class MeteoTable:
    z: np.ndarray      # km
    data: np.ndarray   # (nprop, rows)

    def struct(self) -> Meteo:
        self._z = np.ascontiguousarray(self.z, np.float64)
        self._d = np.ascontiguousarray(self.data, np.float64)
        m = Meteo()
        m.rows, m.nprop = len(self._z), self._d.shape[0]
        m.z = self._z.ctypes.data_as(C.POINTER(C.c_double))
        m.data = self._d.ctypes.data_as(C.POINTER(C.c_double))
        return m


def meteo_read(path: str, rows: int, cols: int) -> MeteoTable:
    z = np.zeros(rows)
    d = np.zeros((cols - 1, rows))
    check(lib().psph_meteo_read(path.encode(), rows, cols, z.ctypes.data_as(C.POINTER(C.c_double)),
                                d.ctypes.data_as(C.POINTER(C.c_double))))
    return MeteoTable(z, d)


def meteo_standard(t0=288.0, p0=101325.0, z_top_km=40.0, dz_km=0.25) -> MeteoTable:
    rows = C.c_int32()
    check(lib().psph_meteo_standard(t0, p0, z_top_km, dz_km, None, None, 0, C.byref(rows)))
    n = rows.value
    z = np.zeros(n)
    d = np.zeros((6, n))
    check(lib().psph_meteo_standard(t0, p0, z_top_km, dz_km, z.ctypes.data_as(C.POINTER(C.c_double)),
                                    d.ctypes.data_as(C.POINTER(C.c_double)), n, C.byref(rows)))
    return MeteoTable(z, d)


def gen_plume(cfg: Config, met: MeteoTable, seed: int, jitter: float = 0.01) -> Particles:
    n = C.c_size_t()
    ms = met.struct()
    check(lib().psph_gen_plume(C.byref(cfg), C.byref(ms), seed, jitter, None, 0, C.byref(n)))
    P = Particles(n.value)
    s = P.soa()
    check(lib().psph_gen_plume(C.byref(cfg), C.byref(ms), seed, jitter, C.byref(s), n.value, C.byref(n)))
    return P.take(n.value)


def gen_shock_tube(cfg: Config, left, right, mid=0.0, n_nosave_l=40, n_nosave_r=20, n_ghost=35,
                   equal_mass=True) -> Particles:
    L = (C.c_double * 3)(*left)
    R = (C.c_double * 3)(*right)
    n = C.c_size_t()
    check(lib().psph_gen_shock_tube(C.byref(cfg), L, R, mid, n_nosave_l, n_nosave_r, n_ghost,
                                    int(equal_mass), None, 0, C.byref(n)))
    P = Particles(n.value)
    s = P.soa()
    check(lib().psph_gen_shock_tube(C.byref(cfg), L, R, mid, n_nosave_l, n_nosave_r, n_ghost,
                                    int(equal_mass), C.byref(s), n.value, C.byref(n)))
    return P


def keys(cfg: Config, x, y, z, add_step=0):
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    n = len(x)
    k = [np.zeros(n, np.uint32) for _ in range(3)]
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    up = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
    check(lib().psph_keys(C.byref(cfg), dp(x), dp(y), dp(z), n, add_step, up(k[0]), up(k[1]), up(k[2])))
    return k


# ---------------------------------------------------------------------------
# the solver: one device context per instance
# ---------------------------------------------------------------------------
class Solver:
    """Device state of one GPU's whole domain and the reference's step stages."""

    def __init__(self, cfg: Config, capacity: int):
        self.cfg = cfg.copy()
        self._ctx = C.c_void_p()
        check(lib().psph_create(C.byref(self.cfg), capacity, C.byref(self._ctx)))
        self.capacity = capacity

    def close(self):
        if self._ctx:
            lib().psph_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state -------------------------------------------------------------
    def load(self, P: Particles, t0: float = 0.0, step0: int = 0):
        s = P.soa()
        check(lib().psph_load(self._ctx, C.byref(s), len(P), t0, step0))

    def config(self) -> Config:
        c = Config()
        check(lib().psph_get_config(self._ctx, C.byref(c)))
        return c

    def read(self) -> Particles:
        n = C.c_size_t()
        check(lib().psph_read(self._ctx, None, 0, C.byref(n)))
        P = Particles(n.value)
        s = P.soa()
        check(lib().psph_read(self._ctx, C.byref(s), n.value, C.byref(n)))
        return P

    def device_view(self) -> dict:
        """Zero-copy torch tensors (HIP device) over the context's particle
        arrays, keyed like Particles, in the context's current order
        (psph_device_view).  They stay valid until the next load, sort
        (neighbour search, step, run) or adjoint run.  Work that torch queues
        on them must be finished (torch.cuda.synchronize) before the next
        library call, which runs on the library's own stream."""
        import torch
        from .dlpack import KEEP, as_tensor
        s, n, st = SoA(), C.c_size_t(), C.c_void_p()
        check(lib().psph_device_view(self._ctx, C.byref(s), C.byref(n), C.byref(st)))
        dev = torch.cuda.current_device()
        keep = KEEP
        out = {}
        for k in ALL_FIELDS:
            ptr = C.cast(getattr(s, k), C.c_void_p).value
            dt = torch.float64 if k in D_FIELDS else (torch.uint32 if k in U_FIELDS else torch.int32)
            out[k] = as_tensor(ptr, n.value, dt, dev, keep)
        return out

    def count(self):
        a, b = C.c_size_t(), C.c_size_t()
        check(lib().psph_count(self._ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def time(self):
        t, s = C.c_double(), C.c_int64()
        check(lib().psph_time(self._ctx, C.byref(t), C.byref(s)))
        return t.value, s.value

    # -- stages under the reference's names (particler.h) ------------------
    def timestep(self) -> float:
        dt = C.c_double()
        check(lib().psph_timestep(self._ctx, C.byref(dt)))
        return dt.value

    def incrtime(self, dt: float) -> float:
        d = C.c_double(dt)
        check(lib().psph_advance_time(self._ctx, C.byref(d)))
        return d.value

    def search_neighs_consth(self):
        check(lib().psph_neighbors(self._ctx))

    def adaptive_sml(self):
        check(lib().psph_adapt_h(self._ctx))

    def mom_engr_update(self, dt: float):
        check(lib().psph_rhs(self._ctx, dt))

    def smooth_density(self):
        check(lib().psph_density(self._ctx))

    def smooth_velocity(self):
        check(lib().psph_vsmooth(self._ctx))

    def update_pos(self, dt: float):
        check(lib().psph_update_pos(self._ctx, dt))

    def add_new_erupt(self, dt: float) -> int:
        n = C.c_int64()
        check(lib().psph_inject(self._ctx, dt, C.byref(n)))
        return n.value

    def apply_bcond(self):
        check(lib().psph_apply_bcond(self._ctx))

    def domain_init(self, met, seed: int, jitter: float = 0.01):
        """Domain adaptation set-up after load (psph_domain_init): the
        atmosphere, seed and jitter of the generated air lattice."""
        check(lib().psph_domain_init(self._ctx, C.byref(met.struct()), seed, jitter))

    def step(self) -> StepOut:
        o = StepOut()
        check(lib().psph_step(self._ctx, C.byref(o)))
        return o

    def run(self, nsteps: int) -> StepOut:
        o = StepOut()
        check(lib().psph_run(self._ctx, nsteps, C.byref(o)))
        return o

    # -- parity / diagnostics ------------------------------------------------
    def neighbors(self):
        """CSR neighbour lists in read order (offsets, idx), sorted per particle."""
        n, _ = self.count()
        off = np.zeros(n + 1, np.int64)
        check(lib().psph_get_neighbors(self._ctx, off.ctypes.data_as(C.POINTER(C.c_int64)), n + 1, None, 0))
        idx = np.zeros(max(int(off[-1]), 1), np.uint32)
        check(lib().psph_get_neighbors(self._ctx, off.ctypes.data_as(C.POINTER(C.c_int64)), n + 1,
                                       idx.ctypes.data_as(C.POINTER(C.c_uint32)), len(idx)))
        return off, idx[: off[-1]]

    def conservation(self) -> np.ndarray:
        out = np.zeros(7)
        check(lib().psph_conservation(self._ctx, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    STAGES = ["timestep", "neighbors", "adapt_h", "rhs", "density", "vsmooth", "update_pos", "inject"]

    def profile(self, enable: bool = True):
        """Enable (and clear) per-stage HIP-event timing on the library's stream."""
        check(lib().psph_profile(self._ctx, int(enable)))

    def profile_read(self) -> dict:
        ms = np.zeros(8)
        n = np.zeros(8, np.int64)
        check(lib().psph_profile_read(self._ctx, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                      n.ctypes.data_as(C.POINTER(C.c_int64))))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.STAGES)}

    def diag_height(self) -> float:
        H = C.c_double()
        check(lib().psph_diag_height(self._ctx, C.byref(H)))
        return H.value

    # -- adjoint -----------------------------------------------------------
    def adjoint_run(self, nsteps: int, ckpt_every: int, grads: bool = True) -> dict:
        """Runs nsteps steps and the reverse sweep (psph_adjoint_run).

        Returns H, its derivatives with respect to the vent velocity, vent
        temperature, gas mass fraction ng0 and solid mass fraction 1 - ng0,
        the timings, and (grads=True) dH/dv and dH/dsv of the starting state
        as (n, 3) arrays in read order."""
        o = AdjointOut()
        n0, _ = self.count() if grads else (0, 0)
        gv = np.zeros((max(n0, 1), 3))
        gs = np.zeros((max(n0, 1), 3))
        ptr = (lambda a: a.ctypes.data_as(C.POINTER(C.c_double))) if grads else (lambda a: None)
        check(lib().psph_adjoint_run(self._ctx, nsteps, ckpt_every, C.byref(o), ptr(gv), ptr(gs), n0))
        r = {k: getattr(o, k) for k, _ in AdjointOut._fields_ if k != "reserved"}
        if grads:
            r["grad_v"] = gv[:o.n_particles]
            r["grad_sv"] = gs[:o.n_particles]
        return r

    def replay(self, nsteps: int):
        """nsteps steps with the frozen schedule of the last adjoint_run."""
        check(lib().psph_replay_run(self._ctx, nsteps))

    def set_vent(self, vent_v: float, vent_T: float, ng0: float):
        check(lib().psph_set_vent(self._ctx, vent_v, vent_T, ng0))
        self.cfg.vent_v, self.cfg.vent_T, self.cfg.ng0 = vent_v, vent_T, ng0
        check(lib().psph_config_vent(C.byref(self.cfg)))
