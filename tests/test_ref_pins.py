"""The oracle and the product against the reference's own code, for the
parts of the path whose source files compile alone (no MPI, HDF5 or generated
header); `make -C oracle ref` builds them from where they lie into
oracle/_ref:

* particle keys, THSFC3d (hilbert.c:329-359): the oracle restatement and the
  product's key function return the reference's words bit for bit, on random
  positions, on dyadic points where the curve changes octant at every level
  and their one-ulp neighbours, on the faces of the cube and outside it
  (clamped);
* the atmosphere interpolation, Meteo::interpolate (meteo.cc:64-143): the
  oracle's value equals the reference's bit for bit at the table's heights,
  between them, below and above the table, for every property.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import psph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libref_hilbert.so")
REF_METEO = os.path.join(ROOT, "oracle", "_ref", "libref_meteo.so")
GOLD = os.path.join(ROOT, "tests", "golden")

pytestmark = pytest.mark.skipif(not (os.path.exists(REF) and os.path.exists(REF_METEO)),
                                reason="reference sources not present: oracle/_ref not built")


def ref_lib():
    L = C.CDLL(REF)
    L.THSFC3d.restype = None
    L.THSFC3d.argtypes = [C.POINTER(C.c_double), C.c_uint, C.POINTER(C.c_uint), C.POINTER(C.c_uint)]
    return L


def ref_key(L, normc, add_step):
    c = (C.c_double * 3)(*normc)
    nkey = C.c_uint(3)                      # TKEYLENGTH (constant.h:16-17), as add_air.cc:607 passes it
    key = (C.c_uint * 3)(0, 0, 0)
    L.THSFC3d(c, add_step, C.byref(nkey), key)
    return tuple(int(k) for k in key)


def positions():
    rng = np.random.default_rng(11)
    pts = [rng.uniform(0, 1, 3) for _ in range(3000)]
    # dyadic points and their neighbours one ulp away: octant changes at every level
    for lev in (1, 2, 3, 5, 8, 13, 21, 31):
        for _ in range(40):
            p = rng.integers(0, 2 ** lev, 3) / 2.0 ** lev
            pts.append(p)
            pts.append(np.nextafter(p, 2.0))
            pts.append(np.nextafter(p, -1.0))
    pts += [np.zeros(3), np.ones(3), np.array([1.0, 0.0, 1.0]), np.array([-0.5, 0.3, 1.7]),
            np.array([2.0, -3.0, 0.999999999]), np.array([1 - 2 ** -33, 2 ** -33, 0.5])]
    return np.array(pts)


def test_oracle_keys_are_the_reference_keys():
    L = ref_lib()
    bad = []
    for n, p in enumerate(positions()):
        step = n % 7
        want = ref_key(L, p, step)
        got = tuple(int(k) for k in oracle.thsfc3d(p, step))
        if got != want:
            bad.append((p.tolist(), step, got, want))
    assert not bad, bad[:5]


def test_product_keys_are_the_reference_keys():
    """psph_keys normalises by the extended domain; with it set to the unit
    cube the normalised coordinate is the position itself."""
    L = ref_lib()
    cfg = psph.default_config(3)
    for k in range(3):
        cfg.ext_lo[k], cfg.ext_hi[k] = 0.0, 1.0
    P = positions()
    k0, k1, k2 = psph.keys(cfg, P[:, 0], P[:, 1], P[:, 2], 9)
    want = np.array([ref_key(L, p, 9) for p in P], dtype=np.uint64)
    assert np.array_equal(np.stack([k0, k1, k2], 1).astype(np.uint64), want)


# ---------------------------------------------------------------- Meteo::interpolate
def ref_meteo(table_z, table_data):
    """Meteo object of the reference over a table in its file layout (rows of
    z followed by the properties, meteo.cc:22-57)."""
    L = C.CDLL(REF_METEO)
    L.ref_meteo_new.restype = C.c_void_p
    L.ref_meteo_new.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int]
    L.ref_meteo_interp.restype = C.c_double
    L.ref_meteo_interp.argtypes = [C.c_void_p, C.c_double, C.c_int]
    nr, npr = table_data.shape[1], table_data.shape[0]
    mat = np.ascontiguousarray(np.column_stack([table_z, table_data.T]), np.float64)
    h = L.ref_meteo_new(mat.ctypes.data_as(C.POINTER(C.c_double)), nr, npr)
    return lambda z0, prop: L.ref_meteo_interp(h, z0, prop)


def heights(z):
    rng = np.random.default_rng(5)
    mids = 0.5 * (z[1:] + z[:-1])
    return np.concatenate([z, mids, rng.uniform(z[0], z[-1], 300), np.nextafter(z, np.inf),
                           np.nextafter(z, -np.inf), [z[0] - 1.0, z[-1] + 3.0, -50.0, 1e4]])


@pytest.mark.parametrize("table", ["file", "standard"])
def test_oracle_meteo_interpolation_is_the_reference(table):
    if table == "file":
        met = psph.meteo_read(os.path.join(GOLD, "Meteo_Profiles-W-V.dat"), 47, 6)
    else:
        met = psph.meteo_standard()
    ref = ref_meteo(met.z, met.data)
    m = oracle.meteo(met)
    bad = []
    for z0 in heights(np.asarray(met.z)):
        for prop in range(met.data.shape[0]):
            got = oracle.lib().orc_meteo_interp(C.byref(m), float(z0), prop)
            want = ref(float(z0), prop)
            if got != want and not (np.isnan(got) and np.isnan(want)):
                bad.append((float(z0), prop, got, want))
    oracle.lib().orc_meteo_free(C.byref(m))
    assert not bad, bad[:5]
