"""The particle keys against the reference's own THSFC3d (hilbert.c:329-359),
compiled from its source file into oracle/_ref by `make -C oracle ref` (the
file needs no MPI, HDF5 or generated header).  Both the oracle restatement and
the product's key function must return the reference's words bit for bit,
on random positions, on dyadic points where the curve changes octant at
every level, on the faces of the cube and outside it (clamped).
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import psph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libref_hilbert.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference sources not present: oracle/_ref not built")


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
