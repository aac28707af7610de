"""Output formats on the CPU: H5Part steps (h5part_writer.cc:40-444,
hdf5calls.cc:288-346) and the 1D CSV table (h5part_writer.cc:448-489).

The file layout is pinned with the HDF5 command-line dump (h5dump, a reader
independent of this library): group names, dataset names in the reference's
order, HDF5 types, rank-1 dataspaces, contiguous storage; the values are
pinned by reading the raw bytes at the storage offsets h5dump reports.  The
CSV format is pinned against the C++ stream's default double format, written
out by hand for values that exercise its branches.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import psph
import psph.io as pio

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
needs_h5dump = pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not in this image")

# reference dataset order (h5part_writer.cc:179-252) without the DEBUG extras
ORDER = ["ID", "x", "y", "z", "Rho", "Vx", "Vy", "Vz", "engr", "mssfrc", "prss", "phase", "bctp", "guest"]
FIELD = {"x": "x", "y": "y", "z": "z", "Rho": "rho", "Vx": "vx", "Vy": "vy", "Vz": "vz", "engr": "e",
         "mssfrc": "xi", "prss": "p", "phase": "phase", "bctp": "bc", "mymass": "m", "sml": "h",
         "Involved": "involved"}


def particles(n=40, seed=3):
    rng = np.random.default_rng(seed)
    P = psph.Particles(n)
    for k in psph.D_FIELDS:
        P[k][:] = rng.standard_normal(n) * 10.0 ** rng.integers(-6, 7, n)
    P["bc"][:] = rng.choice([100, 99, 0, 1, 2], n)
    P["phase"][:] = rng.choice([1, 2], n)
    P["involved"][:] = rng.integers(0, 3, n)
    return P


def dump_layout(path):
    """{group: [(dataset, type, dims, offset, size)]} in h5dump's listing order"""
    out = subprocess.run([H5DUMP, "-p", "-H", path], capture_output=True, text=True, check=True).stdout
    groups = {}
    for g, body in re.findall(r'GROUP "(Step#\d+)" \{(.*?)\n   \}', out, flags=re.S):
        ds = re.findall(r'DATASET "([^"]+)" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+SIMPLE \{ \( (\d+) \) / \( (\d+) \) \}'
                        r'\s*STORAGE_LAYOUT \{\s*CONTIGUOUS\s*SIZE (\d+)\s*OFFSET (\d+)', body)
        groups[g] = [(d, t, int(n), int(mx), int(sz), int(off)) for d, t, n, mx, sz, off in ds]
    return groups


def raw(path, off, n, is_int):
    with open(path, "rb") as f:
        f.seek(off)
        return np.frombuffer(f.read(n * (4 if is_int else 8)), "<i4" if is_int else "<f8")


@needs_h5dump
def test_h5part_layout_and_bytes(tmp_path):
    P = particles()
    path = str(tmp_path / "pvplot000.h5part")
    pio.h5part_write(path, P, 0, True)
    pio.h5part_write(path, P, 1, False, which=pio.SHOW)
    lay = dump_layout(path)
    assert sorted(lay) == ["Step#0", "Step#1"]
    for g, sel in (("Step#0", P["bc"] == 100), ("Step#1", np.isin(P["bc"], (100, 99)))):
        m = int(sel.sum())
        assert 0 < m < len(P)
        names = {d[0]: d for d in lay[g]}
        assert sorted(names) == sorted(ORDER)
        for name in ORDER:
            _, typ, n, mx, size, off = names[name]
            is_int = name in ("ID", "phase", "bctp", "guest")
            assert typ == ("H5T_STD_I32LE" if is_int else "H5T_IEEE_F64LE"), name
            assert n == mx == m and size == m * (4 if is_int else 8), name
            got = raw(path, off, m, is_int)
            if name == "ID":
                want = np.arange(m)
            elif name == "guest":
                want = np.ones(m)
            else:
                want = P[FIELD[name]][sel]
            assert np.array_equal(got, want), name


def test_h5part_reader_roundtrip_and_order(tmp_path):
    P = particles(200, seed=7)
    path = str(tmp_path / "a.h5part")
    pio.h5part_write(path, P, 0, True, extras=pio.INVOLVED | pio.PID | pio.MASS | pio.SML, rank=5)
    pio.h5part_write(path, P, 1, False, which=pio.ALL)
    assert pio.h5part_steps(path) == 2
    sel = P["bc"] == 100
    s0 = pio.h5part_step(path, 0, ORDER + ["Involved", "myprocess", "mymass", "sml"])
    for k, f in FIELD.items():
        assert np.array_equal(s0[k], P[f][sel]), k
    assert np.array_equal(s0["myprocess"], np.full(sel.sum(), 5))
    s1 = pio.h5part_step(path, 1)
    assert np.array_equal(s1["bctp"], P["bc"]) and np.array_equal(s1["x"], P["x"])
    assert np.array_equal(s1["ID"], np.arange(len(P)))


@needs_h5dump
def test_h5part_debug_extras_follow_id(tmp_path):
    """DEBUG build order: ID, Involved, myprocess, mymass, sml, then x ..."""
    P = particles()
    path = str(tmp_path / "d.h5part")
    pio.h5part_write(path, P, 0, True, extras=pio.INVOLVED | pio.PID | pio.MASS | pio.SML)
    ds = dump_layout(path)["Step#0"]
    by_off = [d[0] for d in sorted(ds, key=lambda d: d[5])]      # data blocks are allocated in write order
    assert by_off == ["ID", "Involved", "myprocess", "mymass", "sml"] + ORDER[1:]


def test_h5part_dim1_and_empty(tmp_path):
    P = particles()
    path = str(tmp_path / "one.h5part")
    pio.h5part_write(path, P, 0, True, dim=1)
    s = pio.h5part_step(path, 0)
    m = int((P["bc"] == 100).sum())
    for k in ("y", "z", "Vy", "Vz"):
        assert np.array_equal(s[k], np.zeros(m)), k
    assert np.array_equal(s["x"], P["x"][P["bc"] == 100])
    Q = particles()
    Q["bc"][:] = 2                                      # no real particle: empty datasets
    pio.h5part_write(path, Q, 1, False)
    assert pio.h5part_read(path, 1, "x").size == 0 and pio.h5part_steps(path) == 2


def test_h5part_errors(tmp_path):
    P = particles()
    path = str(tmp_path / "e.h5part")
    with pytest.raises(psph.PsphError) as e:
        pio.h5part_write(path, P, 0, False)             # append to a file that does not exist
    assert e.value.code == -6
    pio.h5part_write(path, P, 0, True)
    with pytest.raises(psph.PsphError) as e:
        pio.h5part_write(path, P, 0, False)             # the group exists (the reference exits)
    assert e.value.code == -6
    with pytest.raises(psph.PsphError) as e:
        pio.h5part_write(path, P, 1, False, which=7)
    assert e.value.code == -1
    with pytest.raises(psph.PsphError) as e:
        pio.h5part_read(path, 0, "nope")
    assert e.value.code == -6
    pio.h5part_write(path, P, 0, True)                  # create truncates
    assert pio.h5part_steps(path) == 1


def test_csv_format(tmp_path):
    P = psph.Particles(3)
    P["bc"][:] = [100, 2, 100]
    P["rho"][:] = [1.0, 5.0, 1234567.0]
    P["x"][:] = [0.1, 0.0, -2.5e-5]
    P["vx"][:] = [100.0, 0.0, 0.000123456789]
    P["e"][:] = [2.5e5, 0.0, 1e20]
    P["xi"][:] = [0.0, 0.0, 1.0 / 3.0]
    P["p"][:] = [101325.0, 0.0, -0.0]
    P["phase"][:] = [1, 2, 2]
    P["involved"][:] = [0, 1, 2]
    P["h"][:] = [0.0125, 1.0, 7.0]
    P["m"][:] = [1e-3, 1.0, 2.0]
    path = str(tmp_path / "pvplot_out.csv")
    pio.csv_write(path, P)
    pio.csv_write(path, P)                              # appends a second table
    want = ("RHO,X,U,E,KS,P,PHASE,BCTP,GUEST_FLAG,INVOLVE_FLAG,smlenpmass\n"
            "1,0.1,100,250000,0,101325,1,100,1,0,0.0125,0.001\n"
            "1.23457e+06,-2.5e-05,0.000123457,1e+20,0.333333,-0,2,100,1,2,7,2\n")
    assert open(path).read() == want * 2
    # the same rows through std::ostream's operator<< (default stream state)
    gxx = shutil.which("g++")
    if gxx:
        src = tmp_path / "row.cc"
        src.write_text(
            "#include <iostream>\nint main() {\n"
            "  double r[2][9] = {{1.0, 0.1, 100.0, 2.5e5, 0.0, 101325.0, 1.0, 0.0125, 1e-3},\n"
            "                    {1234567.0, -2.5e-5, 0.000123456789, 1e20, 1.0 / 3.0, -0.0, 1.0, 7.0, 2.0}};\n"
            "  int ph[2] = {1, 2}, inv[2] = {0, 2};\n"
            "  std::cout << \"RHO\" << \",\" << \"X\" << \",\" << \"U\" << \",\" << \"E\" << \",\" << \"KS\" << \",\" << \"P\""
            " << \",\" << \"PHASE\" << \",\" << \"BCTP\" << \",\" << \"GUEST_FLAG\" << \",\" << \"INVOLVE_FLAG\" << \",\""
            " << \"smlen\" << \"pmass\" << std::endl;\n"
            "  for (int i = 0; i < 2; i++)\n"
            "    std::cout << r[i][0] << \",\" << r[i][1] << \",\" << r[i][2] << \",\" << r[i][3] << \",\" << r[i][4]"
            " << \",\" << r[i][5] << \",\" << ph[i] << \",\" << 100 << \",\" << r[i][6] << \",\" << inv[i] << \",\""
            " << r[i][7] << \",\" << r[i][8] << std::endl;\n}\n")
        exe = tmp_path / "row"
        subprocess.run([gxx, str(src), "-o", str(exe)], check=True)
        assert subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout == want
