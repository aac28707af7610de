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


# ------------------------------------------------------------ grid files
# GH5_bucketstruct (hdf5calls.cc:175-203): scalar members at the offsets of
# BucketStruct (buckstr.h:14-26, KEYLENGTH 2, NEIGH_SIZE 27)
BUCKET_MEMBERS = [("Key", "H5T_STD_U32LE", 0), ("neighs", "H5T_STD_U32LE", 8),
                  ("buckettype", "H5T_STD_I32LE", 224), ("myproc", "H5T_STD_I32LE", 228),
                  ("neigh_proc", "H5T_STD_I32LE", 232), ("xcoord", "H5T_IEEE_F64LE", 368),
                  ("ycoord", "H5T_IEEE_F64LE", 384), ("zcoord", "H5T_IEEE_F64LE", 400),
                  ("elev", "H5T_IEEE_F64LE", 416)]
RECORD = 448


def grid_layout(path):
    out = subprocess.run([H5DUMP, "-p", "-H", path], capture_output=True, text=True, check=True).stdout
    ds = {}
    for name, body in re.findall(r'DATASET "([^"]+)" \{(.*?)\n   \}', out, flags=re.S):
        n = int(re.search(r"SIMPLE \{ \( (\d+) \)", body).group(1))
        size = int(re.search(r"SIZE (\d+)", body).group(1))
        off = int(re.search(r"OFFSET (\d+)", body).group(1))
        members = re.findall(r'(H5T_\w+) "(\w+)";', body)
        ds[name] = (n, size, off, [(m, t) for t, m in members])
    return ds


def grid_case(tmp_path):
    from helpers import small_plume_config
    cfg = small_plume_config()
    path = str(tmp_path / "funky0000.h5")
    pio.grid_write(path, cfg)
    return cfg, path


@needs_h5dump
def test_grid_file_layout(tmp_path):
    """The preprocessor's file: three datasets, the compound record of
    GH5_bucketstruct (member names and types in its order, 448-byte records),
    and each member's value at BucketStruct's offset in the raw bytes."""
    import oracle
    cfg, path = grid_case(tmp_path)
    L = grid_layout(path)
    assert set(L) == {"Buckets", "hashtable_constants", "partition_table"}
    n, size, off, members = L["Buckets"]
    assert [(m, t) for m, t, _ in BUCKET_MEMBERS] == members
    assert size == n * RECORD
    nb = [int(round((cfg.ext_hi[k] - cfg.ext_lo[k]) / cfg.bucket_edge)) for k in range(3)]
    assert n == nb[0] * nb[1] * nb[2]
    ht = raw(path, L["hashtable_constants"][2], 6, False)
    assert list(ht) == [cfg.ext_lo[0], cfg.ext_hi[0], cfg.ext_lo[1], cfg.ext_hi[1], cfg.ext_lo[2], cfg.ext_hi[2]]
    pn, psize, poff, _ = L["partition_table"]
    assert pn == 4 * n and psize == 16 * n
    ptab = np.frombuffer(open(path, "rb").read()[poff:poff + psize], "<u4").reshape(n, 4)
    assert np.array_equal(ptab[:, :2], ptab[:, 2:])
    blob = open(path, "rb").read()[off:off + size]

    def member(r, name):
        o = r * RECORD + dict((m, o) for m, _, o in BUCKET_MEMBERS)[name]
        if name in ("Key", "neighs"):
            return int(np.frombuffer(blob[o:o + 4], "<u4")[0])
        if name in ("buckettype", "myproc", "neigh_proc"):
            return int(np.frombuffer(blob[o:o + 4], "<i4")[0])
        return float(np.frombuffer(blob[o:o + 8], "<f8")[0])

    cell = lambda v, k: int(round((v - cfg.ext_lo[k]) / cfg.bucket_edge))
    keys_by_cell = {}
    for r in range(n):
        b = (cell(member(r, "xcoord"), 0), cell(member(r, "ycoord"), 1), cell(member(r, "zcoord"), 2))
        # type: the oracle's restatement of determine_bucket_type (preprocess.cc:380-440)
        assert member(r, "buckettype") == oracle.bucket_type(cfg, *b)[0], b
        assert member(r, "myproc") == 0 and member(r, "elev") == cfg.dom_lo[2]
        # the key of the bucket centre (HSFC3d of the centre normalised to the hash-table domain)
        c = [cfg.ext_lo[k] + (b[k] + 0.5) * cfg.bucket_edge for k in range(3)]
        k0, _, _ = psph.keys(cfg, np.array([c[0]]), np.array([c[1]]), np.array([c[2]]))
        assert member(r, "Key") == int(k0[0]), b
        assert member(r, "Key") == int(ptab[r, 0])
        keys_by_cell[b] = member(r, "Key")
        # first neighbour (i-1, j-1, k-1): outside the grid -> key 0, process -1
        inside = all(v > 0 for v in b)
        assert member(r, "neigh_proc") == (0 if inside else -1)
        if not inside:
            assert member(r, "neighs") == 0
    for r in range(n):
        b = (cell(member(r, "xcoord"), 0), cell(member(r, "ycoord"), 1), cell(member(r, "zcoord"), 2))
        if all(v > 0 for v in b):
            assert member(r, "neighs") == keys_by_cell[(b[0] - 1, b[1] - 1, b[2] - 1)]
    # partition order (PartiHead::operator<): by key, a lowest-layer bucket by
    # the key of the bucket above it and right after that bucket
    order = []
    for r in range(n):
        b = (cell(member(r, "xcoord"), 0), cell(member(r, "ycoord"), 1), cell(member(r, "zcoord"), 2))
        if b[2] == 0:
            top = keys_by_cell[(b[0], b[1], 1)]
            order.append((top, 1))
        else:
            order.append((member(r, "Key"), 0))
    assert [o[0] for o in order] == sorted(o[0] for o in order)


def test_grid_file_read_and_check(tmp_path):
    cfg, path = grid_case(tmp_path)
    g = pio.grid_read(path)
    assert list(g.ext_lo) == list(cfg.ext_lo) and list(g.ext_hi) == list(cfg.ext_hi)
    assert g.bucket_edge == cfg.bucket_edge and g.nproc == 1 and g.flat == 1 and g.ground == cfg.dom_lo[2]
    nb = [int(round((cfg.ext_hi[k] - cfg.ext_lo[k]) / cfg.bucket_edge)) for k in range(3)]
    assert list(g.nb) == nb and g.n_buckets == nb[0] * nb[1] * nb[2] and g.n_partition == 4 * g.n_buckets
    assert sum(g.n_type) == g.n_buckets and g.n_type[0] == 0
    pio.grid_check(path, cfg)
    # another deck's domain is refused
    from helpers import small_plume_config
    other = small_plume_config(half=1000.0)
    with pytest.raises(psph.PsphError) as e:
        pio.grid_check(path, other)
    assert e.value.code == psph.PSPH_EINVAL


def test_grid_file_terrain_and_processes_refused(tmp_path):
    """A varying ground elevation (the preprocessor's GIS mode) or buckets of
    another process: the file is read, the check refuses it (PSPH_ENOTSUP)."""
    cfg, path = grid_case(tmp_path)
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not in this image")
    n, size, off, _ = grid_layout(path)["Buckets"]
    data = bytearray(open(path, "rb").read())
    for field, o, value, fmt in (("elev", 416, 12.5, "<f8"), ("myproc", 228, 1, "<i4")):
        bad = bytearray(data)
        at = off + 3 * RECORD + o
        v = np.array([value], fmt).tobytes()
        bad[at:at + len(v)] = v
        p = str(tmp_path / f"bad_{field}.h5")
        open(p, "wb").write(bytes(bad))
        g = pio.grid_read(p)
        assert (g.flat == 0) if field == "elev" else (g.nproc == 2)
        with pytest.raises(psph.PsphError) as e:
            pio.grid_check(p, cfg)
        assert e.value.code == psph.PSPH_ENOTSUP, field
