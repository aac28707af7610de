"""C ABI surface (CPU): the library loads, exports every entry point the
header declares, its struct layouts match the bindings, and the host-only
entry points (defaults, deck reader, meteo) behave as the reference does."""
import os
import re

import numpy as np
import pytest

import psph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def header_functions(name="psph.h"):
    src = open(os.path.join(ROOT, "include", name)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(psph_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = psph.lib()
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_io_library_exports_every_declared_symbol():
    import psph.io as pio
    L = pio.io_lib()
    names = header_functions("psph_io.h")
    assert len(names) >= 6
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts():
    import ctypes as C
    L = psph.lib()
    assert L.psph_abi_sizeof(0) == C.sizeof(psph.Config)
    assert L.psph_abi_sizeof(1) == C.sizeof(psph.SoA)
    assert L.psph_abi_sizeof(2) == C.sizeof(psph.StepOut)


def test_defaults_are_the_reference_switches():
    c = psph.default_config(3)
    # options.h 3D, constant.h, parameters.h:889-1043
    assert (c.lans, c.energy_smooth, c.pghost_contrib, c.list_all) == (1, 1, 1, 0)
    assert (c.cutoff, c.cutoff2, c.list_factor) == (3.0, 8.0, 4.25)
    assert (c.alpha, c.beta, c.ata, c.eps_half, c.prandtl) == (0.3, 0.6, 0.01, 0.4, 0.85)
    assert (c.num_loop, c.sml_update_int, c.c_smooth, c.eta, c.thresh) == (2, 50, 2.0, 1.5, 1e-5)
    assert (c.cfl, c.g, c.e_smooth_ratio, c.msfrc_thresh) == (0.8, 9.80665, 4.0, 1e-6)
    assert (c.cvs, c.cvg, c.cva, c.rg, c.ra, c.ng0) == (1100, 1348, 717, 462, 287, 0.026)
    assert (c.vent_r, c.vent_v, c.vent_T, c.vent_p, c.num_erupt_particles) == (120, 75.0, 1273.15, 0.831e5, 1080)
    assert c.p_table_size == 4000000
    d = psph.default_config(1)
    assert (d.lans, d.energy_smooth, d.pghost_contrib, d.list_all) == (0, 0, 2, 1)
    assert (d.cutoff2, d.alpha, d.beta, d.cfl, d.eta, d.num_loop, d.g) == (12.0, 1.0, 2.0, 0.1, 1.0, 5, 0.0)
    with pytest.raises(psph.PsphError):
        psph.default_config(2)


def test_domain_matches_preprocessor():
    c = psph.default_config(3)
    c.h_air = 15.0
    c.dom_lo[:] = [-400, -400, 1500]
    c.dom_hi[:] = [400, 400, 2100]
    psph.finalize_config(c)
    # preprocess.cc:88-118: extend by 1.5 buckets of 8 h, round up to whole buckets
    assert c.bucket_edge == 120.0
    assert list(c.ext_lo) == [-580.0, -580.0, 1320.0]
    assert list(c.ext_hi) == [-580.0 + 10 * 120, -580.0 + 10 * 120, 1320.0 + 8 * 120]


def test_read_deck_weak_plume():
    c = psph.default_config(3)
    path, rows, cols = psph.read_deck(os.path.join(GOLD, "simulation_ex_weak_back_sml15.data"), c)
    assert c.t_end == 100000.0 and c.max_steps == 100000000 and c.t_output == 1.0
    assert (c.erupt_start, c.erupt_end, c.h_air) == (0.0, 7200.0, 15.0)
    assert list(c.dom_lo) == [-400.0, -400.0, 1500.0] and list(c.dom_hi) == [400.0, 400.0, 2100.0]
    assert (rows, cols) == (47, 6) and path.endswith("Meteo_Profiles-W-V.dat")


def test_read_deck_goes_through_float():
    """dataread.cc parses every value with stof: 0.17 becomes float(0.17)."""
    c = psph.default_config(1)
    psph.read_deck(os.path.join(GOLD, "simulation_1D_shocktube_Sod.data"), c)
    assert c.t_end == float(np.float32(0.17)) and c.t_end != 0.17
    assert c.h_air == float(np.float32(0.003))
    assert c.dom_lo[0] == float(np.float32(-0.4))


def test_read_deck_errors():
    c = psph.default_config(3)
    with pytest.raises(psph.PsphError) as e:
        psph.read_deck("/nonexistent/simulation.data", c)
    assert e.value.code == -6


def test_meteo_read_and_standard():
    m = psph.meteo_read(os.path.join(GOLD, "Meteo_Profiles-W-V.dat"), 47, 6)
    assert m.z[0] == 1.4 and m.data.shape == (5, 47)
    assert m.data[0, 0] == 1.120 and m.data[1, 0] == 863.203
    s = psph.meteo_standard()
    assert s.z[0] == 0.0 and s.data[1, 0] == pytest.approx(1013.25)
    assert s.data[2, 44] == pytest.approx(288.0 - 6.5 * 11.0)          # 11 km: tropopause
    assert s.data[0, 0] == pytest.approx(101325.0 / (287.0 * 288.0))


def test_create_without_gpu_fails_loudly():
    """No HIP device here: creating a context must raise, never fall back."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    c = psph.default_config(3)
    c.h_air = 100.0
    psph.finalize_config(c)
    with pytest.raises(psph.PsphError):
        psph.Solver(c, 1000)
