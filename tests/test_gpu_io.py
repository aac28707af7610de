"""Output of a running context: H5Part steps and the 1D CSV table written
from device state (psph_write_h5part, psph_write_csv), and the step result's
output flag (TimeProps::ifoutput, properties.h:105-120, where the reference's
loop writes, main.cc:473-477 and 743-750)."""
import sys

import numpy as np
import pytest

import psph
import psph.io as pio
from helpers import small_plume, sod

pytestmark = pytest.mark.gpu

FIELD = {"x": "x", "y": "y", "z": "z", "Rho": "rho", "Vx": "vx", "Vy": "vy", "Vz": "vz", "engr": "e",
         "mssfrc": "xi", "prss": "p", "phase": "phase", "bctp": "bc"}


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def test_h5part_of_a_run(tmp_path):
    cfg, met, P = small_plume(sml_update_int=3)
    say(f"h5part of a plume run: {len(P)} particles")
    path = str(tmp_path / "pvplot000.h5part")
    W = pio.H5PartWriter(path)
    Wshow = pio.H5PartWriter(str(tmp_path / "pvplot_show000.h5part"), which=pio.SHOW)
    snaps = []
    with psph.Solver(cfg, len(P) + 20000) as G:
        G.load(P)
        for s in range(3):
            W.write(G)
            Wshow.write(G)
            snaps.append(G.read())
            G.step()
    assert pio.h5part_steps(path) == 3 and pio.h5part_steps(Wshow.path) == 3
    for k, Q in enumerate(snaps):
        real = Q["bc"] == 100
        s = pio.h5part_step(path, k)
        assert np.array_equal(s["ID"], np.arange(real.sum()))
        for name, f in FIELD.items():
            assert np.array_equal(s[name], Q[f][real]), (k, name)
        show = np.isin(Q["bc"], (100, 99))
        assert np.array_equal(pio.h5part_read(Wshow.path, k, "x"), Q["x"][show])
    # the steps differ: the writer read the state current at each call
    assert not np.array_equal(pio.h5part_read(path, 0, "Vz"), pio.h5part_read(path, 2, "Vz"))


def test_output_flag_follows_ifoutput():
    cfg, met, P = small_plume(sml_update_int=50)
    with psph.Solver(cfg, len(P) + 20000) as G:
        G.load(P)
        dt0 = G.timestep()
    for t_out, expect in ((1.0e9, "never"), (0.25 * dt0, "every"), (2.5 * dt0, "some")):
        cfg.t_output = t_out
        say(f"output flag, t_output = {t_out:.4g} s")
        flags, times = [], []
        with psph.Solver(cfg, len(P) + 20000) as G:
            G.load(P)
            for s in range(8):
                o = G.step()
                flags.append(int(o.output))
                times.append(o.time)
        flags, times = np.array(flags), np.array(times)
        if expect == "never":
            assert flags.sum() == 0
            continue
        # the m-th flagged step ends on the m-th output time (dt is clipped to
        # reach it); a step that lands an ulp short of it is flagged one later
        m = np.arange(1, flags.sum() + 1)
        assert np.allclose(times[flags == 1], t_out * m, rtol=1e-12, atol=0)
        reached = int(np.floor(times[-1] / t_out * (1 + 1e-12)))
        assert reached - 1 <= flags.sum() <= reached
        if expect == "every":
            assert flags.sum() >= 6
        else:
            assert 0 < flags.sum() < len(flags)


def test_csv_of_sod(tmp_path):
    cfg, P = sod()
    say(f"csv of sod: {len(P)} particles")
    path = str(tmp_path / "pvplot_out.csv")
    with psph.Solver(cfg, len(P) + 16) as G:
        G.load(P)
        for s in range(5):
            G.step()
        pio.write_csv(G, path)
        Q = G.read()
    lines = open(path).read().splitlines()
    assert lines[0] == "RHO,X,U,E,KS,P,PHASE,BCTP,GUEST_FLAG,INVOLVE_FLAG,smlenpmass"
    real = Q["bc"] == 100
    rows = np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])
    assert rows.shape == (real.sum(), 12)
    for col, f in enumerate(("rho", "x", "vx", "e", "xi", "p")):
        assert np.allclose(rows[:, col], Q[f][real], rtol=6e-6, atol=1e-300), f
    assert np.array_equal(rows[:, 7], np.full(real.sum(), 100.0))
    # the 1D H5Part step of the same state
    h5 = str(tmp_path / "pvplot000.h5part")
    with psph.Solver(cfg, len(P) + 16) as G:
        G.load(Q)
        pio.H5PartWriter(h5).write(G)
    s = pio.h5part_step(h5, 0)
    assert np.array_equal(s["x"], Q["x"][real]) and not s["y"].any() and not s["Vz"].any()
