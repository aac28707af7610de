"""psph_device_view: zero-copy device arrays of a context, and the torch
wrapper's on-device initial-velocity path.

Bars: bit-exact (the view is the state itself, and adding dv0 on the device
is the same FP64 addition as adding it on the host).
"""
import sys

import numpy as np
import pytest
import torch

import psph
from helpers import small_plume
from psph.torch_adjoint import plume_height

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def by_key(P):
    """Particles sorted by their (key0, key1, key2) triple."""
    o = np.lexsort((P["key2"], P["key1"], P["key0"]))
    return {k: np.asarray(v)[o] for k, v in (P.a.items() if hasattr(P, "a") else P.items())}


def test_view_after_load_is_the_loaded_state():
    cfg, met, P = small_plume()
    say(f"device view after load: {len(P)} particles")
    with psph.Solver(cfg, len(P) + 1000) as G:
        G.load(P)
        V = G.device_view()
        for k in psph.ALL_FIELDS:
            got = V[k].cpu().numpy()
            assert got.shape == (len(P),), k
            assert np.array_equal(got, P[k]), k
        # a write through the view is the state the next read returns
        V["vx"].add_(1.5)
        V["e"].mul_(2.0)
        torch.cuda.synchronize()
        del V
        Q = by_key(G.read())
        R = by_key(P)
        assert np.array_equal(Q["vx"], R["vx"] + 1.5)
        assert np.array_equal(Q["e"], R["e"] * 2.0)
        assert np.array_equal(Q["vy"], R["vy"])


def test_view_after_steps_matches_read():
    cfg, met, P = small_plume()
    say(f"device view after 3 steps: {len(P)} particles")
    with psph.Solver(cfg, len(P) + 20000) as G:
        G.load(P)
        G.run(3)
        n_total, _ = G.count()
        V = G.device_view()
        D = {k: v.cpu().numpy() for k, v in V.items()}
        del V
        live = D["bc"] >= 0
        D = by_key({k: v[live] for k, v in D.items()})
        R = by_key(G.read())
        assert len(R["x"]) == n_total
        for k in psph.ALL_FIELDS:
            assert np.array_equal(D[k], R[k]), k


def test_torch_function_device_dv0_equals_host_dv0():
    cfg, met, P = small_plume(sml_update_int=10)
    th0 = [cfg.vent_v, cfg.vent_T, 1.0 - cfg.ng0]
    rng = np.random.default_rng(7)
    dv = 1e-3 * rng.standard_normal((len(P), 3))
    say(f"torch function, dv0 on host and on device: {len(P)} particles, 8 steps")
    out = []
    for dev in ("cpu", "cuda"):
        th = [torch.tensor(float(x), dtype=torch.float64, requires_grad=True) for x in th0]
        dv0 = torch.tensor(dv, dtype=torch.float64, device=dev, requires_grad=True)
        H = plume_height(*th, dv0, cfg=cfg, met=met, seed=20261202, jitter=0.01, nsteps=8, ckpt_every=3)
        H.backward()
        assert dv0.grad.device.type == dev
        out.append((H.item(), [t.grad.item() for t in th], dv0.grad.cpu().numpy()))
    (h0, g0, d0), (h1, g1, d1) = out
    assert h0 == h1, (h0, h1)
    assert g0 == g1, (g0, g1)
    assert np.array_equal(d0, d1)
