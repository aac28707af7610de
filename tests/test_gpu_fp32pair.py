"""FP32-pair forms (pair arithmetic in FP32 on exact FP64 pair predicates,
FP64 accumulation / integration / EOS; PSPH_FP32_EXP also takes the weights'
exponentials from the FP32 instruction, the adaptive h's included) against
the FP64 CPU oracle.

Bars: the same pairs as the FP64 form (identical pair counts and particle
sets), and every field within 1e-5 relative L1 of the FP64 oracle."""
import sys

import numpy as np
import pytest

import oracle
import psph
from helpers import FLOAT_FIELDS_3D, FLOAT_FIELDS_1D, EXACT_FIELDS, rel_l1, small_plume, sod

pytestmark = pytest.mark.gpu
REL_L1_FP32 = 1e-5


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("form", ["PSPH_FP32_PAIR", "PSPH_FP32_EXP"])
def test_fp32_pair_plume_steps(form):
    """8 steps with the adaptive h at steps 4 and 8."""
    cfg, met, P = small_plume(sml_update_int=4)
    say(f"{form} plume: {len(P)} particles, 8 steps")
    S = oracle.State(cfg, P)
    c32 = cfg.copy()
    c32.precision = getattr(psph, form)
    with psph.Solver(c32, len(P) + 20000) as G, psph.Solver(cfg, len(P) + 20000) as G64:
        G.load(P)
        G64.load(P)
        for s in range(8):
            S.step()
            o32 = G.step()
            o64 = G64.step()
            assert o32.n_pairs == o64.n_pairs
            assert o32.n_injected == o64.n_injected
        A, B = G.read(), S.particles()
        assert len(A) == len(B)
        for k in EXACT_FIELDS:
            assert np.array_equal(A[k], B[k]), k
        errs = {k: rel_l1(A[k], B[k]) for k in FLOAT_FIELDS_3D}
        say("rel L1: " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()))
        bad = {k: v for k, v in errs.items() if not v <= REL_L1_FP32}
        assert not bad, bad


@pytest.mark.parametrize("form", ["PSPH_FP32_PAIR", "PSPH_FP32_EXP"])
def test_fp32_pair_sod(form):
    cfg, P = sod()
    S = oracle.State(cfg, P)
    S.neighbors()
    S.adaptive_sml()
    c32 = cfg.copy()
    c32.precision = getattr(psph, form)
    with psph.Solver(c32, len(P) + 16) as G:
        G.load(P)
        G.search_neighs_consth()
        G.adaptive_sml()
        for s in range(40):
            S.step()
            G.step()
        A, B = G.read(), S.particles()
        errs = {k: rel_l1(A[k], B[k]) for k in FLOAT_FIELDS_1D}
        say("sod rel L1: " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()))
        bad = {k: v for k, v in errs.items() if not v <= REL_L1_FP32}
        assert not bad, bad


def test_fp32_pure_plume_steps():
    """The FP32-pure form (RHS and density pair arithmetic all FP32): the same
    pairs as the FP64 form, within the looser bar its header states
    (velocities ~1e-4 relative L1 over long runs; 1e-4 here)."""
    cfg, met, P = small_plume(sml_update_int=4)
    say(f"fp32-pure plume: {len(P)} particles, 8 steps")
    cp = cfg.copy()
    cp.precision = psph.PSPH_FP32_PURE
    with psph.Solver(cp, len(P) + 20000) as G, psph.Solver(cfg, len(P) + 20000) as G64:
        G.load(P)
        G64.load(P)
        for s in range(8):
            op = G.step()
            o64 = G64.step()
            assert op.n_pairs == o64.n_pairs
            assert op.n_injected == o64.n_injected
        A, B = G.read(), G64.read()
        assert len(A) == len(B)
        for k in EXACT_FIELDS:
            assert np.array_equal(A[k], B[k]), k
        errs = {k: rel_l1(A[k], B[k]) for k in FLOAT_FIELDS_3D}
        say("rel L1 vs FP64: " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()))
        bad = {k: v for k, v in errs.items() if not v <= 1e-4}
        assert not bad, bad
