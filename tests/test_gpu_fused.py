"""Fused time step (lbm_fused_*) against the CPU oracle on identical inputs.

FP64: bit-identical rho, u, g, g_up, g_eq and drag/lift after the run
(strict evaluation order on both sides, no FMA).  The chunked run loop is
exercised with chunks that split output intervals, graph replays of several
lengths and a stop inside a chunk (checkpoint rollback)."""
import numpy as np
import pytest

from tests.cases import (run_gpu, run_oracle, small_array, small_cavity, small_channel,
                         small_step)

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "g", "g_up", "g_eq")


def _same(a, b, what):
    for k in FIELDS:
        assert a[k].shape == b[k].shape, (what, k)
        if not np.array_equal(a[k], b[k]):
            d = np.abs(a[k] - b[k])
            idx = np.unravel_index(np.argmax(d), d.shape)
            raise AssertionError(f"{what}: {k} differs, max {d.max():.3e} at {idx}")
    if a["drag"] is not None:
        assert np.array_equal(a["drag"], b["drag"]), (what, "drag")


@pytest.mark.parametrize("make", [small_cavity, small_channel, small_step, small_array])
@pytest.mark.parametrize("chunk", [7, 16])
def test_fused_matches_oracle(make, chunk):
    print(f"case {make.__name__} chunk={chunk}", flush=True)
    g = run_gpu(make, 60, chunk=chunk)[0]
    o = run_oracle(make, 60)[0]
    _same(g, o, make.__name__)


@pytest.mark.parametrize("make", [small_cavity, small_channel, small_step])
def test_per_call_matches_oracle(make):
    """The per-call kernels (the lattice methods, run with fused=False)."""
    print(f"case {make.__name__} per-call", flush=True)
    g = run_gpu(make, 30, fused=False)[0]
    o = run_oracle(make, 30)[0]
    _same(g, o, make.__name__ + " per-call")


def test_stop_inside_chunk_rolls_back():
    """check_stop ends the run mid-chunk: the state is restored and re-advanced."""
    def make():
        app = small_channel()
        app.stop_at = 37
        return app
    g = run_gpu(make, 100, chunk=16)[0]
    o = run_oracle(make, 100)[0]
    assert g["drag"].shape[0] == 38
    _same(g, o, "stop at 37")


def test_single_iteration():
    g = run_gpu(small_cavity, 1)[0]
    o = run_oracle(small_cavity, 1)[0]
    _same(g, o, "one iteration")
