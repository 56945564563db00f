"""The allocation-free gradient clr_grad_noalloc against clr_grad.

clr_grad_noalloc runs the same FFT launches as clr_grad with its spectra in
a device array that every call on a GPU shares, so two calls in a row (the
second on a smaller grid, which leaves the first call's spectra behind in
that array) must each equal clr_grad's result bit for bit."""
import ctypes

import numpy as np
import pytest

import claire_amd
from claire_amd import Claire, synth


def _grad_ctx(f):
    c = Claire(f.shape, nt=1, regnorm="h1sn", beta=(1e-2, 0.0, 0.0))
    try:
        return c.grad(f)
    finally:
        c.close()


@pytest.mark.gpu
def test_grad_noalloc_equals_grad_on_two_calls():
    for shape, seed in (((64, 64, 64), 3), ((32, 16, 48), 4)):
        print(f"grad_noalloc shape={shape} seed={seed}", flush=True)
        f = synth.gaussian_blobs(shape, count=6, seed=seed).astype(np.float32)
        g = claire_amd.grad_noalloc(f)
        ref = _grad_ctx(f)
        assert g.shape == ref.shape
        assert np.array_equal(g, ref), float(np.abs(g - ref).max())
        assert np.linalg.norm(g) > 0


@pytest.mark.gpu
def test_grad_noalloc_repeats_bit_for_bit_on_a_stream():
    shape = (48, 96, 32)
    f = synth.gaussian_blobs(shape, count=5, seed=5).astype(np.float32)
    a = claire_amd.grad_noalloc(f)
    b = claire_amd.grad_noalloc(f)
    assert np.array_equal(a, b)
    assert np.array_equal(a, _grad_ctx(f))


def test_grad_noalloc_rejects_grids_beyond_its_scratch():
    lib = claire_amd.lib()
    nx = (ctypes.c_int * 3)(256, 256, 256)
    rc = lib.clr_grad_noalloc(nx, ctypes.c_void_p(16), ctypes.c_void_p(16), None)
    assert rc == 2
    assert b"scratch" in lib.clr_last_error()
    nx = (ctypes.c_int * 3)(64, 64, 64)
    assert lib.clr_grad_noalloc(nx, None, ctypes.c_void_p(16), None) == 1
