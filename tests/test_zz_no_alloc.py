"""tmx_integrate_noalloc (the handle-free, allocation-free integral) against tmx_integrate.

Two calls in a row on the one GPU, the second on a smaller mesh, each checked against the
handle entry point on the same values (1e-14 relative: the handle's lumped mass is M 1 summed
element by element, here the closed form (hx hy / 6) * triangles is formed per vertex); the
result lands in a device double on the caller's stream.  A mesh beyond the static buffer's
maximum is refused before any device work.
"""

import ctypes

import numpy as np
import pytest
import torch

from tests.cases import CANTILEVER, gpu_problem, smoothed_density
from topomax_amd import _lib

pytestmark = pytest.mark.gpu


def noalloc_integral(lib, nx, ny, width, height, values):
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.tmx_integrate_noalloc(nx, ny, width, height, ctypes.c_void_p(values.data_ptr()),
                                   ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
    _lib.check(rc)
    torch.cuda.synchronize()
    return float(out.item())


def test_noalloc_integral_matches_handle_twice_shrinking():
    lib = _lib.load()
    results = []
    for nx, ny, seed in ((300, 100, 1), (45, 15, 2)):  # the second call on a smaller mesh
        gp = gpu_problem(CANTILEVER, nx, ny)
        values = torch.as_tensor(smoothed_density(nx, ny, seed=seed, sigma=2.0),
                                 device="cuda:0")
        ref = gp.integrate(values)
        got = noalloc_integral(lib, nx, ny, 3.0, 1.0, values)
        print(nx, ny, got, ref, abs(got - ref) / abs(ref), flush=True)
        assert abs(got - ref) <= 1e-14 * abs(ref)
        # the closed form of the integral of the constant 1 is the area
        ones = torch.ones((nx + 1) * (ny + 1), dtype=torch.float64, device="cuda:0")
        assert abs(noalloc_integral(lib, nx, ny, 3.0, 1.0, ones) - 3.0) < 1e-13
        results.append(got)
    assert results[0] != results[1]


def test_noalloc_integral_refuses_oversized_mesh():
    lib = _lib.load()
    values = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    out = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    rc = lib.tmx_integrate_noalloc(3000, 3000, 1.0, 1.0, ctypes.c_void_p(values.data_ptr()),
                                   ctypes.c_void_p(out.data_ptr()), None)
    assert rc == _lib.TMX_ERR_INVALID
    assert b"exceed the maximum" in lib.tmx_last_error()
    assert float(out.item()) == 0.0  # nothing was launched
