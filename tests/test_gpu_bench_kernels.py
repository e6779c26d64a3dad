"""tmx_bench_kernel, the timing entry point behind bench.py's per-kernel figures and roofline:
every kernel id runs, reports a positive duration and the documented algorithmic bytes, and
leaves the next objective + gradient unchanged (kernel 5 runs on the PCG's own vectors)."""

import ctypes

import pytest
import torch

from tests.cases import CANTILEVER, gpu_problem, smoothed_density
from topomax_amd import _lib

pytestmark = pytest.mark.gpu


def bench(gp, kernel, precision, reps=3):
    h = gp._h
    us, nbytes = ctypes.c_double(), ctypes.c_double()
    h.stream()
    prec = {"fp64": _lib.TMX_FP64, "fp32": _lib.TMX_FP32}[precision]
    rc = h.lib.tmx_bench_kernel(h.ptr, kernel, prec, reps, ctypes.byref(us), ctypes.byref(nbytes))
    return rc, us.value, nbytes.value


def test_every_kernel_id():
    nx, ny = 120, 40
    nv = (nx + 1) * (ny + 1)
    gp = gpu_problem(CANTILEVER, nx, ny, precision="mixed", rtol=1e-10)
    gp.set_penalization(3)
    rho = torch.as_tensor(smoothed_density(nx, ny, 5), device="cuda:0")
    J0, w0, _ = gp.objective_and_gradient(rho)
    for kernel, prec in ((0, "fp32"), (1, "fp32"), (0, "fp64"), (1, "fp64"), (2, "fp64"),
                         (3, "fp32"), (4, "fp32"), (6, "fp32"), (3, "fp64"), (4, "fp64"),
                         (6, "fp64"), (5, "fp64")):
        rc, us, nbytes = bench(gp, kernel, prec)
        print(kernel, prec, rc, us, nbytes, flush=True)
        assert rc == _lib.TMX_OK, rc
        assert us > 0 and nbytes > 0
        if kernel == 5:
            assert nbytes == 96 * nv
        if kernel == 0:
            assert nbytes == 5 * (4 if prec == "fp32" else 8) * nv
    rc, _, _ = bench(gp, 5, "fp32")
    assert rc == _lib.TMX_ERR_INVALID
    rc, _, _ = bench(gp, 7, "fp64")
    assert rc == _lib.TMX_ERR_INVALID
    J1, w1, _ = gp.objective_and_gradient(rho)
    assert J1 == J0
    assert torch.equal(w1, w0)
