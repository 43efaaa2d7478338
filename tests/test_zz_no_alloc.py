"""zfr_residual_norm_static, the residual norm that allocates nothing (its block partials
live in a fixed device array), against zfr_residual_norm with a caller-provided work
buffer: the same kernels over the same grid, so the norms must be bit-identical.  Two
calls in a row, the second on a smaller element block, for every res_type; the norms are
also checked against numpy on the host (src/solver.cpp:7005-7026 defines them)."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from zefr_amd import _lib


def _desc(nDims, order, E, metric):
    N = order + 1
    nS = N ** nDims
    return _lib.EleDesc(nDims=nDims, order=order, nVars=nDims + 2, nEles=E, nSpts=nS,
                        nFpts=2 * nDims * N ** (nDims - 1), metric=metric.data_ptr())


def _host_norm(divF, metric, nDims, res_type):
    detJ = metric[:, nDims * nDims, :][:, None, :]
    x = divF / detJ
    if res_type == 0:
        return np.abs(x).max(axis=(0, 2))
    if res_type == 1:
        return np.abs(x).sum(axis=(0, 2))
    return (x * x).sum(axis=(0, 2))


@pytest.mark.gpu
@pytest.mark.parametrize("nDims,order,sizes", [(3, 3, (4096, 300)), (2, 4, (20000, 37))])
def test_static_norm_matches_work_buffer_norm(nDims, order, sizes):
    import torch

    L = _lib.lib()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(7 + nDims)
    N = order + 1
    nS, V, nM = N ** nDims, nDims + 2, nDims * nDims + 1
    work = torch.zeros(1024 * V, dtype=torch.float64, device=dev)
    for E in sizes:  # the second call runs on a smaller block than the first
        print(f"no-alloc norm: nDims={nDims} order={order} E={E}", flush=True)
        divF_h = rng.standard_normal((E, V, nS))
        metric_h = rng.standard_normal((E, nM, nS))
        metric_h[:, nDims * nDims, :] = 0.5 + rng.random((E, nS))  # detJ > 0
        divF = torch.from_numpy(divF_h).to(dev)
        metric = torch.from_numpy(metric_h).to(dev)
        d = _desc(nDims, order, E, metric)
        for res_type in (0, 1, 2):
            ref = torch.full((V,), -1.0, dtype=torch.float64, device=dev)
            got = torch.full((V,), -2.0, dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            assert L.zfr_residual_norm(C.byref(d), C.c_void_p(divF.data_ptr()), res_type,
                                       C.c_void_p(work.data_ptr()), C.c_void_p(ref.data_ptr()), None) == 0, \
                _lib.last_error()
            assert L.zfr_residual_norm_static(C.byref(d), C.c_void_p(divF.data_ptr()), res_type,
                                              C.c_void_p(got.data_ptr()), None) == 0, _lib.last_error()
            torch.cuda.synchronize()
            r, g = ref.cpu().numpy(), got.cpu().numpy()
            np.testing.assert_array_equal(g, r)
            host = _host_norm(divF_h, metric_h, nDims, res_type)
            np.testing.assert_allclose(g, host, rtol=1e-12, atol=0)


def test_static_norm_rejects_bad_arguments_before_launching():
    L = _lib.lib()
    buf = C.c_void_p(0x1000)  # never dereferenced: the checks fail first
    d = _lib.EleDesc(nDims=3, order=3, nVars=5, nEles=8, nSpts=64, nFpts=96, metric=buf)
    assert L.zfr_residual_norm_static(C.byref(d), buf, 3, buf, None) != 0
    assert "res_type" in _lib.last_error()
    assert L.zfr_residual_norm_static(C.byref(d), buf, 0, None, None) != 0
    assert "missing buffer" in _lib.last_error()
    d.nVars = 9
    assert L.zfr_residual_norm_static(C.byref(d), buf, 0, buf, None) != 0
    assert "nVars" in _lib.last_error()
