"""na_gemm_wgrad_static, the weight-gradient entry point that allocates nothing (its
partial products live in a static device array), against na_gemm's weight-gradient
path with an allocated partial-product workspace: the same row splits summed in the
same order, so the results are bitwise equal.  Two calls in a row, the second with a
smaller product, so the second call reuses the array the first one filled."""
import ctypes

import pytest
import torch

from neuralangelo_amd import _lib

pytestmark = pytest.mark.gpu


def _args(dpre, X, dW):
    """dW[n, k] += sum_m dpre[m, n] X[m, k]: bf16 operands read transposed, fp32 out."""
    M, N = dpre.shape
    K = X.shape[1]
    a = _lib.GemmArgs()
    a.M, a.N, a.K = N, K, M
    a.A, a.lda, a.trans_a = dpre.data_ptr(), N, 1
    a.B, a.ldb, a.trans_b = X.data_ptr(), K, 1
    a.in_dtype = _lib.NA_BF16
    a.epilogue = _lib.NA_EPI_ATOMIC
    a.C, a.ldc, a.c_dtype = dW.data_ptr(), K, _lib.NA_F32
    a.split_k = 1
    return a


def _with_workspace(dpre, X):
    dW = torch.zeros(dpre.shape[1], X.shape[1], device=dpre.device)
    a = _args(dpre, X, dW)
    n = _lib.call("na_gemm_wgrad_workspace", ctypes.byref(a))
    assert n > 0
    ws = torch.empty(n, device=dpre.device)
    a.wg_partials, a.wg_partials_len = ws.data_ptr(), n
    _lib.call("na_gemm", ctypes.byref(a), _lib.stream_of(dpre))
    torch.cuda.synchronize()
    return dW


def _static(dpre, X):
    dW = torch.zeros(dpre.shape[1], X.shape[1], device=dpre.device)
    a = _args(dpre, X, dW)
    _lib.call("na_gemm_wgrad_static", ctypes.byref(a), _lib.stream_of(dpre))
    torch.cuda.synchronize()
    return dW


def test_static_scratch_matches_workspace_path(cuda):
    g = torch.Generator().manual_seed(0)
    cases = [(200_000, 256, 280), (5_000, 256, 144)]  # second call smaller
    for M, N, K in cases:
        print(f"wgrad static: M={M} N={N} K={K}", flush=True)
        dpre = (torch.randn(M, N, generator=g) * 0.1).to(torch.bfloat16).to(cuda)
        X = torch.randn(M, K, generator=g).to(torch.bfloat16).to(cuda)
        ref = _with_workspace(dpre, X)
        got = _static(dpre, X)
        assert torch.equal(got, ref)
        exact = dpre.float().t().cpu().double() @ X.float().cpu().double()
        assert (got.cpu().double() - exact).abs().max() <= 1e-4 * exact.abs().max()


def test_static_scratch_rejects_what_it_cannot_hold(cuda):
    dpre = torch.zeros(64, 256, dtype=torch.bfloat16, device=cuda)
    X = torch.zeros(64, 256, dtype=torch.bfloat16, device=cuda)
    dW = torch.zeros(256, 256, device=cuda)
    a = _args(dpre, X, dW)
    a.epilogue = _lib.NA_EPI_STORE  # not a weight-gradient product
    with pytest.raises(RuntimeError, match="weight-gradient"):
        _lib.call("na_gemm_wgrad_static", ctypes.byref(a), _lib.stream_of(dpre))
    # a product whose partials cannot fit: 4096 x 8192 outputs need at least
    # 33.5 M floats whatever the row split (the operands are sized for the
    # product, so even a launch would stay in bounds)
    dpre = torch.zeros(64, 4096, dtype=torch.bfloat16, device=cuda)
    X = torch.zeros(64, 8192, dtype=torch.bfloat16, device=cuda)
    dW = torch.zeros(4096, 8192, device=cuda)
    big = _args(dpre, X, dW)
    assert big.M * big.N > _lib.NA_WGRAD_STATIC_MAX_FLOATS
    with pytest.raises(RuntimeError, match="static"):
        _lib.call("na_gemm_wgrad_static", ctypes.byref(big), _lib.stream_of(dpre))
    assert dW.abs().max().item() == 0.0
