"""Dense-layer kernels (GEMM + epilogues, narrow layers, weight norm) vs float64 torch on the CPU."""
import ctypes

import pytest
import torch

from neuralangelo_amd import _lib
from neuralangelo_amd import functional as Fn
from neuralangelo_amd._lib import NA_ACT_RELU, NA_ACT_SIGMOID, NA_ACT_SOFTPLUS

pytestmark = pytest.mark.gpu

BETA = 100.0


def _act(v, act):
    if act == NA_ACT_RELU:
        return torch.relu(v)
    if act == NA_ACT_SOFTPLUS:
        return torch.where(v * BETA > 20, v, torch.log1p(torch.exp(v * BETA)) / BETA)
    if act == NA_ACT_SIGMOID:
        return torch.sigmoid(v)
    return v


def _dact(y, act):
    if act == NA_ACT_RELU:
        return (y > 0).double()
    if act == NA_ACT_SOFTPLUS:
        return torch.where(y * BETA > 20, torch.ones_like(y), -torch.expm1(-y * BETA))
    if act == NA_ACT_SIGMOID:
        return y * (1 - y)
    return torch.ones_like(y)


def _rand(shape, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(*shape, generator=g) * scale


@pytest.mark.parametrize("dt", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(1000, 200, 136), (257, 256, 256), (64, 24, 280)])
def test_gemm_layouts(cuda, dt, ta, tb, M, N, K):
    if ta and M % 8:
        M += 8 - M % 8  # a transposed operand's leading dimension must stay 16-byte aligned
    A = _rand((M, K), 1).to(dt)
    B = _rand((N, K), 2).to(dt)
    ref = A.double() @ B.double().t()
    Ad = (A.t().contiguous() if ta else A).to(cuda)
    Bd = (B.t().contiguous() if tb else B).to(cuda)
    C = torch.empty(M, N, device=cuda)
    Fn.gemm(Ad, M if ta else K, ta, Bd, N if tb else K, tb, M, N, K, C, N)
    err = (C.cpu().double() - ref).abs().max().item()
    tol = 1e-4 * (K ** 0.5) if dt == torch.float32 else 1e-5 * (K ** 0.5)
    assert err <= tol, err


@pytest.mark.parametrize("dt", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("act", [NA_ACT_SOFTPLUS, NA_ACT_RELU, NA_ACT_SIGMOID])
def test_gemm_bias_activation_epilogue(cuda, dt, act):
    M, N, K = 777, 256, 136
    A = _rand((M, K), 3, 0.1).to(dt)
    B = _rand((N, K), 4, 0.1).to(dt)
    bias = _rand((N,), 5, 0.05)
    ref = _act(A.double() @ B.double().t() + bias.double(), act)
    for cdt in (torch.float32, dt):
        C = torch.empty(M, N, dtype=cdt, device=cuda)
        Fn.gemm(A.to(cuda), K, 0, B.to(cuda), K, 0, M, N, K, C, N, bias=bias.to(cuda), act=act)
        got = C.cpu().double()
        rel = 2 ** -7 if cdt == torch.bfloat16 else 1e-5
        assert ((got - ref).abs() <= rel * ref.abs() + 1e-5).all()


def test_gemm_dact_colsum_rank1(cuda):
    M, N, K = 500, 256, 256
    A = _rand((M, K), 6, 0.1)
    W = _rand((K, N), 7, 0.1)  # B(n,k) = W[k, n] (trans_b)
    Y = torch.nn.functional.softplus(_rand((M, N), 8, 0.01), beta=BETA)
    r1a, r1b = _rand((M,), 9), _rand((N,), 10)
    v = A.double() @ W.double() + r1a.double()[:, None] * r1b.double()[None]
    ref = v * _dact(Y.double(), NA_ACT_SOFTPLUS)
    C = torch.empty(M, N, device=cuda)
    cs = torch.zeros(N, device=cuda)
    Fn.gemm(A.to(cuda), K, 0, W.to(cuda), N, 1, M, N, K, C, N, dact=NA_ACT_SOFTPLUS,
            dact_y=Y.to(cuda), ld_dact=N, colsum=cs, r1_a=r1a.to(cuda), r1_b=r1b.to(cuda))
    assert (C.cpu().double() - ref).abs().max() < 1e-4
    assert (cs.cpu().double() - ref.sum(0)).abs().max() < 1e-3


@pytest.mark.parametrize("dt", [torch.float32, torch.bfloat16])
def test_weight_grad_split_k(cuda, dt):
    M, N, K = 20000, 256, 136
    dpre = _rand((M, N), 11, 0.1).to(dt)
    X = _rand((M, K), 12, 0.1).to(dt)
    ref = dpre.double().t() @ X.double()
    dW = torch.zeros(N, K, device=cuda)
    Fn.linear_weight_grad(dpre.to(cuda), X.to(cuda), K, K, dW)
    err = (dW.cpu().double() - ref).abs().max().item()
    assert err < 1e-3 * (M ** 0.5) * 0.01 + 1e-3, err


@pytest.mark.parametrize("M,N,K", [(20000, 256, 136), (3001, 256, 280), (100, 24, 256),
                                   (128 * 900 + 17, 256, 256), (5000, 280, 96)])
def test_weight_grad_split_partials(cuda, M, N, K):
    """Weight gradient with the partial-product workspace (split-ordered sums, no
    atomics): float64 reference, the atomic path, bitwise-equal repeats, and
    accumulation into a non-zero dW."""
    dpre = _rand((M, N), 40, 0.1).to(torch.bfloat16)
    X = _rand((M, K), 41, 0.1).to(torch.bfloat16)
    init = _rand((N, K), 42, 0.01)
    ref = init.double() + dpre.double().t() @ X.double()
    dd, Xd = dpre.to(cuda), X.to(cuda)
    outs = []
    for partials in (True, True, False):
        dW = init.to(cuda)
        Fn.gemm(dd, N, 1, Xd, K, 1, N, K, M, dW, K, epilogue=Fn.NA_EPI_ATOMIC,
                wgrad_partials=partials)
        outs.append(dW.cpu())
    tol = 1e-5 * M ** 0.5 + 1e-5
    assert torch.equal(outs[0], outs[1])
    for o in outs:
        assert (o.double() - ref).abs().max().item() < tol
    assert _lib.call("na_gemm_wgrad_workspace", ctypes.byref(_wgrad_args(dd, Xd, M, N, K))) > 0


def _wgrad_args(dpre, X, M, N, K):
    a = _lib.GemmArgs()
    a.M, a.N, a.K = N, K, M
    a.A, a.lda, a.trans_a = dpre.data_ptr(), N, 1
    a.B, a.ldb, a.trans_b = X.data_ptr(), K, 1
    a.in_dtype, a.epilogue, a.c_dtype = _lib.NA_BF16, Fn.NA_EPI_ATOMIC, _lib.NA_F32
    a.split_k = 1
    return a


def test_narrow_forward_backward(cuda):
    M, K, N = 3000, 256, 3
    X = torch.relu(_rand((M, K), 13, 0.1))
    W = _rand((N, K), 14, 0.1)
    b = _rand((N,), 15, 0.1)
    out = torch.empty(M, N, device=cuda)
    Xd, Wd, bd = X.to(cuda), W.to(cuda), b.to(cuda)
    _lib.call("na_linear_narrow_fwd", _lib.ptr(Xd), K, 0, M, K, _lib.ptr(Wd), _lib.ptr(bd), N,
              NA_ACT_SIGMOID, _lib.ptr(out), _lib.stream_of(out))
    ref = torch.sigmoid(X.double() @ W.double().t() + b.double())
    assert (out.cpu().double() - ref).abs().max() < 1e-6
    dpre = _rand((M, N), 16)
    dpre_d = dpre.to(cuda)
    dW = torch.zeros(N, K, device=cuda)
    db = torch.zeros(N, device=cuda)
    dX = _rand((M, K), 17).to(cuda)
    dX0 = dX.clone()
    cs = torch.zeros(K, device=cuda)
    _lib.call("na_linear_narrow_bwd", _lib.ptr(Xd), K, 0, M, K, _lib.ptr(Wd), N,
              _lib.ptr(dpre_d), _lib.ptr(dW), _lib.ptr(db), _lib.ptr(dX), K, 0, 1,
              _lib.ptr(Xd), K, 0, NA_ACT_RELU, BETA, _lib.ptr(cs), _lib.stream_of(dW))
    refW = dpre.double().t() @ X.double()
    refdX = (dX0.cpu().double() + dpre.double() @ W.double()) * (X.double() > 0)
    assert (dW.cpu().double() - refW).abs().max() < 1e-3
    assert (db.cpu().double() - dpre.double().sum(0)).abs().max() < 1e-3
    assert (dX.cpu().double() - refdX).abs().max() < 1e-5
    assert (cs.cpu().double() - refdX.sum(0)).abs().max() < 1e-3


@pytest.mark.parametrize("M,K,N,dt", [(3001, 256, 1, torch.bfloat16),
                                      (128 * 900, 256, 1, torch.bfloat16),
                                      (2999, 384, 1, torch.bfloat16),
                                      (4097, 256, 3, torch.bfloat16),
                                      (1000, 136, 4, torch.float32)])
def test_narrow_backward_x_only(cuda, M, K, N, dt):
    """Rows whose only operand is X (softplus derivative read from X itself, no dX
    accumulation; the SDF head backward): against the general path, taken when Y is
    a separate copy of X, and float64.  Rows visit the same lanes in the same order
    on both paths, so dX agrees bitwise; the dW-only call (dX null) is covered too."""
    X = torch.nn.functional.softplus(_rand((M, K), 30, 0.02), beta=BETA).to(dt)
    W = _rand((N, K), 31, 0.1)
    dpre = _rand((M, N), 32)
    Xd, Wd, dpre_d = X.to(cuda), W.to(cuda), dpre.to(cuda)
    Ycopy = Xd.clone()
    code = _lib.dtype_code(dt)
    outs = []
    for Y in (Xd, Ycopy):
        dW = torch.zeros(N, K, device=cuda)
        db = torch.zeros(N, device=cuda)
        cs = torch.zeros(K, device=cuda)
        dX = torch.full((M, K), float("nan"), dtype=dt, device=cuda)
        _lib.call("na_linear_narrow_bwd", _lib.ptr(Xd), K, code, M, K, _lib.ptr(Wd), N,
                  _lib.ptr(dpre_d), _lib.ptr(dW), _lib.ptr(db), _lib.ptr(dX), K, code, 0,
                  _lib.ptr(Y), K, code, NA_ACT_SOFTPLUS, BETA, _lib.ptr(cs), _lib.stream_of(dW))
        outs.append((dW.cpu(), db.cpu(), dX.cpu(), cs.cpu()))
    (dW, db, dX, cs), (dW_g, db_g, dX_g, cs_g) = outs
    assert torch.equal(dX.view(torch.int16 if dt == torch.bfloat16 else torch.int32),
                       dX_g.view(torch.int16 if dt == torch.bfloat16 else torch.int32))
    Xr = X.double()
    refW = dpre.double().t() @ Xr
    refdX = (dpre.double() @ W.double()) * _dact(Xr, NA_ACT_SOFTPLUS)
    tol_w = 1e-5 * M ** 0.5 + 1e-4
    assert (dW.double() - refW).abs().max() < tol_w
    assert (dW.double() - dW_g.double()).abs().max() < tol_w
    assert (db.double() - dpre.double().sum(0)).abs().max() < tol_w
    tol_x = (2 ** -7 if dt == torch.bfloat16 else 1e-6) * refdX.abs() + 1e-6
    assert ((dX.double() - refdX).abs() <= tol_x).all()
    assert (cs.double() - refdX.sum(0)).abs().max() < tol_w
    assert (cs.double() - cs_g.double()).abs().max() < tol_w
    # weight gradients only (no input gradient, no derivative)
    dW2 = torch.zeros(N, K, device=cuda)
    db2 = torch.zeros(N, device=cuda)
    _lib.call("na_linear_narrow_bwd", _lib.ptr(Xd), K, code, M, K, _lib.ptr(Wd), N,
              _lib.ptr(dpre_d), _lib.ptr(dW2), _lib.ptr(db2), None, K, code, 0,
              None, K, code, 0, BETA, None, _lib.stream_of(dW2))
    assert (dW2.cpu().double() - refW).abs().max() < tol_w
    assert (db2.cpu().double() - dpre.double().sum(0)).abs().max() < tol_w


def test_weight_norm_matches_torch(cuda):
    v = _rand((256, 131), 18)
    g = _rand((256, 1), 19).abs() + 0.5
    lin = torch.nn.utils.weight_norm(torch.nn.Linear(131, 256).double())
    with torch.no_grad():
        lin.weight_v.copy_(v.double())
        lin.weight_g.copy_(g.double())
    vd = v.to(cuda).requires_grad_(True)
    gd = g.to(cuda).requires_grad_(True)
    W = Fn.weight_norm(vd, gd)
    Wref = torch._weight_norm(lin.weight_v, lin.weight_g, 0)
    assert (W.detach().cpu().double() - Wref.detach()).abs().max() < 1e-6
    dW = _rand((256, 131), 20)
    W.backward(dW.to(cuda))
    gv, gg = torch.autograd.grad((Wref * dW.double()).sum(), [lin.weight_v, lin.weight_g])
    assert (vd.grad.cpu().double() - gv).abs().max() < 1e-5
    assert (gd.grad.cpu().double() - gg).abs().max() < 1e-5


def test_act_grad(cuda):
    M, N = 999, 256
    dy = _rand((M, 280), 21)
    y = torch.nn.functional.softplus(_rand((M, 280), 22, 0.02), beta=BETA)
    cs = torch.zeros(N, device=cuda)
    out = Fn.act_grad(dy.to(cuda), y.to(cuda), NA_ACT_SOFTPLUS, torch.float32, colsum=cs, n_cols=N)
    ref = dy[:, :N].double() * _dact(y[:, :N].double(), NA_ACT_SOFTPLUS)
    assert (out.cpu().double() - ref).abs().max() < 1e-5
    assert (cs.cpu().double() - ref.sum(0)).abs().max() < 1e-3


@pytest.mark.parametrize("M", [3001, 128 * 900])
def test_act_grad_bf16_rows(cuda, M):
    """bf16 gradient and activation rows with a wider row stride (the vectorised path):
    bf16 and fp32 output, column sums over ragged and large row counts."""
    N = 256
    dy = _rand((M, 280), 26).to(torch.bfloat16)
    y = torch.nn.functional.softplus(_rand((M, 280), 27, 0.02), beta=BETA).to(torch.bfloat16)
    ref = dy[:, :N].double() * _dact(y[:, :N].double(), NA_ACT_SOFTPLUS)
    for odt in (torch.bfloat16, torch.float32):
        cs = torch.zeros(N, device=cuda)
        out = Fn.act_grad(dy.to(cuda), y.to(cuda), NA_ACT_SOFTPLUS, odt, colsum=cs, n_cols=N)
        tol = (2 ** -7 if odt == torch.bfloat16 else 1e-6) * ref.abs() + 1e-6
        assert ((out.cpu().double() - ref).abs() <= tol).all(), odt
        assert (cs.cpu().double() - ref.sum(0)).abs().max() < 1e-4 * M ** 0.5 + 1e-3


def test_gemm_rejects_misaligned(cuda):
    A = torch.zeros(10, 6, device=cuda)
    C = torch.zeros(10, 10, device=cuda)
    with pytest.raises(RuntimeError, match="multiple"):
        Fn.gemm(A, 6, 0, A, 6, 0, 10, 10, 6, C, 10)
    with pytest.raises(RuntimeError, match="split_k"):
        Fn.gemm(torch.zeros(8, 8, device=cuda), 8, 0, torch.zeros(8, 8, device=cuda), 8, 0, 8, 8, 8,
                C, 10, split_k=2)


@pytest.mark.parametrize("M,N,K,lda", [(128 * 300, 256, 144, 144), (128 * 517, 256, 256, 256),
                                       (128 * 211, 256, 280, 280), (128 * 3, 256, 256, 264),
                                       (128, 256, 136, 136), (128 * 401, 128, 256, 256),
                                       (128 * 7, 128, 144, 152)])
def test_gemm_weights_stationary_layers(cuda, M, N, K, lda):
    """256- and 128-wide layers on whole 128-row tiles (the weights-stationary path:
    weights in registers, activation stages only, register epilogue): plain,
    bias + ReLU and bias + softplus against float64 on the same bf16 inputs, bf16
    and fp32 C.  K = 136, 144 and 280 end in a partly filled 32-wide stage;
    lda = 264 and 152 pad rows."""
    bf = torch.bfloat16
    A = _rand((M, lda), 31).to(bf)
    B = (_rand((N, K), 32) * 0.1).to(bf)
    b = _rand((N,), 33, 0.5)
    z = A[:, :K].double() @ B.double().t()
    for act in (0, NA_ACT_RELU, NA_ACT_SOFTPLUS):
        ref = z if act == 0 else _act(z + b.double(), act)
        for cdt in (bf, torch.float32):
            C = torch.full((M, N), float("nan"), dtype=cdt, device=cuda)
            if act == 0:
                Fn.gemm(A.to(cuda), lda, 0, B.to(cuda), K, 0, M, N, K, C, N)
            else:
                Fn.gemm(A.to(cuda), lda, 0, B.to(cuda), K, 0, M, N, K, C, N, bias=b.to(cuda), act=act)
            got = C.cpu().double()
            tol = (2 ** -7 if cdt == bf else 1e-5) * ref.abs() + 1e-4
            assert torch.isfinite(got).all(), (act, cdt)
            assert ((got - ref).abs() <= tol).all(), (act, cdt, (got - ref).abs().max().item())


@pytest.mark.parametrize("M", [128 * 613, 128 * 2])
def test_gemm_weights_stationary_derivative(cuda, M):
    """Backward layers on the weights-stationary path: (dY . W + r1_a r1_b) * act'(Y)
    from a bf16 activation Y, ReLU and softplus, with and without the rank-1 term,
    column sums accumulated over row tiles, bf16 and fp32 C, against float64."""
    bf = torch.bfloat16
    N = K = 256
    A = _rand((M, K), 41, 0.3).to(bf)
    B = (_rand((N, K), 42) * 0.1).to(bf)
    r1a, r1b = _rand((M,), 43), _rand((N,), 44)
    z = A.double() @ B.double().t()
    for dact in (NA_ACT_RELU, NA_ACT_SOFTPLUS):
        pre = _rand((M, N), 45, 0.02 if dact == NA_ACT_SOFTPLUS else 1.0)
        Y = (_act(pre, dact)).to(bf)
        d = _dact(Y.double(), dact)
        for rank1 in (False, True):
            ref = (z + (r1a.double()[:, None] * r1b.double()[None] if rank1 else 0.0)) * d
            for cdt in (bf, torch.float32):
                C = torch.full((M, N), float("nan"), dtype=cdt, device=cuda)
                cs = torch.zeros(N, device=cuda)
                kw = dict(r1_a=r1a.to(cuda), r1_b=r1b.to(cuda)) if rank1 else {}
                Fn.gemm(A.to(cuda), K, 0, B.to(cuda), K, 0, M, N, K, C, N, dact=dact,
                        dact_y=Y.to(cuda), ld_dact=N, colsum=cs, **kw)
                got = C.cpu().double()
                tol = (2 ** -7 if cdt == bf else 1e-5) * ref.abs() + 1e-4
                assert ((got - ref).abs() <= tol).all(), (dact, rank1, cdt)
                # column sums of the fp32 values before the output cast
                assert (cs.cpu().double() - ref.sum(0)).abs().max() < 1e-4 * M ** 0.5 + 1e-3


@pytest.mark.parametrize("M,N,K", [(70001, 256, 144), (140001, 128, 256), (3001, 200, 280),
                                   (999, 24, 136), (1000, 300, 256), (128 * 500, 256, 144),
                                   (128 * 300, 256, 256), (128 * 8, 128, 144)])
def test_gemm_nn_bf16_fused_epilogues(cuda, M, N, K):
    """bf16 NN products with every store-epilogue term the MLPs use, at row counts that
    give each workgroup several tiles, ragged M and N, and N below / above one tile;
    the whole-tile shapes (M a multiple of 128, N = 256 or 128) take the
    weights-stationary path, increment mode included when the stored rows fill
    whole tiles."""
    bf = torch.bfloat16
    A = _rand((M, K), 20, 0.1).to(bf)
    B = _rand((N, K), 21, 0.1).to(bf)
    bias = _rand((N,), 22, 0.05)
    hw = _rand((N,), 23)
    acc = A.double() @ B.double().t()
    nparts = Fn.head_parts(N)
    Ad, Bd = A.to(cuda), B.to(cuda)

    # forward: bias + softplus, pre-activation out, fused head, bf16 C
    z = acc + bias.double()
    y = _act(z, NA_ACT_SOFTPLUS)
    C = torch.empty(M, N, dtype=bf, device=cuda)
    zo = torch.empty(M, N, device=cuda)
    ho = torch.full((nparts, M), float("nan"), device=cuda)
    Fn.gemm(Ad, K, 0, Bd, K, 0, M, N, K, C, N, bias=bias.to(cuda), act=NA_ACT_SOFTPLUS,
            z_out=zo, head_w=hw.to(cuda), head_out=ho)
    assert ((C.cpu().double() - y).abs() <= 2 ** -7 * y.abs() + 1e-5).all()
    assert (zo.cpu().double() - z).abs().max() < 1e-5
    head = ho.cpu().double().sum(0)
    assert (head - y @ hw.double()).abs().max() < 1e-3

    # backward: relu derivative from a bf16 activation, column sums, fp32 C
    Y = torch.relu(_rand((M, N), 24)).to(bf)
    ref = acc * (Y.double() > 0)
    C32 = torch.empty(M, N, device=cuda)
    cs = torch.zeros(N, device=cuda)
    Fn.gemm(Ad, K, 0, Bd, K, 0, M, N, K, C32, N, dact=NA_ACT_RELU, dact_y=Y.to(cuda), ld_dact=N,
            colsum=cs)
    assert (C32.cpu().double() - ref).abs().max() < 1e-5
    assert (cs.cpu().double() - ref.sum(0)).abs().max() < 1e-5 * M ** 0.5 + 1e-4

    # increment mode: x = softplus(zc[m % rows] + acc), head over softplus increments
    rows = (M + 3) // 4
    zc = _rand((rows, N), 25, 0.02)
    z0 = zc.double()[torch.arange(M) % rows]
    y1 = _act(z0 + acc, NA_ACT_SOFTPLUS)
    dsp = y1 - _act(z0, NA_ACT_SOFTPLUS)
    ho.fill_(float("nan"))
    Fn.gemm(Ad, K, 0, Bd, K, 0, M, N, K, C, N, act=NA_ACT_SOFTPLUS, zc=zc.to(cuda), zc_rows=rows,
            head_w=hw.to(cuda), head_out=ho)
    assert ((C.cpu().double() - y1).abs() <= 2 ** -7 * y1.abs() + 1e-5).all()
    assert (ho.cpu().double().sum(0) - dsp @ hw.double()).abs().max() < 1e-3


@pytest.mark.parametrize("M,N,K", [(60001, 256, 280), (50000, 256, 256), (20000, 256, 144),
                                   (777, 64, 136), (40000, 296, 24), (33, 256, 256)])
def test_weight_grad_row_reduction(cuda, M, N, K):
    """dW[n, k] += sum_m dpre[m, n] X[m, k] in bf16 with fp32 accumulation: every
    output-tile width, ragged row ranges, N above one tile, and accumulation into a
    non-zero output."""
    bf = torch.bfloat16
    dpre = _rand((M, N), 30, 0.1).to(bf)
    X = _rand((M, K), 31, 0.1).to(bf)
    dW0 = _rand((N, K), 32)
    ref = dW0.double() + dpre.double().t() @ X.double()
    dW = dW0.to(cuda)
    Fn.linear_weight_grad(dpre.to(cuda), X.to(cuda), K, K, dW)
    err = (dW.cpu().double() - ref).abs().max().item()
    assert err < 1e-3 * (M ** 0.5) * 0.01 + 1e-4, err


def test_increment_epilogue_large_steps_and_linear_region(cuda):
    """Increment mode where |beta * dz| >= 1 and beta * z0 lies in or near the linear
    region (17..25), where sigmoid(beta z0) rounds to 1 in float32: values and
    increments stay finite and match float64."""
    bf = torch.bfloat16
    M, N, K = 4096, 128, 16
    v = torch.linspace(-1.0, 1.0, M)
    A = torch.zeros(M, K)
    A[:, 0] = v
    B = torch.zeros(N, K)
    B[:, 0] = 1.0
    A, B = A.to(bf), B.to(bf)
    acc = A.double() @ B.double().t()  # acc[m, n] = v_m (bf16-rounded)
    rows = 512
    z0 = torch.linspace(-0.3, 0.3, N)[None, :].repeat(rows, 1)
    z0[:, ::4] = torch.linspace(0.17, 0.25, N // 4)[None, :]  # beta z0 in 17..25
    z0d = z0.double()[torch.arange(M) % rows]
    y1 = _act(z0d + acc, NA_ACT_SOFTPLUS)
    dsp = y1 - _act(z0d, NA_ACT_SOFTPLUS)
    C = torch.empty(M, N, dtype=torch.float32, device=cuda)
    hw = torch.ones(N)
    ho = torch.empty(Fn.head_parts(N), M, device=cuda)
    Fn.gemm(A.to(cuda), K, 0, B.to(cuda), K, 0, M, N, K, C, N, act=NA_ACT_SOFTPLUS,
            zc=z0.to(cuda), zc_rows=rows, head_w=hw.to(cuda), head_out=ho)
    got = C.cpu().double()
    assert torch.isfinite(got).all() and torch.isfinite(ho).all()
    assert ((got - y1).abs() <= 1e-5 * y1.abs() + 1e-6).all()
    assert (ho.cpu().double().sum(0) - dsp.sum(1)).abs().max() < 1e-4
