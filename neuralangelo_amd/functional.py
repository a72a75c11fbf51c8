"""Autograd functions of the hot path, each backed by the gfx950 C-ABI kernels.

Module boundaries follow the reference: one function per MLP
(MLPforNeuralSDF, projects/neuralangelo/utils/mlp.py:55-69, together with the
hash-grid encoding of projects/neuralangelo/utils/modules.py:76-95; and
MLPwithSkipConnection + NeuralRGB input, modules.py:209-220), and one per
renderer utility (NeuS alpha, model.py:275-298; compositing, render.py:87-112;
4-tap combine, modules.py:157-175; eikonal/curvature, misc.py:74-89).

Buffers in HBM (P points):
  SDF input   X0 [P, K0p]  = [enc(L*F) | x(3) | 0 pad]  (compute dtype)
  hidden      H0 [P, 256]  softplus output (compute dtype)
  colour in   FB [P, 280]  = [feat(256) | x | SH16 | normal | 0 0]; the SDF
              feature layer writes its output straight into columns 0..255.
  colour hid  H1..H4 [P, 256] relu outputs
Compute dtype is float32 or bfloat16 (MFMA inputs); accumulation, the SDF
head, weight gradients and every per-ray quantity are float32.
"""
import ctypes

import torch
import torch.nn.functional as torch_F

from . import _lib
from ._lib import (NA_ACT_NONE, NA_ACT_RELU, NA_ACT_SIGMOID, NA_ACT_SOFTPLUS, NA_EPI_ATOMIC,
                   NA_EPI_STORE)

SOFTPLUS_BETA = 100.0
RGB_LD = 280  # 256 features + 3 + 16 + 3, padded to a multiple of 8


# Optional HIP-event timing of named launches on the launching stream: a dict
# name -> list that collects (start, end) torch.cuda.Event pairs while set.
LAUNCH_EVENTS = None
# Optional record of the table-gradient scatter's inputs: while a list, each launch
# of the centre + taps scatter appends (points, eps, dy_centre, dy_taps, dy_shared,
# ld, increments) (used to count the table entries a launch touches and to tune
# its launch on the step's own inputs).
SCATTER_CAPTURE = None


class _timed:
    def __init__(self, name):
        self.pair = None
        if LAUNCH_EVENTS is not None and name in LAUNCH_EVENTS:
            self.pair = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            LAUNCH_EVENTS[name].append(self.pair)

    def __enter__(self):
        if self.pair:
            self.pair[0].record()

    def __exit__(self, *exc):
        if self.pair:
            self.pair[1].record()


def _pad8(k):
    return (k + 7) // 8 * 8


def _code(t):
    return _lib.dtype_code(t.dtype)


def _s(t):
    return _lib.stream_of(t)


# ---------------------------------------------------------------------------
# GEMM helper
# ---------------------------------------------------------------------------
def gemm(A, lda, trans_a, B, ldb, trans_b, M, N, K, C, ldc, epilogue=NA_EPI_STORE, bias=None,
         act=NA_ACT_NONE, beta=SOFTPLUS_BETA, dact=NA_ACT_NONE, dact_y=None, ld_dact=0,
         colsum=None, split_k=1, r1_a=None, r1_b=None, z_out=None, zc=None, zc_rows=0,
         head_w=None, head_out=None, wgrad_partials=True):
    """C[m,n] = epi(sum_k A(m,k) B(n,k)) via na_gemm (see include/neuralangelo_amd.h).
    When the weight-gradient path runs the call, wgrad_partials gives it a workspace
    for per-split partial products (split-ordered sums instead of atomics)."""
    if A.dtype != B.dtype:
        raise RuntimeError(f"gemm operands must share a dtype ({A.dtype} vs {B.dtype})")
    a = _lib.GemmArgs()
    a.M, a.N, a.K = int(M), int(N), int(K)
    a.A, a.lda, a.trans_a = A.data_ptr(), int(lda), int(trans_a)
    a.B, a.ldb, a.trans_b = B.data_ptr(), int(ldb), int(trans_b)
    a.in_dtype = _code(A)
    a.epilogue = epilogue
    a.C, a.ldc, a.c_dtype = C.data_ptr(), int(ldc), _code(C)
    a.bias = bias.data_ptr() if bias is not None else None
    a.act, a.act_beta = act, beta
    a.dact = dact
    if dact_y is not None:
        a.dact_y, a.ld_dact, a.dact_dtype = dact_y.data_ptr(), int(ld_dact), _code(dact_y)
    a.r1_a = r1_a.data_ptr() if r1_a is not None else None
    a.r1_b = r1_b.data_ptr() if r1_b is not None else None
    a.colsum = colsum.data_ptr() if colsum is not None else None
    a.split_k = int(split_k)
    if z_out is not None:
        a.z_out, a.ld_z = z_out.data_ptr(), int(z_out.stride(0))
    if zc is not None:
        a.zc, a.ld_zc, a.zc_rows = zc.data_ptr(), int(zc.stride(0)), int(zc_rows)
    if head_w is not None:
        a.head_w, a.head_out = head_w.data_ptr(), head_out.data_ptr()
    ws = None
    if wgrad_partials and epilogue == NA_EPI_ATOMIC:
        n = _lib.call("na_gemm_wgrad_workspace", ctypes.byref(a))
        if n > 0:
            ws = torch.empty(n, dtype=torch.float32, device=C.device)
            a.wg_partials, a.wg_partials_len = ws.data_ptr(), n
    _lib.call("na_gemm", ctypes.byref(a), _s(A))


def _zero_block(dev, *shapes):
    """Zero-filled fp32 tensors of the given shapes carved from one buffer: one fill
    launch for a group of gradient accumulators.  Each tensor starts on a 256-byte
    boundary (sizes rounded up to 64 floats) for the vectorised kernels."""
    sizes = []
    for s in shapes:
        n = 1
        for d in s:
            n *= int(d)
        sizes.append(n)
    buf = torch.zeros(sum((n + 63) // 64 * 64 for n in sizes), device=dev)
    out, o = [], 0
    for s, n in zip(shapes, sizes):
        out.append(buf[o:o + n].view(*s))
        o += (n + 63) // 64 * 64
    return out


def head_parts(N):
    """Number of partial sums per row written by a fused GEMM head over N columns
    (one per 32 columns of the 128-column tiles)."""
    return 4 * ((N + 127) // 128)


def head_sum(parts, bias=None):
    """[nparts, M] partial head sums -> [M] (plus bias[0])."""
    nparts, M = parts.shape
    out = torch.empty(M, device=parts.device)
    _lib.call("na_head_sum", _lib.ptr(parts), nparts, M, _lib.ptr(bias), _lib.ptr(out), _s(parts))
    return out


def _t(W):
    """W^T, contiguous: input-gradient GEMMs then read both operands along k."""
    return W.t().contiguous()


def _split_for(rows):
    """split-K factor for weight-gradient GEMMs reducing over `rows` points."""
    return max(1, min(512, rows // 2048))


def linear_weight_grad(dpre, X, ldx, K, dW):
    """dW[n, k] += sum_m dpre[m, n] X[m, k]  (dpre [M, N] contiguous)."""
    M, N = dpre.shape
    if M == 0:
        return
    gemm(dpre, N, 1, X, ldx, 1, N, K, M, dW, dW.stride(0), epilogue=NA_EPI_ATOMIC,
         split_k=_split_for(M))


def act_grad(dy, y, act, out_dtype, colsum=None, n_cols=None):
    """dy * act'(y) over the first n_cols columns (row strides taken from the tensors)."""
    M = dy.shape[0]
    N = dy.shape[1] if n_cols is None else n_cols
    out = torch.empty(M, N, dtype=out_dtype, device=dy.device)
    _lib.call("na_act_grad", _lib.ptr(dy), dy.stride(0), _code(dy), _lib.ptr(y), y.stride(0),
              _code(y), M, N, act, SOFTPLUS_BETA, _lib.ptr(out), N, _code(out),
              _lib.ptr(colsum), _s(dy))
    return out


# ---------------------------------------------------------------------------
# weight normalisation (torch.nn.utils.weight_norm, dim=0)
# ---------------------------------------------------------------------------
class WeightNorm(torch.autograd.Function):
    """W = g * v / ||v||_row  (v [N,K], g [N,1] or [N])."""

    @staticmethod
    def forward(ctx, v, g):
        _lib.require_cuda(v, g)
        v = v.contiguous()
        gf = g.reshape(-1).contiguous()
        N, K = v.shape
        W = torch.empty_like(v)
        _lib.call("na_weight_norm_fwd", _lib.ptr(v), _lib.ptr(gf), N, K, _lib.ptr(W), None, 0, 0,
                  _s(v))
        ctx.save_for_backward(v, gf)
        ctx.g_shape = g.shape
        return W

    @staticmethod
    def backward(ctx, dW):
        v, gf = ctx.saved_tensors
        N, K = v.shape
        dv, dg = torch.empty_like(v), torch.empty_like(gf)  # written, not accumulated
        _lib.call("na_weight_norm_bwd", _lib.ptr(v), _lib.ptr(gf), _lib.ptr(dW.contiguous()), N, K,
                  _lib.ptr(dv), _lib.ptr(dg), _s(v))
        return dv, dg.view(ctx.g_shape)


def weight_norm(v, g):
    return WeightNorm.apply(v, g)


# ---------------------------------------------------------------------------
# SDF network: encoder + MLPforNeuralSDF
# ---------------------------------------------------------------------------
class SDFWeights:
    """Layer weights of one step in kernel layout (built from the module's tensors).

    W0 reference columns are [x(3), enc(LF)]; the kernel layout puts the encoder
    block first (aligned):  fp32: [enc | x | 0 pad];  bf16: [enc | W_hi | W_hi |
    W_lo | 0 pad] against the input [enc | x_hi | x_lo | x_hi | 0], which carries
    the point and its weights to ~16 mantissa bits through the bf16 MFMA.
    """

    def __init__(self, W0, b0, W1, b1, ws, bs, n_enc, cdt):
        self.n_enc = n_enc
        self.split = cdt == torch.bfloat16
        self.k0p = sdf_input_width(n_enc, cdt)
        self.cdt = cdt
        dev = W0.device
        n = n_enc
        Wx = W0[:, :3].float()
        if self.split:
            Whi = Wx.to(cdt)
            Wlo = (Wx - Whi.float()).to(cdt)
            W0p = torch.zeros(W0.shape[0], self.k0p, dtype=cdt, device=dev)
            W0p[:, :n] = W0[:, 3:].to(cdt)
            W0p[:, n:n + 3] = Whi
            W0p[:, n + 3:n + 6] = Whi
            W0p[:, n + 6:n + 9] = Wlo
            self.W0p = W0p.contiguous()
        else:
            W0p = torch.zeros(W0.shape[0], self.k0p, dtype=torch.float32, device=dev)
            W0p[:, :n] = W0[:, 3:]
            W0p[:, n:n + 3] = Wx
            self.W0p = W0p.to(cdt).contiguous()
        self.b0 = b0.contiguous().float()
        self.W1 = W1.to(cdt).contiguous() if W1 is not None else None
        self.b1 = b1.contiguous().float() if b1 is not None else None
        self.ws = ws.reshape(1, -1).contiguous().float()
        self.bs = bs.reshape(-1).contiguous().float()
        self.ws_shape, self.bs_shape = ws.shape, bs.shape
        self.hidden = W0.shape[0]

    def increment_weights(self, ld_d):
        """First-layer weights against the tap-increment rows [hi part and point | lo
        part of the encoder columns]: [W0p | W_enc | 0] of width ld_d (bf16)."""
        key = (ld_d, self.W0p.data_ptr(), self.W0p._version)
        if getattr(self, "_inc_key", None) != key:
            n, k0p = self.n_enc, self.k0p
            W = torch.zeros(self.W0p.shape[0], ld_d, dtype=self.W0p.dtype, device=self.W0p.device)
            W[:, :k0p] = self.W0p
            W[:, k0p:k0p + n] = self.W0p[:, :n]
            self._inc, self._inc_key = W, key
        return self._inc

    def unpermute_dW0(self, dW0p):
        """[hidden, K0p] kernel-layout gradient -> reference layout [hidden, 3+LF].
        With the split layout d/dW_x = sum dpre (x_hi + x_lo): columns hi + lo."""
        n = self.n_enc
        dWx = dW0p[:, n:n + 3] + dW0p[:, n + 3:n + 6] if self.split else dW0p[:, n:n + 3]
        return torch.cat([dWx, dW0p[:, :n]], dim=1)

    def build_input(self, x):
        return reference_input_to_kernel(x, self.n_enc, self.cdt)


def sdf_input_width(n_enc, cdt):
    """Row width of the SDF input buffer: enc + point (+ split point for bf16), padded to 8."""
    return _pad8(n_enc + (9 if cdt == torch.bfloat16 else 3))


def reference_input_to_kernel(x, n_enc, cdt):
    """Reference-order SDF input [P, 3+LF] -> kernel-layout X0 [P, K0p] with torch ops
    (differentiable: the bf16 casts pass gradients through unchanged, so the point
    gradient is the sum of its hi and lo columns' gradients)."""
    enc, pt = x[:, 3:], x[:, :3]
    if cdt == torch.bfloat16:
        hi = pt.to(cdt)
        lo = (pt.float() - hi.float()).to(cdt)
        X0 = torch.cat([enc.to(cdt), hi, lo, hi], dim=1)
    else:
        X0 = torch.cat([enc, pt], dim=1).to(cdt)
    k0p = sdf_input_width(n_enc, cdt)
    return torch_F.pad(X0, (0, k0p - X0.shape[1])) if X0.shape[1] < k0p else X0


def sdf_input(enc, points, cdt, k0p):
    """X0 [P, K0p] = [enc | x | 1 | 0] for world points [P,3]."""
    P = points.shape[0]
    X0 = torch.empty(P, k0p, dtype=cdt, device=points.device)
    table = enc.table_for_forward()
    vmin, vmax = enc.vol_range
    _lib.call("na_sdf_input_fwd", ctypes.byref(enc.desc), _lib.ptr(points), P, float(vmin),
              float(vmax), _lib.ptr(table), _code(table), int(enc.active_levels), _lib.ptr(X0),
              _code(X0), k0p, _s(points))
    return X0


def sdf_forward_nograd(enc, w, points):
    """SDF value only, no autograd state (sampling passes)."""
    P = points.shape[0]
    X0 = sdf_input(enc, points, w.cdt, w.k0p)
    H0 = torch.empty(P, w.hidden, dtype=w.cdt, device=points.device)
    gemm(X0, w.k0p, 0, w.W0p, w.k0p, 0, P, w.hidden, w.k0p, H0, w.hidden, bias=w.b0,
         act=NA_ACT_SOFTPLUS)
    sdf = torch.empty(P, 1, dtype=torch.float32, device=points.device)
    _lib.call("na_linear_narrow_fwd", _lib.ptr(H0), w.hidden, _code(H0), P, w.hidden,
              _lib.ptr(w.ws), _lib.ptr(w.bs), 1, NA_ACT_NONE, _lib.ptr(sdf), _s(points))
    return sdf


def _sdf_mlp_forward(X0, w, with_feat):
    """H0 = softplus(X0 W0p^T + b0), sdf = H0 ws^T + bs, FB[:, :hid] = softplus(H0 W1^T + b1)."""
    P, dev, cdt = X0.shape[0], X0.device, w.cdt
    H0 = torch.empty(P, w.hidden, dtype=cdt, device=dev)
    gemm(X0, w.k0p, 0, w.W0p, w.k0p, 0, P, w.hidden, w.k0p, H0, w.hidden, bias=w.b0,
         act=NA_ACT_SOFTPLUS)
    sdf = torch.empty(P, 1, dtype=torch.float32, device=dev)
    _lib.call("na_linear_narrow_fwd", _lib.ptr(H0), w.hidden, _code(H0), P, w.hidden,
              _lib.ptr(w.ws), _lib.ptr(w.bs), 1, NA_ACT_NONE, _lib.ptr(sdf), _s(X0))
    FB = None
    if with_feat:
        FB = torch.empty(P, RGB_LD, dtype=cdt, device=dev)
        gemm(H0, w.hidden, 0, w.W1, w.hidden, 0, P, w.hidden, w.hidden, FB, RGB_LD, bias=w.b1,
             act=NA_ACT_SOFTPLUS)
    return H0, sdf, FB


class _SDFGrads:
    """Weight-gradient accumulators of the SDF MLP (kernel layout, fp32)."""

    def __init__(self, w, dev, with_feat):
        hid = w.hidden
        shapes = [(hid, w.k0p), (hid,), (1, hid), (1,)] + ([(hid, hid), (hid,)] if with_feat else [])
        z = _zero_block(dev, *shapes)
        self.dW0p, self.db0, self.dws, self.dbs = z[:4]
        self.dW1, self.db1 = (z[4], z[5]) if with_feat else (None, None)


def _sdf_mlp_backward(X0, H0, FB, w, dsdf, dfeat, acc, dpre0=None, wgrad=True):
    """Backward through the SDF MLP; weight gradients accumulate into `acc`.
    Returns the first layer's pre-activation gradient dpre0 [P, hid] (compute dtype,
    written into `dpre0` when given).  wgrad=False leaves the first layer's weight
    gradient to the caller (which may still add to dpre0)."""
    P, hid, cdt, dev = X0.shape[0], w.hidden, w.cdt, X0.device
    dsdf = (torch.zeros(P, 1, device=dev) if dsdf is None else dsdf).reshape(P, 1)
    dsdf = dsdf.contiguous().float()
    has_feat = FB is not None and dfeat is not None
    if dpre0 is None:
        dpre0 = torch.empty(P, hid, dtype=cdt, device=dev)
    if has_feat:
        dpre1 = act_grad(dfeat, FB, NA_ACT_SOFTPLUS, cdt, colsum=acc.db1, n_cols=hid)
        linear_weight_grad(dpre1, H0, hid, hid, acc.dW1)
        # dpre0 = (dpre1 . W1 + dsdf ws) * softplus'(H0), db0 = column sums: one GEMM epilogue
        gemm(dpre1, hid, 0, _t(w.W1), hid, 0, P, hid, hid, dpre0, hid, dact=NA_ACT_SOFTPLUS,
             dact_y=H0, ld_dact=hid, r1_a=dsdf.view(-1), r1_b=w.ws.view(-1), colsum=acc.db0)
        # dws = dsdf^T H0, dbs = sum dsdf (no input gradient from this call)
        _lib.call("na_linear_narrow_bwd", _lib.ptr(H0), hid, _code(H0), P, hid, _lib.ptr(w.ws), 1,
                  _lib.ptr(dsdf), _lib.ptr(acc.dws), _lib.ptr(acc.dbs), None, hid, _code(dpre0), 0,
                  None, hid, _code(H0), NA_ACT_NONE, SOFTPLUS_BETA, None, _s(X0))
    else:
        # dpre0 = (dsdf ws) * softplus'(H0); dws, dbs; db0 = column sums of dpre0
        _lib.call("na_linear_narrow_bwd", _lib.ptr(H0), hid, _code(H0), P, hid, _lib.ptr(w.ws), 1,
                  _lib.ptr(dsdf), _lib.ptr(acc.dws), _lib.ptr(acc.dbs), _lib.ptr(dpre0), hid,
                  _code(dpre0), 0, _lib.ptr(H0), hid, _code(H0), NA_ACT_SOFTPLUS,
                  SOFTPLUS_BETA, _lib.ptr(acc.db0), _s(X0))
    if wgrad:
        linear_weight_grad(dpre0, X0, w.k0p, w.k0p, acc.dW0p)
    return dpre0


def _encoder_grad(dpre0, w):
    """dX (encoder columns) = dpre0 . W0p[:, :n_enc]  ->  [P, n_enc] in the compute dtype
    (the hash-grid scatter reads bf16 or fp32 rows and accumulates the table in fp32)."""
    P = dpre0.shape[0]
    dX = torch.empty(P, w.n_enc, dtype=w.cdt, device=dpre0.device)
    gemm(dpre0, w.hidden, 0, _t(w.W0p), w.hidden, 0, P, w.n_enc, w.hidden, dX, w.n_enc)
    return dX


class SDFNetFunction(torch.autograd.Function):
    """SDF MLP (mlp.py:55-69), optionally fused with the encoder (modules.py:76-95).

    enc given:  inp = world points [P,3]; the encoder builds X0 in-kernel and the
                table gradient is returned for `table_params`.
    enc None:   inp = X0 [P, K0p] already in kernel layout (compute dtype); the
                input gradient dX0 [P, K0p] is returned.
    Outputs: sdf [P,1] float32 and, with_feat, feat [P,256] as a view of a
    [P,280] colour-input buffer (columns 256.. are filled by the colour net).
    """

    @staticmethod
    def forward(ctx, inp, table_params, W0, b0, W1, b1, ws, bs, enc, cdt, with_feat, n_enc):
        _lib.require_cuda(inp)
        w = SDFWeights(W0, b0, W1 if with_feat else None, b1 if with_feat else None, ws, bs,
                       n_enc, cdt)
        dev = inp.device
        if enc is not None:
            points = inp.contiguous().float()
            X0 = sdf_input(enc, points, cdt, w.k0p)
        else:
            points = torch.empty(0, device=dev)
            X0 = inp.contiguous()
            if X0.dtype != cdt or X0.shape[1] != w.k0p:
                raise RuntimeError("prebuilt SDF input must be [P, K0p] in the compute dtype")
        H0, sdf, FB = _sdf_mlp_forward(X0, w, with_feat)
        ctx.save_for_backward(points, X0, H0, FB if FB is not None else torch.empty(0, device=dev))
        ctx.w, ctx.enc, ctx.with_feat = w, enc, with_feat
        if with_feat:
            return sdf, FB[:, :w.hidden]
        return sdf, None

    @staticmethod
    def backward(ctx, dsdf, dfeat):
        points, X0, H0, FB = ctx.saved_tensors
        w, enc = ctx.w, ctx.enc
        P, dev = X0.shape[0], X0.device
        acc = _SDFGrads(w, dev, ctx.with_feat)
        dpre0 = _sdf_mlp_backward(X0, H0, FB if ctx.with_feat else None, w, dsdf, dfeat, acc)
        dW0 = w.unpermute_dW0(acc.dW0p)
        dtable = dinp = None
        if enc is not None and ctx.needs_input_grad[1]:
            dX = _encoder_grad(dpre0, w)
            dtable = enc.table_grad_buffer()
            vmin, vmax = enc.vol_range
            if enc.sorted_grad():
                ws = enc.sorted_workspace(P, 1, dev)
                _lib.call("na_sdf_input_bwd_sorted", ctypes.byref(enc.desc), _lib.ptr(points), P,
                          float(vmin), float(vmax), _lib.ptr(dX), _code(dX), w.n_enc,
                          _lib.ptr(dtable), int(enc.active_levels), _lib.ptr(ws), ws.numel(),
                          _s(X0))
            else:
                _lib.call("na_sdf_input_bwd", ctypes.byref(enc.desc), _lib.ptr(points), P,
                          float(vmin), float(vmax), _lib.ptr(dX), _code(dX), w.n_enc,
                          _lib.ptr(dtable), int(enc.active_levels), _s(X0))
        if enc is None and ctx.needs_input_grad[0]:
            dinp = torch.empty(P, w.k0p, dtype=torch.float32, device=dev)
            gemm(dpre0, w.hidden, 0, _t(w.W0p), w.hidden, 0, P, w.k0p, w.hidden, dinp, w.k0p)
        return (dinp, dtable, dW0, acc.db0, acc.dW1, acc.db1, acc.dws.view(w.ws_shape),
                acc.dbs.view(w.bs_shape), None, None, None, None)


class SDFCentreTapsFunction(torch.autograd.Function):
    """The SDF field at the sample points and at their four tetrahedral taps
    (modules.py:84-95 for the centre, modules.py:157-175 for the taps) as one node.

    Forward: sdf [P,1], feat [P,256] (view of the colour-input buffer) and
    s_taps [4,P].  Backward runs both MLP backward passes into shared weight
    accumulators and scatters the five encoder gradients of each point with one
    merged kernel (na_sdf_taps_bwd)."""

    @staticmethod
    def forward(ctx, points, table_params, W0, b0, W1, b1, ws, bs, enc, cdt, n_enc, eps):
        _lib.require_cuda(points)
        w = SDFWeights(W0, b0, W1, b1, ws, bs, n_enc, cdt)
        pts = points.contiguous().float()
        P = pts.shape[0]
        X0c = sdf_input(enc, pts, cdt, w.k0p)
        H0c, sdf, FB = _sdf_mlp_forward(X0c, w, True)
        tp = taps_points(pts, eps).view(-1, 3)
        X0t = sdf_input(enc, tp, cdt, w.k0p)
        H0t, s_t, _ = _sdf_mlp_forward(X0t, w, False)
        del tp
        ctx.save_for_backward(pts, X0c, H0c, FB, X0t, H0t)
        ctx.w, ctx.enc, ctx.eps = w, enc, float(eps)
        return sdf, FB[:, :w.hidden], s_t.view(4, P)

    @staticmethod
    def backward(ctx, dsdf, dfeat, ds_taps):
        pts, X0c, H0c, FB, X0t, H0t = ctx.saved_tensors
        return _centre_taps_backward(ctx, pts, X0c, H0c, FB, X0t, H0t, dsdf, dfeat,
                                     ds_taps) + (None,)


def taps_increment_width(k0p, n_enc):
    """Row width of the tap-increment buffer D: the centre layout's k0p columns (hi
    part and point offset) plus the lo part of the encoder columns, padded to 5, 8 or
    9 stages of 32 columns (the weights-stationary kernel's K range)."""
    w = (k0p + n_enc + 31) // 32 * 32
    for stages in (5, 8, 9):
        if w <= 32 * stages:
            return 32 * stages
    raise NotImplementedError(f"increment input of {k0p + n_enc} columns exceeds 288")


class SDFCentreTapsDeltaFunction(torch.autograd.Function):
    """BF16 form of SDFCentreTapsFunction with the tap SDFs as increments.

    Forward: sdf [P,1], feat [P,256] and ds [4,P] = sdf(x + eps k_i) - sdf(x).  The
    encoder writes each tap input's difference from the centre input,
    D = X(x + eps k_i) - X(x), in fp32 and stores it as two bf16 parts (rows
    point-major, 4 per point); the tap product multiplies both parts with the
    first-layer weights, so dz keeps about 16 mantissa bits relative to its own
    size, and its epilogue forms softplus(z_c + dz) - softplus(z_c) from the
    centre's fp32 pre-activation z_c.  The four increments' second-order sum (the
    Hessian trace) thus keeps float32-like precision.
    Backward in the same increment form (na_taps_increment_bwd reads the fp32 dz_i the
    forward kept, or recomputes them): with dds_i = a + b_i (a the taps' mean, the Hessian-trace coefficient),
    the tap rows carry ws (b_i sp'(z_c + dz_i) + a dsp'_i) paired with D_i, a shared
    row a ws sp'(z_c) pairs with sum_i D_i, and the centre row receives its own
    gradient plus sum_i dds_i ws dsp'_i (dsp'_i = sp'(z_c + dz_i) - sp'(z_c)); the
    encoder scatter gives the taps' corners g_i + g_shared and the centre's
    g_c - sum_i (g_i + g_shared), in fp32.  Every bf16 row holds a term of its own
    natural size: no difference of nearly equal terms is ever rounded to bf16."""

    @staticmethod
    def forward(ctx, points, table_params, W0, b0, W1, b1, ws, bs, enc, n_enc, eps):
        _lib.require_cuda(points)
        cdt = torch.bfloat16
        w = SDFWeights(W0, b0, W1, b1, ws, bs, n_enc, cdt)
        if w.hidden not in (128, 256):
            raise NotImplementedError("the tap-increment product needs a hidden width of 128 or 256")
        pts = points.contiguous().float()
        P, dev, hid = pts.shape[0], pts.device, w.hidden
        P32 = (P + 31) // 32 * 32  # 4 * P32 rows: whole 128-row tiles
        ld_d = taps_increment_width(w.k0p, n_enc)
        X0c = torch.empty(P, w.k0p, dtype=cdt, device=dev)
        D = torch.empty(4 * P32, ld_d, dtype=cdt, device=dev)
        if P32 > P:
            D[4 * P:].zero_()
        Dsum = torch.empty(P, w.k0p, dtype=cdt, device=dev)
        table = enc.table_for_forward()
        vmin, vmax = enc.vol_range
        _lib.call("na_sdf_taps_increment_input", ctypes.byref(enc.desc), _lib.ptr(pts), P,
                  float(vmin), float(vmax), float(eps), _lib.ptr(table), _code(table),
                  int(enc.active_levels), _lib.ptr(X0c), w.k0p, _lib.ptr(D), ld_d, _lib.ptr(Dsum),
                  _s(pts))
        nparts = head_parts(hid)
        H0c = torch.empty(P, hid, dtype=cdt, device=dev)
        Zc = torch.empty(P32, hid, dtype=torch.float32, device=dev)
        if P32 > P:
            Zc[P:].zero_()
        parts_c = torch.empty(nparts, P, device=dev)
        gemm(X0c, w.k0p, 0, w.W0p, w.k0p, 0, P, hid, w.k0p, H0c, hid, bias=w.b0,
             act=NA_ACT_SOFTPLUS, z_out=Zc, head_w=w.ws.view(-1), head_out=parts_c)
        sdf = head_sum(parts_c, w.bs).view(P, 1)
        FB = torch.empty(P, RGB_LD, dtype=cdt, device=dev)
        gemm(H0c, hid, 0, w.W1, hid, 0, P, hid, hid, FB, RGB_LD, bias=w.b1, act=NA_ACT_SOFTPLUS)
        W0d = w.increment_weights(ld_d)
        parts_t = torch.empty(hid // 32, 4 * P32, device=dev)
        # the split backward reads the increments dz = D W0d^T: kept from here when a
        # backward will run, instead of recomputed there
        keep = any(ctx.needs_input_grad) and _lib.call(
            "na_tune_get", b"taps_increment_bwd.fused") == 0
        dz = torch.empty(4 * P32, hid, dtype=torch.float32, device=dev) if keep else None
        _lib.call("na_taps_increment_fwd", _lib.ptr(D), ld_d, 4 * P32, ld_d, _lib.ptr(W0d), ld_d,
                  hid, _lib.ptr(Zc), hid, _lib.ptr(w.ws), float(SOFTPLUS_BETA), _lib.ptr(parts_t),
                  _lib.ptr(dz), _s(pts))
        ds = head_sum(parts_t)[:4 * P].view(P, 4).t().contiguous()
        ctx.save_for_backward(pts, X0c, H0c, FB, D, Dsum, Zc, W0d)
        ctx.dz = dz
        ctx.w, ctx.enc, ctx.eps, ctx.ld_d = w, enc, float(eps), ld_d
        return sdf, FB[:, :hid], ds

    @staticmethod
    def backward(ctx, dsdf, dfeat, dds):
        pts, X0c, H0c, FB, D, Dsum, Zc, W0d = ctx.saved_tensors
        w, enc, ld_d = ctx.w, ctx.enc, ctx.ld_d
        P, dev, hid = pts.shape[0], pts.device, w.hidden
        P32 = Zc.shape[0]
        acc = _SDFGrads(w, dev, True)
        # the centre's own pre-activation gradient and the taps' shared rows, in
        # P32-row buffers (the tap product covers whole 128-row tiles)
        dpre_c = torch.empty(P32, hid, dtype=w.cdt, device=dev)
        q = torch.empty(P32, hid, dtype=w.cdt, device=dev)
        if P32 > P:
            dpre_c[P:].zero_()
        _sdf_mlp_backward(X0c, H0c, FB, w, dsdf, dfeat, acc, dpre0=dpre_c[:P], wgrad=False)
        dpre_t = torch.empty(4 * P32, hid, dtype=w.cdt, device=dev)
        dds_pm = torch.zeros(P32, 4, device=dev)
        if dds is not None:
            dds_pm[:P] = dds.t()
        # the tap increments dz = D W0d^T: kept by the forward, or recomputed in the
        # fused form's epilogue, or (split form without a kept copy) recomputed into a
        # transient fp32 buffer
        fused = _lib.call("na_tune_get", b"taps_increment_bwd.fused") != 0
        dz, held = ctx.dz, 1
        ctx.dz = None
        if dz is None and not fused:
            dz, held = torch.empty(4 * P32, hid, dtype=torch.float32, device=dev), 0
        _lib.call("na_taps_increment_bwd", _lib.ptr(D), ld_d, 4 * P32, ld_d, _lib.ptr(W0d), ld_d,
                  hid, _lib.ptr(Zc), hid, _lib.ptr(w.ws), float(SOFTPLUS_BETA), _lib.ptr(dds_pm),
                  _lib.ptr(dpre_t), hid, _lib.ptr(dpre_c), hid, _lib.ptr(q), hid,
                  _lib.ptr(acc.db0), _lib.ptr(acc.dws), _lib.ptr(dz) if dz is not None else None,
                  held if dz is not None else 0, _s(pts))
        del dz
        # first-layer weights: centre rows x X0c, tap rows x D (hi part and point
        # offset), shared rows x sum_t D_t
        linear_weight_grad(dpre_c[:P], X0c, w.k0p, w.k0p, acc.dW0p)
        linear_weight_grad(dpre_t, D, ld_d, w.k0p, acc.dW0p)
        linear_weight_grad(q[:P], Dsum, w.k0p, w.k0p, acc.dW0p)
        dtable = None
        if ctx.needs_input_grad[1]:
            dXc = _encoder_grad(dpre_c[:P], w)
            dXt = _encoder_grad(dpre_t, w)
            dXq = _encoder_grad(q[:P], w)
            dtable = enc.table_grad_buffer()
            _taps_scatter(enc, pts, ctx.eps, dXc, dXt, dXq, w.n_enc, dtable, 1)
        dW0 = w.unpermute_dW0(acc.dW0p)
        return (None, dtable, dW0, acc.db0, acc.dW1, acc.db1, acc.dws.view(w.ws_shape),
                acc.dbs.view(w.bs_shape), None, None, None)


def _taps_scatter(enc, pts, eps, dXc, dXt, dXq, ld, dtable, increments):
    """Table gradient of one centre + four-tap point set (na_sdf_taps_bwd), or its
    sorted, bitwise-repeatable variant when the encoder asks for it."""
    P = pts.shape[0]
    vmin, vmax = enc.vol_range
    args = (ctypes.byref(enc.desc), _lib.ptr(pts), P, float(vmin), float(vmax), eps,
            _lib.ptr(dXc), _lib.ptr(dXt), _lib.ptr(dXq) if dXq is not None else None, _code(dXc),
            ld, _lib.ptr(dtable), int(enc.active_levels), increments)
    if SCATTER_CAPTURE is not None:
        SCATTER_CAPTURE.append((pts, eps, dXc, dXt, dXq, ld, increments))
    if enc.sorted_grad():
        ws = enc.sorted_workspace(P, 5, pts.device)
        _lib.call("na_sdf_taps_bwd_sorted", *args, _lib.ptr(ws), ws.numel(), _s(pts))
        return
    with _timed("na_sdf_taps_bwd"):
        _lib.call("na_sdf_taps_bwd", *args, _s(pts))


def _centre_taps_backward(ctx, pts, X0c, H0c, FB, X0t, H0t, dsdf, dfeat, ds_taps):
    """Shared backward of the centre + tap SDF nodes (gradient of sdf(x) and of the
    four sdf(x + eps k_i)); returns the gradients of (points, table, W0, b0, W1, b1,
    ws, bs, enc, cdt/n_enc, n_enc/eps)."""
    w, enc = ctx.w, ctx.enc
    P, dev = pts.shape[0], pts.device
    acc = _SDFGrads(w, dev, True)
    dpre_c = _sdf_mlp_backward(X0c, H0c, FB, w, dsdf, dfeat, acc)
    dpre_t = _sdf_mlp_backward(X0t, H0t, None, w,
                               ds_taps.reshape(4 * P, 1) if ds_taps is not None else None,
                               None, acc)
    dtable = None
    if ctx.needs_input_grad[1]:
        dXc = _encoder_grad(dpre_c, w)
        dXt = _encoder_grad(dpre_t, w)
        dtable = enc.table_grad_buffer()
        _taps_scatter(enc, pts, ctx.eps, dXc, dXt, None, w.n_enc, dtable, 0)
    dW0 = w.unpermute_dW0(acc.dW0p)
    return (None, dtable, dW0, acc.db0, acc.dW1, acc.db1, acc.dws.view(w.ws_shape),
            acc.dbs.view(w.bs_shape), None, None, None)


class TapsDelta(torch.autograd.Function):
    """ds [4,n] increments, sc [n] centre SDF, sdf [n] reference input -> gradient
    [n,3] and hessian [n,3] (or None), as TapsCombine (modules.py:163-175)."""

    @staticmethod
    def forward(ctx, ds, sc, sdf, eps, eps_sq, training):
        ds = ds.contiguous().float()
        n = ds.shape[1]
        dev = ds.device
        grad = torch.empty(n, 3, device=dev)
        hess = torch.empty(n, 3, device=dev) if training else None
        sc_ = sc.reshape(-1).contiguous().float() if training else None
        sdf_ = sdf.reshape(-1).contiguous().float() if training else None
        _lib.call("na_taps_delta_fwd", _lib.ptr(ds), _lib.ptr(sc_), _lib.ptr(sdf_), n, float(eps),
                  float(eps_sq), _lib.ptr(grad), _lib.ptr(hess), _s(ds))
        ctx.n, ctx.eps, ctx.eps_sq, ctx.training = n, eps, eps_sq, training
        if hess is None:
            return grad, torch.empty(0, device=dev)
        return grad, hess

    @staticmethod
    def backward(ctx, dgrad, dhess):
        n = ctx.n
        dev = dgrad.device if dgrad is not None else dhess.device
        dds = torch.empty(4, n, device=dev)
        dsc = torch.empty(n, device=dev) if ctx.training else None
        dsdf = torch.empty(n, device=dev) if ctx.training else None
        dg = dgrad.contiguous().float() if dgrad is not None else None
        dh = dhess.contiguous().float() if (dhess is not None and ctx.training) else None
        _lib.call("na_taps_delta_bwd", n, float(ctx.eps), float(ctx.eps_sq), _lib.ptr(dg),
                  _lib.ptr(dh), _lib.ptr(dds), _lib.ptr(dsc), _lib.ptr(dsdf), _lib.stream_of(dds))
        return dds, dsc, dsdf, None, None, None


# ---------------------------------------------------------------------------
# colour network: input builder + MLPwithSkipConnection + sigmoid
# ---------------------------------------------------------------------------
class RGBWeights:
    """Colour MLP weights in kernel layout.  Reference input columns are
    [x(3), SH(16), normal(3), feat(256)]; kernel columns [feat, x, SH, normal, 0, 0]."""

    def __init__(self, Ws, bs, cdt, n_feat=256):
        self.cdt = cdt
        self.n_feat = n_feat
        W0 = Ws[0]
        W0p = torch.zeros(W0.shape[0], RGB_LD, dtype=torch.float32, device=W0.device)
        W0p[:, :n_feat] = W0[:, 22:]
        W0p[:, n_feat:n_feat + 22] = W0[:, :22]
        self.W = [W0p.to(cdt).contiguous()] + [W.to(cdt).contiguous() for W in Ws[1:-1]]
        self.W_last = Ws[-1].contiguous().float()
        self.b = [b.contiguous().float() for b in bs]
        self.hidden = W0.shape[0]

    def increment_weights(self, ld_d):
        """First-layer weights against the tap-increment rows [hi part and point | lo
        part of the encoder columns]: [W0p | W_enc | 0] of width ld_d (bf16)."""
        key = (ld_d, self.W0p.data_ptr(), self.W0p._version)
        if getattr(self, "_inc_key", None) != key:
            n, k0p = self.n_enc, self.k0p
            W = torch.zeros(self.W0p.shape[0], ld_d, dtype=self.W0p.dtype, device=self.W0p.device)
            W[:, :k0p] = self.W0p
            W[:, k0p:k0p + n] = self.W0p[:, :n]
            self._inc, self._inc_key = W, key
        return self._inc

    def unpermute_dW0(self, dW0p):
        n = self.n_feat
        return torch.cat([dW0p[:, n:n + 22], dW0p[:, :n]], dim=1)


class RGBNetFunction(torch.autograd.Function):
    """Colour MLP (nerf_util.py:189-196) with the NeuralRGB input (modules.py:209-220).

    build_input 1 or 2: feat [P,256] (ideally the SDF function's view of its
      [P,280] buffer), points [P,3], ray_unit [R,3] (n_per_ray samples per ray)
      and normals_src [P,3] -> the input row [feat | x | SH16 | normal | 0 0] is
      assembled on the GPU; 1: normals_src are raw gradients, normalised in the
      kernel (F.normalize), 2: normals_src are used as given.
    build_input 0: feat is the whole prebuilt input [P,280] in kernel layout.
    final_act: NA_ACT_SIGMOID (NeuralRGB) or NA_ACT_NONE (bare MLP).
    """

    @staticmethod
    def forward(ctx, feat, points, ray_unit, normals_src, n_per_ray, build_input, final_act, cdt,
                *wb):
        n_lin = len(wb) // 2
        Ws, bs = list(wb[:n_lin]), list(wb[n_lin:])
        w = RGBWeights(Ws, bs, cdt)
        P = feat.shape[0]
        dev = feat.device
        hid = w.hidden
        if build_input:
            if feat.stride(0) == RGB_LD and feat.dtype == cdt and feat.stride(1) == 1:
                FB = torch.as_strided(feat, (P, RGB_LD), (RGB_LD, 1))
            else:
                FB = torch.empty(P, RGB_LD, dtype=cdt, device=dev)
                FB[:, :w.n_feat] = feat
            points = points.contiguous().float()
            ray_unit = ray_unit.contiguous().float()
            normals_src = normals_src.contiguous().float()
            _lib.call("na_rgb_input_fwd", _lib.ptr(points), _lib.ptr(ray_unit), int(n_per_ray),
                      _lib.ptr(normals_src), int(build_input == 1), P, w.n_feat, _lib.ptr(FB),
                      _code(FB), RGB_LD, _s(feat))
        else:
            FB = feat.contiguous()
            if FB.shape[1] != RGB_LD or FB.dtype != cdt:
                raise RuntimeError("prebuilt colour input must be [P, 280] in the compute dtype")
            normals_src = torch.empty(0, device=dev)
        Hs = []
        X, ldx, K = FB, RGB_LD, RGB_LD
        for li in range(n_lin - 1):
            H = torch.empty(P, hid, dtype=cdt, device=dev)
            gemm(X, ldx, 0, w.W[li], K, 0, P, hid, K, H, hid, bias=w.b[li], act=NA_ACT_RELU)
            Hs.append(H)
            X, ldx, K = H, hid, hid
        out = torch.empty(P, 3, dtype=torch.float32, device=dev)
        _lib.call("na_linear_narrow_fwd", _lib.ptr(X), hid, _code(X), P, hid, _lib.ptr(w.W_last),
                  _lib.ptr(w.b[-1]), 3, int(final_act), _lib.ptr(out), _s(feat))
        ctx.save_for_backward(FB, normals_src, out, *Hs)
        ctx.w, ctx.n_lin, ctx.build, ctx.final_act = w, n_lin, build_input, final_act
        return out

    @staticmethod
    def backward(ctx, dout):
        FB, normals_src, out, *Hs = ctx.saved_tensors
        w, n_lin = ctx.w, ctx.n_lin
        P, hid, cdt, dev = FB.shape[0], w.hidden, w.cdt, FB.device
        dpre3 = dout.float()
        if ctx.final_act == NA_ACT_SIGMOID:
            dpre3 = dpre3 * out * (1 - out)
        dpre3 = dpre3.contiguous()
        dWs = [None] * n_lin
        # every weight/bias gradient of the colour MLP in one zeroed block
        z = _zero_block(dev, *[(b.shape[0],) for b in w.b], (3, hid),
                        *[(hid, hid) for _ in range(n_lin - 2)], (hid, RGB_LD))
        dbs = z[:n_lin]
        dWs[-1] = z[n_lin]
        dW_mid, dW0p = z[n_lin + 1:-1], z[-1]
        # last hidden layer's pre-activation gradient, written in the compute dtype
        dpre = torch.empty(P, hid, dtype=cdt, device=dev)
        _lib.call("na_linear_narrow_bwd", _lib.ptr(Hs[-1]), hid, _code(Hs[-1]), P, hid,
                  _lib.ptr(w.W_last), 3, _lib.ptr(dpre3), _lib.ptr(dWs[-1]), _lib.ptr(dbs[-1]),
                  _lib.ptr(dpre), hid, _code(dpre), 0, _lib.ptr(Hs[-1]), hid, _code(Hs[-1]),
                  NA_ACT_RELU, SOFTPLUS_BETA, _lib.ptr(dbs[-2]), _s(FB))
        for li in range(n_lin - 2, 0, -1):
            X = Hs[li - 1]
            dWs[li] = dW_mid[li - 1]
            linear_weight_grad(dpre, X, hid, hid, dWs[li])
            dprev = torch.empty(P, hid, dtype=cdt, device=dev)
            gemm(dpre, hid, 0, _t(w.W[li]), hid, 0, P, hid, hid, dprev, hid, dact=NA_ACT_RELU,
                 dact_y=X, ld_dact=hid, colsum=dbs[li - 1])
            dpre = dprev
        linear_weight_grad(dpre, FB, RGB_LD, RGB_LD, dW0p)
        dWs[0] = w.unpermute_dW0(dW0p)
        # input gradient in the compute dtype: the feature columns go back to the SDF
        # node as a view, in the dtype of its output, with no cast or copy
        dFB = torch.empty(P, RGB_LD, dtype=cdt, device=dev)
        gemm(dpre, hid, 0, _t(w.W[0]), hid, 0, P, RGB_LD, hid, dFB, RGB_LD)
        dnormals = None
        if ctx.build:
            dnormals = torch.empty(P, 3, dtype=torch.float32, device=dev)
            _lib.call("na_rgb_input_bwd", _lib.ptr(normals_src), int(ctx.build == 1), P, w.n_feat,
                      _lib.ptr(dFB), _code(dFB), RGB_LD, _lib.ptr(dnormals), 0, _s(FB))
            dfeat = dFB[:, :w.n_feat]
        else:
            dfeat = dFB
        return (dfeat, None, None, dnormals, None, None, None, None, *dWs, *dbs)


# ---------------------------------------------------------------------------
# 4-tap numerical gradient
# ---------------------------------------------------------------------------
def taps_points(x, eps):
    x = x.contiguous().float()
    n = x.shape[0]
    out = torch.empty(4, n, 3, dtype=torch.float32, device=x.device)
    _lib.call("na_taps_points", _lib.ptr(x), n, float(eps), _lib.ptr(out), _s(x))
    return out


class TapsCombine(torch.autograd.Function):
    """s_taps [4, n], sdf [n] -> gradient [n,3], hessian [n,3] (or None)."""

    @staticmethod
    def forward(ctx, s_taps, sdf, eps, eps_sq, training):
        s_taps = s_taps.contiguous().float()
        n = s_taps.shape[1]
        grad = torch.empty(n, 3, device=s_taps.device)
        hess = torch.empty(n, 3, device=s_taps.device) if training else None
        sdf_c = sdf.contiguous().float() if training else None
        _lib.call("na_taps_combine_fwd", _lib.ptr(s_taps), _lib.ptr(sdf_c), n, float(eps),
                  float(eps_sq), _lib.ptr(grad), _lib.ptr(hess), _s(s_taps))
        ctx.n, ctx.eps, ctx.eps_sq, ctx.training = n, eps, eps_sq, training
        if hess is None:
            return grad, torch.empty(0, device=s_taps.device)
        return grad, hess

    @staticmethod
    def backward(ctx, dgrad, dhess):
        n = ctx.n
        dev = dgrad.device if dgrad is not None else dhess.device
        ds = torch.empty(4, n, device=dev)
        dsdf = torch.empty(n, device=dev) if ctx.training else None
        dg = dgrad.contiguous().float() if dgrad is not None else None
        dh = dhess.contiguous().float() if (dhess is not None and ctx.training) else None
        _lib.call("na_taps_combine_bwd", n, float(ctx.eps), float(ctx.eps_sq), _lib.ptr(dg),
                  _lib.ptr(dh), _lib.ptr(ds), _lib.ptr(dsdf), 0, _lib.stream_of(ds))
        return ds, dsdf, None, None, None


# ---------------------------------------------------------------------------
# NeuS alpha
# ---------------------------------------------------------------------------
class NeusAlpha(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ray_unit, sdfs, gradients, dists, far, outside, s_var, progress, anneal_end,
                outside_val):
        R, N = sdfs.shape
        ray_unit = ray_unit.contiguous().float()
        sdfs = sdfs.contiguous().float()
        gradients = gradients.contiguous().float()
        dists = dists.contiguous().float()
        far = far.contiguous().float() if far is not None else None
        out_u8 = outside.contiguous().to(torch.uint8) if outside is not None else None
        alpha = torch.empty(R, N, device=sdfs.device)
        _lib.call("na_neus_alpha_fwd", _lib.ptr(ray_unit), _lib.ptr(sdfs), _lib.ptr(gradients),
                  _lib.ptr(dists), _lib.ptr(far), _lib.ptr(out_u8), float(outside_val),
                  _lib.ptr(s_var), float(progress), float(anneal_end), R, N, _lib.ptr(alpha),
                  _s(sdfs))
        ctx.save_for_backward(ray_unit, sdfs, gradients, dists,
                              far if far is not None else torch.empty(0, device=sdfs.device),
                              out_u8 if out_u8 is not None else torch.empty(0, device=sdfs.device),
                              s_var)
        ctx.has_far, ctx.has_out = far is not None, out_u8 is not None
        ctx.progress, ctx.anneal_end, ctx.outside_val = progress, anneal_end, outside_val
        return alpha

    @staticmethod
    def backward(ctx, dalpha):
        ray_unit, sdfs, gradients, dists, far, out_u8, s_var = ctx.saved_tensors
        R, N = sdfs.shape
        dsdf = torch.empty_like(sdfs)
        dgrad = torch.empty_like(gradients)
        ds_var = torch.zeros(1, device=sdfs.device)
        _lib.call("na_neus_alpha_bwd", _lib.ptr(ray_unit), _lib.ptr(sdfs), _lib.ptr(gradients),
                  _lib.ptr(dists), _lib.ptr(far) if ctx.has_far else None,
                  _lib.ptr(out_u8) if ctx.has_out else None, float(ctx.outside_val),
                  _lib.ptr(s_var), float(ctx.progress), float(ctx.anneal_end), R, N,
                  _lib.ptr(dalpha.contiguous().float()), _lib.ptr(dsdf), _lib.ptr(dgrad),
                  _lib.ptr(ds_var), _s(sdfs))
        return (None, dsdf, dgrad, None, None, None, ds_var.view_as(s_var), None, None, None)


# ---------------------------------------------------------------------------
# alpha compositing + composite
# ---------------------------------------------------------------------------
class Composite(torch.autograd.Function):
    """alphas [R,N], q [R,N,K] (or None) -> weights [R,N], out [R,K]."""

    @staticmethod
    def forward(ctx, alphas, q):
        R, N = alphas.shape
        alphas = alphas.contiguous().float()
        K = 0 if q is None else q.shape[-1]
        qc = q.contiguous().float() if q is not None else None
        weights = torch.empty(R, N, device=alphas.device)
        out = torch.empty(R, K, device=alphas.device)
        _lib.call("na_composite_fwd", _lib.ptr(alphas), _lib.ptr(qc), R, N, K, _lib.ptr(weights),
                  _lib.ptr(out) if K else None, _s(alphas))
        ctx.save_for_backward(alphas, qc if qc is not None else torch.empty(0, device=alphas.device),
                              weights)
        ctx.K = K
        return weights, out

    @staticmethod
    def backward(ctx, dweights, dout):
        alphas, qc, weights = ctx.saved_tensors
        R, N = alphas.shape
        K = ctx.K
        dalpha = torch.empty_like(alphas)
        dq = torch.empty(R, N, K, device=alphas.device) if (K and ctx.needs_input_grad[1]) else None
        _lib.call("na_composite_bwd", _lib.ptr(alphas), _lib.ptr(qc) if K else None,
                  _lib.ptr(weights), R, N, K,
                  _lib.ptr(dout.contiguous().float()) if (dout is not None and K) else None,
                  _lib.ptr(dweights.contiguous().float()) if dweights is not None else None,
                  _lib.ptr(dalpha), _lib.ptr(dq), _s(alphas))
        return dalpha, dq


class WeightedSum(torch.autograd.Function):
    """quantities [..., N, K] x weights [..., N, 1] -> [..., K]  (render.composite)."""

    @staticmethod
    def forward(ctx, q, w):
        N, K = q.shape[-2], q.shape[-1]
        lead = q.shape[:-2]
        qc = q.reshape(-1, N, K).contiguous().float()
        wc = w.reshape(-1, N).contiguous().float()
        R = qc.shape[0]
        out = torch.empty(R, K, device=qc.device)
        _lib.call("na_weighted_sum_fwd", _lib.ptr(qc), _lib.ptr(wc), R, N, K, _lib.ptr(out), _s(qc))
        ctx.save_for_backward(qc, wc)
        ctx.shapes = (q.shape, w.shape)
        return out.view(*lead, K)

    @staticmethod
    def backward(ctx, dout):
        qc, wc = ctx.saved_tensors
        R, N, K = qc.shape
        dq = torch.empty_like(qc) if ctx.needs_input_grad[0] else None
        dw = torch.empty_like(wc) if ctx.needs_input_grad[1] else None
        _lib.call("na_weighted_sum_bwd", _lib.ptr(qc), _lib.ptr(wc), R, N, K,
                  _lib.ptr(dout.reshape(R, K).contiguous().float()), _lib.ptr(dq), _lib.ptr(dw),
                  _s(qc))
        return (dq.view(ctx.shapes[0]) if dq is not None else None,
                dw.view(ctx.shapes[1]) if dw is not None else None)


# ---------------------------------------------------------------------------
# generic dense layer and the background NeRF pieces
# ---------------------------------------------------------------------------
class LinearAct(torch.autograd.Function):
    """Y = act(X W^T + b) on the GEMM kernels (bias and activation in the epilogue).
    X [P, K] compute dtype (K a multiple of 8), W [N, K] float32, b [N]; Y [P, N] in
    out_dtype.  The output width is padded to a multiple of 8 inside (zero rows of W),
    so every operand of the backward products has aligned rows.  Backward: act'(Y)
    with the bias gradient in one pass, the weight gradient on the row-reduction
    kernel, the input gradient on the NN GEMM."""

    @staticmethod
    def forward(ctx, X, W, b, act, out_dtype):
        P, K = X.shape
        N = W.shape[0]
        N8 = _pad8(N)
        Wc = W.to(X.dtype)
        bc = b.float()
        if N8 != N:
            Wc = torch_F.pad(Wc, (0, 0, 0, N8 - N))
            bc = torch_F.pad(bc, (0, N8 - N))
        Wc, bc = Wc.contiguous(), bc.contiguous()
        Y = torch.empty(P, N8, dtype=out_dtype, device=X.device)
        gemm(X, K, 0, Wc, K, 0, P, N8, K, Y, N8, bias=bc, act=act)
        ctx.save_for_backward(X, Wc, Y)
        ctx.act, ctx.N = act, N
        return Y[:, :N] if N8 != N else Y

    @staticmethod
    def backward(ctx, dY):
        X, Wc, Y = ctx.saved_tensors
        P, K = X.shape
        N, N8 = ctx.N, Wc.shape[0]
        if N8 != N:
            d8 = torch.zeros(P, N8, dtype=dY.dtype, device=dY.device)
            d8[:, :N] = dY
            dY = d8
        db = torch.zeros(N8, device=X.device)
        if ctx.act == NA_ACT_NONE:
            dpre = dY.to(X.dtype).contiguous()
            db += dY.float().sum(0)
        else:
            dpre = act_grad(dY.contiguous(), Y, ctx.act, X.dtype, colsum=db, n_cols=N8)
        dW = torch.zeros(N8, K, device=X.device)
        linear_weight_grad(dpre, X, K, K, dW)
        dX = None
        if ctx.needs_input_grad[0]:
            dX = torch.empty(P, K, dtype=X.dtype, device=X.device)
            gemm(dpre, N8, 0, _t(Wc), N8, 0, P, K, N8, dX, K)
        return dX, dW[:N], db[:N], None, None


def bg_encode(points, levels, cdt, width):
    """Background MLP input rows [P, width] (compute dtype): [q | Fourier(q) | 0]."""
    P = points.shape[0]
    out = torch.empty(P, width, dtype=cdt, device=points.device)
    _lib.call("na_bg_encode", _lib.ptr(points.contiguous().float()), P, int(levels), _lib.ptr(out),
              _code(out), width, _s(points))
    return out


class BgAlpha(torch.autograd.Function):
    """MLP output rows [R*N, ld] (column 0 = raw density), dists [R, N] -> alphas [R, N]
    and densities [R, N] (softplus of column 0)."""

    @staticmethod
    def forward(ctx, out, dists, R, N):
        o = out.float()
        dists = dists.contiguous().float()
        alpha = torch.empty(R, N, device=out.device)
        density = torch.empty(R, N, device=out.device)
        _lib.call("na_bg_alpha_fwd", _lib.ptr(o), o.stride(0), _lib.ptr(dists), R, N,
                  _lib.ptr(density), _lib.ptr(alpha), _s(o))
        ctx.save_for_backward(o, dists)
        ctx.RN = (R, N)
        ctx.shape, ctx.dtype = out.shape, out.dtype
        return alpha, density

    @staticmethod
    def backward(ctx, dalpha, ddensity):
        o, dists = ctx.saved_tensors
        R, N = ctx.RN
        dout = torch.zeros(ctx.shape, device=o.device)
        _lib.call("na_bg_alpha_bwd", _lib.ptr(o), o.stride(0), _lib.ptr(dists), R, N,
                  _lib.ptr(dalpha.contiguous().float()) if dalpha is not None else None,
                  _lib.ptr(ddensity.contiguous().float()) if ddensity is not None else None,
                  _lib.ptr(dout), dout.stride(0), _s(o))
        return dout.to(ctx.dtype), None, None, None


# ---------------------------------------------------------------------------
# packed (ragged) rays: ray r owns samples offsets[r] .. offsets[r+1]-1
# ---------------------------------------------------------------------------
def check_offsets(offsets, S):
    """Validate packed-ray CSR offsets against the sample count S: int32 [R+1] on the
    GPU, offsets[0] == 0, offsets[-1] == S, non-decreasing.  The kernels index samples
    and rays through these values, so a violation raises here (one device-to-host
    read of three flags)."""
    if offsets.dtype != torch.int32 or offsets.dim() != 1 or offsets.shape[0] < 1:
        raise RuntimeError("packed offsets must be a 1-D int32 tensor [R+1]")
    ok = torch.stack([offsets[0] == 0, offsets[-1] == S,
                      (offsets[1:] >= offsets[:-1]).all()]).cpu()
    if not bool(ok.all()):
        raise RuntimeError(f"packed offsets must start at 0, end at the sample count {S} and "
                           f"not decrease (first/last/monotone checks: {ok.tolist()})")
    _lib.require_cuda(offsets)


def packed_ray_index(offsets, S, validate=True):
    """offsets int32 [R+1] -> int32 [S], the ray of every packed sample.  validate:
    check the offsets against S first (check_offsets)."""
    if validate:
        check_offsets(offsets, S)
    R = offsets.shape[0] - 1
    ray_idx = torch.empty(S, dtype=torch.int32, device=offsets.device)
    _lib.call("na_packed_ray_index", _lib.ptr(offsets), R, _lib.ptr(ray_idx), _s(offsets))
    return ray_idx


def _offsets(offsets):
    if offsets.dtype != torch.int32:
        raise RuntimeError("packed offsets must be int32")
    return offsets.contiguous()


class CompositePacked(torch.autograd.Function):
    """alphas [S], q [S,K] (or None), offsets int32 [R+1] -> weights [S], out [R,K]."""

    @staticmethod
    def forward(ctx, alphas, q, offsets):
        off = _offsets(offsets)
        R = off.shape[0] - 1
        alphas = alphas.contiguous().float()
        K = 0 if q is None else q.shape[-1]
        qc = q.contiguous().float() if q is not None else None
        weights = torch.empty_like(alphas)
        out = torch.empty(R, K, device=alphas.device)
        _lib.call("na_composite_packed_fwd", _lib.ptr(alphas), _lib.ptr(qc), _lib.ptr(off), R, K,
                  _lib.ptr(weights), _lib.ptr(out) if K else None, _s(alphas))
        ctx.save_for_backward(alphas, qc if qc is not None else torch.empty(0, device=alphas.device),
                              weights, off)
        ctx.K = K
        return weights, out

    @staticmethod
    def backward(ctx, dweights, dout):
        alphas, qc, weights, off = ctx.saved_tensors
        R, K = off.shape[0] - 1, ctx.K
        dalpha = torch.empty_like(alphas)
        dq = torch.empty(alphas.shape[0], K, device=alphas.device) \
            if (K and ctx.needs_input_grad[1]) else None
        _lib.call("na_composite_packed_bwd", _lib.ptr(alphas), _lib.ptr(qc) if K else None,
                  _lib.ptr(weights), _lib.ptr(off), R, K,
                  _lib.ptr(dout.contiguous().float()) if (dout is not None and K) else None,
                  _lib.ptr(dweights.contiguous().float()) if dweights is not None else None,
                  _lib.ptr(dalpha), _lib.ptr(dq), _s(alphas))
        return dalpha, dq, None


class WeightedSumPacked(torch.autograd.Function):
    """quantities [S,K] x weights [S] per packed ray -> [R,K]."""

    @staticmethod
    def forward(ctx, q, w, offsets, ray_idx):
        off = _offsets(offsets)
        R = off.shape[0] - 1
        qc = q.contiguous().float()
        wc = w.reshape(-1).contiguous().float()
        K = qc.shape[-1]
        out = torch.empty(R, K, device=qc.device)
        _lib.call("na_weighted_sum_packed_fwd", _lib.ptr(qc), _lib.ptr(wc), _lib.ptr(off), R, K,
                  _lib.ptr(out), _s(qc))
        ctx.save_for_backward(qc, wc, ray_idx)
        ctx.wshape = w.shape
        return out

    @staticmethod
    def backward(ctx, dout):
        qc, wc, ray_idx = ctx.saved_tensors
        S, K = qc.shape
        dq = torch.empty_like(qc) if ctx.needs_input_grad[0] else None
        dw = torch.empty_like(wc) if ctx.needs_input_grad[1] else None
        _lib.call("na_weighted_sum_packed_bwd", _lib.ptr(qc), _lib.ptr(wc), _lib.ptr(ray_idx), S, K,
                  _lib.ptr(dout.contiguous().float()), _lib.ptr(dq), _lib.ptr(dw), _s(qc))
        return dq, (dw.view(ctx.wshape) if dw is not None else None), None, None


class NeusAlphaPacked(torch.autograd.Function):
    """Packed NeuS alphas: ray_unit [R,3], sdfs [S], gradients [S,3], dists [S], far [R]
    (or None: 1e10), outside [R] (or None) -> alphas [S]."""

    @staticmethod
    def forward(ctx, ray_unit, sdfs, gradients, dists, far, outside, offsets, ray_idx, s_var,
                progress, anneal_end, outside_val):
        off = _offsets(offsets)
        S = sdfs.shape[0]
        ray_unit = ray_unit.contiguous().float()
        sdfs = sdfs.reshape(S).contiguous().float()
        gradients = gradients.reshape(S, 3).contiguous().float()
        dists = dists.reshape(S).contiguous().float()
        far = far.reshape(-1).contiguous().float() if far is not None else None
        out_u8 = outside.contiguous().to(torch.uint8) if outside is not None else None
        alpha = torch.empty(S, device=sdfs.device)
        _lib.call("na_neus_alpha_packed_fwd", _lib.ptr(ray_unit), _lib.ptr(sdfs),
                  _lib.ptr(gradients), _lib.ptr(dists), _lib.ptr(far), _lib.ptr(out_u8),
                  float(outside_val), _lib.ptr(s_var), float(progress), float(anneal_end),
                  _lib.ptr(off), _lib.ptr(ray_idx), S, _lib.ptr(alpha), _s(sdfs))
        empty = torch.empty(0, device=sdfs.device)
        ctx.save_for_backward(ray_unit, sdfs, gradients, dists,
                              far if far is not None else empty,
                              out_u8 if out_u8 is not None else empty, s_var, off, ray_idx)
        ctx.has_far, ctx.has_out = far is not None, out_u8 is not None
        ctx.progress, ctx.anneal_end, ctx.outside_val = progress, anneal_end, outside_val
        return alpha

    @staticmethod
    def backward(ctx, dalpha):
        ray_unit, sdfs, gradients, dists, far, out_u8, s_var, off, ray_idx = ctx.saved_tensors
        S = sdfs.shape[0]
        dsdf = torch.empty_like(sdfs)
        dgrad = torch.empty_like(gradients)
        ds_var = torch.zeros(1, device=sdfs.device)
        _lib.call("na_neus_alpha_packed_bwd", _lib.ptr(ray_unit), _lib.ptr(sdfs),
                  _lib.ptr(gradients), _lib.ptr(dists), _lib.ptr(far) if ctx.has_far else None,
                  _lib.ptr(out_u8) if ctx.has_out else None, float(ctx.outside_val),
                  _lib.ptr(s_var), float(ctx.progress), float(ctx.anneal_end), _lib.ptr(off),
                  _lib.ptr(ray_idx), S, _lib.ptr(dalpha.contiguous().float()), _lib.ptr(dsdf),
                  _lib.ptr(dgrad), _lib.ptr(ds_var), _s(sdfs))
        return (None, dsdf, dgrad, None, None, None, None, None, ds_var.view_as(s_var), None,
                None, None)


class NeusCompositePacked(torch.autograd.Function):
    """Packed NeuS alphas and compositing in one persistent launch each way
    (na_neus_composite_packed_fwd/bwd): ray_unit [R,3], sdfs [S], gradients [S,3],
    dists [S], far [R] (or None), outside [R] (or None), offsets [R+1], s_var, q [S,K]
    (K <= 4) -> alphas [S], weights [S], out [R,K].  The same values as
    NeusAlphaPacked followed by CompositePacked."""

    @staticmethod
    def forward(ctx, ray_unit, sdfs, gradients, dists, far, outside, offsets, s_var, q,
                progress, anneal_end, outside_val):
        off = _offsets(offsets)
        R, S = off.shape[0] - 1, sdfs.shape[0]
        dev = sdfs.device
        ctx.shapes = (sdfs.shape, gradients.shape, q.shape)
        ray_unit = ray_unit.contiguous().float()
        sdfs = sdfs.reshape(S).contiguous().float()
        gradients = gradients.reshape(S, 3).contiguous().float()
        dists = dists.reshape(S).contiguous().float()
        far = far.reshape(-1).contiguous().float() if far is not None else None
        out_u8 = outside.contiguous().to(torch.uint8) if outside is not None else None
        qc = q.reshape(S, -1).contiguous().float()
        K = qc.shape[1]
        alpha = torch.empty(S, device=dev)
        weights = torch.empty(S, device=dev)
        out = torch.empty(R, K, device=dev)
        queue = torch.zeros(1, dtype=torch.int32, device=dev)
        with _timed("na_neus_composite_packed_fwd"):
            _lib.call("na_neus_composite_packed_fwd", _lib.ptr(ray_unit), _lib.ptr(sdfs),
                      _lib.ptr(gradients), _lib.ptr(dists), _lib.ptr(far), _lib.ptr(out_u8),
                      float(outside_val), _lib.ptr(s_var), float(progress), float(anneal_end),
                      _lib.ptr(off), R, _lib.ptr(qc), K, _lib.ptr(alpha), _lib.ptr(weights),
                      _lib.ptr(out), _lib.ptr(queue), _s(sdfs))
        empty = torch.empty(0, device=dev)
        ctx.save_for_backward(ray_unit, sdfs, gradients, dists,
                              far if far is not None else empty,
                              out_u8 if out_u8 is not None else empty, s_var, off, qc, alpha,
                              weights)
        ctx.has_far, ctx.has_out = far is not None, out_u8 is not None
        ctx.progress, ctx.anneal_end, ctx.outside_val = progress, anneal_end, outside_val
        return alpha, weights, out

    @staticmethod
    def backward(ctx, dalpha, dweights, dout):
        ray_unit, sdfs, gradients, dists, far, out_u8, s_var, off, qc, alpha, weights = \
            ctx.saved_tensors
        R, S, K = off.shape[0] - 1, sdfs.shape[0], qc.shape[1]
        dev = sdfs.device
        dsdf = torch.empty_like(sdfs)
        dgrad = torch.empty_like(gradients)
        dq = torch.empty(S, K, device=dev) if ctx.needs_input_grad[8] else None
        ds_var = torch.zeros(1, device=dev)
        queue = torch.zeros(1, dtype=torch.int32, device=dev)
        dout_c = dout.contiguous().float() if dout is not None else None
        dw_c = dweights.contiguous().float() if dweights is not None else None
        da_c = dalpha.contiguous().float() if dalpha is not None else None
        with _timed("na_neus_composite_packed_bwd"):
            _lib.call("na_neus_composite_packed_bwd", _lib.ptr(ray_unit), _lib.ptr(sdfs),
                      _lib.ptr(gradients), _lib.ptr(dists),
                      _lib.ptr(far) if ctx.has_far else None,
                      _lib.ptr(out_u8) if ctx.has_out else None, float(ctx.outside_val),
                      _lib.ptr(s_var), float(ctx.progress), float(ctx.anneal_end), _lib.ptr(off),
                      R, _lib.ptr(qc), K, _lib.ptr(alpha), _lib.ptr(weights), _lib.ptr(dout_c),
                      _lib.ptr(dw_c), _lib.ptr(da_c), _lib.ptr(dsdf), _lib.ptr(dgrad),
                      _lib.ptr(dq), _lib.ptr(ds_var), _lib.ptr(queue), _s(sdfs))
        s_shape, g_shape, q_shape = ctx.shapes
        return (None, dsdf.view(s_shape), dgrad.view(g_shape), None, None, None, None,
                ds_var.view_as(s_var), dq.view(q_shape) if dq is not None else None, None, None,
                None)


# ---------------------------------------------------------------------------
# eikonal + curvature losses
# ---------------------------------------------------------------------------
class SDFLosses(torch.autograd.Function):
    """gradients [n,3], hessians [n,3] (or None), outside [R] -> (eikonal, curvature) means."""

    @staticmethod
    def forward(ctx, gradients, hessians, outside, n_per_ray):
        n = gradients.shape[0]
        g = gradients.contiguous().float()
        h = hessians.contiguous().float() if hessians is not None else None
        o = outside.contiguous().to(torch.uint8) if outside is not None else None
        sums = torch.zeros(2, device=g.device)
        _lib.call("na_sdf_losses_fwd", _lib.ptr(g), _lib.ptr(h), _lib.ptr(o), n, int(n_per_ray),
                  _lib.ptr(sums), _s(g))
        ctx.save_for_backward(g, h if h is not None else torch.empty(0, device=g.device),
                              o if o is not None else torch.empty(0, device=g.device))
        ctx.has_h, ctx.has_o, ctx.n_per_ray = h is not None, o is not None, n_per_ray
        means = sums / max(n, 1)
        return means[0], means[1]

    @staticmethod
    def backward(ctx, deik, dcurv):
        g, h, o = ctx.saved_tensors
        n = g.shape[0]
        dev = g.device
        w = torch.stack([deik if deik is not None else torch.zeros((), device=dev),
                         dcurv if dcurv is not None else torch.zeros((), device=dev)]).float()
        dg = torch.empty_like(g)
        dh = torch.empty_like(h) if ctx.has_h else None
        _lib.call("na_sdf_losses_bwd", _lib.ptr(g), _lib.ptr(h) if ctx.has_h else None,
                  _lib.ptr(o) if ctx.has_o else None, n, int(ctx.n_per_ray), _lib.ptr(w),
                  _lib.ptr(dg), 0, _lib.ptr(dh), _s(g))
        return dg, dh, None, None
