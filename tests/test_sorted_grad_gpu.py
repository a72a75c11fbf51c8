"""Sorted (deterministic) table-gradient reduction: bitwise repeatable, equal bit for
bit to a CPU restatement of its documented summation order (per table entry, the
contributions of (point, evaluation, corner) in that order, fp32 fma), and within
float tolerance of the atomic scatters."""
import ctypes

import numpy as np
import pytest
import torch

from neuralangelo_amd import _lib
from neuralangelo_amd.encoding import Encoding

pytestmark = pytest.mark.gpu

TAP_DIRS = np.array([[1, -1, -1], [-1, -1, 1], [-1, 1, -1], [1, 1, 1]], dtype=np.float32)


def _enc(F, log2_T=12, levels=8, base=8, cuda=None):
    enc = Encoding(3, dict(otype="HashGrid", n_levels=levels, n_features_per_level=F,
                           log2_hashmap_size=log2_T, base_resolution=base, per_level_scale=1.6))
    return enc.to(cuda) if cuda is not None else enc


def _fma32(a, b, c):
    """float32 fma: the product of two floats is exact in float64, and the sum rounds
    once to 53 bits before the float32 rounding (double rounding is possible only at
    exact float32 ties of the 53-bit result, far below what these sizes reach)."""
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def _fixed_order(n_params, F, rows, ids, w, g):
    """Per table row, sum w * g over its contributions in increasing id order (fp32
    fma from zero), vectorised over rows.  rows/ids/w [M], g [M, F] (fp32)."""
    out = np.zeros((n_params // F, F), dtype=np.float32)
    if rows.size == 0:
        return out.reshape(-1)
    order = np.lexsort((ids, rows))
    rows, w, g = rows[order], w[order], g[order]
    head = np.r_[True, rows[1:] != rows[:-1]]
    seg = np.cumsum(head) - 1
    starts = np.flatnonzero(head)
    pos = np.arange(rows.size) - starts[seg]
    acc = np.zeros((starts.size, F), dtype=np.float32)
    for k in range(int(pos.max()) + 1):
        sel = pos == k
        s = seg[sel]
        acc[s] = _fma32(w[sel][:, None], g[sel], acc[s])
    out[rows[starts]] = acc
    return out.reshape(-1)


def _corners(enc, u):
    """Global table rows and weights [n, L, 8] of unit-cube points (the encoder's own
    corner kernel, pinned against the oracle elsewhere)."""
    idx, w = enc.corners(u)
    return idx.cpu().numpy().astype(np.int64), w.cpu().numpy()


def _contributions(enc, u_per_eval, g_per_eval, active):
    """Contribution lists of E evaluations: u_per_eval [E] of [n,3] unit points,
    g_per_eval [E] of [n, L*F] fp32 gradients; id = (i * E + e) * 8 + c per level."""
    E = len(u_per_eval)
    F = enc.n_features_per_level
    rows, ids, ws, gs = [], [], [], []
    for e, (u, g) in enumerate(zip(u_per_eval, g_per_eval)):
        idx, w = _corners(enc, u)
        n = u.shape[0]
        for lv in range(active):
            gl = g[:, lv * F:(lv + 1) * F]
            keep = np.any(gl != 0, axis=1)
            i = np.repeat(np.arange(n)[keep], 8)
            c = np.tile(np.arange(8), int(keep.sum()))
            rows.append(idx[keep, lv, :].reshape(-1))
            ids.append((i * E + e) * 8 + c)
            ws.append(w[keep, lv, :].reshape(-1))
            gs.append(np.repeat(gl[keep], 8, axis=0))
    cat = np.concatenate
    return cat(rows), cat(ids), cat(ws).astype(np.float32), cat(gs).astype(np.float32)


@pytest.mark.parametrize("F,dt,active", [(2, torch.float32, 8), (4, torch.bfloat16, 8),
                                         (8, torch.float32, 5)])
def test_encoder_sorted_grad_is_the_fixed_order_sum(cuda, F, dt, active):
    print("sorted encoder bwd", F, dt, active, flush=True)
    enc = _enc(F, cuda=cuda)
    enc.active_levels = active
    g = torch.Generator().manual_seed(3)
    n = 3000
    x = torch.rand(n, 3, generator=g)
    x[1000:1400] = x[:400]  # repeated points: long runs of one entry
    dy = torch.randn(n, enc.n_output_dims, generator=g).to(dt)
    dy[::5] = 0
    xg, dyg = x.to(cuda), dy.to(cuda)
    enc.deterministic = True
    a = torch.zeros(enc.n_params, device=cuda)
    b = torch.zeros(enc.n_params, device=cuda)
    enc.backward_into(xg, dyg, a)
    enc.backward_into(xg, dyg, b)
    enc.deterministic = False
    atomic = torch.zeros(enc.n_params, device=cuda)
    enc.backward_into(xg, dyg, atomic)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    rows, ids, w, gg = _contributions(enc, [xg], [dy.float().numpy()], active)
    ref = _fixed_order(enc.n_params, F, rows, ids, w, gg)
    got = a.cpu().numpy()
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    scale = atomic.abs().max().item()
    assert (a - atomic).abs().max().item() <= 1e-5 * scale


@pytest.mark.parametrize("inc,dt", [(0, torch.float32), (1, torch.float32), (1, torch.bfloat16)])
def test_taps_sorted_grad_is_the_fixed_order_sum(cuda, inc, dt):
    """Centre + four taps (E = 5) in both layouts, against the same fixed-order sum of
    the per-evaluation gradients (increment layout: g_t = dyt + dyq, centre dyc - sum)."""
    print("sorted taps bwd", inc, dt, flush=True)
    F, active, eps = 4, 8, 3e-3
    enc = _enc(F, cuda=cuda)
    LF = enc.n_output_dims
    g = torch.Generator().manual_seed(11)
    n = 2000
    pts = torch.rand(n, 3, generator=g) * 3.6 - 1.8  # world coordinates in [-2, 2]^3
    dyc = torch.randn(n, LF, generator=g).to(dt)
    dyt = torch.randn(4 * n, LF, generator=g).to(dt)
    dyq = torch.randn(n, LF, generator=g).to(dt) if inc else None
    dyc[::7] = 0
    f32 = lambda t: t.float().numpy()  # noqa: E731
    if inc:
        q = f32(dyq)
        gt = [f32(dyt)[t::4] + q for t in range(4)]
        gc = f32(dyc) - ((gt[0] + gt[1]) + (gt[2] + gt[3]))
    else:
        gt = [f32(dyt)[t * n:(t + 1) * n] for t in range(4)]
        gc = f32(dyc)
    xn = pts.numpy()
    us = [(xn + 2.0) * np.float32(0.25)]
    for t in range(4):
        xt = xn + TAP_DIRS[t] * np.float32(eps)
        us.append((xt + 2.0) * np.float32(0.25))
    rows, ids, w, gg = _contributions(enc, [torch.from_numpy(u).to(cuda) for u in us],
                                      [gc] + gt, active)
    ref = _fixed_order(enc.n_params, F, rows, ids, w, gg)
    # device copies held in names: the kernels read them after the calls return
    ptsg, dycg, dytg = pts.to(cuda), dyc.to(cuda), dyt.to(cuda)
    dyqg = dyq.to(cuda) if inc else None
    s = _lib.stream_of(ptsg)
    args = (ctypes.byref(enc.desc), _lib.ptr(ptsg), n, -2.0, 2.0, float(eps), _lib.ptr(dycg),
            _lib.ptr(dytg), _lib.ptr(dyqg) if inc else None, _lib.dtype_code(dt), LF)
    outs, wss = [], []
    for _ in range(2):
        o = torch.zeros(enc.n_params, device=cuda)
        ws = enc.sorted_workspace(n, 5, cuda)
        _lib.call("na_sdf_taps_bwd_sorted", *args, _lib.ptr(o), active, inc, _lib.ptr(ws),
                  ws.numel(), s)
        outs.append(o)
        wss.append(ws)
    atomic = torch.zeros(enc.n_params, device=cuda)
    _lib.call("na_sdf_taps_bwd", *args, _lib.ptr(atomic), active, inc, s)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    got = outs[0].cpu().numpy()
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    scale = atomic.abs().max().item()
    assert (outs[0] - atomic).abs().max().item() <= 2e-5 * scale


def test_sorted_workspace_is_checked(cuda):
    enc = _enc(4, cuda=cuda)
    x = torch.rand(100, 3, device=cuda)
    dy = torch.randn(100, enc.n_output_dims, device=cuda)
    gr = torch.zeros(enc.n_params, device=cuda)
    need = _lib.call("na_hashgrid_sorted_workspace", ctypes.byref(enc.desc), 100, 1)
    small = torch.empty(max(need - 256, 1), dtype=torch.uint8, device=cuda)
    with pytest.raises(RuntimeError, match="workspace"):
        _lib.call("na_hashgrid_bwd_sorted", ctypes.byref(enc.desc), _lib.ptr(x), 100, _lib.ptr(dy),
                  _lib.NA_F32, _lib.ptr(gr), 4, _lib.ptr(small), small.numel(), _lib.stream_of(x))
    with pytest.raises(RuntimeError, match="evals"):
        _lib.call("na_hashgrid_sorted_workspace", ctypes.byref(enc.desc), 100, 3)


def test_deterministic_train_step_table_gradient(cuda):
    """A BF16 train step of the model with the sorted reduction: the table gradient is
    bitwise equal over two runs on the same inputs (the atomic scatter's is not in
    general), and close to the atomic one; torch.use_deterministic_algorithms selects
    the sorted path too."""
    from neuralangelo_amd.model import Model, default_config
    from tests.helpers import randomize_model

    torch.manual_seed(0)
    m = Model(default_config(log2_T=15), compute_dtype=torch.bfloat16)
    randomize_model(m, 2)
    m = m.to(cuda).train()
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    enc = m.neural_sdf.tcnn_encoding
    gen = torch.Generator().manual_seed(9)
    center = torch.randn(256, 3, generator=gen)
    center = center / center.norm(dim=-1, keepdim=True) * 3.0
    ray = -center / 3.0 + 0.3 * torch.randn(256, 3, generator=gen)
    center, ray = center.to(cuda), (ray / ray.norm(dim=-1, keepdim=True)).to(cuda)

    def table_grad():
        m.zero_grad(set_to_none=True)
        out = m.render_rays(center, ray)
        loss = out["rgb"].square().mean() + out["gradients"].norm(dim=-1).mean()
        loss.backward()
        return enc.params.grad.detach().clone()

    enc.deterministic = True
    a, b = table_grad(), table_grad()
    enc.deterministic = False
    atomic = table_grad()
    assert torch.equal(a, b)
    cos = (a.double() @ atomic.double() / (a.double().norm() * atomic.double().norm())).item()
    assert cos > 0.99999, cos
    torch.use_deterministic_algorithms(True, warn_only=True)
    try:
        assert enc.sorted_grad()
        c = table_grad()
    finally:
        torch.use_deterministic_algorithms(False)
    assert torch.equal(a, c)
