"""Merged centre + 4-tap encoder backward (na_sdf_taps_bwd) against five separate
scatters, and the fused SDF node against the unfused forward + compute_gradients."""
import ctypes

import pytest
import torch

from neuralangelo_amd import _lib
from neuralangelo_amd import functional as Fn
from neuralangelo_amd.encoding import Encoding
from neuralangelo_amd.model import Model, default_config
from tests.helpers import randomize_model

pytestmark = pytest.mark.gpu


def _enc(log2_T, F, cuda):
    return Encoding(3, dict(otype="HashGrid", n_levels=16, n_features_per_level=F,
                            log2_hashmap_size=log2_T, base_resolution=32,
                            per_level_scale=2 ** 0.4)).to(cuda)


def _ray_points(n, g, spacing):
    """n points as 100-point runs along random rays (consecutive samples of a ray, as
    the renderer orders them), `spacing` apart."""
    rays = (n + 99) // 100
    o = torch.rand(rays, 1, 3, generator=g) * 1.6 - 0.8
    d = torch.nn.functional.normalize(torch.randn(rays, 1, 3, generator=g), dim=-1)
    t = torch.arange(100).view(1, 100, 1) * spacing
    return (o + d * t).reshape(-1, 3)[:n].contiguous()


@pytest.mark.parametrize("eps,active,dt,F,layout", [
    (1 / 2048 / 3 ** 0.5, 16, torch.float32, 8, "random"),
    (1 / 2048 / 3 ** 0.5, 16, torch.bfloat16, 8, "random"),
    (1 / 256 / 3 ** 0.5, 11, torch.float32, 4, "random"),
    (0.05, 16, torch.float32, 2, "random"),
    (1 / 2048 / 3 ** 0.5, 16, torch.float32, 8, "rays"),
    (1 / 2048 / 3 ** 0.5, 16, torch.bfloat16, 8, "rays_dense"),
    (0.05, 16, torch.float32, 4, "rays")])
def test_merged_scatter_equals_five_scatters(cuda, eps, active, dt, F, layout):
    """Random points, and runs of close samples along rays whose boxes coincide for
    several consecutive points (sums carried across points before the atomics)."""
    print("taps bwd", eps, active, dt, F, layout, flush=True)
    enc = _enc(16, F, cuda)
    n, LF = 4001, 16 * F
    g = torch.Generator().manual_seed(0)
    if layout == "random":
        pts = (torch.rand(n, 3, generator=g) * 3.8 - 1.9).to(cuda)
    else:
        pts = _ray_points(n, g, 2e-3 if layout == "rays" else 1e-4).to(cuda)
    dyc = torch.randn(n, LF, generator=g).to(cuda).to(dt)
    dyt = torch.randn(4 * n, LF, generator=g).to(cuda).to(dt)
    dyc[::7] = 0  # zero rows are skipped
    s = _lib.stream_of(pts)
    ref = torch.zeros(enc.n_params, device=cuda)
    tp = Fn.taps_points(pts, eps).view(-1, 3)
    for x, dy in ((pts, dyc), (tp, dyt)):
        _lib.call("na_sdf_input_bwd", ctypes.byref(enc.desc), _lib.ptr(x), x.shape[0], -2.0, 2.0,
                  _lib.ptr(dy), _lib.dtype_code(dy.dtype), LF, _lib.ptr(ref), active, s)
    got = torch.zeros(enc.n_params, device=cuda)
    _lib.call("na_sdf_taps_bwd", ctypes.byref(enc.desc), _lib.ptr(pts), n, -2.0, 2.0, float(eps),
              _lib.ptr(dyc), _lib.ptr(dyt), None, _lib.dtype_code(dt), LF, _lib.ptr(got), active,
              0, s)
    torch.cuda.synchronize()
    scale = ref.abs().max().item()
    assert (got - ref).abs().max().item() <= 2e-5 * scale
    assert torch.equal(got != 0, ref != 0) or ((got != 0) ^ (ref != 0)).sum() < 1e-4 * ref.numel()
    off = enc.offsets[active] * F
    if off < got.numel():
        assert got[off:].abs().max().item() == 0.0  # inactive levels untouched


def _sdf_model(cuda, cdt, state=None):
    torch.manual_seed(0)
    m = Model(default_config(log2_T=15), compute_dtype=cdt)
    randomize_model(m, 2)
    if state is not None:
        m.load_state_dict(state)
    m = m.to(cuda).train()
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    return m


def test_bf16_increment_node_against_fp32(cuda):
    """The BF16 increment-form node against the FP32 node on the same weights:
    values (sdf, gradient, hessian) and parameter gradients of a loss on them."""
    m32 = _sdf_model(cuda, torch.float32)
    m16 = _sdf_model(cuda, torch.bfloat16, {k: v.cpu() for k, v in m32.state_dict().items()})
    g = torch.Generator().manual_seed(1)
    pts = (torch.rand(3000, 3, generator=g) * 1.6 - 0.8).to(cuda)
    w_grad = torch.randn(3000, 3, generator=g).to(cuda)
    out = {}
    for name, m in (("fp32", m32), ("bf16", m16)):
        sm = m.neural_sdf
        sdf, feat, taps, delta = sm.forward_with_taps(pts)
        grad, hess = sm.gradients_from_taps(taps, sdf, sdf, True, delta)
        loss = sdf.sum() + (grad * w_grad).sum() + ((grad.norm(dim=-1) - 1) ** 2).sum()
        params = [sm.tcnn_encoding.params] + list(sm.mlp.parameters())
        out[name] = (sdf.detach(), grad.detach(), hess.detach(),
                     torch.autograd.grad(loss, params))
    s32, g32, h32, p32 = out["fp32"]
    s16, g16, h16, p16 = out["bf16"]
    e_sdf = (s16 - s32).abs().max().item()
    e_grad = ((g16 - g32).norm(dim=-1).max() / g32.norm(dim=-1).mean()).item()
    # the FP32 hessian is itself dominated by cancellation (the reference's arithmetic),
    # so only its scale is compared
    e_hess = ((h16 - h32).abs().sum(-1)).median().item()
    print(f"\nbf16 vs fp32: sdf {e_sdf:.2e}  grad err max/mean|g| {e_grad:.2e}  "
          f"hess median |diff| {e_hess:.2e}")
    assert e_sdf < 1e-3 and e_grad < 1e-1
    for a, b in zip(p16, p32):
        a, b = a.double().flatten(), b.double().flatten()
        if b.norm() == 0:  # parameters the loss does not reach (W1, b1: no feature term)
            assert a.norm() == 0
            continue
        cos = (a @ b / (a.norm() * b.norm() + 1e-300)).item()
        print(f"  param grad cosine {cos:.5f}")
        assert cos > 0.99, cos


def test_fused_node_matches_unfused_path(cuda):
    cdt = torch.float32
    torch.manual_seed(0)
    m = Model(default_config(log2_T=15), compute_dtype=cdt)
    randomize_model(m, 2)
    m = m.to(cuda).train()
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    sdf_mod = m.neural_sdf
    g = torch.Generator().manual_seed(1)
    pts = (torch.rand(3000, 3, generator=g) * 1.6 - 0.8).to(cuda)
    w_feat = torch.randn(3000, 256, generator=g).to(cuda)
    w_grad = torch.randn(3000, 3, generator=g).to(cuda)

    def loss_of(sdf, feat, grad, hess):
        return (sdf.sum() + (feat.float() * w_feat).sum() * 1e-2 + (grad * w_grad).sum()
                + 1e-3 * hess.abs().sum())

    params = [sdf_mod.tcnn_encoding.params] + list(sdf_mod.mlp.parameters())
    sdf, feat = sdf_mod.forward(pts)
    grad, hess = sdf_mod.compute_gradients(pts, training=True, sdf=sdf)
    ref = torch.autograd.grad(loss_of(sdf, feat, grad, hess), params)
    sdf2, feat2, taps, delta = sdf_mod.forward_with_taps(pts)
    grad2, hess2 = sdf_mod.gradients_from_taps(taps, sdf2, sdf2, True, delta)
    assert torch.equal(sdf, sdf2) and torch.equal(grad, grad2) and torch.equal(hess, hess2)
    got = torch.autograd.grad(loss_of(sdf2, feat2, grad2, hess2), params)
    for a, b in zip(got, ref):
        a, b = a.double().flatten(), b.double().flatten()
        cos = (a @ b / (a.norm() * b.norm() + 1e-300)).item()
        assert cos > 0.999999, cos
        # the hessian term divides by eps^2: summation-order differences reach ~1e-2
        assert (a - b).abs().max() <= 2e-2 * b.abs().max() + 1e-12


@pytest.mark.parametrize("dt", [torch.float32, torch.bfloat16])
def test_increment_layout_scatter(cuda, dt):
    """The increment layout (point-major tap rows, a shared row per point, the centre's
    own row) scatters what the plain layout does with the taps' gradients g_t = dyt +
    dyq (tap-major) and the centre's dyc - sum_t g_t."""
    enc = _enc(16, 8, cuda)
    n, LF, eps = 3001, 128, 2e-3
    g = torch.Generator().manual_seed(5)
    pts = _ray_points(n, g, 2e-3).to(cuda)
    dyc = torch.randn(n, LF, generator=g).to(dt)
    dyt = torch.randn(n, 4, LF, generator=g).to(dt)
    dyq = torch.randn(n, LF, generator=g).to(dt)
    gt = dyt.float() + dyq.float()[:, None, :]
    plain_c = (dyc.float() - gt.sum(1)).to(cuda)
    plain_t = gt.transpose(0, 1).reshape(4 * n, LF).contiguous().to(cuda)
    s = _lib.stream_of(pts)
    ref = torch.zeros(enc.n_params, device=cuda)
    _lib.call("na_sdf_taps_bwd", ctypes.byref(enc.desc), _lib.ptr(pts), n, -2.0, 2.0, float(eps),
              _lib.ptr(plain_c), _lib.ptr(plain_t), None, _lib.NA_F32, LF, _lib.ptr(ref), 16, 0, s)
    got = torch.zeros(enc.n_params, device=cuda)
    dyt_pm = dyt.reshape(4 * n, LF).contiguous().to(cuda)
    dyc_d, dyq_d = dyc.to(cuda), dyq.to(cuda)  # held until the launch has run
    _lib.call("na_sdf_taps_bwd", ctypes.byref(enc.desc), _lib.ptr(pts), n, -2.0, 2.0, float(eps),
              _lib.ptr(dyc_d), _lib.ptr(dyt_pm), _lib.ptr(dyq_d),
              _lib.dtype_code(dt), LF, _lib.ptr(got), 16, 1, s)
    torch.cuda.synchronize()
    scale = ref.abs().max().item()
    assert (got - ref).abs().max().item() <= 2e-5 * scale
