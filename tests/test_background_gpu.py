"""Background NeRF on the GPU (modules.py:232-311, model.py:129-141,192-205) against
the oracle: the encoding kernel, the network (values and parameter gradients, FP32 and
BF16) and a full object + background render."""
import pytest
import torch

from neuralangelo_amd import functional as Fn
from neuralangelo_amd.model import Model, default_config
from oracle import field as of
from oracle import render as orr

pytestmark = pytest.mark.gpu


def _points(n, seed):
    g = torch.Generator().manual_seed(seed)
    d = torch.nn.functional.normalize(torch.randn(n, 3, generator=g), dim=-1)
    r = 1.0 + torch.rand(n, 1, generator=g) * 50.0  # outside the unit sphere
    return d * r, torch.nn.functional.normalize(torch.randn(n, 3, generator=g), dim=-1)


def _encode32(p):
    """[q | PE(q)] evaluated in float32 as torch does (CPU)."""
    n = p.norm(dim=-1, keepdim=True)
    q = torch.cat([p / n, 1.0 / n], dim=-1)
    return torch.cat([q, of.positional_encoding(q, 10)], dim=-1)


def test_encoding_kernel(cuda):
    p, _ = _points(20000, 0)
    X = Fn.bg_encode(p.to(cuda), 10, torch.float32, 88).cpu()
    ref = _encode32(p)
    assert X.shape == (20000, 88) and (X[:, 84:] == 0).all()
    # q agrees to float32 rounding (|p| by sqrtf vs torch.norm may differ in the last
    # bit); a code at frequency 2^l pi then moves by up to ~2 ulp(q) * 2^l pi
    tol = torch.full((84,), 2e-6)
    for d in range(4):
        for s in range(2):
            for l in range(10):
                tol[4 + d * 20 + s * 10 + l] = 2e-6 + 3e-7 * (2 ** l) * 3.1416
    assert ((X[:, :84] - ref).abs() <= tol).all()


def _model(cuda, cdt, seed=0):
    torch.manual_seed(seed)
    m = Model(default_config(log2_T=14, background=True), compute_dtype=cdt)
    with torch.no_grad():
        for lin in list(m.background_nerf.mlp_feat.linears) + list(m.background_nerf.mlp_rgb.linears):
            lin.bias.normal_(0, 0.1)
    return m.to(cuda)


def _oracle(m, p, rays):
    bg = m.background_nerf
    Wf = [l.weight.detach().cpu().double().requires_grad_(True) for l in bg.mlp_feat.linears]
    bf = [l.bias.detach().cpu().double().requires_grad_(True) for l in bg.mlp_feat.linears]
    Wr = [l.weight.detach().cpu().double().requires_grad_(True) for l in bg.mlp_rgb.linears]
    br = [l.bias.detach().cpu().double().requires_grad_(True) for l in bg.mlp_rgb.linears]
    inp = _encode32(p).double()
    # the network after the (float32) input, in float64
    h = inp
    for li, (W, b) in enumerate(zip(Wf, bf)):
        if li == 4:
            h = torch.cat([h, inp], dim=-1)
        h = h @ W.t() + b
        if li != 7:
            h = h.relu()
    density = torch.nn.functional.softplus(h[..., 0])
    x = torch.cat([h[..., 1:].relu(), of.spherical_harmonics3(rays.double())], dim=-1)
    for li, (W, b) in enumerate(zip(Wr, br)):
        x = x @ W.t() + b
        x = x.relu() if li == 0 else x.sigmoid()
    return x, density, Wf + bf + Wr + br


@pytest.mark.parametrize("cdt", [torch.float32, torch.bfloat16])
def test_background_network_against_oracle(cuda, cdt):
    m = _model(cuda, cdt)
    p, rays = _points(4096, 1)
    rgb, density = m.background_nerf(p.to(cuda), rays.to(cuda), None)
    rgb_r, dens_r, params_r = _oracle(m, p, rays)
    tol = 1e-4 if cdt == torch.float32 else 3e-2
    assert (rgb.detach().cpu().double() - rgb_r.detach()).abs().max() < tol
    assert ((density.detach().cpu().double() - dens_r.detach()).abs()
            <= tol * dens_r.detach().abs() + tol).all()
    g = torch.Generator().manual_seed(2)
    w_rgb, w_d = torch.randn(4096, 3, generator=g), torch.randn(4096, generator=g)
    loss = (rgb * w_rgb.to(cuda)).sum() + (density * w_d.to(cuda)).sum()
    bg = m.background_nerf
    mine = [l.weight for l in bg.mlp_feat.linears] + [l.bias for l in bg.mlp_feat.linears] + \
        [l.weight for l in bg.mlp_rgb.linears] + [l.bias for l in bg.mlp_rgb.linears]
    grads = torch.autograd.grad(loss, mine)
    ((rgb_r * w_rgb.double()).sum() + (dens_r * w_d.double()).sum()).backward()
    for a, b in zip(grads, params_r):
        a, b = a.cpu().double().reshape(-1), b.grad.reshape(-1)
        if b.norm() == 0:
            continue
        cos = (a @ b) / (a.norm() * b.norm())
        assert cos > (0.9999 if cdt == torch.float32 else 0.99), cos.item()


def test_render_with_background_composites_both(cuda):
    """render_rays with the background: the object samples (from the same call) and the
    oracle's background samples, alphas and colours composited together."""
    m = _model(cuda, torch.float32, 3).train()
    m.progress = 1.0
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    g = torch.Generator().manual_seed(4)
    R = 256
    c = torch.nn.functional.normalize(torch.randn(R, 3, generator=g), dim=-1) * 3.0
    v = torch.nn.functional.normalize(-c + torch.randn(R, 3, generator=g) * 0.3, dim=-1)
    rb = torch.rand(R, 32, generator=g)
    out = m.render_rays(c.to(cuda), v.to(cuda), stratified=False, rands_bg=rb.to(cuda))
    obj = m.render_rays_object(c.to(cuda), v.to(cuda), out["near"][:, 0], out["far"][:, 0],
                               out["outside"][:, 0], None, dists=out["dists"][:, :96, 0])
    far = out["far"].detach().cpu().double()
    dbg = orr.sample_dists_background(far, 32, rb)
    pts = (c.double()[:, None, :] + v.double()[:, None, :] * dbg[..., None]).reshape(-1, 3)
    rgb_b, dens_b, _ = _oracle(m, pts.float(), v[:, None, :].expand(R, 32, 3).reshape(-1, 3))
    a_b = orr.volume_rendering_alphas_dist(dens_b.detach().view(R, 32), dbg[..., None])
    alphas = torch.cat([obj["alphas"].detach().cpu().double(), a_b], dim=1)
    rgbs = torch.cat([obj["rgbs"].detach().cpu().double().view(R, -1, 3),
                      rgb_b.detach().view(R, 32, 3)], dim=1)
    w = orr.alpha_compositing_weights(alphas)
    rgb_ref = orr.composite(rgbs, w)
    assert out["weights"].shape == (R, 128, 1)
    assert (out["rgb"].detach().cpu().double() - rgb_ref).abs().max() < 1e-4
    assert (out["dists"][:, 96:, 0].cpu().double() - dbg).abs().max() < 1e-3 * dbg.max()
    out["rgb"].sum().backward()
    gb = m.background_nerf.mlp_feat.linears[0].weight.grad
    assert gb is not None and torch.isfinite(gb).all() and gb.abs().max() > 0
