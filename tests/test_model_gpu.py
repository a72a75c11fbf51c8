"""Full object render (training step) on the HIP path vs the CPU oracle."""
import pytest
import torch

from neuralangelo_amd.model import Model, default_config
from oracle import render as orr
from tests.helpers import oracle_field, oracle_params, randomize_model

pytestmark = pytest.mark.gpu


def _rays(R, seed):
    g = torch.Generator().manual_seed(seed)
    c = torch.randn(R, 3, generator=g)
    c = c / c.norm(dim=-1, keepdim=True) * 3.0
    d = -c / 3.0 + 0.3 * torch.randn(R, 3, generator=g)
    return c, d / d.norm(dim=-1, keepdim=True)


def _cos(a, b):
    a, b = a.reshape(-1).double(), b.reshape(-1).double()
    return (a @ b / (a.norm() * b.norm() + 1e-300)).item()


@pytest.fixture(scope="module")
def setup(cuda):
    torch.manual_seed(0)
    m = Model(default_config(log2_T=14), compute_dtype=torch.float32)
    randomize_model(m, 1)
    m = m.to(cuda).train()
    m.progress = 0.05
    return m


def test_sampling_matches_oracle(setup, cuda):
    m = setup
    R = 128
    c, v = _rays(R, 2)
    g = torch.Generator().manual_seed(3)
    u = torch.rand(R, 64, generator=g)
    near, far, outside = m.get_dist_bounds(c.to(cuda), v.to(cuda))
    # the reference signature (model.py:233): center, ray_unit, near, far, stratified
    dists = m.sample_dists_all(c.to(cuda), v.to(cuda), near, far, stratified=True,
                               rands=u.to(cuda))[..., 0].cpu()
    near, far = near[:, 0], far[:, 0]
    P = oracle_params(m, torch.float32)
    fld, _ = oracle_field(m, P, dtype=torch.float32)
    on, of, oo = orr.get_dist_bounds(c, v)
    ref = orr.sample_dists_all(c, v, near.cpu()[:, None], far.cpu()[:, None], fld.sdf, 64, 16, 2,
                               u)[..., 0]
    assert dists.shape == ref.shape == (R, 96)
    assert (dists[:, 1:] >= dists[:, :-1]).all()
    # SDF values agree to ~1e-6, so resampled positions agree except at rare near-ties
    close = (dists - ref).abs() <= 1e-4 * ref.abs() + 1e-5
    assert close.float().mean() > 0.99


def test_train_step_matches_oracle(setup, cuda):
    m = setup
    R = 96
    c, v = _rays(R, 4)
    g = torch.Generator().manual_seed(5)
    u = torch.rand(R, 64, generator=g)
    target = torch.rand(R, 3, generator=g)
    with torch.no_grad():
        dists = m.render_rays(c.to(cuda), v.to(cuda), rands=u.to(cuda))["dists"][..., 0]
    m.zero_grad(set_to_none=True)
    out = m.render_rays(c.to(cuda), v.to(cuda), rands=u.to(cuda), dists=dists)
    loss_ren = torch.nn.functional.l1_loss(out["rgb"], target.to(cuda)) * 3
    from neuralangelo_amd import misc
    eik, curv = misc.sdf_losses(out["gradients"], out["hessians"], out["outside"])
    loss = loss_ren + 0.1 * eik + 5e-4 * curv
    loss.backward()
    # oracle (float64) on the same samples
    P = oracle_params(m)
    fld, table = oracle_field(m, P)
    sv = m.s_var.detach().cpu().double().requires_grad_(True)
    o = orr.render_rays(fld, P, c.double(), v.double(), sv, m.progress,
                        m.neural_sdf.normal_eps, True, dists=dists.cpu().double()[..., None])
    l_ref, parts = orr.train_loss(o, target.double())
    assert abs(loss_ren.item() - parts["render"].item()) < 1e-5
    assert abs(eik.item() - parts["eikonal"].item()) < 1e-3 * abs(parts["eikonal"].item())
    # FP32 vs FP64: the 4-tap gradient divides float32 SDF rounding (~6e-8) by
    # eps (~2.8e-4); it enters alpha through the cosine term, then the colour
    assert (out["rgb"].detach().cpu().double() - o["rgb"].detach()).abs().max() < 2e-4
    assert (out["alphas"].detach().cpu().double() - o["alphas"].detach()).abs().max() < 1e-3
    names = ["W0v", "b0", "W1v", "ws", "R0v", "R2v", "R4v"]
    refs = torch.autograd.grad(l_ref, [table, sv] + [P[k] for k in names], allow_unused=True)
    sdf_mod, rgb_mod = m.neural_sdf, m.neural_rgb
    got = [sdf_mod.tcnn_encoding.params.grad, m.s_var.grad, sdf_mod.mlp.linears[0].weight_v.grad,
           sdf_mod.mlp.linears[0].bias.grad, sdf_mod.mlp.linears[1].weight_v.grad,
           sdf_mod.mlp.linear_sdf.weight.grad, rgb_mod.mlp.linears[0].weight_v.grad,
           rgb_mod.mlp.linears[2].weight_v.grad, rgb_mod.mlp.linears[4].weight_v.grad]
    for name, a, b in zip(["table", "s_var"] + names, got, refs):
        cs = _cos(a.cpu(), b)
        assert cs > 0.999, (name, cs)


def test_reference_signature_operators(setup, cuda):
    """Model.get_dist_bounds / compute_neus_alphas / sample_dists_hierarchical with the
    reference's [B,R,N,...] shapes (model.py:207-298) against the oracle."""
    m = setup
    B, R, N = 2, 24, 40
    g = torch.Generator().manual_seed(9)
    c, v = _rays(B * R, 9)
    c, v = c.view(B, R, 3), v.view(B, R, 3)
    near, far, outside = m.get_dist_bounds(c.to(cuda), v.to(cuda))
    on, of, oo = orr.get_dist_bounds(c, v)
    assert torch.allclose(near.cpu(), on, atol=1e-6) and torch.equal(outside.cpu(), oo)
    dists = torch.sort(torch.rand(B, R, N, 1, generator=g) * 2 + 1, dim=2).values
    sdfs = torch.randn(B, R, N, 1, generator=g) * 0.3
    grads = torch.randn(B, R, N, 3, generator=g)
    dfar = dists[:, :, -1:, :] + 0.1
    a = m.compute_neus_alphas(v.to(cuda), sdfs.to(cuda), grads.to(cuda), dists.to(cuda),
                              dfar.to(cuda), progress=0.05)
    ref = orr.compute_neus_alphas(v, sdfs, grads, dists, dfar, m.s_var.detach().cpu(), 0.05,
                                  m.anneal_end)
    assert a.shape == (B, R, N) and (a.cpu() - ref).abs().max() < 2e-6
    fine = m.sample_dists_hierarchical(dists.to(cuda), sdfs.to(cuda), 64.0)
    ref = orr.sample_dists_hierarchical(dists, sdfs, 64.0, 16)
    assert fine.shape == (B, R, 16, 1)
    assert (fine.cpu() - ref).abs().max() < 1e-5


def test_sampling_weights_follow_parameter_updates(cuda):
    """The sampling passes reuse their kernel-layout weights between calls; an in-place
    parameter update (as the optimizer makes) must reach the next call."""
    torch.manual_seed(0)
    m = Model(default_config(log2_T=14), compute_dtype=torch.float32)
    randomize_model(m, 5)
    m = m.to(cuda)
    sdf = m.neural_sdf
    x = torch.rand(1000, 3, device=cuda) * 1.6 - 0.8
    s0 = sdf.sdf_nograd(x)
    assert torch.equal(sdf.sdf_nograd(x), s0)
    bias = [p for n, p in sdf.mlp.named_parameters() if n.endswith("bias")][-1]
    with torch.no_grad():
        bias.add_(0.25)
    s1 = sdf.sdf_nograd(x)
    with torch.no_grad():
        ref = sdf(x, with_feat=False)[0].reshape(s1.shape)
    assert (s1 - ref).abs().max().item() < 1e-5
    assert (s1 - s0).abs().max().item() > 1e-3
