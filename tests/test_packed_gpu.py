"""Packed (ragged) rays on the GPU: compositing, weighted sums and NeuS alphas at the
packed-batch configuration's size (65,536 rays, 0..192 samples each) against the
oracle's packed restatements, equal counts against the dense kernels, and a packed
render (fwd + bwd) against dense renders of the same rays."""
import pytest
import torch

from neuralangelo_amd import functional as Fn
from neuralangelo_amd import render as nr
from neuralangelo_amd.model import Model, default_config
from oracle import render as orr
from tests.helpers import randomize_model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    c, v, off, d = orr.ragged_batch(65536, seed=11)
    return c, v, off, d


def test_packed_compositing_full_batch(cuda, batch):
    _, _, off, _ = batch
    S = int(off[-1])
    print("packed compositing S =", S, flush=True)
    g = torch.Generator().manual_seed(0)
    a = torch.rand(S, generator=g)
    a[::13] = 0.0
    a[::29] = 1.0
    q = torch.randn(S, 3, generator=g)
    offd = off.to(cuda)
    ad = a.to(cuda).requires_grad_(True)
    qd = q.to(cuda).requires_grad_(True)
    w, out = nr.composite_from_alphas_packed(ad, qd, offd)
    a64 = a.double().requires_grad_(True)
    q64 = q.double().requires_grad_(True)
    w_ref = orr.alpha_compositing_weights_packed(a64, off)
    out_ref = orr.composite_packed(q64, w_ref, off)
    assert (w.detach().cpu().double() - w_ref.detach()).abs().max() < 1e-5
    assert (out.detach().cpu().double() - out_ref.detach()).abs().max() < 1e-5
    # empty rays composite to zero
    empty = (off[1:] == off[:-1])
    assert empty.any() and (out.detach().cpu()[empty] == 0).all()
    gw = torch.randn(S, 1, generator=g)
    go = torch.randn(out.shape[0], 3, generator=g)
    ((w * gw.to(cuda)).sum() + (out * go.to(cuda)).sum()).backward()
    ((w_ref * gw.double()).sum() + (out_ref * go.double()).sum()).backward()
    assert (ad.grad.cpu().double() - a64.grad).abs().max() < 1e-4
    assert (qd.grad.cpu().double() - q64.grad).abs().max() < 1e-5


def test_packed_weighted_sum(cuda, batch):
    _, _, off, _ = batch
    S = int(off[-1])
    g = torch.Generator().manual_seed(1)
    q = torch.randn(S, 3, generator=g)
    w = torch.rand(S, 1, generator=g)
    offd = off.to(cuda)
    qd, wd = q.to(cuda).requires_grad_(True), w.to(cuda).requires_grad_(True)
    out = nr.composite_packed(qd, wd, offd)
    q64, w64 = q.double().requires_grad_(True), w.double().requires_grad_(True)
    ref = orr.composite_packed(q64, w64, off)
    assert (out.detach().cpu().double() - ref.detach()).abs().max() < 1e-4
    go = torch.randn(out.shape, generator=g)
    (out * go.to(cuda)).sum().backward()
    (ref * go.double()).sum().backward()
    assert (qd.grad.cpu().double() - q64.grad).abs().max() < 1e-6
    assert (wd.grad.cpu().double() - w64.grad).abs().max() < 1e-5


@pytest.mark.parametrize("progress", [0.02, 1.0])
def test_packed_neus_alphas(cuda, batch, progress):
    c, v, off, d = batch
    R, S = off.numel() - 1, int(off[-1])
    g = torch.Generator().manual_seed(2)
    sdf = torch.randn(S, 1, generator=g) * 0.05
    grad = torch.randn(S, 3, generator=g)
    _, far, _ = orr.get_dist_bounds(c, v)
    m = Model(default_config(log2_T=15)).to(cuda)
    sd, gd = sdf.to(cuda).requires_grad_(True), grad.to(cuda).requires_grad_(True)
    a = m.compute_neus_alphas_packed(v.to(cuda), sd, gd, d.to(cuda)[:, None], off.to(cuda),
                                     far.to(cuda), progress=progress)
    s64 = m.s_var.detach().cpu().double().requires_grad_(True)
    sdf64, grad64 = sdf.double().requires_grad_(True), grad.double().requires_grad_(True)
    ref = orr.compute_neus_alphas_packed(v.double(), sdf64, grad64, d.double()[:, None],
                                         far.double(), off, s64, progress, m.anneal_end)
    assert (a.detach().cpu().double() - ref.detach()).abs().max() < 2e-5
    ga = torch.randn(S, generator=g)
    (a * ga.to(cuda)).sum().backward()
    (ref * ga.double()).sum().backward()
    for got, want in ((sd.grad, sdf64.grad), (gd.grad, grad64.grad), (m.s_var.grad, s64.grad)):
        assert (got.cpu().double() - want).abs().max() <= 1e-3 * want.abs().max() + 1e-6


def test_equal_counts_match_dense_kernels(cuda):
    R, N = 777, 96
    g = torch.Generator().manual_seed(3)
    off = (torch.arange(R + 1, dtype=torch.int32) * N).to(cuda)
    a = torch.rand(R, N, generator=g).to(cuda)
    q = torch.randn(R, N, 3, generator=g).to(cuda)
    w1, o1 = Fn.Composite.apply(a, q)
    w2, o2 = Fn.CompositePacked.apply(a.reshape(-1), q.reshape(-1, 3), off)
    assert torch.equal(w1.reshape(-1), w2) and torch.equal(o1, o2)


def _model(cuda):
    torch.manual_seed(0)
    m = Model(default_config(log2_T=15), compute_dtype=torch.bfloat16)
    randomize_model(m, 4)
    m = m.to(cuda).train()
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    return m


def test_packed_render_matches_dense_groups(cuda):
    """A ragged batch rendered packed equals the same rays rendered densely, one group
    of rays per sample count (values, and the parameter gradients of an L1 loss)."""
    c, v, off, d = orr.ragged_batch(2048, seed=7, max_coarse=24, n_fine=8, per_unit=40.0)
    counts = (off[1:] - off[:-1]).long()
    m = _model(cuda)
    cd, vd, dd = c.to(cuda), v.to(cuda), d.to(cuda)
    out = m.render_rays_packed(cd, vd, dd, off.to(cuda))
    params = [p for p in m.parameters() if p.requires_grad]
    tgt = torch.rand(c.shape[0], 3, generator=torch.Generator().manual_seed(8)).to(cuda)
    loss = (out["rgb"] - tgt).abs().sum()
    gp = torch.autograd.grad(loss, params, allow_unused=True)
    rgb_ref = torch.zeros(c.shape[0], 3, device=cuda)
    loss_ref = 0
    for n in counts.unique().tolist():
        if n == 0:
            continue
        rays = (counts == n).nonzero()[:, 0]
        starts = off[:-1][rays].long()
        idx = starts[:, None] + torch.arange(n)[None, :]
        with torch.no_grad():
            near, far, outside = m.get_dist_bounds(cd[rays.to(cuda)], vd[rays.to(cuda)])
        o = m.render_rays_object(cd[rays.to(cuda)], vd[rays.to(cuda)], near[:, 0], far[:, 0],
                                 outside[:, 0], None, dists=dd[idx.to(cuda)])
        rgb_ref[rays.to(cuda)] = o["rgb"].detach()
        loss_ref = loss_ref + (o["rgb"] - tgt[rays.to(cuda)]).abs().sum()
    gr = torch.autograd.grad(loss_ref, params, allow_unused=True)
    assert (out["rgb"].detach() - rgb_ref).abs().max().item() < 1e-5
    for a, b in zip(gp, gr):
        if a is None or b is None:
            assert a is None and b is None
            continue
        scale = b.abs().max().item()
        if scale == 0:
            continue
        assert (a - b).abs().max().item() <= 2e-3 * scale, (a - b).abs().max().item() / scale


def test_packed_train_step_full_size(cuda, batch):
    """One BF16 train step over the full ragged batch (65,536 rays, 0..192 samples):
    packed render, L1 + eikonal + curvature, backward to table, MLPs and s_var, AdamW.
    Size-independent checks: empty rays render black, per-ray weight sums lie in
    [0, 1], every gradient is finite and the table gradient is non-zero."""
    from neuralangelo_amd import misc
    from neuralangelo_amd.optim import AdamW
    c, v, off, d = batch
    torch.manual_seed(0)
    m = Model(default_config(log2_T=22, F=8), compute_dtype=torch.bfloat16).to(cuda).train()
    m.progress = 1.0
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    opt = AdamW(m.parameters(), lr=1e-3, weight_decay=1e-2)
    enc = m.neural_sdf.tcnn_encoding
    opt.set_lowp_shadow(enc.params, enc)
    offd = off.to(cuda)
    tgt = torch.rand(c.shape[0], 3, generator=torch.Generator().manual_seed(9)).to(cuda)
    # two steps: the geometric init zeroes the encoder columns of the first SDF layer,
    # so the table gradient is zero in the first step and non-zero from the second on
    for step in range(2):
        out = m.render_rays_packed(c.to(cuda), v.to(cuda), d.to(cuda), offd)
        eik, curv = misc.sdf_losses(out["gradients"], out["hessians"], out["outside"])
        loss = (out["rgb"] - tgt).abs().mean() * 3 + 0.1 * eik + 5e-4 * curv
        opt.zero_grad(set_to_none=True)
        loss.backward()
        torch.cuda.synchronize()
        print("packed step", step, "loss", loss.item(), flush=True)
        assert torch.isfinite(loss)
        empty = (off[1:] == off[:-1]).to(cuda)
        assert (out["rgb"].detach()[empty] == 0).all()
        wsum = nr.composite_packed(1.0, out["weights"].detach().view(-1, 1), offd)
        assert (wsum >= -1e-6).all() and (wsum <= 1 + 1e-5).all()
        for p in m.parameters():
            if p.grad is not None:
                assert torch.isfinite(p.grad).all()
        table_grad_max = enc.params.grad.abs().max().item()
        opt.step()  # (clears the table gradient in its update pass)
    assert table_grad_max > 0
    assert torch.isfinite(enc.params).all()


@pytest.mark.parametrize("shape", [(64, 1), (256, 16), (128, 4)])
def test_persistent_fused_launch_matches_separate_kernels(cuda, batch, shape):
    """NeuS alphas + compositing in one persistent launch each way (a wave per ray,
    lanes over samples, rays taken in chunks from a counter) against the separate
    one-lane-per-ray kernels: alphas bit for bit (the same per-sample terms); weights,
    colour, dsdf, dgrad, drgb (with a direct dalpha's share) and s_var's gradient to
    float rounding (the transmittance and its backward recurrence run as wave scans),
    at several workgroup sizes and grids; the alphas also against the oracle."""
    from neuralangelo_amd import _lib
    rb, wpc = shape
    c, v, off, d = batch
    R, S = off.numel() - 1, int(off[-1])
    print("persistent packed", rb, wpc, R, S, flush=True)
    g = torch.Generator().manual_seed(8)
    sdf = torch.randn(S, generator=g) * 0.05
    grad = torch.randn(S, 3, generator=g)
    q = torch.rand(S, 3, generator=g)
    _, far, outside = orr.get_dist_bounds(c, v)
    outside = outside.reshape(-1).clone()
    outside[::97] = True  # some rays outside the sphere
    m = Model(default_config(log2_T=15)).to(cuda)
    offd, vd, dd = off.to(cuda), v.to(cuda), d.to(cuda)
    fard, outd = far.reshape(-1).to(cuda), outside.to(cuda)
    go = torch.randn(R, 3, generator=g).to(cuda)
    gw = torch.randn(S, generator=g).to(cuda)
    ga = torch.randn(S, generator=g).to(cuda)

    def run(fused):
        sd = sdf.to(cuda).requires_grad_(True)
        gd = grad.to(cuda).requires_grad_(True)
        qd = q.to(cuda).requires_grad_(True)
        m.s_var.grad = None
        if fused:
            a, w, out = Fn.NeusCompositePacked.apply(vd, sd, gd, dd, fard, outd, offd, m.s_var, qd,
                                                     0.5, 0.1, 1000.0)
        else:
            ridx = Fn.packed_ray_index(offd, S)
            a = Fn.NeusAlphaPacked.apply(vd, sd, gd, dd, fard, outd, offd, ridx, m.s_var, 0.5, 0.1,
                                         1000.0)
            w, out = Fn.CompositePacked.apply(a, qd, offd)
        ((out * go).sum() + (w * gw).sum() + (a * ga).sum()).backward()
        return [t.detach().clone() for t in (a, w, out, sd.grad, gd.grad, qd.grad)], \
            m.s_var.grad.detach().clone()

    old = (_lib.call("na_tune_set", b"packed.rays_per_wg", rb),
           _lib.call("na_tune_set", b"packed.wg_per_cu", wpc))
    try:
        got, gs = run(True)
    finally:
        _lib.call("na_tune_set", b"packed.rays_per_wg", old[0])
        _lib.call("na_tune_set", b"packed.wg_per_cu", old[1])
    want, ws = run(False)
    assert torch.equal(got[0], want[0]), "alpha"
    for name, x, y in zip(("weights", "rgb", "dsdf", "dgrad", "drgb"), got[1:], want[1:]):
        err = (x - y).abs().max().item()
        assert err <= 1e-5 * y.abs().max().item() + 1e-7, (name, err)
    assert (gs - ws).abs().item() <= 1e-4 * ws.abs().item() + 1e-6
    s64 = m.s_var.detach().cpu().double()
    ref = orr.compute_neus_alphas_packed(v.double(), sdf.double()[:, None], grad.double(),
                                         d.double()[:, None], far.double(), off, s64, 0.5,
                                         m.anneal_end)
    inside = ~outside[torch.repeat_interleave(torch.arange(R), off[1:] - off[:-1])]
    assert (got[0].cpu().double()[inside] - ref.detach()[inside]).abs().max() < 2e-5
