"""Pins the rendering oracle (oracle/render.py) against closed forms, brute force and
library routines (CPU only)."""
import math

import numpy as np
import pytest
import torch

from oracle import render as orr

D = torch.float64


def test_sequential_reductions_match_torch():
    x = torch.rand(7, 33, generator=torch.Generator().manual_seed(0), dtype=D)
    assert torch.allclose(orr.seq_sum(x), x.sum(-1), rtol=1e-14)
    assert torch.allclose(orr.seq_cumsum(x), x.cumsum(-1), rtol=1e-14)
    assert torch.allclose(orr.seq_cumprod(x), x.cumprod(-1), rtol=1e-13)


def test_sphere_intersection_analytic():
    c = torch.tensor([[0.0, 0.0, -3.0], [0.0, 0.5, -3.0], [0.0, 2.0, -3.0], [0.0, 0.0, 0.2]], dtype=D)
    v = torch.tensor([[0.0, 0.0, 1.0]] * 4, dtype=D)
    near, far, outside = orr.get_dist_bounds(c, v)
    h = math.sqrt(1 - 0.25)
    assert torch.allclose(near[:2, 0], torch.tensor([2.0, 3 - h], dtype=D))
    assert torch.allclose(far[:2, 0], torch.tensor([4.0, 3 + h], dtype=D))
    assert outside[:, 0].tolist() == [False, False, True, False]
    assert near[2, 0] == 1.0 and far[2, 0] == 1.2                 # missed ray
    assert near[3, 0] == 0.0 and abs(far[3, 0] - 0.8) < 1e-15     # inside: near clamped


def test_sample_dists_midpoints_and_strata():
    near = torch.tensor([[1.0], [2.0]], dtype=D)
    far = torch.tensor([[3.0], [2.5]], dtype=D)
    d = orr.sample_dists(near, far, 4)[..., 0]
    assert torch.allclose(d, torch.tensor([[1.25, 1.75, 2.25, 2.75],
                                           [2.0625, 2.1875, 2.3125, 2.4375]], dtype=D))
    u = torch.rand(2, 64, generator=torch.Generator().manual_seed(1), dtype=D)
    d = orr.sample_dists(near, far, 64, u)[..., 0]
    k = torch.arange(64, dtype=D)
    lo = near + k / 64 * (far - near)
    hi = near + (k + 1) / 64 * (far - near)
    assert ((d >= lo - 1e-15) & (d <= hi + 1e-15)).all()


def test_compositing_weights_closed_forms():
    a = 0.3
    w = orr.alpha_compositing_weights(torch.full((1, 10), a, dtype=D))[..., 0]
    assert torch.allclose(w, a * (1 - a) ** torch.arange(10, dtype=D))
    assert abs(w.sum().item() - (1 - (1 - a) ** 10)) < 1e-15
    al = torch.zeros(1, 6, dtype=D)
    al[0, 2] = 1.0
    al[0, 4] = 0.7
    w = orr.alpha_compositing_weights(al)[0, :, 0]
    assert w.tolist() == [0.0, 0.0, 1.0, 0.0, 0.0, 0.0]  # all weight on the opaque sample


def test_composite_is_weighted_sum():
    g = torch.Generator().manual_seed(2)
    q = torch.rand(3, 5, 4, generator=g, dtype=D)
    w = torch.rand(3, 5, 1, generator=g, dtype=D)
    ref = torch.stack([sum(q[r, i] * w[r, i, 0] for i in range(5)) for r in range(3)])
    assert torch.allclose(orr.composite(q, w), ref)


def _brute_pdf_samples(bins, weights, nf):
    """Inverse-CDF sampling written out with Python scalars (searchsorted right=True)."""
    out = []
    for b, wt in zip(bins[..., 0].tolist(), weights.tolist()):
        s = max(sum(abs(x) for x in wt), 1e-12)
        cdf = [0.0]
        for x in wt:
            cdf.append(cdf[-1] + x / s)
        row = []
        for j in range(nf):
            u = (j + 0.5) / nf
            idx = sum(1 for c in cdf if c <= u)
            lo, hi = max(idx - 1, 0), min(idx, len(cdf) - 1)
            t = (u - cdf[lo]) / (cdf[hi] - cdf[lo] + 1e-8)
            row.append(b[lo] + t * (b[hi] - b[lo]))
        out.append(row)
    return torch.tensor(out, dtype=D)


def test_pdf_sampling_uniform_delta_and_brute_force():
    nb, nf = 17, 16
    bins = torch.linspace(0, 1, nb, dtype=D).view(1, nb, 1)
    d = orr.sample_dists_from_pdf(bins, torch.ones(1, nb - 1, dtype=D), nf)[0, :, 0]
    assert torch.allclose(d, (torch.arange(nf, dtype=D) + 0.5) / nf, atol=1e-7)
    w = torch.zeros(1, nb - 1, dtype=D)
    w[0, 5] = 2.0
    d = orr.sample_dists_from_pdf(bins, w, nf)[0, :, 0]
    assert ((d >= bins[0, 5, 0]) & (d <= bins[0, 6, 0])).all()        # collapsed into bin 5
    assert torch.allclose(d, bins[0, 5, 0] + (torch.arange(nf, dtype=D) + 0.5) / nf / 16,
                          atol=1e-7)
    g = torch.Generator().manual_seed(3)
    bins = torch.sort(torch.rand(20, nb, generator=g, dtype=D) * 3 + 1, dim=-1).values[..., None]
    w = torch.rand(20, nb - 1, generator=g, dtype=D) ** 3
    w[3] = 0.0  # all-zero weights row
    d = orr.sample_dists_from_pdf(bins, w, nf)[..., 0]
    assert torch.allclose(d, _brute_pdf_samples(bins, w, nf), rtol=0, atol=1e-12)


def test_neus_alpha_telescopes_on_a_plane():
    """A plane crossed head-on (sdf = s0 - d, gradient = -v): the per-section
    transmittances telescope: 1 - alpha_i = (Phi(next_i) + e)/(Phi(prev_i) + e) with
    next_i = prev_{i+1}, so prod(1 - alpha_i) = (Phi(s(d_N + h)) + e) / (Phi(s(d_0 - h)) + e),
    Phi = sigmoid(inv_s .), e = 1e-5, once the cosine anneal is complete."""
    n, dd = 32, 0.05
    dists = (1.0 + dd * torch.arange(n, dtype=D)).view(1, n, 1)
    s0 = 1.8
    sdfs = s0 - dists
    v = torch.tensor([[0.0, 0.0, 1.0]], dtype=D)
    grads = -v.view(1, 1, 3).expand(1, n, 3)
    far = dists[:, -1:, :] + dd
    s_var = torch.tensor(math.log(40.0), dtype=D)
    alpha = orr.compute_neus_alphas(v, sdfs, grads, dists, far, s_var, 1.0, 0.1)
    T = torch.prod(1 - alpha)
    sig = lambda x: 1 / (1 + math.exp(-40.0 * x))  # noqa: E731
    expect = ((sig(s0 - (dists[0, -1, 0].item() + dd / 2)) + 1e-5)
              / (sig(s0 - (dists[0, 0, 0].item() - dd / 2)) + 1e-5))
    assert abs(T.item() - expect) < 1e-12
    assert ((alpha >= 0) & (alpha <= 1)).all()


def test_hierarchical_alpha_telescopes_on_a_plane():
    """model.py:251-264 on sdf = 0.9 - d: the robust cos estimate is
    c = -dd/(dd + 1e-5), so neighbouring sections miss each other by
    delta = dd (1 - |c|) ~ 1e-5 and the product of (1 - alpha) equals the
    telescoped ratio (Phi(mid_last + c dd/2) + e)/(Phi(mid_0 - c dd/2) + e) up to a
    factor exp(+-(n-2) inv_s delta) (|d log(Phi + e)/dx| <= inv_s)."""
    n, dd, inv_s = 24, 0.04, 64.0
    dists = (0.5 + dd * torch.arange(n, dtype=D)).view(1, n, 1)
    sdfs = 0.9 - dists
    a = orr.hierarchical_alphas(dists, sdfs, inv_s)
    T = torch.prod(1 - a).item()
    sig = lambda x: 1 / (1 + math.exp(-inv_s * x))  # noqa: E731
    s_ = sdfs[0, :, 0]
    c = -dd / (dd + 1e-5)
    first = ((s_[0] + s_[1]) / 2).item() - c * dd / 2
    last = ((s_[-2] + s_[-1]) / 2).item() + c * dd / 2
    expect = (sig(last) + 1e-5) / (sig(first) + 1e-5)
    delta = dd * (1 - abs(c))
    assert abs(math.log(T) - math.log(expect)) <= (n - 2) * inv_s * delta * 1.01
    # and a sign or cosine mistake is far outside that band
    assert (n - 2) * inv_s * delta < 0.02


def test_iter_cos_anneal_endpoints():
    c = torch.tensor([-1.0, -0.5, 0.0, 0.5, 1.0], dtype=D)
    assert torch.allclose(orr.get_iter_cos(c, 0.0, 0.1), -(-c * 0.5 + 0.5).relu())
    assert torch.allclose(orr.get_iter_cos(c, 0.2, 0.1), -(-c).relu())


def test_losses_hand_values():
    g = torch.tensor([[[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 0.0], [float("nan"), 0.0, 0.0]]],
                     dtype=D)
    outside = torch.tensor([[False]])
    assert abs(orr.eikonal_loss(g, outside).item() - (0 + 1 + 1 + 0) / 4) < 1e-15
    assert orr.eikonal_loss(g, torch.tensor([[True]])).item() == 0.0
    h = torch.tensor([[[1.0, -3.0, 0.5], [0.0, 0.0, 0.0], [float("inf"), 0.0, 0.0],
                       [2.0, 2.0, 2.0]]], dtype=D)
    assert abs(orr.curvature_loss(h, outside).item() - (1.5 + 0 + 0 + 6) / 4) < 1e-15


def test_camera_rays():
    """camera.py:283-311: pixel centres through the inverse intrinsics, cam -> world."""
    f, W, H = 20.0, 4, 3
    intr = torch.tensor([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1]], dtype=torch.float32)
    pose = torch.tensor([[1.0, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]])
    c, r = orr.get_center_and_ray(pose, intr, (H, W))
    assert torch.allclose(c, torch.zeros_like(c))
    px = torch.tensor([0.5, 0.5])
    assert torch.allclose(r[0], torch.tensor([(px[0] - W / 2) / f, (px[1] - H / 2) / f, 1.0]))
    # rotate 90 deg about z and translate: centre = -R^T t
    R = torch.tensor([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]])
    t = torch.tensor([0.5, -1.0, 3.0])
    pose = torch.cat([R, t[:, None]], 1)
    c, r = orr.get_center_and_ray(pose, intr, (H, W))
    assert torch.allclose(c[0], -R.t() @ t)
    cam = torch.tensor([(0.5 - W / 2) / f, (0.5 - H / 2) / f, 1.0])
    assert torch.allclose(r[0], R.t() @ cam, atol=1e-6)


def test_stable_merge_keeps_every_sample_in_order():
    g = torch.Generator().manual_seed(4)
    c = torch.randn(16, 3, generator=g, dtype=D)
    c = c / c.norm(dim=-1, keepdim=True) * 3
    v = -c / 3
    near, far, _ = orr.get_dist_bounds(c, v)
    sdf_fn = lambda p: p.norm(dim=-1, keepdim=True) - 0.5  # noqa: E731
    d = orr.sample_dists_all(c, v, near, far, sdf_fn, 64, 16, 2)[..., 0]
    assert d.shape == (16, 96) and (d[:, 1:] >= d[:, :-1]).all()
    coarse = orr.sample_dists(near, far, 64)[..., 0]
    for r in range(16):  # the coarse samples are a sub-multiset of the merged ones
        merged = d[r].tolist()
        for x in coarse[r].tolist():
            merged.remove(x)
    # fine samples concentrate around the sphere hit (d = 2.5 along these rays)
    assert ((d - 2.5).abs() < 0.1).sum(-1).min() >= 16


def _plane_field(seed=0, hidden=256, L=4, F=2, b=(0.3, -1.2, 2.0)):
    """A field whose SDF is the plane s(x) = a.x + c0 and whose colour is the constant
    sigmoid(b): zero hash table; first-layer biases of 2 with |W0 x| < 1 inside the
    unit ball, so beta z > 20 and the thresholded softplus is the identity; colour
    layers with weight_g = 0 (zero weights) and last bias b."""
    from oracle import field as of
    from oracle import hashgrid as hg
    g = torch.Generator().manual_seed(seed)
    lt = hg.level_table(L, F, 10, 4, 1.5)
    n_enc = L * F
    W0 = torch.zeros(hidden, 3 + n_enc, dtype=D)
    W0[:, :3] = (torch.rand(hidden, 3, generator=g, dtype=D) * 2 - 1) * 0.3
    W0[:, 3:] = torch.randn(hidden, n_enc, generator=g, dtype=D) * 0.1
    b0 = torch.full((hidden,), 2.0, dtype=D)
    ws = torch.randn(1, hidden, generator=g, dtype=D) * 0.2
    a = (ws @ W0[:, :3])[0]
    c0 = -0.15
    bs = torch.tensor([c0], dtype=D) - (ws @ b0)
    t = dict(W0v=W0, W0g=W0.norm(dim=1, keepdim=True), b0=b0,
             W1v=torch.randn(hidden, hidden, generator=g, dtype=D),
             W1g=torch.ones(hidden, 1, dtype=D), b1=torch.zeros(hidden, dtype=D), ws=ws, bs=bs)
    dims = [3 + 16 + 3 + hidden] + [hidden] * 4 + [3]
    for i in range(5):
        t[f"R{i}v"] = torch.randn(dims[i + 1], dims[i], generator=g, dtype=D)
        t[f"R{i}g"] = torch.zeros(dims[i + 1], 1, dtype=D)
        t[f"R{i}b"] = torch.zeros(dims[i + 1], dtype=D)
    t["R4b"] = torch.tensor(b, dtype=D)
    P = of.FieldParams(t)
    table = torch.zeros(lt["n_params"], dtype=D)
    return of.SDFField(lt, table, P, dtype=D), P, a, c0, torch.tensor(b, dtype=D)


@pytest.mark.parametrize("training", [False, True])
def test_render_rays_closed_form_on_a_plane(training):
    """The composed oracle render (model.py:124-190) on a plane SDF with a constant colour
    network, against closed forms that use only the plane, near/far and the NeuS ratio:
    4-tap gradient = a, Hessian = 0, sdfs = a.(c + d v) + c0 (1000 outside).  With the
    64 uniform midpoint samples (spacing h, no resampling) and progress 1, a ray heading
    into the plane has 1 - alpha_i = (Phi(s(d_i + h/2)) + e) / (Phi(s(d_i - h/2)) + e)
    (Phi = sigmoid(inv_s .), e = 1e-5), so the transmittance telescopes up to the last
    section, which ends at far (half a spacing after the last sample):
    T = (Phi(s(far - h)) + e) / (Phi(s(near)) + e) * (Phi(s(far - h/4)) + e) /
    (Phi(s(far - 3h/4)) + e); a ray heading away keeps T = 1.  Then rgb = sigmoid(b)(1 - T),
    the composited gradient is a (1 - T) and depth = sum_i d_i w_i with the telescoped
    weights w_i = (Phi(s(d_i - h/2)) - Phi(s(d_i + h/2))) / (Phi(s(near)) + e) before the
    last sample."""
    fld, P, a, c0, b = _plane_field()
    g = torch.Generator().manual_seed(3)
    R, N = 400, 64
    c = torch.randn(R, 3, generator=g, dtype=D)
    c = c / c.norm(dim=-1, keepdim=True) * 2.5
    v = -c / 2.5 + 0.45 * torch.randn(R, 3, generator=g, dtype=D)
    v = v / v.norm(dim=-1, keepdim=True)
    inv_s = 30.0
    s_var = torch.tensor(math.log(inv_s), dtype=D)
    o = orr.render_rays(fld, P, c, v, s_var, 1.0, 1.0 / 64, training, n_coarse=N, n_hier=0)
    near, far, outside = orr.get_dist_bounds(c, v)
    inside = ~outside[:, 0]
    assert inside.any() and outside.any()
    # SDF values, gradients and Hessian
    pts = c[:, None, :] + v[:, None, :] * o["dists"]
    plane = (pts * a).sum(-1) + c0
    assert torch.allclose(o["sdfs"][inside], plane[inside], rtol=0, atol=1e-12)
    assert (o["sdfs"][~inside] == 1000.0).all()
    assert torch.allclose(o["gradients"], a.expand_as(o["gradients"]), rtol=0, atol=1e-9)
    if training:
        # inside rays: 0; outside rays see the filled value (model.py:163, :166) in the
        # centre term, as in the reference
        assert o["hessians"][inside].abs().max().item() < 1e-6
        assert (o["hessians"][~inside].abs() > 1e6).all()
    # closed-form transmittance
    h = (far - near)[:, 0] / N
    sig = lambda x: torch.sigmoid(inv_s * x)  # noqa: E731
    s_of = lambda d: (c * a).sum(-1) + d * (v * a).sum(-1) + c0  # noqa: E731
    cos = (v * a).sum(-1)
    fr = far[:, 0]
    T = ((sig(s_of(fr - h)) + 1e-5) / (sig(s_of(near[:, 0])) + 1e-5)
         * (sig(s_of(fr - h / 4)) + 1e-5) / (sig(s_of(fr - 3 * h / 4)) + 1e-5))
    T = torch.where(cos < 0, T, torch.ones_like(T))
    opacity = 1 - T
    w = o["weights"][..., 0]
    assert torch.allclose(w.sum(-1)[inside], opacity[inside], rtol=0, atol=1e-9)
    rgb = torch.sigmoid(b) * opacity[:, None]
    assert torch.allclose(o["rgb"][inside], rgb[inside], rtol=0, atol=1e-9)
    assert (cos[inside] < 0).any() and (cos[inside] > 0).any()
    if not training:
        assert torch.allclose(o["opacity"][inside, 0], opacity[inside], rtol=0, atol=1e-9)
        assert torch.allclose(o["gradient"][inside], a * opacity[inside, None], rtol=0, atol=1e-9)
        d = o["dists"][..., 0]
        p0 = sig(s_of(near[:, 0])) + 1e-5
        s2 = lambda t: (c * a).sum(-1)[:, None] + t * cos[:, None] + c0  # noqa: E731
        wi = (sig(s2(d - h[:, None] / 2)) - sig(s2(d + h[:, None] / 2))) / p0[:, None]
        pl, nl = sig(s_of(fr - 3 * h / 4)), sig(s_of(fr - h / 4))
        wi[:, -1] = (pl - nl) / (pl + 1e-5) * (sig(s_of(fr - h)) + 1e-5) / p0
        depth = (d * wi).sum(-1)
        sel = inside & (cos < 0)
        assert torch.allclose(o["depth"][sel, 0], depth[sel], rtol=0, atol=1e-9)
