"""The packed (ragged-ray) restatements of the oracle against the per-ray dense
oracle: every ray evaluated on its own (no padding), and equal counts against the
dense [R, N] form."""
import torch

from oracle import render as orr


def _ragged(R=37, seed=0):
    g = torch.Generator().manual_seed(seed)
    counts = torch.randint(0, 12, (R,), generator=g)
    counts[3] = 0
    counts[5] = 1
    off = torch.zeros(R + 1, dtype=torch.int32)
    off[1:] = counts.cumsum(0).to(torch.int32)
    return g, counts, off


def test_packed_compositing_equals_per_ray_oracle():
    g, counts, off = _ragged()
    S = int(off[-1])
    a = torch.rand(S, generator=g, dtype=torch.float64)
    a[::7] = 1.0
    q = torch.randn(S, 3, generator=g, dtype=torch.float64)
    w = orr.alpha_compositing_weights_packed(a, off)
    out = orr.composite_packed(q, w, off)
    for r in range(counts.numel()):
        b, e = int(off[r]), int(off[r + 1])
        if e == b:
            assert (out[r] == 0).all()
            continue
        wr = orr.alpha_compositing_weights(a[b:e][None])[0]
        assert torch.equal(w[b:e], wr)
        assert torch.allclose(out[r], orr.composite(q[b:e][None], wr[None])[0], rtol=0, atol=1e-14)


def test_packed_neus_alphas_equal_per_ray_oracle():
    g, counts, off = _ragged(seed=1)
    R, S = counts.numel(), int(off[-1])
    v = torch.nn.functional.normalize(torch.randn(R, 3, generator=g, dtype=torch.float64), dim=-1)
    sdf = torch.randn(S, 1, generator=g, dtype=torch.float64) * 0.05
    grad = torch.randn(S, 3, generator=g, dtype=torch.float64)
    d = torch.sort(torch.rand(S, generator=g, dtype=torch.float64))[0][:, None] + 1.0
    far = torch.full((R, 1), 2.5, dtype=torch.float64)
    s_var = torch.tensor(3.0, dtype=torch.float64)
    got = orr.compute_neus_alphas_packed(v, sdf, grad, d, far, off, s_var, 0.05, 0.1)
    for r in range(R):
        b, e = int(off[r]), int(off[r + 1])
        if e == b:
            continue
        ref = orr.compute_neus_alphas(v[r:r + 1], sdf[b:e][None], grad[b:e][None], d[b:e][None],
                                      far[r].view(1, 1, 1), s_var, 0.05, 0.1)[0]
        # float64; torch's vectorised CPU reductions differ in the last bit by batch shape
        assert torch.allclose(got[b:e], ref, rtol=0, atol=1e-12)


def test_packed_with_equal_counts_is_the_dense_form():
    g = torch.Generator().manual_seed(2)
    R, N = 9, 13
    off = torch.arange(R + 1, dtype=torch.int32) * N
    a = torch.rand(R * N, generator=g, dtype=torch.float64)
    assert torch.equal(orr.alpha_compositing_weights_packed(a, off).view(R, N, 1),
                       orr.alpha_compositing_weights(a.view(R, N)))


def test_ragged_batch_counts_span_the_configured_range():
    c, v, off, d = orr.ragged_batch(2048, seed=5)
    counts = (off[1:] - off[:-1]).long()
    assert counts.min() == 0 and counts.max() == 192
    assert ((counts == 0) | (counts >= 32)).all()  # resampled rays carry the 32 fine samples
    near, far, outside = orr.get_dist_bounds(c, v)
    ridx = torch.repeat_interleave(torch.arange(2048), counts)
    assert (d >= near[ridx, 0] - 1e-6).all() and (d <= far[ridx, 0] + 1e-6).all()
    for r in range(0, 2048, 97):
        seg = d[int(off[r]):int(off[r + 1])]
        assert (seg[1:] >= seg[:-1]).all()
