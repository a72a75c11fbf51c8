"""HIP hash-grid encoder vs the CPU oracle (indices bit-exact, values within tolerance)."""
import numpy as np
import pytest
import torch

from neuralangelo_amd.encoding import Encoding, growth_rate
from oracle import hashgrid as hg

pytestmark = pytest.mark.gpu

G = growth_rate(5, 11, 16)


def _cfg(log2_T, F, levels=16, base=32, scale=G):
    return dict(otype="HashGrid", n_levels=levels, n_features_per_level=F,
                log2_hashmap_size=log2_T, base_resolution=base, per_level_scale=scale)


def _points(n, seed, lo=0.0, hi=1.0):
    g = torch.Generator().manual_seed(seed)
    x = torch.rand(n, 3, generator=g) * (hi - lo) + lo
    return x.float()


def _edge_points():
    vals = [0.0, 1.0, 0.5, 1e-7, 1 - 1e-7, 0.25, 0.75, -0.01, 1.01]
    pts = torch.tensor([[a, b, c] for a in vals for b in vals[:4] for c in vals[4:]],
                       dtype=torch.float32)
    return pts


@pytest.mark.parametrize("log2_T,F", [(22, 8), (19, 4), (14, 2)])
def test_corner_indices_and_weights_bit_exact(cuda, log2_T, F):
    enc = Encoding(3, _cfg(log2_T, F)).to(cuda)
    x = torch.cat([_points(3001, 1), _edge_points()])
    idx, w = enc.corners(x.to(cuda))
    lt = hg.level_table(16, F, log2_T, 32, G)
    idx_o, w_o = hg.corners(x.numpy(), lt)
    np.testing.assert_array_equal(idx.cpu().numpy().astype(np.int64), idx_o)
    np.testing.assert_array_equal(w.cpu().numpy(), w_o)


@pytest.mark.parametrize("log2_T,F,tdt", [(22, 8, torch.float32), (19, 4, torch.float32),
                                          (22, 8, torch.bfloat16), (14, 2, torch.float32),
                                          (12, 1, torch.float32)])
def test_forward_matches_oracle(cuda, log2_T, F, tdt):
    enc = Encoding(3, _cfg(log2_T, F), seed=7, dtype=tdt)
    with torch.no_grad():  # table values of realistic (trained) magnitude
        g = torch.Generator().manual_seed(3)
        enc.params.copy_(torch.randn(enc.n_params, generator=g) * 0.1)
    enc = enc.to(cuda)
    x = torch.cat([_points(5003, 2), _edge_points()])
    out = enc(x.to(cuda)).float().cpu()
    lt = hg.level_table(16, F, log2_T, 32, G)
    p_cpu = enc.params.detach().cpu()
    tdt_o = torch.bfloat16 if tdt == torch.bfloat16 else None
    ref32 = hg.encode(x, p_cpu.double(), lt, dtype=torch.float64, table_dtype=tdt_o,
                      weights32=True)
    ref64 = hg.encode(x, p_cpu.double(), lt, dtype=torch.float64, table_dtype=tdt_o)
    bound = hg.abs_bound(x, p_cpu, lt).double()
    out = out.double()
    if tdt == torch.float32:
        # same float32 weights as the oracle: only accumulation order differs
        err = (out - ref32).abs()
        assert (err <= 1e-5 * bound + 1e-30).all(), float((err / (bound + 1e-30)).max())
        # vs FP64 weights: float32 rounding of pos gives |dt| <= ulp(pos)/2 <= 2^-13
        # (pos < 2^12); d(weight)/dt is within [-1, 1], so each of the 3 dims moves
        # the interpolant by at most |dt| * sum over the 8 corners of |v|
        err64 = (out - ref64).abs()
        idx, _ = hg.corners(x.numpy(), lt)
        tab = p_cpu.double().abs().view(-1, F)
        if tdt_o is not None:
            tab = p_cpu.to(tdt_o).double().abs().view(-1, F)
        csum = tab[torch.from_numpy(idx).view(-1)].view(-1, 16, 8, F).sum(2).reshape(-1, 16 * F)
        assert (err64 <= 3 * 2.0 ** -13 * csum + 1e-5 * bound + 1e-30).all()
    else:  # BF16 output rounding (2^-9 relative) on top of exact BF16 table values
        err = (out - ref32).abs()
        assert (err <= 2 ** -8 * ref32.abs() + 1e-5 * bound + 1e-30).all()


def test_active_levels_zero_fill(cuda):
    enc = Encoding(3, _cfg(19, 4)).to(cuda)
    enc.active_levels = 5
    out = enc(_points(1000, 3).to(cuda))
    assert (out[:, 5 * 4:] == 0).all()
    assert (out[:, :5 * 4] != 0).any()


def test_empty_input(cuda):
    enc = Encoding(3, _cfg(19, 4)).to(cuda)
    out = enc(torch.zeros(0, 3, device=cuda))
    assert out.shape == (0, 64)


@pytest.mark.parametrize("log2_T,F,active", [(22, 8, 16), (19, 4, 16), (19, 4, 9), (14, 2, 16)])
def test_backward_matches_oracle_autograd(cuda, log2_T, F, active):
    enc = Encoding(3, _cfg(log2_T, F), seed=11).to(cuda)
    enc.active_levels = active
    x = torch.cat([_points(4000, 4), _edge_points()])
    g = torch.Generator().manual_seed(5)
    dy = torch.randn(x.shape[0], 16 * F, generator=g)
    out = enc(x.to(cuda))
    out.backward(dy.to(cuda))
    grad = enc.params.grad.cpu().double()
    lt = hg.level_table(16, F, log2_T, 32, G)
    p = enc.params.detach().cpu().double().requires_grad_(True)
    ref_out = hg.encode(x, p, lt, active_levels=active, dtype=torch.float64, weights32=True)
    (gref,) = torch.autograd.grad((ref_out * dy.double()).sum(), p)
    # scale of each entry's gradient: sum of |w * dy| contributions
    lt_w = hg.corners(x.numpy(), lt)
    idx, w = torch.from_numpy(lt_w[0]), torch.from_numpy(lt_w[1]).double()
    mask = torch.zeros(16, dtype=torch.float64); mask[:active] = 1
    contrib = (w[..., None] * mask[None, :, None, None] *
               dy.double().view(-1, 16, 1, F).abs())  # [P,L,8,F]
    scale = torch.zeros(lt["n_entries"], F, dtype=torch.float64)
    scale.index_add_(0, idx.view(-1), contrib.reshape(-1, F))
    scale = scale.view(-1)
    err = (grad - gref).abs()
    assert (err <= 1e-5 * scale + 1e-30).all(), float((err / (scale + 1e-30)).max())


def test_bf16_table_refreshes_after_update(cuda):
    enc = Encoding(3, _cfg(14, 2), dtype=torch.bfloat16).to(cuda)
    x = _points(100, 6).to(cuda)
    a = enc(x).float()
    with torch.no_grad():
        enc.params.add_(0.5)
    b = enc(x).float()
    assert not torch.equal(a, b)
