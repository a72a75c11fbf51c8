"""The background NeRF restatement of the oracle pinned by closed forms: positional
encoding layout, the inverted-sphere input reaching the skip layer, density and colour
activations, background sample placement and volume-rendering alphas."""
import math

import numpy as np
import torch

from oracle import field as of
from oracle import render as orr


def test_positional_encoding_layout():
    x = torch.tensor([[0.25, -0.5]], dtype=torch.float64)
    got = of.positional_encoding(x, 3)[0]
    want = []
    for v in (0.25, -0.5):
        want += [math.sin(v * math.pi * 2 ** l) for l in range(3)]
        want += [math.cos(v * math.pi * 2 ** l) for l in range(3)]
    assert torch.allclose(got, torch.tensor(want, dtype=torch.float64), atol=1e-15)


def _zero_net(levels=2, hid=8, hid_rgb=4):
    d_in = 4 + 8 * levels
    dims = [d_in] + [hid] * 7 + [hid + 1]
    Wf, bf = [], []
    for li in range(8):
        k_in = dims[li] + (d_in if li == 4 else 0)
        Wf.append(torch.zeros(dims[li + 1], k_in, dtype=torch.float64))
        bf.append(torch.zeros(dims[li + 1], dtype=torch.float64))
    Wr = [torch.zeros(hid_rgb, hid + 16, dtype=torch.float64), torch.zeros(3, hid_rgb, dtype=torch.float64)]
    br = [torch.zeros(hid_rgb, dtype=torch.float64), torch.zeros(3, dtype=torch.float64)]
    return d_in, hid, Wf, bf, Wr, br


def test_skip_layer_reads_the_inverted_sphere_input():
    """Layer 4 reads 1/|p| through the concatenated input (its columns after the
    hidden ones); layers 5-6 pass unit 0 on; the last layer's row 0 is the density."""
    d_in, hid, Wf, bf, Wr, br = _zero_net()
    Wf[4][0, hid + 3] = 1.0  # q[3] = 1 / |p|
    Wf[5][0, 0] = 1.0
    Wf[6][0, 0] = 1.0
    Wf[7][0, 0] = 2.0
    bf[7][0] = -0.1
    br[1][:] = torch.tensor([0.0, 1.0, -2.0], dtype=torch.float64)
    p = torch.tensor([[0.0, 3.0, 4.0], [2.0, 0.0, 0.0]], dtype=torch.float64)
    rays = torch.nn.functional.normalize(torch.randn(2, 3, dtype=torch.float64), dim=-1)
    rgb, density = of.background_nerf(p, rays, Wf, bf, Wr, br, levels=2)
    inv = torch.tensor([0.2, 0.5], dtype=torch.float64)
    assert torch.allclose(density, torch.nn.functional.softplus(2.0 * inv - 0.1), atol=1e-15)
    assert torch.allclose(rgb, torch.sigmoid(br[1]).expand(2, 3), atol=1e-15)


def test_background_samples_and_alphas():
    far = torch.tensor([[2.0], [1.5]], dtype=torch.float64)
    d = orr.sample_dists_background(far, 4)
    inv = 1.0 - (torch.arange(4, dtype=torch.float64) + 0.5) / 4
    assert torch.allclose(d, far / (inv + 1e-5), atol=0)
    assert (d[:, 1:] > d[:, :-1]).all() and (d[:, 0] > far[:, 0]).all()
    sigma = torch.tensor([[0.5, 1.0, 0.0, 2.0]], dtype=torch.float64)
    dd = torch.tensor([[1.0, 1.5, 3.0, 4.0]], dtype=torch.float64)[..., None]
    a = orr.volume_rendering_alphas_dist(sigma, dd)
    want = 1 - np.exp(-np.array([0.5 * 0.5, 1.0 * 1.5, 0.0, 2.0 * (1e10 - 4.0)]))
    assert torch.allclose(a[0], torch.tensor(want), atol=1e-15)
