"""Oracle: SDF field, colour field and numerical gradients (torch CPU).  TEST INFRASTRUCTURE.

Restates, op for op in the stated dtype (float32 or float64):
  MLPforNeuralSDF.forward            projects/neuralangelo/utils/mlp.py:55-69
  geometric init                     projects/neuralangelo/utils/mlp.py:71-84
  weight_norm (dim=0)                torch.nn.utils.weight_norm, mlp.py:42-43
  softplus(beta=100, threshold=20)   base.yaml:46-48 via misc.get_activation (misc.py:92-107)
  NeuralSDF.encode / forward / sdf   projects/neuralangelo/utils/modules.py:68-95
  NeuralSDF.compute_gradients taps=4 projects/neuralangelo/utils/modules.py:157-175
  NeuralRGB.forward (mode idr)       projects/neuralangelo/utils/modules.py:209-220
  MLPwithSkipConnection              projects/nerf/utils/nerf_util.py:158-196
  get_spherical_harmonics levels=3   projects/neuralangelo/utils/spherical_harmonics.py:47-71

Pinned by (tests/test_oracle_field.py): SH constants against their closed forms
(1/(2 sqrt(pi)), sqrt(3)/(2 sqrt(pi)), ...) and SH orthonormality by quadrature;
softplus against the closed form incl. the threshold branch; weight_norm
against torch.nn.utils.weight_norm; the SDF MLP at geometric init against the
sphere SDF it is built to approximate (|x| - 0.5, mlp.py:79-84 with out_bias
0.5); the 4-tap gradient on the analytic sphere (gradient = x/|x|, Hessian
trace = 2/|x| to O(eps^2)); MLP layers against torch.nn.functional.
"""
import math

import numpy as np
import torch
import torch.nn.functional as torch_F

from . import hashgrid as hg

SOFTPLUS_BETA = 100.0


def softplus(x, beta=SOFTPLUS_BETA, threshold=20.0):
    return torch.where(x * beta > threshold, x, torch.log1p(torch.exp(x * beta)) / beta)


def weight_norm(v, g):
    """W = v * (g / ||v||_row)  (torch _weight_norm with dim=0; g of shape [N,1])."""
    norm = v.norm(dim=1, keepdim=True)
    return v * (g / norm)


def sh_constants():
    """Real SH normalisation constants up to degree 3 from their closed forms."""
    pi = math.pi
    c0 = 1 / (2 * math.sqrt(pi))
    c1 = math.sqrt(3) / (2 * math.sqrt(pi))
    c2 = [math.sqrt(15) / (2 * math.sqrt(pi)), -math.sqrt(15) / (2 * math.sqrt(pi)),
          math.sqrt(5) / (4 * math.sqrt(pi)), -math.sqrt(15) / (2 * math.sqrt(pi)),
          math.sqrt(15) / (4 * math.sqrt(pi))]
    c3 = [-math.sqrt(35 / 2) / (4 * math.sqrt(pi)), math.sqrt(105) / (2 * math.sqrt(pi)),
          -math.sqrt(21 / 2) / (4 * math.sqrt(pi)), math.sqrt(7) / (4 * math.sqrt(pi)),
          -math.sqrt(21 / 2) / (4 * math.sqrt(pi)), math.sqrt(105) / (4 * math.sqrt(pi)),
          -math.sqrt(35 / 2) / (4 * math.sqrt(pi))]
    return c0, c1, c2, c3


def spherical_harmonics3(dirs):
    """[..., 3] unit directions -> [..., 16] (spherical_harmonics.py:47-71 ordering)."""
    c0, c1, c2, c3 = sh_constants()
    x, y, z = dirs.unbind(-1)
    xx, yy, zz = x * x, y * y, z * z
    xy, yz, xz = x * y, y * z, x * z
    vals = [torch.full_like(x, c0), -c1 * y, c1 * z, -c1 * x,
            c2[0] * xy, c2[1] * yz, c2[2] * (2.0 * zz - xx - yy), c2[3] * xz, c2[4] * (xx - yy),
            c3[0] * y * (3 * xx - yy), c3[1] * xy * z, c3[2] * y * (4 * zz - xx - yy),
            c3[3] * z * (2 * zz - 3 * xx - 3 * yy), c3[4] * x * (4 * zz - xx - yy),
            c3[5] * z * (xx - yy), c3[6] * x * (xx - 3 * yy)]
    return torch.stack(vals, dim=-1)


class FieldParams:
    """Plain-tensor parameters of NeuralSDF + NeuralRGB (reference layout).

    sdf: W0v [256, 3+LF], W0g [256,1], b0, W1v [256,256], W1g, b1, ws [1,256], bs [1]
    rgb: list of (v, g, b) for 5 weight-normed linears (278->256 x4 ->3)
    """

    def __init__(self, tensors):
        self.t = dict(tensors)

    def __getitem__(self, k):
        return self.t[k]

    def to(self, dtype):
        return FieldParams({k: v.detach().to(dtype).requires_grad_(True) for k, v in self.t.items()})

    def named(self):
        return self.t.items()


def init_field_params(n_enc, seed=0, geometric=True, rgb_layers=4, hidden=256, rgb_in=None,
                      randomize_encoder_cols=False):
    """Initialisation as mlp.py:71-84 (SDF) and nn.Linear defaults (colour), seeded.

    randomize_encoder_cols breaks the geometric-init zero block W0[:,3:] so that
    the encoder receives gradients (step-0 tables otherwise get exact zeros).
    """
    g = torch.Generator().manual_seed(seed)
    t = {}
    k_in = 3 + n_enc
    w0 = torch.randn(hidden, k_in, generator=g) * math.sqrt(2 / hidden)
    if geometric:
        w0[:, 3:] = 0.0
    if randomize_encoder_cols:
        w0[:, 3:] = torch.randn(hidden, n_enc, generator=g) * math.sqrt(2 / hidden)
    t["W0v"] = w0
    t["W0g"] = w0.norm(dim=1, keepdim=True)
    t["b0"] = torch.zeros(hidden)
    w1 = torch.randn(hidden, hidden, generator=g) * math.sqrt(2 / hidden)
    t["W1v"] = w1
    t["W1g"] = w1.norm(dim=1, keepdim=True)
    t["b1"] = torch.zeros(hidden)
    t["ws"] = torch.randn(1, hidden, generator=g) * 1e-4 + math.sqrt(math.pi / hidden)
    t["bs"] = torch.full((1,), -0.5)
    rgb_in = rgb_in or (3 + 16 + 3 + hidden)
    dims = [rgb_in] + [hidden] * rgb_layers + [3]
    for i, (a, b) in enumerate(zip(dims[:-1], dims[1:])):
        bound = 1 / math.sqrt(a)
        v = (torch.rand(b, a, generator=g) * 2 - 1) * bound
        bias = (torch.rand(b, generator=g) * 2 - 1) * bound
        if i == len(dims) - 2:
            bias = torch.zeros(b)
        t[f"R{i}v"] = v
        t[f"R{i}g"] = v.norm(dim=1, keepdim=True)
        t[f"R{i}b"] = bias
    return FieldParams(t)


def sdf_mlp(inp, P, with_sdf=True, with_feat=True):
    """MLPforNeuralSDF.forward (mlp.py:55-69) for layer_dims [K0, 256, 256]."""
    W0 = weight_norm(P["W0v"], P["W0g"])
    h0 = softplus(inp @ W0.t() + P["b0"])
    sdf = h0 @ P["ws"].t() + P["bs"] if with_sdf else None
    feat = None
    if with_feat:
        W1 = weight_norm(P["W1v"], P["W1g"])
        feat = softplus(h0 @ W1.t() + P["b1"])
    return sdf, feat


class SDFField:
    """NeuralSDF with a hash-grid encoder (modules.py:24-178), oracle version."""

    def __init__(self, lt, table, params, vol_range=(-2.0, 2.0), active_levels=None,
                 dtype=torch.float32, table_dtype=None, weights32=None):
        """weights32: trilinear weights from the float32 cell fraction (the kernel
        algorithm; default for float32) or from float64 positions (default for
        float64, the "exact" oracle; cells are float32-derived either way)."""
        self.lt, self.table, self.P = lt, table, params
        self.weights32 = (dtype == torch.float32) if weights32 is None else weights32
        self.vmin, self.vmax = vol_range
        self.active = lt["n_levels"] if active_levels is None else active_levels
        self.dtype = dtype
        self.table_dtype = table_dtype

    def encode(self, x):
        """modules.py:76-95: normalise, encode, c2f mask, cat([x, enc])."""
        shape = x.shape[:-1]
        xn = ((x - self.vmin) / (self.vmax - self.vmin)).reshape(-1, 3)
        if self.dtype != torch.float64:
            xn = xn.float()
        enc = hg.encode(xn, self.table, self.lt, active_levels=self.active, dtype=self.dtype,
                        table_dtype=self.table_dtype, weights32=self.weights32)
        enc = enc.view(*shape, -1)
        return torch.cat([x.to(self.dtype), enc], dim=-1)

    def forward(self, x, with_sdf=True, with_feat=True):
        return sdf_mlp(self.encode(x), self.P, with_sdf, with_feat)

    def sdf(self, x):
        return self.forward(x, with_sdf=True, with_feat=False)[0]

    def compute_gradients(self, x, eps_world, training=False, sdf=None):
        """4-tap numerical gradient (modules.py:157-175); eps_world = normal_eps."""
        eps = np.float64(eps_world) / np.sqrt(3)
        dt = x.dtype
        k1 = torch.tensor([1, -1, -1], dtype=dt)
        k2 = torch.tensor([-1, -1, 1], dtype=dt)
        k3 = torch.tensor([-1, 1, -1], dtype=dt)
        k4 = torch.tensor([1, 1, 1], dtype=dt)
        s1 = self.sdf(x + k1 * eps)
        s2 = self.sdf(x + k2 * eps)
        s3 = self.sdf(x + k3 * eps)
        s4 = self.sdf(x + k4 * eps)
        gradient = (k1 * s1 + k2 * s2 + k3 * s3 + k4 * s4) / (4.0 * eps)
        hessian = None
        if training:
            hxx = ((s1 + s2 + s3 + s4) / 2.0 - 2 * sdf) / eps ** 2
            hessian = torch.cat([hxx, hxx, hxx], dim=-1) / 3.0
        return gradient, hessian


def rgb_mlp(inp, P, n_layers=5):
    """MLPwithSkipConnection.forward (nerf_util.py:189-196), relu hidden, no skip."""
    h = inp
    for i in range(n_layers):
        W = weight_norm(P[f"R{i}v"], P[f"R{i}g"])
        h = h @ W.t() + P[f"R{i}b"]
        if i != n_layers - 1:
            h = torch.relu(h)
    return h


def neural_rgb(points, normals, rays_unit, feats, P):
    """NeuralRGB.forward (modules.py:209-220), mode idr, SH levels 3, no appearance."""
    view_enc = spherical_harmonics3(rays_unit)
    inp = torch.cat([points, view_enc, normals, feats], dim=-1)
    return torch.sigmoid(rgb_mlp(inp, P))


# ---------------------------------------------------------------------------
# background NeRF (projects/neuralangelo/utils/modules.py:232-311)
# ---------------------------------------------------------------------------
def positional_encoding(x, num_freq_bases):
    """nerf_util.py:134-147: per component, sin then cos of x * 2^l * pi, l < L."""
    freq = 2.0 ** torch.arange(num_freq_bases, dtype=x.dtype) * np.pi
    spectrum = x[..., None] * freq
    return torch.stack([spectrum.sin(), spectrum.cos()], dim=-2).view(*x.shape[:-1], -1)


def background_nerf(points, rays_unit, Wf, bf, Wr, br, levels=10, skip=(4,)):
    """BackgroundNeRF.forward (modules.py:264-280) with encode (:282-296): inverted
    sphere q = [p/|p|, 1/|p|], input [q, PE(q)], feature MLP with skip concat
    (nerf_util.py:186-196, ReLU), density = softplus(out[0]) (beta 1), feature =
    relu(out[1:]), colour MLP on [feature, SH16(view)] ending in a sigmoid.
    Wf, bf / Wr, br: the layers' weights and biases (reference shapes)."""
    n = points.norm(dim=-1, keepdim=True)
    q = torch.cat([points / n, 1.0 / n], dim=-1)
    inp = torch.cat([q, positional_encoding(q, levels)], dim=-1)
    h = inp
    for li, (W, b) in enumerate(zip(Wf, bf)):
        if li in skip:
            h = torch.cat([h, inp], dim=-1)
        h = h @ W.t() + b
        if li != len(Wf) - 1:
            h = h.relu()
    density = torch.nn.functional.softplus(h[..., 0])
    feat = h[..., 1:].relu()
    x = torch.cat([feat, spherical_harmonics3(rays_unit)], dim=-1)
    for li, (W, b) in enumerate(zip(Wr, br)):
        x = x @ W.t() + b
        if li != len(Wr) - 1:
            x = x.relu()
    return x.sigmoid(), density
