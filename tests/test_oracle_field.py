"""Pins the field oracle (oracle/field.py) against closed forms, library routines
and the reference's published constants (CPU only)."""
import math

import numpy as np
import torch

from oracle import field as of
from oracle import hashgrid as hg

# real SH constants as printed in projects/neuralangelo/utils/spherical_harmonics.py:16-33
REF_C0 = 0.28209479177387814
REF_C1 = 0.4886025119029199
REF_C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
          0.5462742152960396]
REF_C3 = [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
          -0.4570457994644658, 1.445305721320277, -0.5900435899266435]


def test_sh_constants_match_reference_values():
    c0, c1, c2, c3 = of.sh_constants()
    assert abs(c0 - REF_C0) < 1e-15 and abs(c1 - REF_C1) < 1e-15
    assert np.allclose(c2, REF_C2, rtol=0, atol=1e-15)
    assert np.allclose(c3, REF_C3, rtol=0, atol=1e-15)


def test_sh_basis_is_orthonormal():
    """Gauss-Legendre in cos(theta) x uniform phi integrates degree <= 6 products exactly."""
    nt, nphi = 16, 32
    ct, wt = np.polynomial.legendre.leggauss(nt)
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    C, PH = np.meshgrid(ct, phi, indexing="ij")
    S = np.sqrt(1 - C * C)
    dirs = torch.tensor(np.stack([S * np.cos(PH), S * np.sin(PH), C], -1), dtype=torch.float64)
    Y = of.spherical_harmonics3(dirs).numpy()            # [nt, nphi, 16]
    w = (wt[:, None] * np.full(nphi, 2 * np.pi / nphi)[None, :])
    G = np.einsum("tp,tpi,tpj->ij", w, Y, Y)
    assert np.abs(G - np.eye(16)).max() < 1e-12


def test_sh_sign_convention_on_axes():
    """Degree-1 terms are (-y, z, -x) * C1 (the reference's sign convention)."""
    d = torch.eye(3, dtype=torch.float64)
    Y = of.spherical_harmonics3(d)
    assert torch.allclose(Y[:, 1:4], torch.tensor([[0, 0, -REF_C1], [-REF_C1, 0, 0],
                                                   [0, REF_C1, 0]], dtype=torch.float64))


def test_softplus_matches_torch():
    x = torch.linspace(-1, 1, 20001, dtype=torch.float64)
    ref = torch.nn.functional.softplus(x, beta=100, threshold=20)
    assert torch.allclose(of.softplus(x), ref, rtol=1e-14, atol=1e-16)
    xf = x.float()
    assert torch.equal(of.softplus(xf)[xf * 100 > 20], xf[xf * 100 > 20])  # linear branch exact
    assert abs(of.softplus(torch.zeros(1, dtype=torch.float64)).item() - math.log(2) / 100) < 1e-17


def test_weight_norm_matches_torch():
    g = torch.Generator().manual_seed(0)
    lin = torch.nn.Linear(7, 5).double()
    lin = torch.nn.utils.weight_norm(lin)
    with torch.no_grad():
        lin.weight_v.copy_(torch.randn(5, 7, generator=g, dtype=torch.float64))
        lin.weight_g.copy_(torch.rand(5, 1, generator=g, dtype=torch.float64) + 0.5)
    x = torch.randn(3, 7, generator=g, dtype=torch.float64)
    lin(x)  # recompute weight
    W = of.weight_norm(lin.weight_v.detach(), lin.weight_g.detach())
    assert torch.allclose(W, lin.weight.detach(), rtol=1e-14)


def _torch_sdf_mlp(P, k_in):
    """MLPforNeuralSDF with torch.nn modules (weight_norm + Softplus(beta=100))."""
    l0 = torch.nn.utils.weight_norm(torch.nn.Linear(k_in, 256).double())
    l1 = torch.nn.utils.weight_norm(torch.nn.Linear(256, 256).double())
    ls = torch.nn.Linear(256, 1).double()
    with torch.no_grad():
        l0.weight_v.copy_(P["W0v"]); l0.weight_g.copy_(P["W0g"]); l0.bias.copy_(P["b0"])
        l1.weight_v.copy_(P["W1v"]); l1.weight_g.copy_(P["W1g"]); l1.bias.copy_(P["b1"])
        ls.weight.copy_(P["ws"]); ls.bias.copy_(P["bs"])
    act = torch.nn.Softplus(beta=100)

    def f(x):
        h = act(l0(x))
        return ls(h), act(l1(h))
    return f


def test_sdf_mlp_matches_torch_modules():
    P = of.init_field_params(16, seed=1, randomize_encoder_cols=True).to(torch.float64)
    x = torch.randn(50, 19, generator=torch.Generator().manual_seed(2), dtype=torch.float64)
    s, f = of.sdf_mlp(x, P)
    s2, f2 = _torch_sdf_mlp(P, 19)(x)
    assert torch.allclose(s, s2, rtol=1e-12, atol=1e-14)
    assert torch.allclose(f, f2, rtol=1e-12, atol=1e-14)


def test_geometric_init_approximates_the_sphere():
    """mlp.py:71-84: at init the SDF is close to |x| - 0.5 (encoder columns are zero).

    In expectation over the weights, sqrt(pi/256) * sum_i relu(w_i . x) = |x|; the
    softplus(beta=100) smoothing adds about +0.07 near the origin and flattens the slope
    slightly, so the fitted line is ~0.94 |x| - 0.42 (zero near |x| = 0.45)."""
    g = torch.Generator().manual_seed(4)
    d = torch.randn(2000, 3, generator=g, dtype=torch.float64)
    d = d / d.norm(dim=-1, keepdim=True)
    r = torch.rand(2000, 1, generator=g, dtype=torch.float64) * 0.9 + 0.2
    inp = torch.cat([d * r, torch.zeros(2000, 32, dtype=torch.float64)], -1)
    s = sum(of.sdf_mlp(inp, of.init_field_params(32, seed=k).to(torch.float64),
                       with_feat=False)[0] for k in range(16)) / 16
    A = torch.cat([r, torch.ones_like(r)], 1)
    slope, icpt = torch.linalg.lstsq(A, s).solution.squeeze().tolist()
    assert 0.85 < slope < 1.05 and -0.5 < icpt < -0.38, (slope, icpt)
    far = (r - 0.45).abs() > 0.1
    assert ((s < 0) == (r < 0.45))[far].all()


class _Analytic:
    def __init__(self, f):
        self.sdf = f


def test_four_tap_gradient_exact_on_separable_quadratics():
    """sum_i k_i = 0 and sum_i k_i k_i^T = 4I make the tetrahedral stencil exact for
    f = a.x + sum_d b_d x_d^2: gradient a + 2 b x, Hessian trace 2 sum b."""
    g = torch.Generator().manual_seed(5)
    a = torch.randn(3, generator=g, dtype=torch.float64)
    b = torch.randn(3, generator=g, dtype=torch.float64)

    def f(x):
        return (x @ a + (x * x) @ b)[..., None]
    x = torch.rand(100, 3, generator=g, dtype=torch.float64) * 2 - 1
    grad, hess = of.SDFField.compute_gradients(_Analytic(f), x, 1 / 2048, training=True, sdf=f(x))
    assert torch.allclose(grad, a + 2 * b * x, rtol=0, atol=1e-9)
    assert torch.allclose(hess.sum(-1), torch.full((100,), 2 * b.sum().item(), dtype=torch.float64),
                          rtol=0, atol=1e-5)
    assert torch.allclose(hess[:, 0], hess[:, 1]) and torch.allclose(hess[:, 0], hess[:, 2])


def test_four_tap_gradient_on_sphere():
    g = torch.Generator().manual_seed(6)
    d = torch.randn(200, 3, generator=g, dtype=torch.float64)
    x = d / d.norm(dim=-1, keepdim=True) * (0.3 + torch.rand(200, 1, generator=g,
                                                              dtype=torch.float64))
    f = lambda p: p.norm(dim=-1, keepdim=True) - 0.5  # noqa: E731
    eps = 1 / 512
    grad, hess = of.SDFField.compute_gradients(_Analytic(f), x, eps, training=True, sdf=f(x))
    r = x.norm(dim=-1)
    # the stencil's leading error is eps_t * H_ab in component c ({a,b,c} distinct);
    # for the sphere |H_ab| <= 1/(2r), so |error| <= sqrt(3) eps_t / (2r) + O(eps_t^2)
    et = eps / math.sqrt(3)
    err = (grad - x / r[:, None]).norm(dim=-1)
    assert (err <= math.sqrt(3) * et / (2 * r) * 1.05 + 10 * et ** 2 / r ** 2).all()
    # trace estimate: leading error 2 eps_t f_xyz (third moments of the stencil),
    # f_xyz = 3xyz/r^5 for f = |x|
    fxyz = 3 * x.prod(dim=-1) / r ** 5
    herr = (hess.sum(-1) - 2 / r).abs()
    assert (herr <= 2 * et * fxyz.abs() * 1.05 + 20 * et ** 2 / r ** 3).all()


def test_encode_normalises_the_volume_range():
    """modules.py:82-83: u = (x - vmin)/(vmax - vmin) with range [-2, 2]; [x, enc] order."""
    lt = hg.level_table(4, 2, 12, 16, 1.5)
    g = torch.Generator().manual_seed(7)
    table = torch.rand(lt["n_params"], generator=g) * 2 - 1
    P = of.init_field_params(8, seed=0)
    fld = of.SDFField(lt, table, P)
    x = torch.rand(64, 3, generator=g) * 3.6 - 1.8
    out = fld.encode(x)
    assert torch.equal(out[:, :3], x)
    ref = hg.encode(((x + 2) / 4).float(), table, lt, dtype=torch.float32, weights32=True)
    assert torch.equal(out[:, 3:], ref)


def test_rgb_mlp_matches_torch_modules():
    """MLPwithSkipConnection (nerf_util.py:158-196): 4 x (weight-normed Linear + ReLU),
    weight-normed Linear to 3, sigmoid applied by NeuralRGB (modules.py:219)."""
    P = of.init_field_params(8, seed=5).to(torch.float64)
    dims = [278, 256, 256, 256, 256, 3]
    layers = []
    for i, (a, b) in enumerate(zip(dims[:-1], dims[1:])):
        lin = torch.nn.utils.weight_norm(torch.nn.Linear(a, b).double())
        with torch.no_grad():
            lin.weight_v.copy_(P[f"R{i}v"]); lin.weight_g.copy_(P[f"R{i}g"])
            lin.bias.copy_(P[f"R{i}b"])
        layers.append(lin)
    g = torch.Generator().manual_seed(6)
    pts = torch.randn(20, 3, generator=g, dtype=torch.float64)
    nrm = torch.nn.functional.normalize(torch.randn(20, 3, generator=g, dtype=torch.float64), dim=-1)
    ray = torch.nn.functional.normalize(torch.randn(20, 3, generator=g, dtype=torch.float64), dim=-1)
    feat = torch.rand(20, 256, generator=g, dtype=torch.float64)
    h = torch.cat([pts, of.spherical_harmonics3(ray), nrm, feat], -1)
    for i, lin in enumerate(layers):
        h = lin(h)
        if i < 4:
            h = torch.relu(h)
    assert torch.allclose(of.neural_rgb(pts, nrm, ray, feat, P), torch.sigmoid(h), rtol=1e-12)
