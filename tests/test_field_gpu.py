"""NeuralSDF / NeuralRGB on the HIP kernels vs the CPU oracle (values and gradients)."""
import pytest
import torch
import torch.nn.functional as torch_F

from neuralangelo_amd.model import Model, default_config
from oracle import field as ofield
from tests.helpers import oracle_field, oracle_params, randomize_model

pytestmark = pytest.mark.gpu


def _model(cuda, cdt=torch.float32, log2_T=14, seed=0):
    torch.manual_seed(seed)
    m = Model(default_config(log2_T=log2_T), compute_dtype=cdt)
    randomize_model(m, seed)
    return m.to(cuda)


def _points(n, seed):
    g = torch.Generator().manual_seed(seed)
    return (torch.rand(n, 3, generator=g) * 2 - 1) * 0.9


def _cos(a, b):
    a, b = a.reshape(-1).double(), b.reshape(-1).double()
    return (a @ b / (a.norm() * b.norm() + 1e-300)).item()


@pytest.mark.parametrize("cdt", [torch.float32, torch.bfloat16])
def test_sdf_forward_backward(cuda, cdt):
    m = _model(cuda, cdt)
    sdf_mod = m.neural_sdf
    x = _points(3000, 1)
    P = oracle_params(m)
    fld, table = oracle_field(m, P, table_dtype=torch.bfloat16 if cdt == torch.bfloat16 else None)
    s_ref, f_ref = fld.forward(x.double())
    s, f = sdf_mod.forward(x.to(cuda))
    s_err = (s.detach().cpu().double() - s_ref.detach()).abs().max().item()
    f_err = (f.detach().float().cpu().double() - f_ref.detach()).abs().max().item()
    if cdt == torch.float32:
        assert s_err < 2e-5 and f_err < 2e-5, (s_err, f_err)
    else:
        assert s_err < 5e-3 and f_err < 2e-2, (s_err, f_err)
    g = torch.Generator().manual_seed(2)
    us, uf = torch.randn(3000, 1, generator=g), torch.randn(3000, 256, generator=g) * 0.01
    ((s * us.to(cuda)).sum() + (f.float() * uf.to(cuda)).sum()).backward()
    loss_ref = (s_ref * us.double()).sum() + (f_ref * uf.double()).sum()
    names = ["W0v", "W0g", "b0", "W1v", "W1g", "b1", "ws", "bs"]
    grads_ref = torch.autograd.grad(loss_ref, [table] + [P[k] for k in names])
    l0, l1 = sdf_mod.mlp.linears
    got = [sdf_mod.tcnn_encoding.params.grad, l0.weight_v.grad, l0.weight_g.grad, l0.bias.grad,
           l1.weight_v.grad, l1.weight_g.grad, l1.bias.grad, sdf_mod.mlp.linear_sdf.weight.grad,
           sdf_mod.mlp.linear_sdf.bias.grad]
    thr = 0.99999 if cdt == torch.float32 else 0.999
    for name, a, b in zip(["table"] + names, got, grads_ref):
        c = _cos(a.cpu(), b)
        assert c > thr, (name, c)
        rel = ((a.cpu().double() - b).norm() / (b.norm() + 1e-30)).item()
        assert rel < (1e-4 if cdt == torch.float32 else 5e-2), (name, rel)


def test_numerical_gradient_matches_oracle(cuda):
    m = _model(cuda)
    sdf_mod = m.neural_sdf
    x = _points(2000, 3)
    P = oracle_params(m, torch.float32)
    fld, table = oracle_field(m, P, dtype=torch.float32)
    s_ref = fld.sdf(x)
    g_ref, h_ref = fld.compute_gradients(x, sdf_mod.normal_eps, training=True, sdf=s_ref)
    s = sdf_mod.sdf(x.to(cuda))
    g, h = sdf_mod.compute_gradients(x.to(cuda), training=True, sdf=s)
    # finite differences amplify float32 rounding of the SDF by 1/eps (gradient)
    # and 1/eps^2 (Hessian); compare against the float32 oracle at that scale
    eps = float(sdf_mod.tap_eps()[0])
    assert (g.detach().cpu() - g_ref.detach()).abs().max() < 4e-6 / eps
    assert (h.detach().cpu() - h_ref.detach()).abs().max() < 8e-6 / eps ** 2
    assert _cos(g.detach().cpu(), g_ref.detach()) > 0.9999
    # backward through the taps vs the oracle's autograd
    g.backward(torch.ones_like(g), retain_graph=True)
    gt = sdf_mod.tcnn_encoding.params.grad.cpu()
    (gt_ref,) = torch.autograd.grad(g_ref.sum(), [table])
    assert _cos(gt, gt_ref) > 0.999


def test_rgb_forward_backward(cuda):
    m = _model(cuda)
    x = _points(1536, 4).view(16, 96, 3)
    g = torch.Generator().manual_seed(5)
    rays = torch_F.normalize(torch.randn(16, 3, generator=g), dim=-1)
    grads = torch.randn(16, 96, 3, generator=g)
    feats = torch.rand(16 * 96, 256, generator=g) * 0.1
    P = oracle_params(m)
    rgb_ref = ofield.neural_rgb(x.double(), torch_F.normalize(grads.double(), dim=-1),
                                rays.double()[:, None, :].expand(16, 96, 3), feats.double().view(16, 96, 256), P)
    gd = grads.to(cuda).requires_grad_(True)
    fd = feats.to(cuda).requires_grad_(True)
    rgb = m.neural_rgb.forward(x.to(cuda), gd, rays.to(cuda), fd, None, normals_are_gradients=True)
    assert (rgb.detach().cpu().double() - rgb_ref.detach()).abs().max() < 2e-6
    up = torch.randn(16, 96, 3, generator=g)
    rgb.backward(up.to(cuda))
    gd64 = grads.double().requires_grad_(True)
    fd64 = feats.double().requires_grad_(True)
    rgb_ref2 = ofield.neural_rgb(x.double(), torch_F.normalize(gd64, dim=-1),
                                 rays.double()[:, None, :].expand(16, 96, 3), fd64.view(16, 96, 256), P)
    keys = [f"R{i}{s}" for i in range(5) for s in "vgb"]
    refs = torch.autograd.grad((rgb_ref2 * up.double()).sum(), [gd64, fd64] + [P[k] for k in keys])
    got = [gd.grad, fd.grad] + [getattr(m.neural_rgb.mlp.linears[i], a).grad
                                for i in range(5) for a in ("weight_v", "weight_g", "bias")]
    for name, a, b in zip(["grad", "feat"] + keys, got, refs):
        assert _cos(a.cpu(), b) > 0.99999, name
        assert ((a.cpu().double() - b).norm() / (b.norm() + 1e-30)).item() < 1e-4, name


def test_generic_mlp_api(cuda):
    """MLPforNeuralSDF / MLPwithSkipConnection forward on reference-order inputs."""
    m = _model(cuda)
    P = oracle_params(m)
    g = torch.Generator().manual_seed(6)
    inp = torch.randn(700, 131, generator=g) * 0.1
    s, f = m.neural_sdf.mlp(inp.to(cuda))
    s_ref, f_ref = ofield.sdf_mlp(inp.double(), P)
    assert (s.detach().cpu().double() - s_ref.detach()).abs().max() < 1e-5
    assert (f.detach().cpu().double() - f_ref.detach()).abs().max() < 1e-5
    rin = torch.randn(700, 278, generator=g) * 0.1
    out = m.neural_rgb.mlp(rin.to(cuda))
    out_ref = ofield.rgb_mlp(rin.double(), P)
    assert (out.detach().cpu().double() - out_ref.detach()).abs().max() < 1e-5
