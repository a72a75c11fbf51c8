"""Train-step parity at the 2^22 x 16 x 8 configuration (FP32 and the BF16 mixed-precision
variant) against the FP64 oracle: the full render of 128 rays at the samples the GPU
drew, the loss L1*3 + 0.1 eikonal + 5e-4 curvature (trainer.py:40-48, base.yaml:28-31),
and every parameter gradient of that loss (cosine >= 0.999; >= 0.997 for BF16 on the
rough random field), at the configuration's own table scale and on a rough random
field.  The 4-tap Hessian trace feeding the curvature loss is compared against the
oracle's, and the gradient of the curvature term alone is checked (BF16)."""
import pytest
import torch

from neuralangelo_amd import misc
from neuralangelo_amd.model import Model, default_config
from oracle import render as orr
from tests.helpers import hessian_floor, oracle_field, oracle_params, randomize_model

pytestmark = pytest.mark.gpu

R = 128
NAMES = ["W0v", "W0g", "b0", "W1v", "W1g", "b1", "ws", "bs"] + \
    [f"R{i}{k}" for i in range(5) for k in "vgb"]


def _rays(n, seed):
    g = torch.Generator().manual_seed(seed)
    c = torch.randn(n, 3, generator=g)
    c = c / c.norm(dim=-1, keepdim=True) * 3.0
    d = -c / 3.0 + 0.3 * torch.randn(n, 3, generator=g)
    return c, d / d.norm(dim=-1, keepdim=True)


def _params(m):
    sdf, rgb = m.neural_sdf, m.neural_rgb
    l0, l1 = sdf.mlp.linears
    out = [l0.weight_v, l0.weight_g, l0.bias, l1.weight_v, l1.weight_g, l1.bias,
           sdf.mlp.linear_sdf.weight, sdf.mlp.linear_sdf.bias]
    for lin in rgb.mlp.linears:
        out += [lin.weight_v, lin.weight_g, lin.bias]
    return [sdf.tcnn_encoding.params, m.s_var] + out


def _cos(a, b):
    a, b = a.detach().reshape(-1).cpu().double(), b.detach().reshape(-1).cpu().double()
    return (a @ b / (a.norm() * b.norm() + 1e-300)).item()


_STATES = {}


def _state(table_scale):
    """Model parameters: the configuration's own init scale of the table (1e-4, a smooth
    field near the geometric-init sphere) or a rough random field (0.05)."""
    if table_scale not in _STATES:
        torch.manual_seed(0)
        m = Model(default_config(log2_T=22, F=8), compute_dtype=torch.float32)
        randomize_model(m, 6, table_scale=table_scale)
        _STATES.clear()
        _STATES[table_scale] = {k: v.detach().clone() for k, v in m.state_dict().items()}
    return _STATES[table_scale]


@pytest.fixture(scope="module")
def rays():
    c, v = _rays(R, 21)
    g = torch.Generator().manual_seed(22)
    return c, v, torch.rand(R, 64, generator=g), torch.rand(R, 3, generator=g)


def _oracle(state):
    """FP64 oracle on the FP32 master parameters."""
    m = Model(default_config(log2_T=22, F=8), compute_dtype=torch.float32)
    m.load_state_dict(state)
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    P = oracle_params(m)
    fld, table = oracle_field(m, P)
    return m, P, fld, table


def _gpu_model(state, cdt, cuda):
    m = Model(default_config(log2_T=22, F=8), compute_dtype=cdt)
    m.load_state_dict(state)
    m = m.to(cuda).train()
    m.progress = 1.0
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    return m


def _loss(out, target):
    ren = torch.nn.functional.l1_loss(out["rgb"], target) * 3
    eik, curv = misc.sdf_losses(out["gradients"], out["hessians"], out["outside"])
    return ren, eik, curv


@pytest.mark.parametrize("table_scale", [1e-4, 0.05])
@pytest.mark.parametrize("cdt", [torch.float32, torch.bfloat16])
def test_train_step_matches_fp64_oracle(rays, cuda, cdt, table_scale):
    c, v, u, target = rays
    state = _state(table_scale)
    m = _gpu_model(state, cdt, cuda)
    cd, vd, td = c.to(cuda), v.to(cuda), target.to(cuda)
    with torch.no_grad():
        dists = m.render_rays(cd, vd, rands=u.to(cuda))["dists"][..., 0]
    params = _params(m)
    out = m.render_rays(cd, vd, dists=dists)
    ren, eik, curv = _loss(out, td)
    got = torch.autograd.grad(ren + 0.1 * eik + 5e-4 * curv, params)
    out2 = m.render_rays(cd, vd, dists=dists)
    got_curv = torch.autograd.grad(5e-4 * _loss(out2, td)[2], params, allow_unused=True)
    # oracle at the same samples
    mo, P, fld, table = _oracle(state)
    sv = mo.s_var.detach().double().requires_grad_(True)
    o = orr.render_rays(fld, P, c.double(), v.double(), sv, 1.0, mo.neural_sdf.normal_eps, True,
                        dists=dists.cpu().double()[..., None])
    l_ref, parts = orr.train_loss(o, target.double())
    ref_p = [table, sv] + [P[k] for k in NAMES]
    ref = torch.autograd.grad(l_ref, ref_p, retain_graph=True)
    ref_curv = torch.autograd.grad(5e-4 * parts["curvature"], ref_p, allow_unused=True)
    # values
    inside = ~o["outside"][:, 0]
    e_rgb = (out["rgb"].detach().cpu().double() - o["rgb"]).abs().max().item()
    e_alpha = (out["alphas"].detach().cpu().double() - o["alphas"]).abs().max().item()
    e_loss = abs(ren.item() - parts["render"].item())
    h = out["hessians"].detach()[..., 0].cpu().double()[inside] * 3
    h_ref = o["hessians"].detach()[..., 0][inside] * 3
    dh = (h - h_ref).abs()
    # the reference's own float32 arithmetic at the same sample points (FP32 oracle)
    P32 = oracle_params(mo, torch.float32)
    fld32, _ = oracle_field(mo, P32, dtype=torch.float32)
    xs = (c.double()[:, None] + v.double()[:, None] * dists.cpu().double()[..., None])[inside]
    with torch.no_grad():
        xs = xs.reshape(-1, 3).float()
        _, h32 = fld32.compute_gradients(xs, mo.neural_sdf.normal_eps, training=True,
                                         sdf=fld32.sdf(xs))
    err_ref32 = (h32[:, 0].double().view(h_ref.shape) * 3 - h_ref).abs()
    print(f"\n{cdt} table scale {table_scale}: |rgb err| {e_rgb:.2e}  |alpha err| {e_alpha:.2e}  render loss err "
          f"{e_loss:.2e}  eik {eik.item():.5f}/{parts['eikonal'].item():.5f}  curv "
          f"{curv.item():.4f}/{parts['curvature'].item():.4f}", flush=True)
    print(f"  tr H: median |ref| {h_ref.abs().median().item():.3f}  |err| median "
          f"{dh.median().item():.3f}  p99 {dh.quantile(0.99).item():.3f}  max {dh.max().item():.3f}"
          f"  (float32 oracle p99 {err_ref32.quantile(0.99).item():.3f})")
    cos_all = {n: _cos(a, b) for n, a, b in zip(["table", "s_var"] + NAMES, got, ref)}
    cos_curv = {n: _cos(a, b) for n, a, b in zip(["table", "s_var"] + NAMES, got_curv, ref_curv)
                if a is not None and b is not None and b.abs().max() > 0}
    print("  grad cosines: " + " ".join(f"{k} {x:.5f}" for k, x in cos_all.items()))
    print("  curvature-term grad cosines: " + " ".join(f"{k} {x:.4f}" for k, x in cos_curv.items()),
          flush=True)
    tol = dict(rgb=2e-4, alpha=1e-3, loss=1e-5) if cdt == torch.float32 else \
        dict(rgb=2e-2, alpha=2e-2, loss=2e-3)
    assert e_rgb < tol["rgb"] and e_alpha < tol["alpha"] and e_loss < tol["loss"]
    assert abs(eik.item() - parts["eikonal"].item()) < 2e-3 * abs(parts["eikonal"].item())
    # Hessian trace: within 3 % of |tr H| plus the floor hessian_floor derives (the
    # increments' precision for BF16 at the configuration's table scale, else twice
    # the reference's own float32 arithmetic error at these points)
    floor = hessian_floor(cdt, table_scale, err_ref32)
    over = (dh / (0.03 * h_ref.abs() + floor)).quantile(0.99).item()
    assert over <= 1.0, over
    for name, cs in cos_all.items():
        if cdt == torch.float32 and name == "table":
            # the FP32 table gradient of the curvature term is dominated by the same
            # rounding (amplified by 1 / (2 eps^2) in the backward); BF16 avoids it
            assert cs > 0.95, (name, cs)
            continue
        # BF16 on the rough field: its feature and colour activations span a wide
        # range, and the bf16 rounding of those layers' operands shows at 1e-3
        bound = 0.997 if (cdt == torch.bfloat16 and table_scale > 1e-3) else 0.999
        assert cs > bound, (name, cs)
    if cdt == torch.bfloat16:
        for name, cs in cos_curv.items():
            assert cs > 0.999, ("curvature term", name, cs)
