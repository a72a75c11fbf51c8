"""The full-size train step (8192 rays x 96 samples, 2^22 x 16 x 8, the benchmarked
workload) in BF16 and FP32 with property checks: finite loss and gradients,
per-ray weights summing to at most one, colours in [0, 1], the losses equal to
their definitions on the returned gradients and Hessians (misc.py:74-89), the two
precisions agreeing on the same inputs, and the SDF at a subsample of the step's
own sample points within the north-star bound (1e-3) of the FP64 oracle."""
import pytest
import torch

from neuralangelo_amd import misc
from neuralangelo_amd.model import Model, default_config
from neuralangelo_amd.optim import AdamW
from tests.helpers import oracle_field, oracle_params, randomize_model

pytestmark = pytest.mark.gpu

R = 8192


def _rays():
    g = torch.Generator().manual_seed(31)
    c = torch.randn(R, 3, generator=g)
    c = c / c.norm(dim=-1, keepdim=True) * 3.0
    v = -c / 3.0 + 0.3 * torch.randn(R, 3, generator=g)
    return c, v / v.norm(dim=-1, keepdim=True), torch.rand(R, 3, generator=g)


def _state():
    """The geometric initialisation with the encoder block of the first layer drawn
    (the initialisation zeroes it, which would leave the table gradient zero) and the
    table at the configuration's own scale (1e-4)."""
    torch.manual_seed(0)
    m = Model(default_config(log2_T=22, F=8), compute_dtype=torch.float32)
    randomize_model(m, 4, table_scale=1e-4)
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    return m


def _step(cdt, cuda, state, c, v, target):
    m = Model(default_config(log2_T=22, F=8), compute_dtype=cdt)
    m.load_state_dict(state)
    m = m.to(cuda).train()
    m.progress = 1.0
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    opt = AdamW(m.parameters(), lr=1e-3, weight_decay=1e-2)
    enc = m.neural_sdf.tcnn_encoding
    opt.set_lowp_shadow(enc.params, enc)
    out = m.render_rays(c.to(cuda), v.to(cuda))
    render = torch.nn.functional.l1_loss(out["rgb"], target.to(cuda)) * 3
    eik, curv = misc.sdf_losses(out["gradients"], out["hessians"], out["outside"])
    loss = render + 0.1 * eik + 5e-4 * curv
    opt.zero_grad(set_to_none=True)
    loss.backward()
    grads = {n: p.grad.detach().clone() for n, p in m.named_parameters() if p.grad is not None}
    opt.step()
    return m, out, (loss.item(), eik.item(), curv.item()), grads


def test_full_size_step_properties_and_precisions(cuda):
    c, v, target = _rays()
    m0 = _state()  # the shared initial parameters, and the oracle's
    state = {k: t.detach().clone() for k, t in m0.state_dict().items()}
    fld, _ = oracle_field(m0, oracle_params(m0))
    runs = {}
    for cdt in (torch.float32, torch.bfloat16):
        print("full step", cdt, flush=True)
        m, out, losses, grads = _step(cdt, cuda, state, c, v, target)
        assert all(torch.isfinite(torch.tensor(x)) for x in losses), losses
        for n, gr in grads.items():
            assert torch.isfinite(gr).all(), n
        assert grads["neural_sdf.tcnn_encoding.params"].abs().max() > 0
        w = out["weights"].detach()[..., 0]
        assert w.shape == (R, 96) and (w >= 0).all() and (w.sum(-1) <= 1 + 1e-5).all()
        rgb = out["rgb"].detach()
        assert (rgb >= 0).all() and (rgb <= 1).all()
        # the losses are their definitions (misc.py:74-89) on the returned tensors
        mask = (~out["outside"].detach()).float().expand(R, 96)
        gn = out["gradients"].detach().norm(dim=-1)
        eik = ((((gn - 1) ** 2).nan_to_num(0.0, 0.0, 0.0)) * mask).mean().item()
        lap = out["hessians"].detach().sum(-1).abs().nan_to_num(0.0, 0.0, 0.0)
        curv = (lap * mask).mean().item()
        assert abs(losses[1] - eik) <= 1e-4 * max(eik, 1e-6), (losses[1], eik)
        assert abs(losses[2] - curv) <= 1e-4 * max(curv, 1e-6), (losses[2], curv)
        # the SDF at 4096 of the step's sample points (rays that meet the unit
        # sphere; the others carry the outside value) against the FP64 oracle
        cand = torch.nonzero(mask.reshape(-1) > 0)[:, 0].cpu()
        idx = cand[torch.randperm(cand.numel(), generator=torch.Generator().manual_seed(2))[:4096]]
        d = out["dists"].detach().reshape(R, 96)
        pts = (c.to(cuda)[:, None] + v.to(cuda)[:, None] * d[..., None]).reshape(-1, 3)[idx.to(cuda)]
        with torch.no_grad():
            s_ref = fld.sdf(pts.cpu().double())
        sdfs = out["sdfs"].detach().reshape(-1)[idx.to(cuda)].cpu().double()[:, None]
        e_sdf = (sdfs - s_ref).abs().max().item()
        print(f"  loss {losses[0]:.6f}  eikonal {losses[1]:.6f}  curvature {losses[2]:.4f}  "
              f"|sdf err| vs FP64 {e_sdf:.2e}", flush=True)
        assert e_sdf <= (1e-5 if cdt == torch.float32 else 1e-3), e_sdf
        runs[cdt] = (out["rgb"].detach(), losses, grads)
        del m, out
        torch.cuda.empty_cache()
    (r32, l32, g32), (r16, l16, g16) = runs[torch.float32], runs[torch.bfloat16]
    assert (r32 - r16).abs().max().item() < 2e-2
    assert abs(l32[0] - l16[0]) <= 2e-3 * abs(l32[0])
    for n, a in g32.items():
        b = g16[n]
        a, b = a.double().flatten(), b.double().flatten()
        if a.norm() == 0:
            continue
        cos = (a @ b / (a.norm() * b.norm() + 1e-300)).item()
        print(f"  grad {n}: cos(FP32, BF16) {cos:.5f}", flush=True)
        if n == "neural_sdf.tcnn_encoding.params":
            # The FP32 table gradient of the curvature term carries the reference's
            # float32 rounding of the tap SDFs, amplified by 1 / (2 eps^2) in the
            # backward; the BF16 increment form avoids it.  Against an FP64 oracle of
            # this very step (the same samples, 1.5 M table entries) FP32 reaches a
            # cosine of 0.975 and BF16 0.9999, so the two agree only to ~0.95 here.
            assert cos > 0.9, (n, cos)
            continue
        assert cos > 0.99, (n, cos)
