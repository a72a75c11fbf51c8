"""The reference's Model entry points (projects/neuralangelo/model.py:66-122) on the HIP
path: camera rays of posed cameras, forward(data) on sampled pixels, and
inference(data) on a full image (the 2^19 x 16 x 4 eval render of 1024 rays from one
32 x 32 camera at radius 3, FOV 50 deg) against the FP64 oracle."""
import pytest
import torch
import torch.nn.functional as torch_F

from neuralangelo_amd import camera
from neuralangelo_amd.model import Model, default_config
from oracle import render as orr
from tests.helpers import oracle_field, oracle_params, randomize_model, sphere_cameras

pytestmark = pytest.mark.gpu

SIZE = [32, 32]


def _model(cuda, log2_T, F, seed=None, cdt=torch.float32):
    torch.manual_seed(0)
    m = Model(default_config(log2_T=log2_T, F=F),
              cfg_data=dict(train=dict(image_size=SIZE), val=dict(image_size=SIZE)),
              compute_dtype=cdt)
    if seed is not None:
        randomize_model(m, seed, table_scale=0.02)
    m = m.to(cuda)
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    return m


def test_camera_rays_match_oracle(cuda):
    """get_center_and_ray (camera.py:283-311) for every pixel and for sampled pixels
    (slice_by_ray_idx, nerf_util.py:127-131), against the oracle's restatement."""
    pose, intr = sphere_cameras(3, seed=1)
    g = torch.Generator().manual_seed(2)
    idx = torch.randint(0, SIZE[0] * SIZE[1], (3, 200), generator=g)
    c_all, r_all = camera.get_center_and_ray(pose.to(cuda), intr.to(cuda), SIZE)
    c_sel, r_sel, u_sel = camera.camera_rays(pose.to(cuda), intr.to(cuda), SIZE,
                                             ray_idx=idx.to(cuda))
    assert c_all.shape == r_all.shape == (3, 1024, 3) and c_sel.shape == (3, 200, 3)
    for b in range(3):
        oc, orr_ = orr.get_center_and_ray(pose[b], intr[b], SIZE)
        assert torch.allclose(c_all[b].cpu(), oc, rtol=1e-6, atol=1e-6)
        assert torch.allclose(r_all[b].cpu(), orr_, rtol=1e-5, atol=1e-6)
        assert torch.allclose(c_sel[b].cpu(), oc[idx[b]], rtol=1e-6, atol=1e-6)
        assert torch.allclose(r_sel[b].cpu(), orr_[idx[b]], rtol=1e-5, atol=1e-6)
        assert torch.allclose(u_sel[b].cpu(), torch_F.normalize(orr_[idx[b]], dim=-1), atol=2e-6)
    # the centre ray of a look-at camera points at the origin
    mid = SIZE[1] * (SIZE[0] // 2) + SIZE[1] // 2
    _, _, u = camera.camera_rays(pose[:1].to(cuda), intr[:1].to(cuda), SIZE)
    c0 = -pose[0, :, :3].t() @ pose[0, :, 3]
    toward = -c0 / c0.norm()
    assert torch.dot(u[0, mid].cpu(), toward) > 0.999


def test_forward_renders_the_sampled_pixels(cuda):
    """forward(data) (model.py:66-72) = render_rays on the rays of data["ray_idx"], with the
    reference's output keys and [B,R,...] shapes."""
    m = _model(cuda, 14, 8, seed=1).train()
    pose, intr = sphere_cameras(2, seed=4)
    g = torch.Generator().manual_seed(5)
    ray_idx = torch.randint(0, 1024, (2, 96), generator=g).to(cuda)
    data = dict(pose=pose.to(cuda), intr=intr.to(cuda), idx=torch.arange(2, device=cuda),
                ray_idx=ray_idx)
    torch.manual_seed(11)
    out = m(data)
    for k in ("rgb", "opacity", "outside", "dists", "weights", "gradient", "gradients",
              "hessians"):
        assert k in out
    assert out["rgb"].shape == (2, 96, 3) and out["dists"].shape == (2, 96, 96, 1)
    assert out["gradients"].shape == (2, 96, 96, 3) and out["hessians"].shape == (2, 96, 96, 3)
    assert out["opacity"] is None and out["gradient"] is None  # training mode
    c, _, v = camera.camera_rays(pose.to(cuda), intr.to(cuda), SIZE, ray_idx=ray_idx)
    torch.manual_seed(11)
    ref = m.render_rays(c, v, stratified=True)
    assert torch.equal(out["rgb"], ref["rgb"]) and torch.equal(out["dists"], ref["dists"])
    out["rgb"].sum().backward()
    assert torch.isfinite(m.neural_sdf.tcnn_encoding.params.grad).all()


@pytest.mark.parametrize("seed", [None, 3])
def test_inference_full_image_matches_oracle(cuda, seed):
    """inference(data) (model.py:74-114) on one 32 x 32 image of the 2^19 x 16 x 4 model
    (geometric init, and a randomised variant): rgb, opacity, composited gradient and
    depth against the FP64 oracle rendering the same rays at the same samples; the
    sampled distances against the oracle's own sampling."""
    m = _model(cuda, 19, 4, seed=seed)
    pose, intr = sphere_cameras(1, seed=7)
    data = dict(pose=pose.to(cuda), intr=intr.to(cuda), idx=torch.zeros(1, dtype=torch.long,
                                                                         device=cuda))
    out = m.inference(data)
    assert not m.training
    R = SIZE[0] * SIZE[1]
    assert out["rgb"].shape == (1, R, 3) and out["depth"].shape == (1, R, 1)
    assert out["rgb_map"].shape == (1, 3, 32, 32) and out["normal_map"].shape == (1, 3, 32, 32)
    assert out["opacity_map"].shape == (1, 1, 32, 32) and out["depth_map"].shape == (1, 1, 32, 32)
    c, ray, v = camera.camera_rays(pose.to(cuda), intr.to(cuda), SIZE)
    c, ray, v = c[0].cpu().double(), ray[0].cpu().double(), v[0].cpu().double()
    dists = out["dists"][0].cpu().double()
    P = oracle_params(m)
    fld, _ = oracle_field(m, P)
    with torch.no_grad():
        o = orr.render_rays(fld, P, c, v, m.s_var.detach().cpu().double(), m.progress,
                            m.neural_sdf.normal_eps, False, dists=dists)
        # the sampling itself against the oracle's float32 field (the arithmetic the
        # sampling runs in), as the pdf inversion amplifies tiny SDF differences
        f32, _ = oracle_field(m, oracle_params(m, torch.float32), dtype=torch.float32)
        near, far, _ = orr.get_dist_bounds(c.float(), v.float())
        own = orr.sample_dists_all(c.float(), v.float(), near, far, f32.sdf, 64, 16,
                                   2)[..., 0].double()
    e = dict(rgb=(out["rgb"][0].cpu().double() - o["rgb"]).abs().max().item(),
             opacity=(out["opacity"][0].cpu().double() - o["opacity"]).abs().max().item(),
             gradient=(out["gradient"][0].cpu().double() - o["gradient"]).abs().max().item(),
             depth=(out["depth"][0].cpu().double() - o["depth"] / ray.norm(dim=-1, keepdim=True))
             .abs().max().item())
    print(f"\ninference vs FP64 oracle (seed {seed}): " +
          " ".join(f"{k} {v:.2e}" for k, v in e.items()), flush=True)
    assert o["opacity"].max() > 0.5  # the surface is in view
    assert e["rgb"] < 2e-4 and e["opacity"] < 1e-3 and e["gradient"] < 2e-3 and e["depth"] < 4e-3
    close = (dists[..., 0] - own).abs() <= 1e-4 * own.abs() + 1e-5
    assert close.float().mean() > 0.99
    # the normal map is the composited gradient in camera axes, negated (model.py:81)
    rot = pose[0, :, :3].double()
    nc = -(o["gradient"] @ rot.t())
    assert (out["normal_map"][0].permute(1, 2, 0).reshape(R, 3).cpu().double() - nc).abs().max() < 2e-3
