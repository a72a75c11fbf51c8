"""Oracle: ray sampling, NeuS rendering, compositing and losses (torch CPU).  TEST INFRASTRUCTURE.

Restates:
  intersect_with_sphere                 projects/nerf/utils/nerf_util.py:199-205
  Model.get_dist_bounds                 projects/neuralangelo/model.py:207-213
  sample_dists (u passed in)            projects/nerf/utils/nerf_util.py:20-38
  sample_dists_from_pdf                 projects/nerf/utils/nerf_util.py:41-68
  alpha_compositing_weights / composite projects/nerf/utils/render.py:87-112
  Model.sample_dists_all / _hierarchical projects/neuralangelo/model.py:232-267
  Model.compute_neus_alphas / _get_iter_cos projects/neuralangelo/model.py:275-298
  Model.render_rays_object / render_rays  projects/neuralangelo/model.py:124-190 (no background)
  eikonal_loss / curvature_loss         projects/neuralangelo/utils/misc.py:74-89
  *_packed                              the same functions over packed (ragged) rays, through
                                        the padded dense form (zero alpha / far distance)
  render loss (L1 * 3)                  projects/neuralangelo/trainer.py:43
  get_center_and_ray / points from dist projects/nerf/utils/camera.py:283-320

Canonical reduction order.  The reference evaluates cumsum/cumprod and the
L1 normalisation with torch reductions whose float32 association order is
unspecified on the GPU.  Discrete outputs (searchsorted indices, merge order)
depend on those last bits, so this oracle fixes the order: sums and prefix
sums/products along a ray accumulate sequentially in float32, sample 0 first.
The kernels follow the same order, which makes sample ordering comparable
bit-for-bit on identical inputs.

Pinned by (tests/test_oracle_render.py): sample_dists_from_pdf with a uniform
pdf (evenly spaced samples), a delta pdf (collapsed samples) and brute-force
searchsorted; compositing weights against the closed form for constant alpha
(geometric series, sum w = 1 - (1-a)^N) and alpha = 1 (all weight on one
sample); the NeuS alpha against its closed form for a planar SDF along the
ray; sphere intersection against analytic distances; losses against hand
values; the sort/merge against torch.sort(stable=True).
"""
import numpy as np
import torch
import torch.nn.functional as torch_F


# ---------------------------------------------------------------------------
# canonical sequential reductions
# ---------------------------------------------------------------------------
def seq_sum(x):
    """Sum over the last axis, sequentially from index 0, in x's dtype."""
    acc = torch.zeros_like(x[..., 0])
    for i in range(x.shape[-1]):
        acc = acc + x[..., i]
    return acc


def seq_cumsum(x):
    acc = torch.zeros_like(x[..., 0])
    out = []
    for i in range(x.shape[-1]):
        acc = acc + x[..., i]
        out.append(acc)
    return torch.stack(out, dim=-1)


def seq_cumprod(x):
    acc = torch.ones_like(x[..., 0])
    out = []
    for i in range(x.shape[-1]):
        acc = acc * x[..., i]
        out.append(acc)
    return torch.stack(out, dim=-1)


# ---------------------------------------------------------------------------
# rays
# ---------------------------------------------------------------------------
def intersect_with_sphere(center, ray_unit, radius=1.0):
    ctc = (center * center).sum(dim=-1, keepdim=True)
    ctv = (center * ray_unit).sum(dim=-1, keepdim=True)
    b2_minus_4ac = ctv ** 2 - (ctc - radius ** 2)
    dist_near = -ctv - b2_minus_4ac.sqrt()
    dist_far = -ctv + b2_minus_4ac.sqrt()
    return dist_near, dist_far


def get_dist_bounds(center, ray_unit):
    near, far = intersect_with_sphere(center, ray_unit, radius=1.0)
    near = near.relu()
    outside = near.isnan()
    near = torch.where(outside, torch.ones_like(near), near)
    far = torch.where(outside, torch.full_like(far, 1.2), far)
    return near, far, outside


def sample_dists(near, far, intvs, rands=None):
    """near/far [..., 1]; rands [..., intvs] (stratified) or None (0.5)."""
    shape = near.shape[:-1]
    if rands is None:
        rands = torch.full((*shape, intvs), 0.5, dtype=near.dtype)
    rands = rands + torch.arange(intvs, dtype=near.dtype)
    return (rands / intvs * (far - near) + near)[..., None]  # [..., N, 1]


def get_3D_points_from_dist(center, ray_unit, dist):
    return center[..., None, :] + ray_unit[..., None, :] * dist


def get_center_and_ray(pose, intr, image_size):
    """camera.py:283-311 for one [3,4] world-to-camera pose and [3,3] intrinsics."""
    H, W = image_size
    y_range = torch.arange(H, dtype=torch.float32).add_(0.5)
    x_range = torch.arange(W, dtype=torch.float32).add_(0.5)
    Y, X = torch.meshgrid(y_range, x_range, indexing="ij")
    xy = torch.stack([X, Y], dim=-1).view(-1, 2)
    xy1 = torch.cat([xy, torch.ones_like(xy[:, :1])], dim=-1)
    grid_3D = xy1 @ intr.inverse().transpose(-1, -2)
    R, t = pose[:, :3], pose[:, 3]
    # cam2world: invert [R|t]: x_w = R^T (x_c - t)
    grid_w = (grid_3D - t) @ R
    center_w = (torch.zeros_like(grid_3D) - t) @ R
    return center_w, grid_w - center_w


# ---------------------------------------------------------------------------
# compositing and pdf sampling
# ---------------------------------------------------------------------------
def alpha_compositing_weights(alphas):
    """[..., N] -> [..., N, 1]; exclusive product of (1 - alpha) in sequential order."""
    alphas_front = torch.cat([torch.zeros_like(alphas[..., :1]), alphas[..., :-1]], dim=-1)
    visibility = seq_cumprod(1 - alphas_front)
    return (alphas * visibility)[..., None]


def composite(quantities, weights):
    return (quantities * weights).sum(dim=-2)


def sample_dists_from_pdf(bins, weights, intvs_fine):
    """bins [..., N, 1], weights [..., N-1] -> [..., Nf, 1] (sequential cdf)."""
    denom = seq_sum(weights.abs()).clamp_min(1e-12)[..., None]
    pdf = weights / denom
    cdf = seq_cumsum(pdf)
    cdf = torch.cat([torch.zeros_like(cdf[..., :1]), cdf], dim=-1)
    grid = torch.linspace(0, 1, intvs_fine + 1, dtype=weights.dtype)
    unif = 0.5 * (grid[:-1] + grid[1:]).repeat(*cdf.shape[:-1], 1)
    idx = torch.searchsorted(cdf.contiguous(), unif.contiguous(), right=True)
    low = (idx - 1).clamp(min=0)
    high = idx.clamp(max=cdf.shape[-1] - 1)
    b = bins[..., 0]
    dist_min = b.gather(dim=-1, index=low)
    dist_max = b.gather(dim=-1, index=high)
    cdf_low = cdf.gather(dim=-1, index=low)
    cdf_high = cdf.gather(dim=-1, index=high)
    t = (unif - cdf_low) / (cdf_high - cdf_low + 1e-8)
    return (dist_min + t * (dist_max - dist_min))[..., None]


def hierarchical_alphas(dists, sdfs, inv_s, eps=1e-5):
    """model.py:251-264 (robust cos); dists [..., N, 1], sdfs [..., N, 1] -> [..., N-1]."""
    s = sdfs[..., 0]
    prev_s, next_s = s[..., :-1], s[..., 1:]
    prev_d, next_d = dists[..., :-1, 0], dists[..., 1:, 0]
    mid = (prev_s + next_s) * 0.5
    cos_val = (next_s - prev_s) / (next_d - prev_d + 1e-5)
    prev_cos = torch.cat([torch.zeros_like(cos_val)[..., :1], cos_val[..., :-1]], dim=-1)
    cos_val = torch.stack([prev_cos, cos_val], dim=-1).min(dim=-1).values
    intv = dists[..., 1:, 0] - dists[..., :-1, 0]
    est_prev = mid - cos_val * intv * 0.5
    est_next = mid + cos_val * intv * 0.5
    prev_cdf = (est_prev * inv_s).sigmoid()
    next_cdf = (est_next * inv_s).sigmoid()
    return ((prev_cdf - next_cdf) / (prev_cdf + eps)).clip(0.0, 1.0)


def sample_dists_hierarchical(dists, sdfs, inv_s, intvs_fine):
    alphas = hierarchical_alphas(dists, sdfs, inv_s)
    weights = alpha_compositing_weights(alphas)
    return sample_dists_from_pdf(dists, weights[..., 0], intvs_fine)


def sample_dists_all(center, ray_unit, near, far, sdf_fn, n_coarse, n_fine, n_hier, rands=None):
    """model.py:232-248.  Returns (dists [..., N,1], per-round fine dists, sdfs used)."""
    dists = sample_dists(near, far, n_coarse, rands)
    if n_hier > 0:
        sdfs = sdf_fn(get_3D_points_from_dist(center, ray_unit, dists))
    for h in range(n_hier):
        fine = sample_dists_hierarchical(dists, sdfs, 64 * 2 ** h, n_fine)
        dists = torch.cat([dists, fine], dim=-2)
        dists, sort_idx = dists.sort(dim=-2, stable=True)
        if h != n_hier - 1:
            sdfs_fine = sdf_fn(get_3D_points_from_dist(center, ray_unit, fine))
            sdfs = torch.cat([sdfs, sdfs_fine], dim=-2)
            sdfs = sdfs.gather(dim=-2, index=sort_idx.expand_as(sdfs))
    return dists


# ---------------------------------------------------------------------------
# NeuS alpha
# ---------------------------------------------------------------------------
def get_iter_cos(true_cos, progress, anneal_end):
    anneal_ratio = min(progress / anneal_end, 1.0)
    return -((-true_cos * 0.5 + 0.5).relu() * (1.0 - anneal_ratio) + (-true_cos).relu() * anneal_ratio)


def compute_neus_alphas(ray_unit, sdfs, gradients, dists, dist_far, s_var, progress, anneal_end,
                        eps=1e-5):
    sdfs = sdfs[..., 0]
    inv_s = s_var.exp()
    true_cos = (ray_unit[..., None, :] * gradients).sum(dim=-1)
    iter_cos = get_iter_cos(true_cos, progress, anneal_end)
    if dist_far is None:
        dist_far = torch.empty_like(dists[..., :1, :]).fill_(1e10)
    dists = torch.cat([dists, dist_far], dim=-2)
    dist_intvs = dists[..., 1:, 0] - dists[..., :-1, 0]
    est_prev = sdfs - iter_cos * dist_intvs * 0.5
    est_next = sdfs + iter_cos * dist_intvs * 0.5
    prev_cdf = (est_prev * inv_s).sigmoid()
    next_cdf = (est_next * inv_s).sigmoid()
    return ((prev_cdf - next_cdf) / (prev_cdf + eps)).clip(0.0, 1.0)


# ---------------------------------------------------------------------------
# background samples and volume-rendering alphas
# ---------------------------------------------------------------------------
def sample_dists_background(far, n, rands=None, eps=1e-5):
    """model.py:269-273: inverse distances sample_dists over (1, 0), dists = far / (inv + eps);
    far [R, 1] -> [R, n]."""
    R = far.shape[0]
    u = torch.full((R, n), 0.5, dtype=far.dtype) if rands is None else rands.to(far.dtype)
    inv = (torch.arange(n, dtype=far.dtype)[None, :] + u) / n * (0.0 - 1.0) + 1.0
    return far.reshape(R, 1) / (inv + eps)


def volume_rendering_alphas_dist(densities, dists, dist_far=None):
    """render.py:68-84: densities [..., N], dists [..., N, 1] -> alphas [..., N]."""
    if dist_far is None:
        dist_far = torch.empty_like(dists[..., :1, :]).fill_(1e10)
    d = torch.cat([dists, dist_far], dim=-2)
    intv = d[..., 1:, 0] - d[..., :-1, 0]
    return 1 - (-(densities * intv)).exp()


# ---------------------------------------------------------------------------
# packed (ragged) rays: ray r owns samples offsets[r] .. offsets[r+1]-1
# ---------------------------------------------------------------------------
# The reference renders N samples per ray.  A packed batch is restated through
# that dense form by padding every ray to the longest one:
#   compositing: padding samples carry alpha = 0, which leaves every real
#     sample's transmittance and weight unchanged and adds zero weight;
#   NeuS alphas: padding samples carry the ray's far distance, so the last real
#     sample's interval ends at far exactly as model.py:170-171 sets it.
# Real samples are then read back in packed order.
def pad_packed(x, offsets, fill):
    """[S, ...] packed -> ([R, Nmax, ...] padded with `fill`, mask [R, Nmax])."""
    counts = (offsets[1:] - offsets[:-1]).long()
    R, Nmax = counts.numel(), int(counts.max().item()) if counts.numel() else 0
    Nmax = max(Nmax, 1)
    mask = torch.arange(Nmax)[None, :] < counts[:, None]
    out = torch.empty((R, Nmax) + tuple(x.shape[1:]), dtype=x.dtype)
    if torch.is_tensor(fill):
        out[...] = fill.reshape((R, 1) + (1,) * (x.dim() - 1)).to(x.dtype)
    else:
        out[...] = fill
    out[mask] = x
    return out, mask


def alpha_compositing_weights_packed(alphas, offsets):
    """alphas [S] -> weights [S, 1]."""
    a, mask = pad_packed(alphas, offsets, 0.0)
    return alpha_compositing_weights(a)[mask]


def composite_packed(quantities, weights, offsets):
    """quantities [S, K], weights [S, 1] -> [R, K]."""
    q, _ = pad_packed(quantities, offsets, 0.0)
    w, _ = pad_packed(weights, offsets, 0.0)
    return composite(q, w)


def compute_neus_alphas_packed(ray_unit, sdfs, gradients, dists, dist_far, offsets, s_var,
                               progress, anneal_end, eps=1e-5):
    """ray_unit [R,3], sdfs [S,1], gradients [S,3], dists [S,1], dist_far [R,1] -> alphas [S]."""
    s, mask = pad_packed(sdfs, offsets, 0.0)
    g, _ = pad_packed(gradients, offsets, 0.0)
    d, _ = pad_packed(dists, offsets, dist_far.reshape(-1))
    alphas = compute_neus_alphas(ray_unit, s, g, d, dist_far.reshape(-1, 1, 1), s_var, progress,
                                 anneal_end, eps)
    return alphas[mask]


def ragged_batch(n_rays, seed=0, n_spheres=None, max_coarse=160, n_fine=32, per_unit=160.0):
    """Ragged sample counts in the manner of the packed-batch configuration: rays from
    cameras on a sphere of radius 3 looking at the origin pass through a random union of
    1-8 spheres with radii U(0.1, 0.5); a ray gets floor(per_unit * length inside the
    union) coarse samples (at most max_coarse) plus n_fine resampled ones when it has at
    least 2, i.e. 0..192 samples.  Returns center [R,3], ray_unit [R,3], offsets int32
    [R+1] and sorted dists [S] inside the ray's unit-sphere bounds."""
    g = torch.Generator().manual_seed(seed)
    k = n_spheres or int(torch.randint(1, 9, (1,), generator=g))
    sc = (torch.rand(k, 3, generator=g, dtype=torch.float64) * 1.2 - 0.6)
    sr = torch.rand(k, generator=g, dtype=torch.float64) * 0.4 + 0.1
    c = torch.randn(n_rays, 3, generator=g, dtype=torch.float64)
    c = c / c.norm(dim=-1, keepdim=True) * 3.0
    # three rays in four aim at a sphere (jittered within its radius), the rest anywhere
    pick = torch.randint(0, k, (n_rays,), generator=g)
    jit = (torch.rand(n_rays, 3, generator=g, dtype=torch.float64) * 2 - 1) * sr[pick, None]
    tgt = torch.where(torch.rand(n_rays, 1, generator=g) < 0.75, sc[pick] + jit,
                      torch.rand(n_rays, 3, generator=g, dtype=torch.float64) * 1.6 - 0.8)
    v = tgt - c
    v = v / v.norm(dim=-1, keepdim=True)
    # length of the ray inside the union, by dense stepping (test data, not timed)
    steps = torch.linspace(1.5, 4.5, 601, dtype=torch.float64)
    pts = c[:, None, :] + v[:, None, :] * steps[None, :, None]
    inside = ((pts[:, :, None, :] - sc[None, None]).norm(dim=-1) < sr[None, None]).any(-1)
    length = inside.double().sum(-1) * (steps[1] - steps[0])
    n_c = (length * per_unit).floor().clamp(max=max_coarse).long()
    counts = torch.where(n_c >= 2, n_c + n_fine, n_c)
    offsets = torch.zeros(n_rays + 1, dtype=torch.int32)
    offsets[1:] = counts.cumsum(0).to(torch.int32)
    near, far, _ = get_dist_bounds(c.float(), v.float())
    S = int(offsets[-1])
    ridx = torch.repeat_interleave(torch.arange(n_rays), counts)
    u = torch.rand(S, generator=g)
    d = near[ridx, 0] + u * (far[ridx, 0] - near[ridx, 0])
    # sort within each ray: key = ray index, then distance
    order = torch.argsort(ridx.double() * 16.0 + (d.double() - near[ridx, 0].double()) /
                          (far[ridx, 0] - near[ridx, 0] + 1e-9).double(), stable=True)
    return c.float(), v.float(), offsets, d[order].contiguous()


# ---------------------------------------------------------------------------
# losses
# ---------------------------------------------------------------------------
def eikonal_loss(gradients, outside=None):
    err = (gradients.norm(dim=-1) - 1.0) ** 2
    err = err.nan_to_num(nan=0.0, posinf=0.0, neginf=0.0)
    if outside is not None:
        return (err * (~outside).to(err.dtype)).mean()
    return err.mean()


def curvature_loss(hessian, outside=None):
    lap = hessian.sum(dim=-1).abs()
    lap = lap.nan_to_num(nan=0.0, posinf=0.0, neginf=0.0)
    if outside is not None:
        return (lap * (~outside).to(lap.dtype)).mean()
    return lap.mean()


# ---------------------------------------------------------------------------
# full object render (background disabled)
# ---------------------------------------------------------------------------
def render_rays(field, rgb_params, center, ray_unit, s_var, progress, eps_world, training,
                n_coarse=64, n_fine=16, n_hier=2, anneal_end=0.1, rands=None, outside_val=1000.0,
                dists=None):
    """model.py:124-190 with background disabled; returns the output dict.

    field: oracle.field.SDFField; rgb_params: FieldParams with R* keys.
    dists: optional [R, N, 1] samples to render instead of sampling.
    """
    from . import field as fld
    with torch.no_grad():
        near, far, outside = get_dist_bounds(center, ray_unit)
        if dists is None:
            dists = sample_dists_all(center, ray_unit, near, far, field.sdf, n_coarse, n_fine,
                                     n_hier, rands)
    points = get_3D_points_from_dist(center, ray_unit, dists)
    sdfs, feats = field.forward(points)
    sdfs = torch.where(outside[..., None, :].expand_as(sdfs), torch.full_like(sdfs, outside_val),
                       sdfs)
    rays_unit = ray_unit[..., None, :].expand_as(points).contiguous()
    gradients, hessians = field.compute_gradients(points, eps_world, training=training, sdf=sdfs)
    normals = torch_F.normalize(gradients, dim=-1)
    rgbs = fld.neural_rgb(points, normals, rays_unit, feats, rgb_params)
    alphas = compute_neus_alphas(ray_unit, sdfs, gradients, dists, far[..., None], s_var, progress,
                                 anneal_end)
    weights = alpha_compositing_weights(alphas)
    rgb = composite(rgbs, weights)
    out = dict(rgb=rgb, rgbs=rgbs, sdfs=sdfs[..., 0], dists=dists, alphas=alphas, weights=weights,
               gradients=gradients, hessians=hessians, outside=outside, near=near, far=far)
    if not training:
        out["opacity"] = composite(torch.ones_like(alphas)[..., None], weights)
        out["gradient"] = composite(gradients, weights)
        out["depth"] = composite(dists, weights)
    return out


def train_loss(out, target_rgb, w_eik=0.1, w_curv=5e-4):
    """trainer.py:40-48 + base loss weights (base.yaml:28-31)."""
    render = torch_F.l1_loss(out["rgb"], target_rgb) * 3
    eik = eikonal_loss(out["gradients"], out["outside"])
    curv = curvature_loss(out["hessians"], out["outside"])
    return render + w_eik * eik + w_curv * curv, dict(render=render, eikonal=eik, curvature=curv)
