"""Drop-in Model for the object render path (projects/neuralangelo/model.py:24-298).

Covers the hot path with the background model disabled (the object-only
training and extraction configurations); every per-sample operator runs on the
HIP kernels:
  get_dist_bounds + sample_dists      -> one fused bounds/coarse kernel
  sample_dists_all / _hierarchical    -> SDF passes + one resample/merge kernel per round
  neural_sdf.forward / compute_grads  -> fused encoder + SDF MLP (GEMM kernels), 4-tap combine
  neural_rgb.forward                  -> fused input builder (SH, normalise) + colour MLP
  compute_neus_alphas                 -> NeuS alpha kernel (outside fill inside)
  alpha_compositing_weights+composite -> one fused compositing kernel
"""
from collections import defaultdict

import torch

from . import camera, misc, nerf_util, render
from . import functional as Fn
from .modules import BackgroundNeRF, NeuralRGB, NeuralSDF, _get


def _image_size(cfg_data, split):
    """cfg_data.<split>.image_size ([H, W]) or None."""
    if cfg_data is None:
        return None
    part = _get(cfg_data, split, None)
    return None if part is None else _get(part, "image_size", None)


def background_config():
    """The reference's default background NeRF (base.yaml:84-104)."""
    return dict(enabled=True, white=False,
                mlp=dict(num_layers=8, hidden_dim=256, skip=[4], num_layers_rgb=2,
                         hidden_dim_rgb=128, skip_rgb=[], activ="relu", activ_params={},
                         activ_density="softplus", activ_density_params={}),
                view_dep=True, encoding=dict(type="fourier", levels=10),
                encoding_view=dict(type="spherical", levels=3))


def default_config(log2_T=22, F=8, levels=16, n_coarse=64, n_fine=16, n_hier=2, background=False,
                   n_background=32):
    """The hot-path configuration of base.yaml:38-115 (background off unless asked)."""
    return dict(
        object=dict(
            sdf=dict(
                mlp=dict(num_layers=1, hidden_dim=256, skip=[], activ="softplus",
                         activ_params=dict(beta=100), geometric_init=True, weight_norm=True,
                         out_bias=0.5, inside_out=False),
                encoding=dict(type="hashgrid", levels=levels,
                              hashgrid=dict(min_logres=5, max_logres=11, dict_size=log2_T, dim=F,
                                            range=[-2, 2]),
                              coarse2fine=dict(enabled=True, init_active_level=4, step=5000)),
                gradient=dict(mode="numerical", taps=4)),
            rgb=dict(mlp=dict(num_layers=4, hidden_dim=256, skip=[], activ="relu_", activ_params={},
                              weight_norm=True),
                     mode="idr", encoding_view=dict(type="spherical", levels=3)),
            s_var=dict(init_val=3.0, anneal_end=0.1)),
        background=background_config() if background else dict(enabled=False, white=False),
        render=dict(rand_rays=512, num_samples=dict(coarse=n_coarse, fine=n_fine,
                                                    background=n_background),
                    num_sample_hierarchy=n_hier, stratified=True),
        appear_embed=dict(enabled=False, dim=8),
    )


class Model(torch.nn.Module):
    """The reference's Model (model.py:24-298) with background, object SDF and colour
    fields on the HIP kernels.  Method names, arguments and output keys follow the
    reference; keyword-only extras (`rands`, `dists`, `rands_bg`) let tests fix the
    random stratification or the samples."""

    def __init__(self, cfg_model, cfg_data=None, compute_dtype=torch.float32):
        super().__init__()
        self.cfg_render = _get(cfg_model, "render")
        bg = _get(cfg_model, "background")
        self.white_background = _get(bg, "white", False)
        self.with_background = bool(_get(bg, "enabled", False))
        obj = _get(cfg_model, "object")
        self.anneal_end = _get(_get(obj, "s_var"), "anneal_end")
        self.outside_val = 1000. * (-1 if _get(_get(_get(obj, "sdf"), "mlp"), "inside_out") else 1)
        self.image_size_train = _image_size(cfg_data, "train")
        self.image_size_val = _image_size(cfg_data, "val")
        self.compute_dtype = compute_dtype
        self.neural_sdf = NeuralSDF(_get(obj, "sdf"), compute_dtype=compute_dtype)
        self.neural_rgb = NeuralRGB(_get(obj, "rgb"), feat_dim=_get(_get(_get(obj, "sdf"), "mlp"),
                                                                      "hidden_dim"),
                                    appear_embed=_get(cfg_model, "appear_embed"),
                                    compute_dtype=compute_dtype)
        self.s_var = torch.nn.Parameter(torch.tensor(_get(_get(obj, "s_var"), "init_val"),
                                                     dtype=torch.float32))
        self.background_nerf = (BackgroundNeRF(bg, _get(cfg_model, "appear_embed"),
                                               compute_dtype=compute_dtype)
                                if self.with_background else None)
        self.progress = 1.0

    # -- entry points (model.py:66-122) ------------------------------------------
    def forward(self, data):
        """model.py:66-72: render the rays data["ray_idx"] [B,R] of the cameras
        data["pose"] [B,3,4] / data["intr"] [B,3,3] at the training image size."""
        return self.render_pixels(data["pose"], data["intr"], image_size=self._size("train"),
                                  stratified=_get(self.cfg_render, "stratified", False),
                                  sample_idx=data.get("idx"), ray_idx=data["ray_idx"])

    @torch.no_grad()
    def inference(self, data):
        """model.py:74-90: full validation images plus rgb/opacity/depth/normal maps."""
        self.eval()
        size = self._size("val")
        output = self.render_image(data["pose"], data["intr"], image_size=size, stratified=False,
                                   sample_idx=data.get("idx"))
        rot = data["pose"][..., :3, :3]
        normal_cam = -output["gradient"] @ rot.transpose(-1, -2)
        output.update(rgb_map=misc.to_full_image(output["rgb"], size),
                      opacity_map=misc.to_full_image(output["opacity"], size),
                      depth_map=misc.to_full_image(output["depth"], size),
                      normal_map=misc.to_full_image(normal_cam, size))
        return output

    def render_image(self, pose, intr, image_size, stratified=False, sample_idx=None):
        """model.py:92-114: every pixel, in chunks of render.rand_rays rays per camera;
        in eval mode each chunk adds depth = composite(dists, weights) / |ray|."""
        output = defaultdict(list)
        chunk = _get(self.cfg_render, "rand_rays")
        for center, ray, ray_unit in camera.ray_chunks(pose, intr, image_size, chunk):
            out = self.render_rays(center, ray_unit, sample_idx=sample_idx, stratified=stratified)
            if not self.training:
                dist = render.composite(out["dists"], out["weights"])
                out.update(depth=dist / ray.norm(dim=-1, keepdim=True))
            for key, value in out.items():
                if value is not None:
                    output[key].append(value.detach())
        return {k: torch.cat(v, dim=1) for k, v in output.items()}

    def render_pixels(self, pose, intr, image_size, stratified=False, sample_idx=None,
                      ray_idx=None):
        """model.py:116-122: the selected pixels' rays (one kernel forms the centres and
        unit rays of exactly those pixels), then render_rays."""
        center, _, ray_unit = camera.camera_rays(pose, intr, image_size, ray_idx=ray_idx)
        return self.render_rays(center, ray_unit, sample_idx=sample_idx, stratified=stratified)

    def _size(self, split):
        size = self.image_size_train if split == "train" else self.image_size_val
        if size is None:
            raise RuntimeError(f"Model built without cfg_data: no {split} image size")
        return size

    # -- sampling --------------------------------------------------------------
    def _n_samples(self):
        ns = _get(self.cfg_render, "num_samples")
        return _get(ns, "coarse"), _get(ns, "fine"), _get(self.cfg_render, "num_sample_hierarchy")

    @torch.no_grad()
    def sample_dists_all(self, center, ray_unit, near, far, stratified=False, *, rands=None):
        """model.py:232-248: center, ray_unit [...,3], near, far [...,1] -> sorted dists
        [..., Nc + H*Nf, 1].  `rands` [..., Nc] fixes the stratification offsets."""
        lead = center.shape[:-1]
        c = center.reshape(-1, 3).contiguous().float()
        v = ray_unit.reshape(-1, 3).contiguous().float()
        R = c.shape[0]
        n_coarse, _, _ = self._n_samples()
        if rands is not None:
            rands = rands.reshape(R, n_coarse)
        dists = nerf_util.sample_dists((1, R), (near.reshape(1, R), far.reshape(1, R)), n_coarse,
                                       stratified, device=c.device, rands=rands).view(R, n_coarse)
        return self._hierarchical(c, v, dists).view(*lead, -1, 1)

    def _hierarchical(self, c, v, dists):
        """The hierarchy rounds of model.py:237-247 on flat rays: dists [R, Nc] -> [R, N]."""
        _, n_fine, n_hier = self._n_samples()
        if n_hier > 0:
            pts = c[:, None, :] + v[:, None, :] * dists[..., None]
            sdfs = self.neural_sdf.sdf_nograd(pts)[..., 0]
        for h in range(n_hier):
            merged, fine, pc, pf = nerf_util.hierarchical_round(dists, sdfs, 64 * 2 ** h, n_fine)
            if h != n_hier - 1:
                pts = c[:, None, :] + v[:, None, :] * fine[..., None]
                sdf_fine = self.neural_sdf.sdf_nograd(pts)[..., 0]
                sdfs = nerf_util.scatter_merge(sdfs, pc, sdf_fine, pf)
            dists = merged
        return dists

    # -- reference-signature operators (model.py:207-298) ------------------------
    @torch.no_grad()
    def get_dist_bounds(self, center, ray_unit):
        """model.py:207-213: [..., 3] -> near, far, outside [..., 1] (unit sphere)."""
        lead = center.shape[:-1]
        near, far, outside, _ = nerf_util.bounds_and_coarse(center.reshape(-1, 3),
                                                            ray_unit.reshape(-1, 3), 1)
        return near.view(*lead, 1), far.view(*lead, 1), outside.view(*lead, 1)

    def sample_dists_from_pdf(self, bin, weights):
        """model.py:43-44 partial: fine samples per the configured count."""
        return nerf_util.sample_dists_from_pdf(bin, weights, _get(_get(self.cfg_render,
                                                                       "num_samples"), "fine"))

    @torch.no_grad()
    def sample_dists_hierarchical(self, dists, sdfs, inv_s, robust=True, eps=1e-5):
        """model.py:250-267: dists [...,N,1], sdfs [...,N,1] -> dists_fine [...,Nf,1]."""
        if not robust or eps != 1e-5:
            raise NotImplementedError("only the robust estimator with eps=1e-5 is implemented")
        lead, N = dists.shape[:-2], dists.shape[-2]
        n_fine = _get(_get(self.cfg_render, "num_samples"), "fine")
        _, fine, _, _ = nerf_util.hierarchical_round(dists.reshape(-1, N).float(),
                                                     sdfs.reshape(-1, N).float(), inv_s, n_fine)
        return fine.view(*lead, n_fine, 1)

    def compute_neus_alphas(self, ray_unit, sdfs, gradients, dists, dist_far=None, progress=1.,
                            eps=1e-5):
        """model.py:275-292: ray_unit [...,3], sdfs [...,N,1], gradients [...,N,3],
        dists [...,N,1], dist_far [...,1,1] or None (1e10) -> alphas [...,N]."""
        if eps != 1e-5:
            raise NotImplementedError("the NeuS alpha kernel uses eps = 1e-5")
        lead, N = sdfs.shape[:-2], sdfs.shape[-2]
        R = int(torch.tensor(lead).prod().item()) if len(lead) else 1
        far = dist_far.reshape(R) if dist_far is not None else None
        alphas = Fn.NeusAlpha.apply(ray_unit.reshape(R, 3), sdfs.reshape(R, N),
                                    gradients.reshape(R, N, 3), dists.reshape(R, N), far, None,
                                    self.s_var, float(progress), float(self.anneal_end),
                                    float(self.outside_val))
        return alphas.view(*lead, N)

    def get_appearance_embedding(self, sample_idx, num_rays):
        """model.py:215-227; appearance embeddings are not part of this build."""
        return None, None

    # -- rendering ---------------------------------------------------------------
    def render_rays(self, center, ray_unit, sample_idx=None, stratified=False, *, rands=None,
                    dists=None, rands_bg=None):
        """model.py:124-156: center, ray_unit [..., 3] -> output dict with the reference's
        keys (rgb, opacity, outside, dists, weights, gradient, gradients, hessians) plus
        sdfs, alphas, near and far.  rands [..., Nc] fixes the coarse stratification,
        dists [..., N] replaces the sampling (parity testing), rands_bg [..., Nb] fixes
        the background stratification."""
        lead = center.shape[:-1]
        c = center.reshape(-1, 3).contiguous().float()
        v = ray_unit.reshape(-1, 3).contiguous().float()
        R = c.shape[0]
        with torch.no_grad():
            near, far, outside = self.get_dist_bounds(c, v)
        out = self.render_rays_object(c, v, near, far, outside, None, stratified=stratified,
                                      rands=rands, dists=dists)
        N = out["dists"].shape[1]
        weights, rgb, dists_all = out["weights"], out["rgb"], out["dists"]
        if self.with_background:
            # object and background samples composited together (model.py:129-141)
            bg = self.render_rays_background(c, v, far, None, stratified=stratified,
                                             rands=rands_bg)
            alphas = torch.cat([out["alphas"], bg["alphas"]], dim=1)
            rgbs = torch.cat([out["rgbs"], bg["rgbs"]], dim=1)
            weights, rgb = Fn.Composite.apply(alphas, rgbs.view(R, -1, 3))
            weights = weights[..., None]
            dists_all = torch.cat([out["dists"], bg["dists"]], dim=1)
        if self.white_background:
            rgb = rgb + (1 - weights.sum(dim=1))
        Nw = weights.shape[1]

        def shp(t, tail):
            return t.reshape(*lead, *tail) if t is not None else None
        return dict(
            rgb=shp(rgb, (3,)), opacity=shp(out["opacity"], (1,)), outside=shp(outside, (1,)),
            dists=shp(dists_all, (Nw, 1)), weights=shp(weights, (Nw, 1)),
            gradient=shp(out["gradient"], (3,)), gradients=shp(out["gradients"], (N, 3)),
            hessians=shp(out["hessians"], (N, 3)), sdfs=shp(out["sdfs"], (N,)),
            alphas=shp(out["alphas"], (N,)), near=shp(near, (1,)), far=shp(far, (1,)))

    def sample_dists_background(self, ray_unit, far, stratified=False, eps=1e-5, *, rands=None):
        """model.py:269-273: inverse distances from 1 down to 0 (sample_dists), then
        far / (inv + eps): ray_unit [...,3], far [...,1] -> [..., Nb, 1]."""
        nb = _get(_get(self.cfg_render, "num_samples"), "background")
        lead = ray_unit.shape[:-1]
        R = far.numel()
        inv = nerf_util.sample_dists((1, R), (1, 0), nb, stratified, device=far.device,
                                     rands=rands.reshape(R, nb) if rands is not None else None)
        return (far.reshape(R, 1, 1) / (inv.view(R, nb, 1) + eps)).view(*lead, nb, 1)

    def render_rays_background(self, center, ray_unit, far, app_outside, stratified=False, *,
                               rands=None):
        """model.py:192-205: background samples beyond the unit sphere, the background
        NeRF, and volume-rendering alphas (render.py:68-84).  Flat rays: center,
        ray_unit [R,3], far [R,1] -> rgbs [R,Nb,3], dists [R,Nb,1], alphas [R,Nb]."""
        if app_outside is not None:
            raise NotImplementedError("appearance embeddings are not part of this build")
        with torch.no_grad():
            dists = self.sample_dists_background(ray_unit, far, stratified=stratified,
                                                 rands=rands)[..., 0]
        R, Nb = dists.shape
        points = (center[:, None, :] + ray_unit[:, None, :] * dists[..., None]).reshape(-1, 3)
        rays = ray_unit[:, None, :].expand(R, Nb, 3).reshape(-1, 3)
        out = self.background_nerf.mlp_out(points)
        rgbs = self.background_nerf.rgb_from(out, rays).view(R, Nb, 3)
        alphas, _ = Fn.BgAlpha.apply(out, dists, R, Nb)
        return dict(rgbs=rgbs, dists=dists[..., None], alphas=alphas)

    def render_rays_object(self, center, ray_unit, near, far, outside, app, stratified=False, *,
                           rands=None, dists=None):
        """model.py:158-190: center, ray_unit [...,3], near, far, outside [...,1] ->
        rgbs [...,N,3], sdfs [...,N], dists [...,N,1], alphas [...,N], opacity [...,1]
        and gradient [...,3] (eval mode, else None), gradients / hessians [...,N,3];
        plus rgb [...,3] and weights [...,N,1] of the object samples alone."""
        if app is not None:
            raise NotImplementedError("appearance embeddings are not part of this build")
        lead = center.shape[:-1]
        c = center.reshape(-1, 3).contiguous().float()
        v = ray_unit.reshape(-1, 3).contiguous().float()
        R = c.shape[0]
        far_f = far.reshape(R).contiguous().float()
        outside_f = outside.reshape(R)
        if dists is None:
            dists = self.sample_dists_all(c, v, near.reshape(R, 1), far_f.view(R, 1),
                                          stratified=stratified, rands=rands)
        dists = dists.reshape(R, -1).contiguous().float()
        N = dists.shape[1]
        points = c[:, None, :] + v[:, None, :] * dists[..., None]  # [R,N,3]
        pts = points.reshape(-1, 3)
        # centre SDF + features and the four tap SDFs of the numerical gradient in
        # one node (model.py:160-166 and modules.py:157-175)
        sdfs, feats, taps, delta = self.neural_sdf.forward_with_taps(pts)  # [P,1], [P,256], [4,P]
        # sdfs[outside] = outside_val (model.py:163): the fill blocks the gradient
        fill = outside_f[:, None].expand(R, N).reshape(-1, 1)
        sdfs_filled = torch.where(fill, torch.full_like(sdfs, self.outside_val), sdfs)
        gradients, hessians = self.neural_sdf.gradients_from_taps(taps, sdfs, sdfs_filled,
                                                                  self.training, delta)
        rgbs = self.neural_rgb.forward(points, gradients.view(R, N, 3), v, feats, None,
                                       normals_are_gradients=True)  # [R,N,3]
        alphas = Fn.NeusAlpha.apply(v, sdfs.view(R, N), gradients.view(R, N, 3), dists, far_f,
                                    outside_f, self.s_var, float(self.progress),
                                    float(self.anneal_end), float(self.outside_val))
        weights, rgb = Fn.Composite.apply(alphas, rgbs)
        opacity = gradient = None
        if not self.training:
            opacity = weights.sum(dim=1, keepdim=True)
            gradient = Fn.WeightedSum.apply(gradients.view(R, N, 3), weights[..., None])

        def shp(t, tail):
            return t.reshape(*lead, *tail) if t is not None else None
        return dict(rgbs=shp(rgbs, (N, 3)), sdfs=shp(sdfs_filled, (N,)), dists=shp(dists, (N, 1)),
                    alphas=shp(alphas, (N,)), opacity=shp(opacity, (1,)),
                    gradient=shp(gradient, (3,)), gradients=shp(gradients, (N, 3)),
                    hessians=shp(hessians, (N, 3)), rgb=shp(rgb, (3,)),
                    weights=shp(weights, (N, 1)))

    # -- packed (ragged) rays -----------------------------------------------------
    def compute_neus_alphas_packed(self, ray_unit, sdfs, gradients, dists, offsets, dist_far=None,
                                   ray_idx=None, outside=None, progress=1.):
        """compute_neus_alphas (model.py:275-292) for packed rays: ray r owns samples
        offsets[r] .. offsets[r+1]-1 of sdfs [S,1], gradients [S,3], dists [S,1]; ray_unit
        [R,3], dist_far [R,1] or None -> alphas [S]."""
        S = sdfs.shape[0]
        if ray_idx is None:
            ray_idx = Fn.packed_ray_index(offsets, S)
        far = dist_far.reshape(-1) if dist_far is not None else None
        return Fn.NeusAlphaPacked.apply(ray_unit, sdfs.reshape(S), gradients.reshape(S, 3),
                                        dists.reshape(S), far, outside, offsets, ray_idx,
                                        self.s_var, float(progress), float(self.anneal_end),
                                        float(self.outside_val))

    def render_rays_packed(self, center, ray_unit, dists, offsets, validate=True):
        """render_rays_object (model.py:158-190) over packed rays with uneven sample
        counts: center, ray_unit [R,3]; dists [S] sorted within each ray; offsets int32
        [R+1].  The encoder and the MLPs run on the flat sample list (per-ray
        quantities read through the sample's ray index); the NeuS alphas and the
        compositing run in one persistent launch each way that balances rays of
        uneven sample counts across the chip.  Returns rgb [R,3] and the
        per-sample tensors (gradients/hessians [S,1,3] and outside [S] for the losses).
        validate: check the offsets against the sample count first (one host read)."""
        c = center.reshape(-1, 3).contiguous().float()
        v = ray_unit.reshape(-1, 3).contiguous().float()
        dists = dists.reshape(-1).contiguous().float()
        S = dists.shape[0]
        with torch.no_grad():
            near, far, outside = self.get_dist_bounds(c, v)
        far = far.reshape(-1)
        outside = outside.reshape(-1)
        ray_idx = Fn.packed_ray_index(offsets, S, validate=validate)
        ridx = ray_idx.long()
        v_s = v.index_select(0, ridx)
        points = c.index_select(0, ridx) + v_s * dists[:, None]
        sdfs, feats, taps, delta = self.neural_sdf.forward_with_taps(points)
        out_s = outside.index_select(0, ridx)
        sdfs_filled = torch.where(out_s[:, None], torch.full_like(sdfs, self.outside_val), sdfs)
        gradients, hessians = self.neural_sdf.gradients_from_taps(taps, sdfs, sdfs_filled,
                                                                  self.training, delta)
        rgbs = self.neural_rgb.forward(points, gradients, v_s, feats, None,
                                       normals_are_gradients=True)  # [S,3]
        # NeuS alphas and the compositing of every ray: one persistent launch each way
        alphas, weights, rgb = Fn.NeusCompositePacked.apply(
            v, sdfs, gradients, dists, far, outside, offsets, self.s_var, rgbs,
            float(self.progress), float(self.anneal_end), float(self.outside_val))
        return dict(rgb=rgb, rgbs=rgbs, sdfs=sdfs_filled.view(S), dists=dists, alphas=alphas,
                    weights=weights, ray_idx=ray_idx, outside=out_s,
                    gradients=gradients.view(S, 1, 3),
                    hessians=hessians.view(S, 1, 3) if hessians is not None else None)
