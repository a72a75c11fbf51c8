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

from . import functional as Fn
from . import nerf_util
from .modules import BackgroundNeRF, NeuralRGB, NeuralSDF, _get


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

    def __init__(self, cfg_model, cfg_data=None, compute_dtype=torch.float32):
        super().__init__()
        self.cfg_render = _get(cfg_model, "render")
        bg = _get(cfg_model, "background")
        self.white_background = _get(bg, "white", False)
        self.with_background = bool(_get(bg, "enabled", False))
        obj = _get(cfg_model, "object")
        self.anneal_end = _get(_get(obj, "s_var"), "anneal_end")
        self.outside_val = 1000. * (-1 if _get(_get(_get(obj, "sdf"), "mlp"), "inside_out") else 1)
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

    # -- sampling --------------------------------------------------------------
    def _n_samples(self):
        ns = _get(self.cfg_render, "num_samples")
        return _get(ns, "coarse"), _get(ns, "fine"), _get(self.cfg_render, "num_sample_hierarchy")

    @torch.no_grad()
    def sample_dists_all(self, center, ray_unit, coarse):
        """model.py:232-248 from coarse dists [R, Nc] -> sorted dists [R, N]."""
        _, n_fine, n_hier = self._n_samples()
        dists = coarse
        if n_hier > 0:
            pts = center[:, None, :] + ray_unit[:, None, :] * dists[..., None]
            sdfs = self.neural_sdf.sdf_nograd(pts)[..., 0]
        for h in range(n_hier):
            merged, fine, pc, pf = nerf_util.hierarchical_round(dists, sdfs, 64 * 2 ** h, n_fine)
            if h != n_hier - 1:
                pts = center[:, None, :] + ray_unit[:, None, :] * fine[..., None]
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
        """model.py:229-230 delegate: fine samples per the configured count."""
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

    # -- rendering ---------------------------------------------------------------
    def render_rays(self, center, ray_unit, sample_idx=None, stratified=False, rands=None,
                    dists=None, rands_bg=None):
        """center, ray_unit [R,3] (or [B,R,3]) -> output dict (model.py:124-156).

        rands: optional [R, Nc] stratification offsets; dists: optional [R, N]
        samples to render instead of sampling (parity testing); rands_bg: optional
        [R, Nb] stratification offsets of the background samples."""
        lead = center.shape[:-1]
        c = center.reshape(-1, 3).contiguous().float()
        v = ray_unit.reshape(-1, 3).contiguous().float()
        n_coarse, _, _ = self._n_samples()
        if rands is None and stratified:
            rands = torch.rand(c.shape[0], n_coarse, device=c.device)
        with torch.no_grad():
            near, far, outside, coarse = nerf_util.bounds_and_coarse(c, v, n_coarse, rands)
        out = self.render_rays_object(c, v, near, far, outside, coarse, dists=dists)
        rgb, weights, dists_all = out["rgb"], out["weights"], out["dists"]
        if self.with_background:
            # object and background samples composited together (model.py:129-141)
            bg = self.render_rays_background(c, v, far, stratified=stratified, rands=rands_bg)
            alphas = torch.cat([out["alphas"], bg["alphas"]], dim=1)
            rgbs = torch.cat([out["rgbs"].view(c.shape[0], -1, 3), bg["rgbs"]], dim=1)
            weights, rgb = Fn.Composite.apply(alphas, rgbs)
            dists_all = torch.cat([out["dists"], bg["dists"]], dim=1)
        if self.white_background:
            rgb = rgb + (1 - weights.sum(dim=1, keepdim=False)[..., None])
        R = c.shape[0]
        N = out["dists"].shape[1]
        Nw = weights.shape[1]

        def shp(t, tail):
            return t.view(*lead, *tail)
        res = dict(
            rgb=shp(rgb, (3,)), opacity=shp(out["opacity"], (1,)) if out["opacity"] is not None else None,
            outside=shp(outside, (1,)), dists=shp(dists_all, (Nw, 1)),
            weights=shp(weights, (Nw, 1)),
            gradient=shp(out["gradient"], (3,)) if out["gradient"] is not None else None,
            gradients=shp(out["gradients"], (N, 3)),
            hessians=shp(out["hessians"], (N, 3)) if out["hessians"] is not None else None,
            sdfs=shp(out["sdfs"], (N,)), alphas=shp(out["alphas"], (N,)),
            near=shp(near, (1,)), far=shp(far, (1,)))
        del R
        return res

    def sample_dists_background(self, ray_unit, far, stratified=False, eps=1e-5, rands=None):
        """model.py:269-273: inverse distances (i + u)/Nb from 1 down to 0, dists =
        far / (inv + eps) -> [R, Nb] (u = 0.5, or U(0,1) when stratified)."""
        nb = _get(_get(self.cfg_render, "num_samples"), "background")
        R = ray_unit.reshape(-1, 3).shape[0]
        if rands is None:
            rands = (torch.rand(R, nb, device=ray_unit.device) if stratified else
                     torch.full((R, nb), 0.5, device=ray_unit.device))
        i = torch.arange(nb, device=ray_unit.device, dtype=torch.float32)
        inv = (i[None, :] + rands) / nb * (0.0 - 1.0) + 1.0
        return far.reshape(R, 1) / (inv + eps)

    def render_rays_background(self, center, ray_unit, far, stratified=False, rands=None):
        """model.py:192-205: background samples beyond the unit sphere, the background
        NeRF, and volume-rendering alphas (render.py:68-84)."""
        with torch.no_grad():
            dists = self.sample_dists_background(ray_unit, far, stratified=stratified, rands=rands)
        R, Nb = dists.shape
        points = (center[:, None, :] + ray_unit[:, None, :] * dists[..., None]).reshape(-1, 3)
        rays = ray_unit[:, None, :].expand(R, Nb, 3).reshape(-1, 3)
        out = self.background_nerf.mlp_out(points)
        rgbs = self.background_nerf.rgb_from(out, rays).view(R, Nb, 3)
        alphas, _ = Fn.BgAlpha.apply(out, dists, R, Nb)
        return dict(rgbs=rgbs, dists=dists, alphas=alphas)

    def render_rays_object(self, center, ray_unit, near, far, outside, coarse, dists=None):
        if dists is None:
            dists = self.sample_dists_all(center, ray_unit, coarse)  # [R, N]
        dists = dists.contiguous().float()
        R, N = dists.shape
        points = center[:, None, :] + ray_unit[:, None, :] * dists[..., None]  # [R,N,3]
        pts = points.reshape(-1, 3)
        # centre SDF + features and the four tap SDFs of the numerical gradient in
        # one node (model.py:160-166 and modules.py:157-175)
        sdfs, feats, taps, delta = self.neural_sdf.forward_with_taps(pts)  # [P,1], [P,256], [4,P]
        # sdfs[outside] = outside_val (model.py:163): the fill blocks the gradient
        fill = outside[:, None].expand(R, N).reshape(-1, 1)
        sdfs_filled = torch.where(fill, torch.full_like(sdfs, self.outside_val), sdfs)
        gradients, hessians = self.neural_sdf.gradients_from_taps(taps, sdfs, sdfs_filled,
                                                                  self.training, delta)
        rgbs = self.neural_rgb.forward(points, gradients.view(R, N, 3), ray_unit, feats, None,
                                       normals_are_gradients=True)  # [R,N,3]
        alphas = Fn.NeusAlpha.apply(ray_unit, sdfs.view(R, N), gradients.view(R, N, 3), dists, far,
                                    outside, self.s_var, float(self.progress),
                                    float(self.anneal_end), float(self.outside_val))
        weights, rgb = Fn.Composite.apply(alphas, rgbs)
        opacity = gradient = None
        if not self.training:
            opacity = weights.sum(dim=1, keepdim=True)
            gradient = Fn.WeightedSum.apply(gradients.view(R, N, 3), weights[..., None])
        return dict(rgb=rgb, rgbs=rgbs, sdfs=sdfs_filled.view(R, N), dists=dists, alphas=alphas,
                    weights=weights, opacity=opacity, gradient=gradient,
                    gradients=gradients.view(R, N, 3),
                    hessians=hessians.view(R, N, 3) if hessians is not None else None)

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

    def render_rays_packed(self, center, ray_unit, dists, offsets):
        """render_rays_object (model.py:158-190) over packed rays with uneven sample
        counts: center, ray_unit [R,3]; dists [S] sorted within each ray; offsets int32
        [R+1].  Every kernel of the dense path runs on the packed samples; per-ray
        quantities are read through the sample's ray index.  Returns rgb [R,3] and the
        per-sample tensors (gradients/hessians [S,1,3] and outside [S] for the losses)."""
        c = center.reshape(-1, 3).contiguous().float()
        v = ray_unit.reshape(-1, 3).contiguous().float()
        dists = dists.reshape(-1).contiguous().float()
        S = dists.shape[0]
        with torch.no_grad():
            near, far, outside = self.get_dist_bounds(c, v)
        far = far.reshape(-1)
        outside = outside.reshape(-1)
        ray_idx = Fn.packed_ray_index(offsets, S)
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
        alphas = self.compute_neus_alphas_packed(v, sdfs, gradients, dists, offsets, far,
                                                 ray_idx=ray_idx, outside=outside,
                                                 progress=self.progress)
        weights, rgb = Fn.CompositePacked.apply(alphas, rgbs, offsets)
        return dict(rgb=rgb, rgbs=rgbs, sdfs=sdfs_filled.view(S), dists=dists, alphas=alphas,
                    weights=weights, ray_idx=ray_idx, outside=out_s,
                    gradients=gradients.view(S, 1, 3),
                    hessians=hessians.view(S, 1, 3) if hessians is not None else None)

    def forward(self, data):
        return self.render_rays(data["center"], data["ray_unit"], stratified=True,
                                rands=data.get("rands"))

    @torch.no_grad()
    def inference(self, center, ray_unit, chunk=4096):
        self.eval()
        out = defaultdict(list)
        for i in range(0, center.shape[0], chunk):
            o = self.render_rays(center[i:i + chunk], ray_unit[i:i + chunk], stratified=False)
            for k, v in o.items():
                if v is not None:
                    out[k].append(v)
        return {k: torch.cat(v, dim=0) for k, v in out.items()}
