"""Drop-in NeuralSDF / NeuralRGB (projects/neuralangelo/utils/modules.py:24-229).

Same constructor (a cfg node with the reference's keys), same methods and
state-dict keys (``tcnn_encoding.params``, ``mlp.linears.{0,1}.weight_{g,v}``,
``mlp.linear_sdf.*``), hash-grid encoding + 4-tap numerical gradients only
(the hot-path configuration; base.yaml:53-68).  ``compute_dtype`` selects FP32
or BF16 matrix-core inputs (FP32 masters and accumulation either way).
"""
from functools import partial

import numpy as np
import torch

from . import functional as Fn
from .encoding import Encoding
from .mlp import MLPforNeuralSDF, MLPwithSkipConnection
from ._lib import NA_ACT_NONE, NA_ACT_RELU, NA_ACT_SIGMOID
from .spherical_harmonics import get_spherical_harmonics


def _get(cfg, key, default=None):
    if isinstance(cfg, dict):
        return cfg.get(key, default)
    return getattr(cfg, key, default)


class NeuralSDF(torch.nn.Module):

    def __init__(self, cfg_sdf, compute_dtype=torch.float32):
        super().__init__()
        self.cfg_sdf = cfg_sdf
        self.compute_dtype = compute_dtype
        encoding_dim = self.build_encoding(_get(cfg_sdf, "encoding"))
        input_dim = 3 + encoding_dim
        self.build_mlp(_get(cfg_sdf, "mlp"), input_dim=input_dim)
        self.warm_up_end = 0
        self.anneal_levels = _get(_get(cfg_sdf, "encoding"), "levels")
        self.active_levels = self.anneal_levels
        self.normal_eps = 1.0 / self.resolutions[-1]

    def build_encoding(self, cfg_encoding):
        if _get(cfg_encoding, "type") != "hashgrid":
            raise NotImplementedError("only the hashgrid encoding is on the accelerated path")
        hg = _get(cfg_encoding, "hashgrid")
        l_min, l_max = _get(hg, "min_logres"), _get(hg, "max_logres")
        r_min, r_max = 2 ** l_min, 2 ** l_max
        num_levels = _get(cfg_encoding, "levels")
        self.growth_rate = np.exp((np.log(r_max) - np.log(r_min)) / (num_levels - 1))
        config = dict(otype="HashGrid", n_levels=num_levels, n_features_per_level=_get(hg, "dim"),
                      log2_hashmap_size=_get(hg, "dict_size"), base_resolution=2 ** l_min,
                      per_level_scale=self.growth_rate)
        table_dtype = torch.bfloat16 if self.compute_dtype == torch.bfloat16 else torch.float32
        self.tcnn_encoding = Encoding(3, config, dtype=table_dtype)
        rng = _get(hg, "range", [-2, 2])
        self.tcnn_encoding.vol_range = (float(rng[0]), float(rng[1]))
        self.resolutions = []
        for lv in range(0, num_levels):
            size = np.floor(r_min * self.growth_rate ** lv).astype(int) + 1
            self.resolutions.append(size)
        return _get(hg, "dim") * num_levels

    def build_mlp(self, cfg_mlp, input_dim=3):
        layer_dims = [input_dim] + [_get(cfg_mlp, "hidden_dim")] * _get(cfg_mlp, "num_layers") + \
            [_get(cfg_mlp, "hidden_dim")]
        if _get(cfg_mlp, "activ") != "softplus":
            raise NotImplementedError("SDF MLP activation must be softplus")
        beta = _get(_get(cfg_mlp, "activ_params", {}), "beta", 100)
        if float(beta) != Fn.SOFTPLUS_BETA:
            raise NotImplementedError("softplus beta must be 100")
        self.mlp = MLPforNeuralSDF(layer_dims, skip_connection=_get(cfg_mlp, "skip", []),
                                   use_weightnorm=_get(cfg_mlp, "weight_norm", True),
                                   geometric_init=_get(cfg_mlp, "geometric_init", True),
                                   out_bias=_get(cfg_mlp, "out_bias", 0.5),
                                   invert=_get(cfg_mlp, "inside_out", False),
                                   compute_dtype=self.compute_dtype)
        self._nograd_w = None  # (key, SDFWeights) of the sampling passes

    # -- reference API -------------------------------------------------------
    def forward(self, points_3D, with_sdf=True, with_feat=True):
        shape = points_3D.shape[:-1]
        pts = points_3D.reshape(-1, 3)
        W0, b0, W1, b1, ws, bs = self.mlp.weights()
        self.tcnn_encoding.active_levels = self.active_levels
        sdf, feat = Fn.SDFNetFunction.apply(pts, self.tcnn_encoding.params, W0, b0, W1, b1, ws, bs,
                                            self.tcnn_encoding, self.compute_dtype, with_feat,
                                            self.tcnn_encoding.n_output_dims)
        sdf = sdf.view(*shape, 1) if with_sdf else None
        if feat is not None:
            feat = feat.view(*shape, feat.shape[-1]) if len(shape) == 1 else _keep_view(feat, shape)
        return sdf, feat

    def sdf(self, points_3D):
        return self.forward(points_3D, with_sdf=True, with_feat=False)[0]

    def forward_with_taps(self, points):
        """forward(points) and the four tap SDFs of compute_gradients in one autograd node.

        points [P,3] -> (sdf [P,1], feat [P,256] buffer view, taps [4,P], delta).  FP32:
        taps are sdf(x + eps*k_i) for the tetrahedral k_i (modules.py:157-175), delta
        False.  BF16: taps are the increments sdf(x + eps*k_i) - sdf(x), delta True
        (evaluated without cancellation).  The backward merges the five encoder
        scatters of each point."""
        grad_cfg = _get(self.cfg_sdf, "gradient")
        if _get(grad_cfg, "mode", "numerical") != "numerical" or _get(grad_cfg, "taps", 4) != 4:
            raise NotImplementedError("only numerical 4-tap gradients are supported")
        W0, b0, W1, b1, ws, bs = self.mlp.weights()
        self.tcnn_encoding.active_levels = self.active_levels
        eps, _ = self.tap_eps()
        enc = self.tcnn_encoding
        pts = points.reshape(-1, 3)
        if self.compute_dtype == torch.bfloat16:
            sdf, feat, ds = Fn.SDFCentreTapsDeltaFunction.apply(
                pts, enc.params, W0, b0, W1, b1, ws, bs, enc, enc.n_output_dims, eps)
            return sdf, feat, ds, True
        sdf, feat, s_taps = Fn.SDFCentreTapsFunction.apply(
            pts, enc.params, W0, b0, W1, b1, ws, bs, enc, self.compute_dtype, enc.n_output_dims,
            eps)
        return sdf, feat, s_taps, False

    def gradients_from_taps(self, taps, sdf_centre, sdf, training, delta):
        """The 4-tap combination of compute_gradients (modules.py:163-175) on the taps
        of forward_with_taps; sdf is the value the reference passes in (outside fill
        included), sdf_centre the evaluated centre SDF."""
        eps, eps_sq = self.tap_eps()
        if delta:
            grad, hess = Fn.TapsDelta.apply(taps, sdf_centre.reshape(-1),
                                            sdf.reshape(-1) if training else None, eps, eps_sq,
                                            training)
        else:
            grad, hess = Fn.TapsCombine.apply(taps, sdf.reshape(-1) if training else None, eps,
                                              eps_sq, training)
        return grad, (hess if training else None)

    @torch.no_grad()
    def sdf_nograd(self, points_3D):
        """SDF values without autograd bookkeeping (sampling passes)."""
        shape = points_3D.shape[:-1]
        # the kernel-layout weights are rebuilt only when a parameter changed (its
        # version counter moves on every in-place update, the optimizer's included;
        # the storage pointer catches a swapped .data): the sampling passes of one
        # step share one build
        key = (self.compute_dtype,) + tuple((id(p), p._version, p.data_ptr())
                                            for p in self.mlp.parameters())
        if self._nograd_w is None or self._nograd_w[0] != key:
            W0, b0, W1, b1, ws, bs = [t.detach() for t in self.mlp.weights()]
            self._nograd_w = (key, Fn.SDFWeights(W0, b0, None, None, ws, bs,
                                                 self.tcnn_encoding.n_output_dims,
                                                 self.compute_dtype))
        w = self._nograd_w[1]
        self.tcnn_encoding.active_levels = self.active_levels
        return Fn.sdf_forward_nograd(self.tcnn_encoding, w, points_3D.reshape(-1, 3).contiguous()
                                     .float()).view(*shape, 1)

    def encode(self, points_3D):
        """[x, enc] in reference order (modules.py:76-95)."""
        vmin, vmax = self.tcnn_encoding.vol_range
        pn = (points_3D - vmin) / (vmax - vmin)
        self.tcnn_encoding.active_levels = self.active_levels
        enc = self.tcnn_encoding(pn.reshape(-1, 3)).view(*points_3D.shape[:-1], -1)
        return torch.cat([points_3D, enc.to(points_3D.dtype)], dim=-1)

    def set_active_levels(self, current_iter=None):
        c2f = _get(_get(self.cfg_sdf, "encoding"), "coarse2fine")
        anneal_levels = max((current_iter - self.warm_up_end) // _get(c2f, "step"), 1)
        self.anneal_levels = min(_get(_get(self.cfg_sdf, "encoding"), "levels"), anneal_levels)
        self.active_levels = max(_get(c2f, "init_active_level"), self.anneal_levels)

    def set_normal_epsilon(self):
        c2f = _get(_get(self.cfg_sdf, "encoding"), "coarse2fine")
        if _get(c2f, "enabled", True):
            epsilon_res = self.resolutions[self.anneal_levels - 1]
        else:
            epsilon_res = self.resolutions[-1]
        self.normal_eps = 1. / epsilon_res

    def tap_eps(self):
        """(eps, eps_sq) as float32 values the way modules.py:158-172 evaluates them."""
        eps64 = np.float64(self.normal_eps) / np.sqrt(3)
        return float(np.float32(eps64)), float(np.float32(eps64 ** 2))

    def compute_gradients(self, x, training=False, sdf=None):
        """4-tap numerical gradient and Hessian trace (modules.py:157-175)."""
        grad_mode = _get(_get(self.cfg_sdf, "gradient"), "mode", "numerical")
        taps = _get(_get(self.cfg_sdf, "gradient"), "taps", 4)
        if grad_mode != "numerical" or taps != 4:
            raise NotImplementedError("only numerical 4-tap gradients are supported")
        shape = x.shape[:-1]
        xf = x.reshape(-1, 3)
        eps, eps_sq = self.tap_eps()
        tp = Fn.taps_points(xf, eps)  # [4, n, 3]
        s_taps = self.sdf(tp.view(-1, 3)).view(4, -1)
        sdf_c = sdf.reshape(-1) if (training and sdf is not None) else None
        if training and sdf_c is None:
            raise AssertionError("sdf is required for the Hessian in training mode")
        grad, hess = Fn.TapsCombine.apply(s_taps, sdf_c, eps, eps_sq, training)
        gradient = grad.view(*shape, 3)
        hessian = hess.view(*shape, 3) if training else None
        return gradient, hessian


def _keep_view(feat, shape):
    """Reshape the [P,256] buffer view to [..., 256] without copying (splitting the
    row dimension keeps the buffer's row stride; a plain view keeps autograd from
    materialising a buffer-shaped gradient)."""
    return feat.view(*shape, feat.shape[-1])


class NeuralRGB(torch.nn.Module):

    def __init__(self, cfg_rgb, feat_dim, appear_embed, compute_dtype=torch.float32):
        super().__init__()
        self.cfg_rgb = cfg_rgb
        self.cfg_appear_embed = appear_embed
        self.compute_dtype = compute_dtype
        if _get(appear_embed, "enabled", False):
            raise NotImplementedError("appearance embeddings are outside the accelerated path")
        if _get(cfg_rgb, "mode", "idr") != "idr":
            raise NotImplementedError("only the idr colour mode is supported")
        encoding_view_dim = self.build_encoding(_get(cfg_rgb, "encoding_view"))
        input_dim = 6 + encoding_view_dim + feat_dim
        self.build_mlp(_get(cfg_rgb, "mlp"), input_dim=input_dim)

    def build_encoding(self, cfg_encoding_view):
        if _get(cfg_encoding_view, "type") != "spherical" or _get(cfg_encoding_view, "levels") != 3:
            raise NotImplementedError("view encoding must be spherical harmonics of degree 3")
        self.spherical_harmonic_encoding = partial(get_spherical_harmonics, levels=3)
        return 16

    def build_mlp(self, cfg_mlp, input_dim=3):
        layer_dims = [input_dim] + [_get(cfg_mlp, "hidden_dim")] * _get(cfg_mlp, "num_layers") + [3]
        if _get(cfg_mlp, "activ") not in ("relu", "relu_"):
            raise NotImplementedError("colour MLP activation must be relu")
        self.mlp = MLPwithSkipConnection(layer_dims, skip_connection=_get(cfg_mlp, "skip", []),
                                         use_weightnorm=_get(cfg_mlp, "weight_norm", True),
                                         compute_dtype=self.compute_dtype)

    def forward(self, points_3D, normals, rays_unit, feats, app, normals_are_gradients=False):
        """rgb [..., 3] = sigmoid(MLP([x, SH(view), normal, feat])) (modules.py:209-220).

        rays_unit: one direction per sample ([..., 3] like points_3D) or one per
        ray ([R, 3] with points_3D [R, N, 3]); the SH encoding is evaluated in
        the input kernel.  normals_are_gradients: ``normals`` holds raw SDF
        gradients and the unit normal (F.normalize, eps 1e-12) is formed in the
        same kernel, which also back-propagates through the normalisation."""
        if app is not None:
            raise NotImplementedError("appearance embeddings are outside the accelerated path")
        shape = points_3D.shape[:-1]
        P = int(np.prod(shape))
        if rays_unit.shape[:-1] == shape:
            rays, n_per_ray = rays_unit.reshape(P, 3), 1
        else:
            rays = rays_unit.reshape(-1, 3)
            n_per_ray = P // rays.shape[0]
            if n_per_ray * rays.shape[0] != P:
                raise ValueError("rays_unit must hold one direction per sample or per ray")
        feat2 = _flat_view(feats, P) if _is_buffer_view(feats) else feats.reshape(P, feats.shape[-1])
        Ws, bs = self.mlp.weights()
        mode = 1 if normals_are_gradients else 2
        rgb = Fn.RGBNetFunction.apply(feat2, points_3D.reshape(P, 3), rays, normals.reshape(P, 3),
                                      n_per_ray, mode, NA_ACT_SIGMOID, self.compute_dtype, *Ws, *bs)
        return rgb.view(*shape, 3)

    def encode_view(self, rays_unit):
        return self.spherical_harmonic_encoding(rays_unit)


def _is_buffer_view(t):
    return t.dim() >= 2 and t.stride(-1) == 1 and t.stride(-2) == Fn.RGB_LD


def _flat_view(t, P):
    """[..., C] buffer view -> [P, C] view (the leading dimensions merge exactly)."""
    return t.view(P, t.shape[-1])


class _LinearStack(torch.nn.Module):
    """nerf_util.MLPwithSkipConnection's parameters (nerf_util.py:158-184): plain
    Linear layers (default init, last bias zero) named linears.<i>.weight / bias."""

    def __init__(self, layer_dims, skip_connection):
        super().__init__()
        self.skip_connection = list(skip_connection)
        self.linears = torch.nn.ModuleList()
        pairs = list(zip(layer_dims[:-1], layer_dims[1:]))
        for li, (k_in, k_out) in enumerate(pairs):
            if li in self.skip_connection:
                k_in += layer_dims[0]
            lin = torch.nn.Linear(k_in, k_out)
            if li == len(pairs) - 1:
                with torch.no_grad():
                    lin.bias.zero_()
            self.linears.append(lin)


class BackgroundNeRF(torch.nn.Module):
    """Background NeRF (modules.py:232-311) on the GEMM kernels.

    Points are reparametrised on the inverted sphere, q = [p/|p|, 1/|p|], and
    Fourier-encoded in one kernel into a zero-padded input row; the feature MLP
    (ReLU, skip connection re-reading that row) ends in [density | feature]; the
    colour MLP reads [relu(feature) | SH16(view)] and ends in a sigmoid.  Density
    activation: softplus (beta 1).  Parameters keep the reference's names
    (mlp_feat.linears.<i>, mlp_rgb.linears.<i>)."""

    def __init__(self, cfg_background, appear_embed, compute_dtype=torch.float32):
        super().__init__()
        self.cfg_background = cfg_background
        if _get(appear_embed, "enabled", False):
            raise NotImplementedError("appearance embeddings are outside the accelerated path")
        enc = _get(cfg_background, "encoding")
        encv = _get(cfg_background, "encoding_view")
        mlp = _get(cfg_background, "mlp")
        if _get(enc, "type") != "fourier":
            raise NotImplementedError("background encoding must be fourier")
        if _get(encv, "type") != "spherical" or _get(encv, "levels") != 3:
            raise NotImplementedError("background view encoding must be spherical, degree 3")
        if not _get(cfg_background, "view_dep", True):
            raise NotImplementedError("only the view-dependent background is implemented")
        if _get(mlp, "activ") not in ("relu", "relu_") or _get(mlp, "activ_density") != "softplus":
            raise NotImplementedError("background activations must be relu / softplus")
        if _get(mlp, "activ_density_params", {}) or _get(mlp, "skip_rgb", []):
            raise NotImplementedError("softplus parameters and colour skips are not supported")
        self.compute_dtype = compute_dtype
        self.levels = _get(enc, "levels")
        self.input_dim = 4 + 8 * self.levels
        self.width = (self.input_dim + 7) // 8 * 8
        hid = _get(mlp, "hidden_dim")
        self.hidden = hid
        layer_dims = [self.input_dim] + [hid] * (_get(mlp, "num_layers") - 1) + [hid + 1]
        self.mlp_feat = _LinearStack(layer_dims, _get(mlp, "skip", []))
        layer_dims_rgb = [hid + 16] + [_get(mlp, "hidden_dim_rgb")] * (_get(mlp, "num_layers_rgb") - 1) + [3]
        self.mlp_rgb = _LinearStack(layer_dims_rgb, [])
        self.spherical_harmonic_encoding = partial(get_spherical_harmonics, levels=3)

    def _kernel_weight(self, li, lin, last):
        """Input columns in the kernel layout (the encoded block padded to self.width);
        the 1 + hidden outputs of the last layer padded to a multiple of 8."""
        W, b = lin.weight, lin.bias
        pad_in = self.width - self.input_dim
        if pad_in and (li == 0 or li in self.mlp_feat.skip_connection):
            W = torch.nn.functional.pad(W, (0, pad_in))  # the encoded block is last
        if last:
            n_out = (W.shape[0] + 7) // 8 * 8
            W = torch.nn.functional.pad(W, (0, 0, 0, n_out - W.shape[0]))
            b = torch.nn.functional.pad(b, (0, n_out - b.shape[0]))
        return W, b

    def mlp_out(self, points):
        """points [P,3] -> raw feature-MLP output [P, hidden + 1 padded] float32."""
        X = Fn.bg_encode(points.reshape(-1, 3), self.levels, self.compute_dtype, self.width)
        h = X
        n = len(self.mlp_feat.linears)
        for li, lin in enumerate(self.mlp_feat.linears):
            if li in self.mlp_feat.skip_connection:
                h = torch.cat([h, X], dim=-1)
            last = li == n - 1
            W, b = self._kernel_weight(li, lin, last)
            h = Fn.LinearAct.apply(h, W, b, NA_ACT_NONE if last else NA_ACT_RELU,
                                   torch.float32 if last else self.compute_dtype)
        return h

    def rgb_from(self, out, rays_unit):
        """Colour MLP on relu(feature) and SH16 of the view direction per point."""
        feat = torch.relu(out[:, 1:1 + self.hidden])
        view = self.spherical_harmonic_encoding(rays_unit.reshape(-1, 3).float())
        h = torch.cat([feat, view], dim=-1).to(self.compute_dtype).contiguous()
        n = len(self.mlp_rgb.linears)
        for li, lin in enumerate(self.mlp_rgb.linears):
            last = li == n - 1
            h = Fn.LinearAct.apply(h, lin.weight, lin.bias, NA_ACT_SIGMOID if last else NA_ACT_RELU,
                                   torch.float32 if last else self.compute_dtype)
        return h

    def forward(self, points_3D, rays_unit, app_outside):
        """(rgb [..., 3], density [...]) as modules.py:264-280."""
        if app_outside is not None:
            raise NotImplementedError("appearance embeddings are outside the accelerated path")
        shape = points_3D.shape[:-1]
        out = self.mlp_out(points_3D.reshape(-1, 3).contiguous().float())
        density = torch.nn.functional.softplus(out[:, 0])
        rgb = self.rgb_from(out, rays_unit.expand_as(points_3D))
        return rgb.view(*shape, 3), density.view(shape)
