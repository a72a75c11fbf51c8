"""Drop-in MLP modules with the reference's constructor signatures and state-dict keys.

MLPforNeuralSDF       projects/neuralangelo/utils/mlp.py:18-84
MLPwithSkipConnection projects/nerf/utils/nerf_util.py:158-196

Each weight-normed linear keeps torch.nn.utils.weight_norm's parameter names
(``weight_g`` [N,1], ``weight_v`` [N,K], ``bias``) so reference checkpoints
load unchanged; the forward passes run on the HIP GEMM kernels.  Supported
configurations are the ones the hot path uses: SDF MLP with one hidden layer
(layer_dims [K0, H, H]) and softplus, colour MLP with ReLU and no skip
connections; others raise NotImplementedError.
"""
import numpy as np
import torch
import torch.nn.functional as torch_F

from . import functional as Fn
from ._lib import NA_ACT_NONE, NA_ACT_SIGMOID


class WNLinear(torch.nn.Module):
    """nn.Linear wrapped by weight_norm(dim=0): W = weight_g * weight_v / ||weight_v||_row."""

    def __init__(self, k_in, k_out, weight_norm=True):
        super().__init__()
        lin = torch.nn.Linear(k_in, k_out)  # default init, as the reference's Linear
        self.weight_norm = weight_norm
        if weight_norm:
            self.weight_v = torch.nn.Parameter(lin.weight.detach().clone())
            self.weight_g = torch.nn.Parameter(lin.weight.detach().norm(dim=1, keepdim=True))
        else:
            self.weight = torch.nn.Parameter(lin.weight.detach().clone())
        self.bias = torch.nn.Parameter(lin.bias.detach().clone())
        self.in_features, self.out_features = k_in, k_out

    def set_weight(self, W):
        """Re-initialise from a plain weight (the g,v split keeps W unchanged)."""
        with torch.no_grad():
            if self.weight_norm:
                self.weight_v.copy_(W)
                self.weight_g.copy_(W.norm(dim=1, keepdim=True))
            else:
                self.weight.copy_(W)

    def effective_weight(self):
        if self.weight_norm:
            return Fn.weight_norm(self.weight_v, self.weight_g)
        return self.weight


class MLPforNeuralSDF(torch.nn.Module):

    def __init__(self, layer_dims, skip_connection=[], activ=None, use_layernorm=False,
                 use_weightnorm=False, geometric_init=False, out_bias=0., invert=False,
                 softplus_beta=100.0, compute_dtype=torch.float32):
        super().__init__()
        if len(layer_dims) != 3 or layer_dims[1] != layer_dims[2]:
            raise NotImplementedError("MLPforNeuralSDF: only layer_dims [K0, H, H] (num_layers=1)")
        if skip_connection or use_layernorm:
            raise NotImplementedError("MLPforNeuralSDF: skip connections / layernorm unsupported")
        self.skip_connection = skip_connection
        self.use_layernorm = use_layernorm
        self.softplus_beta = softplus_beta
        self.compute_dtype = compute_dtype
        self.linears = torch.nn.ModuleList()
        pairs = list(zip(layer_dims[:-1], layer_dims[1:]))
        for li, (k_in, k_out) in enumerate(pairs):
            lin = WNLinear(k_in, k_out, weight_norm=use_weightnorm)
            if geometric_init:
                # mlp.py:71-77
                W = torch.empty(k_out, k_in).normal_(0.0, np.sqrt(2 / k_out))
                if li == 0:
                    W[:, 3:] = 0.0
                lin.set_weight(W)
                with torch.no_grad():
                    lin.bias.zero_()
            if li == len(pairs) - 1:
                with torch.no_grad():
                    lin.bias.zero_()
            self.linears.append(lin)
        k_in = layer_dims[-2]
        self.linear_sdf = torch.nn.Linear(k_in, 1)
        if geometric_init:
            # mlp.py:79-84
            torch.nn.init.normal_(self.linear_sdf.weight, mean=np.sqrt(np.pi / k_in), std=0.0001)
            torch.nn.init.constant_(self.linear_sdf.bias, -out_bias)
            if invert:
                with torch.no_grad():
                    self.linear_sdf.weight.mul_(-1)
                    self.linear_sdf.bias.mul_(-1)

    def weights(self):
        l0, l1 = self.linears
        return (l0.effective_weight(), l0.bias, l1.effective_weight(), l1.bias,
                self.linear_sdf.weight, self.linear_sdf.bias)

    def forward(self, input, with_sdf=True, with_feat=True):
        """input [..., K0] in reference column order -> [sdf [...,1] | None, feat [...,H] | None]."""
        shape = input.shape[:-1]
        x = input.reshape(-1, input.shape[-1])
        n_enc = x.shape[1] - 3
        # kernel layout (encoder block first; split point columns in bf16), built with
        # differentiable torch ops so the input gradient flows back through it
        X0 = Fn.reference_input_to_kernel(x, n_enc, self.compute_dtype)
        W0, b0, W1, b1, ws, bs = self.weights()
        sdf, feat = Fn.SDFNetFunction.apply(X0, None, W0, b0, W1, b1, ws, bs, None,
                                            self.compute_dtype, with_feat, n_enc)
        sdf = sdf.view(*shape, 1) if with_sdf else None
        feat = feat.reshape(*shape, -1) if with_feat else None
        return [sdf, feat]


class MLPwithSkipConnection(torch.nn.Module):

    def __init__(self, layer_dims, skip_connection=[], activ=None, use_layernorm=False,
                 use_weightnorm=False, compute_dtype=torch.float32):
        super().__init__()
        if skip_connection or use_layernorm:
            raise NotImplementedError("MLPwithSkipConnection: skip connections / layernorm unsupported")
        if layer_dims[-1] > 4 or len(set(layer_dims[1:-1])) != 1:
            raise NotImplementedError("MLPwithSkipConnection: equal hidden widths and <= 4 outputs")
        self.skip_connection = skip_connection
        self.use_layernorm = use_layernorm
        self.compute_dtype = compute_dtype
        self.linears = torch.nn.ModuleList()
        pairs = list(zip(layer_dims[:-1], layer_dims[1:]))
        for li, (k_in, k_out) in enumerate(pairs):
            lin = WNLinear(k_in, k_out, weight_norm=use_weightnorm)
            if li == len(pairs) - 1:
                with torch.no_grad():
                    lin.bias.zero_()
            self.linears.append(lin)

    def weights(self):
        return ([lin.effective_weight() for lin in self.linears],
                [lin.bias for lin in self.linears])

    def run(self, feat, points, ray_unit, normals_src, n_per_ray, build_input, final_act):
        Ws, bs = self.weights()
        return Fn.RGBNetFunction.apply(feat, points, ray_unit, normals_src, n_per_ray, build_input,
                                       final_act, self.compute_dtype, *Ws, *bs)

    def forward(self, input):
        """input [..., 278] in reference order [x, SH16, normal, feat] -> [..., 3] (no final act)."""
        shape = input.shape[:-1]
        x = input.reshape(-1, input.shape[-1])
        n_feat = x.shape[1] - 22
        FB = torch_F.pad(torch.cat([x[:, 22:], x[:, :22]], dim=1), (0, Fn.RGB_LD - n_feat - 22))
        FB = FB.to(self.compute_dtype)
        out = self.run(FB, None, None, None, 1, False, NA_ACT_NONE)
        return out.view(*shape, 3)
