"""Drop-in volume-rendering utilities (projects/nerf/utils/render.py:87-112) on HIP kernels.

alpha_compositing_weights: w_i = alpha_i * prod_{j<i} (1 - alpha_j); the kernel
forms the transmittance in log space (exp of a float32 prefix sum of
log1p(-alpha)), exact zeros where alpha = 1, and an exact (division-free)
backward recurrence.
"""
import torch

from . import functional as Fn


def _flat(alphas):
    shape = alphas.shape
    return alphas.reshape(-1, shape[-1]), shape


def alpha_compositing_weights(alphas):
    """alphas [..., N] -> weights [..., N, 1]."""
    a, shape = _flat(alphas)
    weights, _ = Fn.Composite.apply(a, None)
    return weights.view(*shape, 1)


def composite_from_alphas(alphas, quantities):
    """Fused alpha_compositing_weights + composite: (weights [...,N,1], out [...,K])."""
    a, shape = _flat(alphas)
    q = quantities.reshape(a.shape[0], a.shape[1], -1)
    weights, out = Fn.Composite.apply(a, q)
    return weights.view(*shape, 1), out.view(*shape[:-1], q.shape[-1])


def composite(quantities, weights):
    """(quantities * weights).sum over samples (render.py:102-112): [..., N, K] x [..., N, 1]."""
    if not torch.is_tensor(quantities):
        quantities = torch.full_like(weights, float(quantities))
    return Fn.WeightedSum.apply(quantities.expand(*weights.shape[:-1], quantities.shape[-1]),
                                weights)


# ---------------------------------------------------------------------------
# packed (ragged) rays: ray r owns samples offsets[r] .. offsets[r+1]-1
# ---------------------------------------------------------------------------
def alpha_compositing_weights_packed(alphas, offsets):
    """alphas [S] of packed rays (offsets int32 [R+1]) -> weights [S, 1]."""
    weights, _ = Fn.CompositePacked.apply(alphas.reshape(-1), None, offsets)
    return weights.view(-1, 1)


def composite_from_alphas_packed(alphas, quantities, offsets):
    """Fused packed alpha_compositing_weights + composite: (weights [S,1], out [R,K])."""
    weights, out = Fn.CompositePacked.apply(alphas.reshape(-1), quantities.reshape(alphas.numel(), -1),
                                            offsets)
    return weights.view(-1, 1), out


def composite_packed(quantities, weights, offsets, ray_idx=None):
    """Per-ray sum of quantities [S,K] x weights [S,1] over packed rays -> [R,K]."""
    S = weights.shape[0]
    if not torch.is_tensor(quantities):
        quantities = torch.full((S, 1), float(quantities), device=weights.device)
    if ray_idx is None:
        ray_idx = Fn.packed_ray_index(offsets, S)
    return Fn.WeightedSumPacked.apply(quantities.reshape(S, -1), weights, offsets, ray_idx)
