"""Real spherical harmonics of view directions, on the GPU
(projects/neuralangelo/utils/spherical_harmonics.py:47-71, get_spherical_harmonics).

The basis is evaluated by ``na_sh_encode``, the same device code that builds the
colour network's input rows, so the standalone encoding and the fused one agree
bit for bit.  Degrees 0..3 are supported; the hot path uses degree 3.
"""
import torch

from . import _lib


def get_spherical_harmonics(dirs, levels):
    """dirs [..., 3] (unit vectors, on the GPU) -> [..., (levels+1)^2] float32."""
    if levels > 3:
        raise NotImplementedError("degree > 3 is not used on the accelerated path")
    _lib.require_cuda(dirs)
    d = dirs.reshape(-1, 3).contiguous().float()
    nb = (levels + 1) ** 2
    out = torch.empty(d.shape[0], nb, device=d.device)
    _lib.call("na_sh_encode", _lib.ptr(d), d.shape[0], int(levels), _lib.ptr(out),
              _lib.stream_of(d))
    return out.view(*dirs.shape[:-1], nb)
