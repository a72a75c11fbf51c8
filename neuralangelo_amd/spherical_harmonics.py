"""Real spherical harmonics of the view direction (projects/neuralangelo/utils/spherical_harmonics.py:47-84).

The colour-MLP input kernel evaluates the degree-3 basis on the GPU inside the
input builder; this torch version serves the standalone drop-in API.  The
constants are the closed forms of the real SH normalisation factors.
"""
import math

import torch

_PI = math.pi
SH_C0 = 1 / (2 * math.sqrt(_PI))
SH_C1 = math.sqrt(3) / (2 * math.sqrt(_PI))
SH_C2 = [math.sqrt(15) / (2 * math.sqrt(_PI)), -math.sqrt(15) / (2 * math.sqrt(_PI)),
         math.sqrt(5) / (4 * math.sqrt(_PI)), -math.sqrt(15) / (2 * math.sqrt(_PI)),
         math.sqrt(15) / (4 * math.sqrt(_PI))]
SH_C3 = [-math.sqrt(35 / 2) / (4 * math.sqrt(_PI)), math.sqrt(105) / (2 * math.sqrt(_PI)),
         -math.sqrt(21 / 2) / (4 * math.sqrt(_PI)), math.sqrt(7) / (4 * math.sqrt(_PI)),
         -math.sqrt(21 / 2) / (4 * math.sqrt(_PI)), math.sqrt(105) / (4 * math.sqrt(_PI)),
         -math.sqrt(35 / 2) / (4 * math.sqrt(_PI))]


def get_spherical_harmonics(dirs, levels):
    if levels > 3:
        raise NotImplementedError("degree > 3 is not used on the accelerated path")
    vals = torch.empty((*dirs.shape[:-1], (levels + 1) ** 2), device=dirs.device, dtype=dirs.dtype)
    vals[..., 0] = SH_C0
    if levels >= 1:
        x, y, z = dirs.unbind(-1)
        vals[..., 1] = -SH_C1 * y
        vals[..., 2] = SH_C1 * z
        vals[..., 3] = -SH_C1 * x
    if levels >= 2:
        xx, yy, zz = x * x, y * y, z * z
        xy, yz, xz = x * y, y * z, x * z
        vals[..., 4] = SH_C2[0] * xy
        vals[..., 5] = SH_C2[1] * yz
        vals[..., 6] = SH_C2[2] * (2.0 * zz - xx - yy)
        vals[..., 7] = SH_C2[3] * xz
        vals[..., 8] = SH_C2[4] * (xx - yy)
    if levels >= 3:
        vals[..., 9] = SH_C3[0] * y * (3 * xx - yy)
        vals[..., 10] = SH_C3[1] * xy * z
        vals[..., 11] = SH_C3[2] * y * (4 * zz - xx - yy)
        vals[..., 12] = SH_C3[3] * z * (2 * zz - 3 * xx - 3 * yy)
        vals[..., 13] = SH_C3[4] * x * (4 * zz - xx - yy)
        vals[..., 14] = SH_C3[5] * z * (xx - yy)
        vals[..., 15] = SH_C3[6] * x * (xx - 3 * yy)
    return vals
