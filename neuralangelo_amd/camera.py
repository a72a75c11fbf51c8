"""Camera rays on the GPU (projects/nerf/utils/camera.py:283-320).

``get_center_and_ray`` and the ray selection of the renderer run in one kernel
(``na_camera_rays``): for every requested pixel it forms the pixel-centre point,
maps it through K^-1 and the inverse world-to-camera pose, and writes the
camera centre, the depth-1 ray and (optionally) its unit direction.
"""
import torch

from . import _lib


def _batched(pose, intr):
    single = pose.dim() == 2
    P = pose.reshape(-1, 3, 4).contiguous().float()
    K = intr.reshape(-1, 3, 3).contiguous().float()
    if K.shape[0] != P.shape[0]:
        K = K.expand(P.shape[0], 3, 3).contiguous()
    return single, P, K


def camera_rays(pose, intr, image_size, ray_idx=None, unit=True):
    """Centres, depth-1 rays and unit rays of the pixels ray_idx [B,R] (int64) of
    pose [B,3,4] / intr [B,3,3] cameras, or of every pixel when ray_idx is None.
    Returns center, ray, ray_unit [B,R,3] (ray_unit None unless `unit`)."""
    _lib.require_cuda(pose, intr)
    single, P, K = _batched(pose, intr)
    B = P.shape[0]
    H, W = int(image_size[0]), int(image_size[1])
    if ray_idx is not None:
        idx = ray_idx.reshape(B, -1).to(device=P.device, dtype=torch.int64).contiguous()
        R = idx.shape[1]
    else:
        idx, R = None, H * W
    center = torch.empty(B, R, 3, device=P.device)
    ray = torch.empty(B, R, 3, device=P.device)
    ray_unit = torch.empty(B, R, 3, device=P.device) if unit else None
    _lib.call("na_camera_rays", _lib.ptr(P), _lib.ptr(K), B, H, W, _lib.ptr(idx), R,
              _lib.ptr(center), _lib.ptr(ray), _lib.ptr(ray_unit), _lib.stream_of(P))
    if single:
        center, ray = center[0], ray[0]
        ray_unit = ray_unit[0] if ray_unit is not None else None
    return center, ray, ray_unit


def get_center_and_ray(pose, intr, image_size):
    """camera.py:283-311: pose [3,4]/[B,3,4], intr [3,3]/[B,3,3] -> center, ray
    [HW,3]/[B,HW,3] (ray has depth 1, not unit length)."""
    center, ray, _ = camera_rays(pose, intr, image_size, unit=False)
    return center, ray


def get_3D_points_from_dist(center, ray_unit, dist, multi=True):
    """camera.py:314-320: center + ray_unit * dist (per sample when `multi`)."""
    if multi:
        center, ray_unit = center[..., None, :], ray_unit[..., None, :]
    return center + ray_unit * dist


def ray_chunks(pose, intr, image_size, num_rays):
    """Every pixel of the [B] cameras in chunks of `num_rays` pixels (the full-image
    mode of nerf_util.py:86-124 without NDC): yields center, ray, ray_unit [B,r,3]."""
    H, W = int(image_size[0]), int(image_size[1])
    B = pose.reshape(-1, 3, 4).shape[0]
    for c0 in range(0, H * W, num_rays):
        c1 = min(c0 + num_rays, H * W)
        idx = torch.arange(c0, c1, device=pose.device, dtype=torch.int64).expand(B, c1 - c0)
        yield camera_rays(pose.reshape(B, 3, 4), intr.reshape(-1, 3, 3), image_size, ray_idx=idx)
