"""Drop-in ray-sampling utilities (projects/nerf/utils/nerf_util.py:20-68,199-205)."""
import ctypes

import torch

from . import _lib


def _per_ray(v, batch_size, num_rays, device):
    """A distance bound as (pointer tensor or None, constant): tensors broadcast to [B*R]."""
    if not torch.is_tensor(v):
        return None, float(v)
    t = v.to(device=device, dtype=torch.float32)
    while t.dim() > 2 and t.shape[-1] == 1:  # [B,R,1,1] -> [B,R]
        t = t.squeeze(-1)
    return torch.broadcast_to(t, (batch_size, num_rays)).contiguous().view(-1), 0.0


def sample_dists(ray_size, dist_range, intvs, stratified, device="cuda", rands=None):
    """[B,R,intvs,1] distances (nerf_util.py:20-38) from ``na_sample_dists``.

    dist_range = (dist_min, dist_max), each a number or a tensor broadcastable to
    [B,R,1,1]; `rands` optionally supplies the stratification offsets [B,R,intvs(,1)]."""
    batch_size, num_rays = ray_size
    dist_min, dist_max = dist_range
    dev = torch.device(device)
    if dev.type == "cuda" and dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    lo, lo_c = _per_ray(dist_min, batch_size, num_rays, dev)
    hi, hi_c = _per_ray(dist_max, batch_size, num_rays, dev)
    R = batch_size * num_rays
    if rands is None and stratified:
        rands = torch.rand(batch_size, num_rays, intvs, 1, device=dev)
    r = rands.reshape(R, intvs).contiguous().float() if rands is not None else None
    out = torch.empty(R, intvs, device=dev)
    _lib.require_cuda(out)
    _lib.call("na_sample_dists", _lib.ptr(lo), _lib.ptr(hi), lo_c, hi_c, _lib.ptr(r), R,
              int(intvs), _lib.ptr(out), _lib.stream_of(out))
    return out.view(batch_size, num_rays, intvs, 1)


def sample_dists_from_pdf(bin, weights, intvs_fine):
    """Inverse-CDF resampling on the GPU (nerf_util.py:41-68).

    bin [..., N, 1], weights [..., N-1] -> [..., Nf, 1].  The CDF is a serial
    float32 prefix per ray, so searchsorted indices are reproducible."""
    _lib.require_cuda(bin, weights)
    N = bin.shape[-2]
    b = bin.reshape(-1, N).contiguous().float()
    w = weights.reshape(-1, N - 1).contiguous().float()
    R = b.shape[0]
    out = torch.empty(R, intvs_fine, device=b.device)
    _lib.call("na_sample_pdf", _lib.ptr(b), _lib.ptr(w), R, N, int(intvs_fine), _lib.ptr(out),
              _lib.stream_of(b))
    return out.view(*bin.shape[:-2], intvs_fine, 1)


def intersect_with_sphere(center, ray_unit, radius=1.0):
    """Sphere crossings (nerf_util.py:199-205) from ``na_sphere_crossings``:
    center, ray_unit [..., 3] -> dist_near, dist_far [..., 1] (NaN where a ray misses)."""
    _lib.require_cuda(center, ray_unit)
    c = center.reshape(-1, 3).contiguous().float()
    v = torch.broadcast_to(ray_unit, center.shape).reshape(-1, 3).contiguous().float()
    R = c.shape[0]
    near = torch.empty(R, device=c.device)
    far = torch.empty(R, device=c.device)
    _lib.call("na_sphere_crossings", _lib.ptr(c), _lib.ptr(v), R, float(radius), _lib.ptr(near),
              _lib.ptr(far), _lib.stream_of(c))
    lead = center.shape[:-1]
    return near.view(*lead, 1), far.view(*lead, 1)


def bounds_and_coarse(center, ray_unit, n_coarse, rands=None):
    """Fused unit-sphere bounds (model.py:207-213) + coarse sampling (nerf_util.py:20-38).

    center, ray_unit [R,3]; rands [R, n_coarse] or None (u = 0.5).
    Returns near [R], far [R], outside [R] (bool), dists [R, n_coarse]."""
    _lib.require_cuda(center, ray_unit)
    c = center.contiguous().float()
    v = ray_unit.contiguous().float()
    R = c.shape[0]
    dev = c.device
    near = torch.empty(R, device=dev)
    far = torch.empty(R, device=dev)
    outside = torch.empty(R, dtype=torch.uint8, device=dev)
    dists = torch.empty(R, n_coarse, device=dev)
    r = rands.contiguous().float() if rands is not None else None
    _lib.call("na_ray_bounds_sample", _lib.ptr(c), _lib.ptr(v), _lib.ptr(r), R, int(n_coarse),
              _lib.ptr(near), _lib.ptr(far), _lib.ptr(outside), _lib.ptr(dists), _lib.stream_of(c))
    return near, far, outside.bool(), dists


def hierarchical_round(dists, sdfs, inv_s, n_fine):
    """One round of model.py:239-247 on the GPU: returns merged [R,N+Nf], fine [R,Nf],
    pos_coarse [R,N], pos_fine [R,Nf] (int32 positions in the merged row)."""
    R, N = dists.shape
    dev = dists.device
    merged = torch.empty(R, N + n_fine, device=dev)
    fine = torch.empty(R, n_fine, device=dev)
    pc = torch.empty(R, N, dtype=torch.int32, device=dev)
    pf = torch.empty(R, n_fine, dtype=torch.int32, device=dev)
    _lib.call("na_hier_resample", _lib.ptr(dists.contiguous()), _lib.ptr(sdfs.contiguous()), R, N,
              float(inv_s), int(n_fine), _lib.ptr(merged), _lib.ptr(fine), _lib.ptr(pc),
              _lib.ptr(pf), _lib.stream_of(dists))
    return merged, fine, pc, pf


def scatter_merge(a, pos_a, b, pos_b):
    R, Na = a.shape
    Nb = b.shape[1]
    out = torch.empty(R, Na + Nb, device=a.device)
    _lib.call("na_scatter_merge", _lib.ptr(a.contiguous()), _lib.ptr(pos_a), Na,
              _lib.ptr(b.contiguous()), _lib.ptr(pos_b), Nb, R, _lib.ptr(out), _lib.stream_of(a))
    return out
