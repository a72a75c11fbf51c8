"""Mesh extraction on the GPU (projects/neuralangelo/utils/mesh.py:24-49,68-148).

The reference walks LatticeGrid blocks of block_res+1 points (overlapping by
one point), evaluates the SDF of each block on the GPU, copies it to the host,
runs marching cubes on the CPU, maps vertices to world space
(V*intv + xyz[0,0,0]), drops faces with a vertex outside the unit sphere and
concatenates the blocks without vertex welding.

Here the SDF is evaluated once over the whole lattice (in slabs of bounded
size; every lattice point once, where the reference evaluates block seams
twice), stays in HBM, and one lattice marching-cubes launch chain produces the
concatenation of all per-block meshes, in block order, with the world
transform fused.  The sphere filter runs on the device over the concatenated
mesh; since it acts per vertex and per face and the order is the block order,
the result equals filtering each block and concatenating.  Returns numpy
(vertices float64 [nv,3], faces int64 [nf,3]).

Not reproduced: trimesh's own processing (vertex merging at 1e-8) and the
largest-connected-component / texture options (trimesh is not available).
"""
import numpy as np
import torch

from .mcubes import lattice_marching_cubes


def lattice_axes(bounds, intv):
    """LatticeGrid coordinates (mesh.py:73-75): float32 arange per axis."""
    return [torch.arange(lo, hi, intv) for lo, hi in bounds]


def iter_blocks(bounds, intv, block_res):
    """Reference block order (mesh.py:82-96): the axes of each block, x-major."""
    xg, yg, zg = lattice_axes(bounds, intv)
    nbx = int(np.ceil(len(xg) / block_res))
    nby = int(np.ceil(len(yg) / block_res))
    nbz = int(np.ceil(len(zg) / block_res))
    for idx in range(nbx * nby * nbz):
        bx = idx // (nby * nbz)
        by = (idx // nbz) % nby
        bz = idx % nbz
        xi, yi, zi = bx * block_res, by * block_res, bz * block_res
        yield (xg[xi:xi + block_res + 1], yg[yi:yi + block_res + 1], zg[zi:zi + block_res + 1])


def filter_outside_sphere(V, F):
    """Keep vertices with |v| < 1 and faces whose three vertices are kept (mesh.py:136-148)."""
    mask = V.norm(dim=-1) < 1.0
    if not bool(mask.any()):
        return V[:0], F[:0]
    idx = torch.cumsum(mask.to(torch.int64), 0) - 1
    fmask = mask[F[:, 0]] & mask[F[:, 1]] & mask[F[:, 2]]
    return V[mask], idx[F[fmask]]


@torch.no_grad()
def lattice_sdf(sdf_func, axes, device, max_points=1 << 24):
    """SDF over the full lattice [X,Y,Z] (float32, on the device), evaluated in x-slabs."""
    xg, yg, zg = (a.to(device) for a in axes)
    nx, ny, nz = len(xg), len(yg), len(zg)
    vol = torch.empty(nx, ny, nz, dtype=torch.float32, device=device)
    step = max(1, max_points // max(ny * nz, 1))
    for x0 in range(0, nx, step):
        x, y, z = torch.meshgrid(xg[x0:x0 + step], yg, zg, indexing="ij")
        xyz = torch.stack([x, y, z], dim=-1)
        vol[x0:x0 + step] = sdf_func(xyz)[..., 0]
    return vol


@torch.no_grad()
def extract_mesh(sdf_func, bounds, intv, block_res=64, device=None, max_points=1 << 24):
    """sdf_func(xyz [...,3] on the GPU) -> [...,1]; the caller passes the negated SDF
    (projects/neuralangelo/scripts/extract_mesh.py:85)."""
    device = device or torch.device("cuda", torch.cuda.current_device())
    axes = lattice_axes(bounds, intv)
    vol = lattice_sdf(sdf_func, axes, device, max_points)
    V, F = lattice_marching_cubes(vol, block_res, 0.0, axes=axes, intv=intv)
    del vol
    if V.shape[0] == 0:
        return np.zeros((0, 3)), np.zeros((0, 3), np.int64)
    V, F = filter_outside_sphere(V, F)
    return V.cpu().numpy(), F.cpu().numpy()
