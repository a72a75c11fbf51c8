"""Marching cubes on the GPU: a drop-in for ``mcubes.marching_cubes(volume, isovalue)``
and the lattice form used by mesh extraction.

Reference call site: projects/neuralangelo/utils/mesh.py:121.  Vertices are
float64 [nv, 3] (array-index coordinates, or world coordinates in the lattice
form), faces int64 [nf, 3], in the order of the sequential algorithm (cells
i-major, per-cell owned edges 6, 5, 10 then boundary edges, triangles in case
table order) — the order the C ABI ``na_mc_count`` / ``na_mc_emit`` define.
"""
import numpy as np
import torch

from . import _lib


def lattice_marching_cubes(volume, block_res=0, isovalue=0.0, axes=None, intv=1.0,
                           workspace=None, block_ranges=False):
    """Concatenated per-block meshes of a lattice volume (CUDA float32 [X,Y,Z]).

    block_res <= 0 treats the volume as one block.  With ``axes`` (three CUDA
    float32 lattice coordinate vectors) vertices are ``local*intv + axes[block origin]``
    (projects/neuralangelo/utils/mesh.py:123), else block-local index coordinates.
    Returns (V [nv,3] float64, F [nf,3] int64) CUDA tensors; F indexes the
    concatenated vertex list.  block_ranges: also return the per-block offsets
    (block_v, block_f: int64 [nblocks + 1] CUDA tensors, x-major block order).
    """
    _lib.require_cuda(volume)
    if volume.dim() != 3:
        raise ValueError(f"expected a 3-D volume, got shape {tuple(volume.shape)}")
    vol = volume.contiguous().float()
    nx, ny, nz = vol.shape
    dev = vol.device
    lib = _lib.load()
    need = lib.na_mc_workspace_size(nx, ny, nz, int(block_res))
    _lib.check(need, "na_mc_workspace_size")
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(int(need), dtype=torch.uint8, device=dev)
    totals = torch.empty(2, dtype=torch.int64, device=dev)
    s = _lib.stream_of(vol)
    _lib.call("na_mc_count", _lib.ptr(vol), nx, ny, nz, int(block_res), float(isovalue),
              _lib.ptr(workspace), int(workspace.numel()), _lib.ptr(totals), s)
    nv, nf = (int(x) for x in totals.tolist())
    V = torch.empty(nv, 3, dtype=torch.float64, device=dev)
    F = torch.empty(nf, 3, dtype=torch.int64, device=dev)
    if nv or nf:
        if axes is not None:
            ax = [a.to(dev).float().contiguous() for a in axes]
            if [a.numel() for a in ax] != [nx, ny, nz]:
                raise ValueError("axis lengths must match the volume shape")
        else:
            ax = [None, None, None]
        _lib.call("na_mc_emit", _lib.ptr(vol), nx, ny, nz, int(block_res), float(isovalue),
                  _lib.ptr(workspace), int(workspace.numel()), _lib.ptr(ax[0]), _lib.ptr(ax[1]),
                  _lib.ptr(ax[2]), float(intv), _lib.ptr(V), _lib.ptr(F), s)
    if not block_ranges:
        return V, F
    nb = lib.na_mc_num_blocks(nx, ny, nz, int(block_res))
    _lib.check(nb, "na_mc_num_blocks")
    bv = torch.zeros(int(nb) + 1, dtype=torch.int64, device=dev)
    bf = torch.zeros(int(nb) + 1, dtype=torch.int64, device=dev)
    if nb:
        _lib.call("na_mc_block_offsets", nx, ny, nz, int(block_res), _lib.ptr(workspace),
                  int(workspace.numel()), _lib.ptr(bv), _lib.ptr(bf), s)
    return V, F, bv, bf


def marching_cubes_gpu(volume, isovalue=0.0, workspace=None):
    """volume: CUDA float32 tensor [X,Y,Z] -> (V [nv,3] f64, F [nf,3] i64) CUDA tensors."""
    return lattice_marching_cubes(volume, 0, isovalue, workspace=workspace)


def marching_cubes(volume, isovalue):
    """mcubes.marching_cubes semantics: numpy in -> numpy out; tensors stay tensors."""
    if torch.is_tensor(volume):
        if volume.is_cuda:
            return marching_cubes_gpu(volume, isovalue)
        V, F = marching_cubes_gpu(volume.cuda(), isovalue)
        return V.cpu(), F.cpu()
    vol = torch.from_numpy(np.ascontiguousarray(volume, dtype=np.float32)).cuda()
    V, F = marching_cubes_gpu(vol, isovalue)
    return V.cpu().numpy(), F.cpu().numpy()
