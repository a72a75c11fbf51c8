"""Oracle: marching cubes + block-wise mesh extraction (CPU).  TEST INFRASTRUCTURE.

marching_cubes  -> the C restatement in oracle/mc_oracle.c (PyMCubes algorithm,
                   projects/neuralangelo/utils/mesh.py:121)
extract_mesh    -> projects/neuralangelo/utils/mesh.py:24-49 (LatticeGrid blocks
                   of block_res+1 points, :68-100), marching_cubes per block
                   (:119-133) with trimesh's construction-time vertex merge
                   (trimesh_process), sphere filter (:136-148), concatenation
                   without vertex welding (:44-46)
largest_component -> filter_largest_cc (:151-158): trimesh's edge-connected split and
                   the component of largest area (trimesh is absent; its split is
                   restated from its documented behaviour)

Pinned by (tests/test_oracle_mesh.py): every one of the 256 single-cube cases
against the case table's crossed-edge sets and triangle counts; watertightness
and orientability on random volumes (each interior edge shared by exactly two
faces with opposite orientation); a sphere volume gives Euler characteristic
2 and vertices within the interpolation error of the sphere.  Exact PyMCubes
vertex order is unpinned (PyMCubes is absent and unpinned).
"""
import ctypes
import os

import numpy as np
import torch

from . import build as _build

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.LIB):
            _build.build()
        lib = ctypes.CDLL(_build.LIB)
        lib.mc_oracle.restype = ctypes.c_int
        lib.mc_oracle.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                  ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        _lib = lib
    return _lib


def marching_cubes(volume, isovalue=0.0):
    """volume [X,Y,Z] -> (vertices float64 [nv,3] index coords, faces int64 [nf,3])."""
    vol = np.ascontiguousarray(np.asarray(volume, dtype=np.float32))
    nx, ny, nz = vol.shape
    lib = _load()
    nv, nf = ctypes.c_int64(), ctypes.c_int64()
    st = lib.mc_oracle(vol.ctypes.data, nx, ny, nz, float(isovalue), None, None,
                       ctypes.byref(nv), ctypes.byref(nf))
    if st:
        raise RuntimeError("mc_oracle failed")
    V = np.zeros((nv.value, 3), np.float64)
    F = np.zeros((nf.value, 3), np.int64)
    st = lib.mc_oracle(vol.ctypes.data, nx, ny, nz, float(isovalue), V.ctypes.data, F.ctypes.data,
                       ctypes.byref(nv), ctypes.byref(nf))
    if st:
        raise RuntimeError("mc_oracle failed")
    return V, F


def lattice_marching_cubes(vol, block_res, isovalue=0.0, axes=None, intv=1.0, counts=False,
                           threads=1):
    """The reference's per-block loop (mesh.py:24-49, :82-96 slicing, :121-123) on a
    stored lattice volume, without the sphere filter: blocks of block_res+1
    points, x-major; each block's V (index coords, or V*intv + axes[block
    origin] when axes are given) and F + running vertex offset, concatenated.
    counts: also return each block's vertex and face count (empty blocks included).
    threads: blocks extracted concurrently (the C routine keeps no shared state);
    the concatenation keeps the sequential block order."""
    vol = np.asarray(vol, np.float32)
    n = vol.shape
    nb = [int(np.ceil(d / block_res)) for d in n]
    origins = [(bx * block_res, by * block_res, bz * block_res)
               for bx in range(nb[0]) for by in range(nb[1]) for bz in range(nb[2])]

    def one(o):
        xi, yi, zi = o
        blk = vol[xi:xi + block_res + 1, yi:yi + block_res + 1, zi:zi + block_res + 1]
        return marching_cubes(blk, isovalue)

    if threads > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max_workers=threads) as ex:
            results = list(ex.map(one, origins))
    else:
        results = [one(o) for o in origins]
    Vs, Fs, off = [], [], 0
    cv, cf = [], []
    for (xi, yi, zi), (V, F) in zip(origins, results):
        cv.append(V.shape[0])
        cf.append(F.shape[0])
        if V.shape[0] == 0:
            continue
        if axes is not None:
            V = V * intv + np.array([axes[0][xi], axes[1][yi], axes[2][zi]], np.float32)
        Vs.append(V)
        Fs.append(F + off)
        off += V.shape[0]
    if not Vs:
        out = np.zeros((0, 3)), np.zeros((0, 3), np.int64)
    else:
        out = np.concatenate(Vs), np.concatenate(Fs)
    return out + (np.array(cv, np.int64), np.array(cf, np.int64)) if counts else out


def largest_component(V, F, C=None):
    """filter_largest_cc (mesh.py:151-158) restated: trimesh's Trimesh.split joins faces
    that share an edge; the component of largest total area is kept (the first on
    ties), with its faces in their original order and its vertices (those it
    references) in their original order.  Union-find over faces, sequential."""
    nf = len(F)
    if nf == 0 or len(V) == 0:
        return (V[:0], F[:0]) + ((C[:0],) if C is not None else ())
    parent = list(range(nf))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    owner = {}
    for f in range(nf):
        a, b, c = (int(x) for x in F[f])
        for e in ((a, b), (b, c), (c, a)):
            e = (min(e), max(e))
            if e in owner:
                ra, rb = find(owner[e]), find(f)
                if ra != rb:
                    parent[max(ra, rb)] = min(ra, rb)
            else:
                owner[e] = f
    roots = [find(f) for f in range(nf)]
    tri = V[F]
    areas = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=-1)
    comp_area, order = {}, []
    for f in range(nf):
        r = roots[f]
        if r not in comp_area:
            comp_area[r] = 0.0
            order.append(r)
        comp_area[r] += areas[f]
    best = order[0]
    for r in order[1:]:
        if comp_area[r] > comp_area[best]:
            best = r
    keep = np.array([r == best for r in roots])
    Fk = F[keep]
    used = np.zeros(len(V), bool)
    used[Fk.reshape(-1)] = True
    remap = np.cumsum(used) - 1
    return (V[used], remap[Fk]) + ((C[used],) if C is not None else ())


def trimesh_process(V, F, C=None, digits=8):
    """trimesh.Trimesh(V, F) construction processing (process=True, validate=False),
    restated from trimesh's documented merge_vertices: vertices whose coordinates agree
    after rounding to `digits` decimals (tol.merge = 1e-8) are merged, and only
    vertices referenced by a face take part -- the rest are dropped (a mesh without
    faces keeps all its vertices).  Kept vertices in order of first occurrence (the
    order trimesh's order-preserving merge gives; trimesh is absent and unpinned,
    so that order is unpinned), faces renumbered, neither dropped nor reordered.
    Sequential: one dictionary lookup per vertex."""
    if len(F) == 0:
        return (V, F) + ((C,) if C is not None else ())
    referenced = np.zeros(len(V), bool)
    referenced[F.reshape(-1)] = True
    scale = 10.0 ** digits
    group, keep = {}, []
    new_id = np.full(len(V), -1, np.int64)
    for i in range(len(V)):
        if not referenced[i]:
            continue
        k = tuple(int(np.round(c * scale)) for c in V[i])
        if k not in group:
            group[k] = len(keep)
            keep.append(i)
        new_id[i] = group[k]
    keep = np.array(keep, np.int64)
    return (V[keep], new_id[F]) + ((C[keep],) if C is not None else ())


def lattice_blocks(bounds, intv, block_res):
    """LatticeGrid (mesh.py:68-100): yields (block index, xyz [bx,by,bz,3] float32)."""
    (x0, x1), (y0, y1), (z0, z1) = bounds
    xg = torch.arange(x0, x1, intv)
    yg = torch.arange(y0, y1, intv)
    zg = torch.arange(z0, z1, intv)
    nbx = int(np.ceil(len(xg) / block_res))
    nby = int(np.ceil(len(yg) / block_res))
    nbz = int(np.ceil(len(zg) / block_res))
    for idx in range(nbx * nby * nbz):
        bx = idx // (nby * nbz)
        by = (idx // nbz) % nby
        bz = idx % nbz
        xi, yi, zi = bx * block_res, by * block_res, bz * block_res
        x, y, z = torch.meshgrid(xg[xi:xi + block_res + 1], yg[yi:yi + block_res + 1],
                                 zg[zi:zi + block_res + 1], indexing="ij")
        yield idx, torch.stack([x, y, z], dim=-1)


def filter_points_outside_bounding_sphere(V, F, C=None):
    """mesh.py:136-148 on plain arrays (colours C follow their vertices)."""
    mask = np.linalg.norm(V, axis=-1) < 1.0
    if not np.any(mask):
        out = np.zeros((0, 3)), np.zeros((0, 3), np.int64)
        return out + ((np.zeros((0, 3), np.uint8),) if C is not None else ())
    indices = np.ones(len(V), dtype=np.int64) * -1
    indices[mask] = np.arange(mask.sum())
    fmask = mask[F[:, 0]] & mask[F[:, 1]] & mask[F[:, 2]]
    out = V[mask], indices[F[fmask]]
    return out + ((C[mask],) if C is not None else ())


def extract_mesh(sdf_func, bounds, intv, block_res=64, texture_func=None, filter_lcc=False,
                 process=True):
    """mesh.py:24-49 with marching_cubes (:119-133): per block, MC on sdf_func(xyz)
    (already negated by the caller, extract_mesh.py:85), V*intv + xyz[0,0,0], colours
    texture_func(V), the Trimesh construction (trimesh_process when process), sphere
    filter (a new Trimesh: processed again), the block's largest component when
    filter_lcc; the blocks concatenated without merging.  Returns (V, F) or, with
    texture_func, (V, F, C)."""
    Vs, Fs, Cs, off = [], [], [], 0
    for _, xyz in lattice_blocks(bounds, intv, block_res):
        sdf = sdf_func(xyz)[..., 0]
        V, F = marching_cubes(sdf.numpy(), 0.0)
        if V.shape[0] == 0:
            continue
        V = V * intv + xyz[0, 0, 0].numpy()
        C = np.asarray(texture_func(V), np.uint8) if texture_func is not None else None
        if process:
            V, F, *rest = trimesh_process(V, F, C)
            C = rest[0] if rest else None
        res = filter_points_outside_bounding_sphere(V, F, C)
        if process:
            res = trimesh_process(*res)
        if filter_lcc and res[0].shape[0] > 0:
            res = largest_component(*res)
        V, F = res[0], res[1]
        if V.shape[0] == 0:
            continue
        Vs.append(V)
        Fs.append(F + off)
        if texture_func is not None:
            Cs.append(res[2])
        off += V.shape[0]
    tail = (np.concatenate(Cs) if Cs else np.zeros((0, 3), np.uint8),) if texture_func else ()
    if not Vs:
        return (np.zeros((0, 3)), np.zeros((0, 3), np.int64)) + tail
    return (np.concatenate(Vs), np.concatenate(Fs)) + tail
