"""Mesh extraction on the GPU (projects/neuralangelo/utils/mesh.py:24-158,
projects/neuralangelo/scripts/extract_mesh.py:80-101).

The reference walks LatticeGrid blocks of block_res+1 points (overlapping by
one point), evaluates the SDF of each block on the GPU, copies it to the host,
runs marching cubes on the CPU, maps vertices to world space
(V*intv + xyz[0,0,0]), colours them (optional), drops faces with a vertex
outside the unit sphere, keeps each block's largest connected component
(optional) and concatenates the blocks without vertex welding.

Here the SDF is evaluated once over the whole lattice (in slabs of bounded
size; every lattice point once, where the reference evaluates block seams
twice), stays in HBM, and one lattice marching-cubes launch chain produces the
concatenation of all per-block meshes, in block order, with the world
transform fused, plus each block's vertex / face range.  Colours come from the
network at the vertices (extract_texture).  The sphere filter runs on the
device over the concatenated mesh: it acts per vertex and per face, so it
equals filtering each block.  The largest-component filter runs per block on
the host (face components joined through shared edges, as trimesh's
Trimesh.split; blocks share no vertices, so components never cross blocks).
Returns numpy arrays (vertices float64 [nv,3], faces int64 [nf,3][, colours
uint8 [nv,3]]).

Not reproduced: trimesh's own processing on construction (merging vertices
that coincide to 1e-8, which only happens when the SDF is exactly 0 at a
lattice point).
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
def extract_texture(xyz, neural_rgb, neural_sdf, appear_embed=None, chunk=1 << 20):
    """Per-vertex colours (mesh.py:53-65): the colour network at the vertex, with the
    unit SDF gradient as normal and its negation as the viewing direction; rgb * 255
    truncated to uint8.  xyz: [N,3] numpy or tensor -> uint8 numpy [N,3]."""
    if appear_embed is not None:
        raise NotImplementedError("appearance embeddings are outside the accelerated path")
    dev = neural_sdf.tcnn_encoding.params.device
    x = torch.as_tensor(np.asarray(xyz) if not torch.is_tensor(xyz) else xyz).to(
        dev, torch.float32).reshape(-1, 3)
    out = torch.empty(x.shape[0], 3, dtype=torch.uint8, device=dev)
    for i in range(0, x.shape[0], chunk):
        p = x[i:i + chunk].contiguous()
        sdfs, feats = neural_sdf.forward(p)
        gradients, _ = neural_sdf.compute_gradients(p, training=False, sdf=sdfs)
        normals = torch.nn.functional.normalize(gradients, dim=-1)
        rgbs = neural_rgb.forward(p, normals, -normals, feats, None)
        out[i:i + chunk] = (rgbs.float() * 255).to(torch.uint8)
    return out.cpu().numpy()


def _face_areas(V, F):
    e1 = V[F[:, 1]] - V[F[:, 0]]
    e2 = V[F[:, 2]] - V[F[:, 0]]
    return 0.5 * np.linalg.norm(np.cross(e1, e2), axis=-1)


def largest_component_per_block(V, F, block_f, C=None):
    """Keep, in every block's face range [block_f[b], block_f[b+1]), the connected
    component of largest area (mesh.py:151-158 applied per block as at mesh.py:130):
    faces are connected through shared edges (trimesh's face adjacency); the kept
    faces keep their order, vertices are renumbered in their original order (trimesh
    submesh).  Blocks whose range is empty stay empty."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    nf = F.shape[0]
    if nf == 0:
        return (V, F) if C is None else (V, F, C)
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    e.sort(axis=1)
    fid = np.tile(np.arange(nf), 3)
    key = e[:, 0].astype(np.int64) * (V.shape[0] + 1) + e[:, 1]
    order = np.argsort(key, kind="stable")
    ks, fs = key[order], fid[order]
    same = ks[1:] == ks[:-1]
    a, b = fs[:-1][same], fs[1:][same]
    g = sp.coo_matrix((np.ones(a.shape[0], np.int8), (a, b)), shape=(nf, nf))
    ncomp, labels = connected_components(g, directed=False)
    area = np.bincount(labels, weights=_face_areas(V, F), minlength=ncomp)
    block_of_face = np.searchsorted(block_f, np.arange(nf), side="right") - 1
    first_face = np.full(ncomp, nf, np.int64)
    np.minimum.at(first_face, labels, np.arange(nf))
    comp_block = block_of_face[first_face]
    # per block: the component of largest area, the first one on ties
    best = {}
    for c in np.argsort(first_face, kind="stable"):
        blk = comp_block[c]
        if blk not in best or area[c] > area[best[blk]]:
            best[blk] = c
    keep_comp = np.zeros(ncomp, bool)
    keep_comp[list(best.values())] = True
    Fk = F[keep_comp[labels]]
    used = np.zeros(V.shape[0], bool)
    used[Fk.reshape(-1)] = True
    remap = np.cumsum(used) - 1
    out = (V[used], remap[Fk])
    return out if C is None else out + (C[used],)


def nondegenerate_faces(V, F, height=1e-8):
    """Mask of faces whose smallest height (2 area / longest edge) exceeds `height`
    (trimesh's nondegenerate_faces with its merge tolerance, as used at
    extract_mesh.py:99)."""
    if F.shape[0] == 0:
        return np.zeros(0, bool)
    t = V[F]
    longest = np.max(np.linalg.norm(t - np.roll(t, 1, axis=1), axis=-1), axis=-1)
    h = 2.0 * _face_areas(V, F) / np.maximum(longest, 1e-300)
    return h > height


def write_ply(path, V, F, C=None):
    """Binary little-endian PLY: vertices as doubles, optional RGBA uint8 colours
    (alpha 255), faces as uint8-counted int32 index lists."""
    V = np.ascontiguousarray(V, dtype="<f8")
    F = np.ascontiguousarray(F, dtype="<i4")
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {V.shape[0]}",
              "property double x", "property double y", "property double z"]
    vdt = [("x", "<f8"), ("y", "<f8"), ("z", "<f8")]
    if C is not None:
        header += ["property uchar red", "property uchar green", "property uchar blue",
                   "property uchar alpha"]
        vdt += [("red", "u1"), ("green", "u1"), ("blue", "u1"), ("alpha", "u1")]
    header += [f"element face {F.shape[0]}", "property list uchar int vertex_indices",
               "end_header"]
    verts = np.empty(V.shape[0], dtype=vdt)
    verts["x"], verts["y"], verts["z"] = V[:, 0], V[:, 1], V[:, 2]
    if C is not None:
        C = np.asarray(C, dtype=np.uint8)
        verts["red"], verts["green"], verts["blue"] = C[:, 0], C[:, 1], C[:, 2]
        verts["alpha"] = C[:, 3] if C.shape[1] > 3 else 255
    faces = np.empty(F.shape[0], dtype=[("n", "u1"), ("v", "<i4", (3,))])
    faces["n"] = 3
    faces["v"] = F
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode("ascii"))
        f.write(verts.tobytes())
        f.write(faces.tobytes())


MERGE_DIGITS = 8  # trimesh's tol.merge = 1e-8, as decimal digits


@torch.no_grad()
def weld_blocks(V, F, block_v, digits=MERGE_DIGITS):
    """The processing trimesh.Trimesh(V, F) applies at construction (process=True,
    mesh.py:124-127): merge_vertices -- vertices whose coordinates agree after rounding
    to `digits` decimals become one -- and removal of vertices no face references;
    done for every block on its own (the blocks are joined afterwards without merging,
    trimesh.util.concatenate at mesh.py:46).  V [nv,3] and F [nf,3] on the GPU, block b
    owning vertices block_v[b] .. block_v[b+1]-1.  Kept vertices stay in the order of
    their first occurrence and faces are renumbered (faces are neither dropped nor
    reordered).  Returns (V, F, keep, block_v) with keep the input index of every kept
    vertex (to carry per-vertex data) and the blocks' new vertex offsets.
    In MC output a vertex repeats only where the field is exactly zero at a lattice
    point (the edges meeting there all place their vertex on it)."""
    nv = V.shape[0]
    dev = V.device
    block_v = torch.as_tensor(block_v, dtype=torch.int64, device=dev)
    if nv == 0 or F.shape[0] == 0:
        empty = torch.zeros(0, dtype=torch.int64, device=dev)
        return V[:0], F[:0], empty, torch.zeros_like(block_v)
    blk = torch.repeat_interleave(torch.arange(block_v.numel() - 1, device=dev),
                                  block_v[1:] - block_v[:-1])
    q = torch.round(V.double() * 10.0 ** digits).to(torch.int64)
    ref = torch.zeros(nv, dtype=torch.bool, device=dev)
    ref[F.reshape(-1)] = True
    idx = torch.nonzero(ref)[:, 0]
    key = torch.cat([blk[idx, None], q[idx]], dim=1)
    _, inv = torch.unique(key, dim=0, return_inverse=True)
    ngroups = int(inv.max().item()) + 1
    first = torch.full((ngroups,), nv, dtype=torch.int64, device=dev)
    first.scatter_reduce_(0, inv, idx, reduce="amin")
    order = torch.argsort(first)
    rank = torch.empty_like(order)
    rank[order] = torch.arange(ngroups, device=dev)
    new_id = torch.full((nv,), -1, dtype=torch.int64, device=dev)
    new_id[idx] = rank[inv]
    keep = first[order]
    counts = torch.bincount(blk[keep], minlength=block_v.numel() - 1)
    new_block_v = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev), torch.cumsum(counts, 0)])
    return V[keep], new_id[F], keep, new_block_v


@torch.no_grad()
def extract_mesh(sdf_func, bounds, intv, block_res=64, texture_func=None, filter_lcc=False,
                 device=None, max_points=1 << 24, process=True):
    """sdf_func(xyz [...,3] on the GPU) -> [...,1]; the caller passes the negated SDF
    (projects/neuralangelo/scripts/extract_mesh.py:85).  texture_func(V [N,3]) -> uint8
    colours [N,3] (e.g. functools.partial(extract_texture, neural_rgb=..., neural_sdf=...)).
    process: trimesh's construction-time processing of each block mesh, as the
    reference builds every block with trimesh.Trimesh (weld_blocks, then the sphere
    filter, whose new Trimesh drops the vertices left without a face); process=False
    returns the marching-cubes output with only the sphere filter's own masking.
    Returns (V, F) or, with texture_func, (V, F, C)."""
    device = device or torch.device("cuda", torch.cuda.current_device())
    axes = lattice_axes(bounds, intv)
    vol = lattice_sdf(sdf_func, axes, device, max_points)
    V, F, block_v, block_f = lattice_marching_cubes(vol, block_res, 0.0, axes=axes, intv=intv,
                                                    block_ranges=True)
    del vol
    C = None
    if texture_func is not None and V.shape[0] > 0:
        C = np.asarray(texture_func(V.float()), dtype=np.uint8)
    if process and V.shape[0] > 0:
        V, F, keep, block_v = weld_blocks(V, F, block_v)
        if C is not None:
            C = C[keep.cpu().numpy()]
    empty = (np.zeros((0, 3)), np.zeros((0, 3), np.int64)) + \
        ((np.zeros((0, 3), np.uint8),) if texture_func is not None else ())
    if V.shape[0] == 0:
        return empty
    mask = V.norm(dim=-1) < 1.0
    if not bool(mask.any()):
        return empty
    fmask = mask[F[:, 0]] & mask[F[:, 1]] & mask[F[:, 2]]
    kept_before = torch.cumsum(fmask.to(torch.int64), 0)
    block_f = torch.cat([torch.zeros(1, dtype=torch.int64, device=F.device), kept_before])[block_f]
    nb = block_f.numel() - 1
    vblock = torch.repeat_interleave(torch.arange(nb, device=V.device),
                                     block_v[1:] - block_v[:-1])[mask]
    V, F = filter_outside_sphere(V, F)
    if C is not None:
        C = C[mask.cpu().numpy()]
    if process:
        # the filtered block is a new Trimesh: a block with faces keeps only the
        # vertices they reference; a block left without faces keeps its vertices
        used = (block_f[1:] - block_f[:-1])[vblock] == 0
        used[F.reshape(-1)] = True
        if not bool(used.all()):
            V, F = V[used], (torch.cumsum(used.to(torch.int64), 0) - 1)[F]
            if C is not None:
                C = C[used.cpu().numpy()]
    Vn, Fn = V.cpu().numpy(), F.cpu().numpy()
    if filter_lcc:
        res = largest_component_per_block(Vn, Fn, block_f.cpu().numpy(), C)
        Vn, Fn = res[0], res[1]
        C = res[2] if C is not None else None
    return (Vn, Fn) if texture_func is None else (Vn, Fn, C)
