"""Pins the marching-cubes case table and the C restatement (CPU only)."""
import os
import re
from collections import Counter

import numpy as np
import pytest

from oracle import mesh as omesh

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]


def _tables():
    text = open(os.path.join(REPO, "neuralangelo_amd", "csrc", "mc_tables.h")).read()
    edge_blk = text[text.index("NA_MC_EDGE_TABLE_INIT"):text.index("NA_MC_TRI_TABLE_INIT")]
    edge = [int(v, 16) for v in re.findall(r"0x[0-9a-f]+", edge_blk)]
    tri_blk = text[text.index("NA_MC_TRI_TABLE_INIT"):text.index("static const uint16_t")]
    rows = re.findall(r"\{([-0-9, ]+)\}", tri_blk)
    tri = [[int(v) for v in r.split(",")] for r in rows]
    return edge, tri


def _crossed(c):
    return {e for e, (a, b) in enumerate(EDGES) if ((c >> a) & 1) != ((c >> b) & 1)}


def test_edge_table_is_the_crossed_edge_set():
    edge, tri = _tables()
    assert len(edge) == 256 and len(tri) == 256
    for c in range(256):
        assert edge[c] == sum(1 << e for e in _crossed(c))


def test_triangle_table_local_topology():
    _, tri = _tables()
    n_tri = 0
    for c in range(256):
        row = tri[c]
        t = row[:row.index(-1)] if -1 in row else row
        assert all(v == -1 for v in row[len(t):])
        assert len(t) % 3 == 0 and len(t) <= 15
        assert set(t) == _crossed(c), c
        tris = [t[i:i + 3] for i in range(0, len(t), 3)]
        n_tri += len(tris)
        directed = Counter()
        for a, b, d in tris:
            for u, v in ((a, b), (b, d), (d, a)):
                directed[(u, v)] += 1
        for (u, v), n in directed.items():
            assert n == 1, c                       # no doubled triangle edge
            if (v, u) not in directed:            # boundary segment: on a cube face
                s = set(EDGES[u]) | set(EDGES[v])
                assert any(s <= set(f) for f in FACES), c
    assert n_tri == 820


def test_triangle_orientation_is_consistent():
    """Every triangle's normal points against the trilinear field's gradient."""
    _, tri = _tables()
    P = np.array(CORNERS, float)
    for c in range(256):
        vals = np.array([-1.0 if (c >> m) & 1 else 1.0 for m in range(8)])
        row = tri[c]
        t = row[:row.index(-1)] if -1 in row else row

        def pt(e):
            a, b = EDGES[e]
            s = (0 - vals[a]) / (vals[b] - vals[a])
            return P[a] + s * (P[b] - P[a])

        def f(x):
            w = [(1 - x[0]) * (1 - x[1]) * (1 - x[2]), x[0] * (1 - x[1]) * (1 - x[2]),
                 x[0] * x[1] * (1 - x[2]), (1 - x[0]) * x[1] * (1 - x[2]),
                 (1 - x[0]) * (1 - x[1]) * x[2], x[0] * (1 - x[1]) * x[2], x[0] * x[1] * x[2],
                 (1 - x[0]) * x[1] * x[2]]
            return float(np.dot(w, vals))
        for i in range(0, len(t), 3):
            p0, p1, p2 = pt(t[i]), pt(t[i + 1]), pt(t[i + 2])
            n = np.cross(p1 - p0, p2 - p0)
            cen = (p0 + p1 + p2) / 3
            g = np.array([f(cen + h) - f(cen - h) for h in np.eye(3) * 1e-5])
            assert np.dot(n, g) < 0, c


@pytest.mark.parametrize("case", range(256))
def test_single_cube_cases(case):
    vol = np.zeros((2, 2, 2), np.float32)
    for m, (x, y, z) in enumerate(CORNERS):
        vol[x, y, z] = -1.0 if (case >> m) & 1 else 1.0
    V, F = omesh.marching_cubes(vol, 0.0)
    assert V.shape[0] == len(_crossed(case))
    # all vertices are edge midpoints (values -1/+1)
    assert np.all(np.isin(np.round(V * 2), [0, 1, 2]))
    _, tri = _tables()
    assert F.shape[0] == sum(1 for v in tri[case][::3] if v >= 0)


def _edge_use(F):
    use = Counter()
    for a, b, c in F:
        for u, v in ((a, b), (b, c), (c, a)):
            use[(int(u), int(v))] += 1
    return use


def test_sphere_is_closed_with_euler_characteristic_two():
    n = 40
    g = np.linspace(-1, 1, n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    vol = (np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 0.6).astype(np.float32)
    V, F = omesh.marching_cubes(vol, 0.0)
    use = _edge_use(F)
    for (u, v), k in use.items():
        assert k == 1 and use.get((v, u)) == 1  # closed, consistently oriented
    n_edges = len(use) // 2
    assert V.shape[0] - n_edges + F.shape[0] == 2
    h = 2.0 / (n - 1)
    r = np.linalg.norm(V * h - 1.0, axis=-1)
    assert np.abs(r - 0.6).max() < 0.25 * h


def test_random_volume_is_watertight_in_the_interior():
    rng = np.random.default_rng(0)
    vol = rng.normal(size=(17, 19, 23)).astype(np.float32)
    vol[0], vol[-1], vol[:, 0], vol[:, -1], vol[:, :, 0], vol[:, :, -1] = 1, 1, 1, 1, 1, 1
    V, F = omesh.marching_cubes(vol, 0.0)
    use = _edge_use(F)
    open_edges = [(u, v) for (u, v) in use if (v, u) not in use]
    assert not open_edges
    assert F.max() < V.shape[0] and len(np.unique(F)) == V.shape[0]


def test_empty_and_degenerate_volumes():
    V, F = omesh.marching_cubes(np.ones((5, 5, 5), np.float32), 0.0)
    assert V.shape == (0, 3) and F.shape == (0, 3)
    V, F = omesh.marching_cubes(np.ones((1, 5, 5), np.float32), 0.0)
    assert V.shape == (0, 3)


def test_lattice_loop_equals_blockwise_extraction():
    """Stored-lattice block loop + global sphere filter == the per-block extract_mesh
    (without trimesh's processing, which the stored-lattice loop does not apply)."""
    import torch

    def sdf(xyz):
        x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
        return -(x * x + y * y + z * z - 0.5)[..., None]
    bounds = ((-1.0, 1.0), (-1.0, 1.0), (-1.0, 1.0))
    intv = 2.0 / 40
    V1, F1 = omesh.extract_mesh(sdf, bounds, intv, block_res=12, process=False)
    axes = [torch.arange(lo, hi, intv) for lo, hi in bounds]
    x, y, z = torch.meshgrid(*axes, indexing="ij")
    vol = sdf(torch.stack([x, y, z], -1))[..., 0].numpy()
    V2, F2 = omesh.lattice_marching_cubes(vol, 12, 0.0, [a.numpy() for a in axes], intv)
    mask = np.linalg.norm(V2, axis=-1) < 1.0
    idx = np.cumsum(mask) - 1
    fm = mask[F2[:, 0]] & mask[F2[:, 1]] & mask[F2[:, 2]]
    assert np.array_equal(V1, V2[mask]) and np.array_equal(F1, idx[F2[fm]])
    # one block spanning the volume == plain marching cubes
    V3, F3 = omesh.lattice_marching_cubes(vol, 1000)
    V4, F4 = omesh.marching_cubes(vol)
    assert np.array_equal(V3, V4) and np.array_equal(F3, F4)


def test_block_loop_threads_keep_the_sequential_order():
    """Blocks extracted concurrently concatenate exactly as the sequential loop."""
    rng = np.random.default_rng(5)
    vol = rng.normal(size=(41, 30, 37)).astype(np.float32)
    V1, F1, c1, f1 = omesh.lattice_marching_cubes(vol, 8, counts=True)
    V4, F4, c4, f4 = omesh.lattice_marching_cubes(vol, 8, counts=True, threads=4)
    assert np.array_equal(V1, V4) and np.array_equal(F1, F4)
    assert np.array_equal(c1, c4) and np.array_equal(f1, f4)


def _zero_sphere(n=15, r2=25):
    """x^2 + y^2 + z^2 - r2 on an integer lattice: exact zeros at the lattice points of
    the sphere, where marching cubes puts the vertices of several edges on one point."""
    g = np.arange(n, dtype=np.float32) - n // 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return (X * X + Y * Y + Z * Z - r2).astype(np.float32)


def _edge_counts(F):
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    e.sort(axis=1)
    return np.unique(e, axis=0, return_counts=True)


def test_trimesh_process_merges_coincident_vertices():
    """The processed sphere has no repeated vertex, the same face geometry, the first
    occurrence order, and its non-degenerate faces form a closed surface of Euler
    characteristic 2 (the merged points collapse faces to zero area, which stay)."""
    V, F = omesh.marching_cubes(_zero_sphere(), 0.0)
    assert len(V) - len(np.unique(V, axis=0)) > 50  # the exact zeros repeat vertices
    V2, F2 = omesh.trimesh_process(V, F)
    assert len(np.unique(V2, axis=0)) == len(V2)
    assert F2.shape == F.shape and np.array_equal(V2[F2], V[F])
    first = np.array([np.flatnonzero((V == v).all(axis=1))[0] for v in V2])
    assert np.all(np.diff(first) > 0)
    nd = (F2[:, 0] != F2[:, 1]) & (F2[:, 1] != F2[:, 2]) & (F2[:, 0] != F2[:, 2])
    assert (~nd).sum() > 0
    edges, counts = _edge_counts(F2[nd])
    assert np.all(counts == 2)
    assert len(np.unique(F2[nd])) - len(edges) + nd.sum() == 2


def test_trimesh_process_against_unique_rows():
    """Against a vectorised restatement (rows rounded to 8 decimals, np.unique,
    first-occurrence order) on random vertices with exact repeats, repeats within the
    rounding (3e-9 apart), near misses (3e-8 apart) and unreferenced vertices."""
    rng = np.random.default_rng(4)
    base = rng.uniform(-1, 1, (300, 3))
    V = np.concatenate([base, base[:40], base[40:80] + 3e-9, base[80:120] + 3e-8,
                        rng.uniform(-1, 1, (30, 3))])
    F = rng.integers(0, 420, (500, 3))  # vertices 420.. are never referenced
    V2, F2 = omesh.trimesh_process(V, F)
    ref = np.zeros(len(V), bool)
    ref[F.reshape(-1)] = True
    idx = np.flatnonzero(ref)
    q = np.round(V[idx] * 1e8).astype(np.int64)
    _, first, inv = np.unique(q, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    new_id = np.full(len(V), -1)
    new_id[idx] = rank[inv.reshape(-1)]
    assert np.array_equal(V2, V[idx[first[order]]])
    assert np.array_equal(F2, new_id[F])
    assert len(V2) == len(np.unique(np.round(V[idx] * 1e8), axis=0))
    # a mesh without faces keeps every vertex
    V3, F3 = omesh.trimesh_process(V, F[:0])
    assert np.array_equal(V3, V) and F3.shape == (0, 3)


def test_extract_mesh_processes_each_block():
    """Blocks are processed on their own: welding never joins two blocks (their seam
    vertices stay duplicated, as trimesh.util.concatenate leaves them), and the
    processed output has no repeated vertex inside a block."""
    import torch
    res, block_res = 32, 16
    intv = 2.0 / res

    def sdf(xyz):  # exact zeros at lattice points of the radius-0.5 sphere
        return (xyz * xyz).sum(-1, keepdim=True) - 0.25

    V, F = omesh.extract_mesh(sdf, ((-1.0, 1.0),) * 3, intv, block_res=block_res)
    V0, F0 = omesh.extract_mesh(sdf, ((-1.0, 1.0),) * 3, intv, block_res=block_res, process=False)
    assert len(V) < len(V0) and len(F) == len(F0)
    assert np.array_equal(V[F], V0[F0])
    # seam vertices (on block faces) appear once per block that holds them
    assert len(np.unique(V, axis=0)) < len(V)
    del torch
