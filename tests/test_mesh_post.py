"""Mesh post-processing on the host: the per-block largest-component filter against the
oracle's union-find restatement of trimesh's split (mesh.py:151-158), hand cases of
the oracle itself, the degenerate-face mask and the PLY writer."""
import numpy as np

from neuralangelo_amd.mesh import largest_component_per_block, nondegenerate_faces, write_ply
from oracle import mesh as om


def _tri(o, s):
    return np.array([o, o + [s, 0, 0], o + [0, s, 0]], float)


def test_oracle_largest_component_hand_cases():
    # component A: two faces sharing an edge (area 2 * 0.5); B: one face (area 2);
    # C: a face touching A only at a vertex (separate under edge adjacency)
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],        # A
                  [5, 0, 0], [7, 0, 0], [5, 2, 0],                    # B
                  [1, 1, 1], [2, 1, 1]], float)                       # C (with vertex 3)
    F = np.array([[0, 1, 2], [1, 3, 2], [4, 5, 6], [3, 7, 8]])
    Vk, Fk = om.largest_component(V, F)
    assert np.array_equal(Vk, V[4:7]) and np.array_equal(Fk, [[0, 1, 2]])
    F2 = np.array([[0, 1, 2], [1, 3, 2], [3, 7, 8]])
    Vk, Fk = om.largest_component(V, F2)  # A (area 1) beats C (area 0.5)
    assert np.array_equal(Fk, [[0, 1, 2], [1, 3, 2]]) and len(Vk) == 4


def _random_blocks(seed, nblocks=5):
    rng = np.random.default_rng(seed)
    Vs, Fs, bf, off = [], [], [0], 0
    for b in range(nblocks):
        nv, nf = int(rng.integers(0, 40)), int(rng.integers(0, 30))
        if nv < 3:
            nv, nf = 0, 0
        V = rng.normal(size=(nv, 3)) + b * 10
        F = rng.integers(0, max(nv, 1), size=(nf, 3)) if nv else np.zeros((0, 3), np.int64)
        F = F[(F[:, 0] != F[:, 1]) & (F[:, 1] != F[:, 2]) & (F[:, 0] != F[:, 2])]
        Vs.append(V)
        Fs.append(F + off)
        off += nv
        bf.append(bf[-1] + len(F))
    return np.concatenate(Vs), np.concatenate(Fs).astype(np.int64), np.array(bf)


def test_per_block_filter_matches_oracle_per_block():
    for seed in range(6):
        V, F, bf = _random_blocks(seed)
        C = (np.arange(len(V) * 3).reshape(-1, 3) % 256).astype(np.uint8)
        Vg, Fg, Cg = largest_component_per_block(V, F, bf, C)
        Vr, Fr, Cr, off = [], [], [], 0
        for b in range(len(bf) - 1):
            Fb = F[bf[b]:bf[b + 1]]
            if len(Fb) == 0:
                continue
            used = np.unique(Fb)
            # the block's own vertex range: every vertex of the block, as trimesh holds them
            lo, hi = used.min(), used.max() + 1
            vb, fb, cb = om.largest_component(V[lo:hi], Fb - lo, C[lo:hi])
            Vr.append(vb)
            Fr.append(fb + off)
            Cr.append(cb)
            off += len(vb)
        assert np.array_equal(Vg, np.concatenate(Vr))
        assert np.array_equal(Fg, np.concatenate(Fr))
        assert np.array_equal(Cg, np.concatenate(Cr))


def test_nondegenerate_faces():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [2, 0, 0], [0.5, 1e-12, 0]], float)
    F = np.array([[0, 1, 2], [0, 1, 3], [0, 3, 4], [1, 1, 2]])
    assert nondegenerate_faces(V, F).tolist() == [True, False, False, False]


def test_write_ply_round_trip(tmp_path):
    V = np.random.default_rng(0).normal(size=(7, 3))
    F = np.array([[0, 1, 2], [2, 3, 4], [4, 5, 6]])
    C = np.arange(21, dtype=np.uint8).reshape(7, 3)
    path = tmp_path / "m.ply"
    write_ply(path, V, F, C)
    raw = path.read_bytes()
    head, body = raw.split(b"end_header\n", 1)
    assert b"element vertex 7" in head and b"element face 3" in head
    vdt = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("r", "u1"), ("g", "u1"),
                    ("b", "u1"), ("a", "u1")])
    verts = np.frombuffer(body[:7 * vdt.itemsize], vdt)
    faces = np.frombuffer(body[7 * vdt.itemsize:], np.dtype([("n", "u1"), ("v", "<i4", (3,))]))
    assert np.array_equal(np.stack([verts["x"], verts["y"], verts["z"]], -1), V)
    assert np.array_equal(np.stack([verts["r"], verts["g"], verts["b"]], -1), C)
    assert (verts["a"] == 255).all() and (faces["n"] == 3).all()
    assert np.array_equal(faces["v"], F)
