"""Pins the hash-grid oracle (CPU only) before it is used to judge the HIP kernels."""
import math
from fractions import Fraction

import numpy as np
import pytest
import torch

from oracle import hashgrid as hg

G = float(np.exp((np.log(2 ** 11) - np.log(2 ** 5)) / 15))  # modules.py:41


def _lt(log2_T=22, F=8):
    return hg.level_table(16, F, log2_T, 32, G)


def test_level_table_default_config():
    lt = _lt()
    assert list(lt["res"]) == [32, 43, 56, 74, 98, 129, 169, 223, 295, 389, 513, 676, 892,
                               1177, 1553, 2049]
    assert list(lt["offset"][:8]) == [0, 32768, 112280, 287896, 693120, 1634312, 3781008,
                                      3781008 + 4194304]
    assert lt["n_entries"] == 45724048
    assert lt["n_params"] == 365792384
    assert list(lt["dense"]) == [True] * 6 + [False] * 10


def test_level_table_small_config():
    lt = _lt(19, 4)
    assert lt["n_entries"] == 6984576
    assert lt["n_params"] == 27938304
    assert list(lt["dense"]) == [True] * 4 + [False] * 12


def test_level_scale_float32_rounding():
    # Independent float64 evaluation of the same float32 step sequence.
    lt = _lt()
    s32 = float(np.float32(G))
    log2s = float(np.float32(math.log2(s32)))
    for lv in range(16):
        a = float(np.float32(lv * log2s))
        e = float(np.float32(2.0 ** a))
        b = float(np.float32(e * 32.0))
        sc = float(np.float32(b - 1.0))
        assert lt["scale"][lv] == np.float32(sc)
    # the well-known consequence: levels 5, 10, 15 round up past the power of two
    assert lt["scale"][5] > 127 and lt["res"][5] == 129
    assert lt["scale"][10] > 511 and lt["res"][10] == 513
    assert lt["scale"][15] > 2047 and lt["res"][15] == 2049


def _fma_exact(a, b, c):
    v = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    # correctly round the rational to float32 by bracketing
    f = np.float32(float(v))
    cand = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    best = min(cand, key=lambda z: (abs(Fraction(float(z)) - v),
                                    int(np.array(z, np.float32).view(np.uint32)) & 1))
    return np.float32(best)


def test_fmaf32_against_rationals():
    rng = np.random.default_rng(0)
    a = rng.uniform(1, 2100, 400).astype(np.float32)
    b = rng.uniform(-0.5, 1.5, 400).astype(np.float32)
    c = np.full(400, 0.5, np.float32)
    got = hg.fmaf32(a, b, c)
    for i in range(400):
        assert got[i] == _fma_exact(a[i], b[i], c[i])


def test_fmaf32_ties_decided_by_low_bits():
    # a*b = 1 + 2^-24 + 2^-60 is just above the float32 midpoint between 1 and
    # 1+2^-23; the double sum rounds to the midpoint; only the exact FMA rounds up.
    a = np.float32(1.0 + 2.0 ** -23)
    b = np.float32(1.0 - 2.0 ** -23)
    # a*b = 1 - 2^-46 exactly; plus c = 2^-24 + ... build a case numerically:
    cases = 0
    rng = np.random.default_rng(1)
    for _ in range(20000):
        a = np.float32(rng.uniform(1, 2))
        b = np.float32(rng.uniform(1, 2))
        p = Fraction(float(a)) * Fraction(float(b))
        # choose c so that a*b + c lands near a float32 midpoint
        c = np.float32(-float(p) + float(np.float32(float(p))))
        got = hg.fmaf32(np.array([a]), np.array([b]), np.array([c]))[0]
        assert got == _fma_exact(a, b, c)
        cases += 1
    assert cases == 20000


def _hash_py(cx, cy, cz, T):
    m = 0xFFFFFFFF
    h = (cx & m) ^ ((cy * 2654435761) & m) ^ ((cz * 805459861) & m)
    return h % T


def test_index_known_answers():
    lt = _lt()
    x = np.array([[0.5, 0.25, 0.75], [0.3, 0.6, 0.45], [0.9999, 0.0001, 0.5]], np.float32)
    idx, w = hg.corners(x, lt)
    for lv in (0, 5, 6, 10, 15):
        sc = float(lt["scale"][lv])
        res = int(lt["res"][lv])
        size = int(lt["size"][lv])
        for p in range(3):
            cells = []
            for d in range(3):
                pos = Fraction(sc) * Fraction(float(x[p, d])) + Fraction(1, 2)
                cells.append(math.floor(pos))
            for c in range(8):
                cc = [cells[d] + ((c >> d) & 1) for d in range(3)]
                if res ** 3 <= size:
                    ref = (cc[0] + cc[1] * res + cc[2] * res * res) % 2 ** 32 % size
                else:
                    ref = _hash_py(*cc, 2 ** 22)
                assert idx[p, lv, c] == int(lt["offset"][lv]) + ref, (lv, p, c)
    # spot values written out by hand for the first point at level 0:
    # pos = 31*0.5+0.5 = 16, 31*0.25+0.5 = 8.25, 31*0.75+0.5 = 23.75
    assert idx[0, 0, 0] == 16 + 8 * 32 + 23 * 32 * 32
    assert np.isclose(w[0, 0, 0], 1.0 * 0.75 * 0.25)


def test_weights_sum_to_one_and_nonnegative():
    lt = _lt()
    x = np.random.default_rng(2).uniform(0, 1, (500, 3)).astype(np.float32)
    _, w = hg.corners(x, lt)
    assert (w >= 0).all()
    np.testing.assert_allclose(w.sum(-1), 1.0, atol=2e-6)


def test_multilinear_reproduction_dense_levels():
    # A table whose entry at lattice point (i,j,k) holds a*i+b*j+c*k+d returns,
    # at dense levels (no wraparound inside the cube), exactly that affine
    # function of the continuous position pos = scale*x+0.5.
    lt = hg.level_table(4, 2, 22, 8, 1.5)
    coef = np.array([0.3, -0.7, 0.2, 0.1])
    params = np.zeros((lt["n_entries"], 2))
    for lv in range(4):
        r = int(lt["res"][lv])
        assert lt["dense"][lv]
        ii, jj, kk = np.meshgrid(np.arange(r), np.arange(r), np.arange(r), indexing="ij")
        lin = ii + jj * r + kk * r * r
        val = coef[0] * ii + coef[1] * jj + coef[2] * kk + coef[3]
        params[lt["offset"][lv] + lin.ravel(), 0] = val.ravel()
        params[lt["offset"][lv] + lin.ravel(), 1] = 2 * val.ravel()
    params = torch.from_numpy(params.ravel())
    rng = np.random.default_rng(3)
    # keep cell+1 inside the lattice at every level (pos < res - 1): no wraparound
    xmax = min((float(lt["res"][lv]) - 1.5) / float(lt["scale"][lv]) for lv in range(4))
    x = rng.uniform(0.0, xmax, (300, 3)).astype(np.float32)
    out = hg.encode(x, params, lt, dtype=torch.float64).numpy().reshape(300, 4, 2)
    for lv in range(4):
        pos = np.float64(lt["scale"][lv]) * x.astype(np.float64) + 0.5
        exp = pos @ coef[:3] + coef[3]
        np.testing.assert_allclose(out[:, lv, 0], exp, atol=1e-9)
        np.testing.assert_allclose(out[:, lv, 1], 2 * exp, atol=1e-9)


def test_brute_force_tiny_grid():
    lt = hg.level_table(2, 1, 8, 3, 2.0)  # res 3 then 6; second level hashed (216>... )
    rng = np.random.default_rng(4)
    params = torch.from_numpy(rng.normal(size=lt["n_params"]))
    x = rng.uniform(0, 1, (50, 3)).astype(np.float32)
    out = hg.encode(x, params, lt, dtype=torch.float64).numpy()
    for p in range(50):
        for lv in range(2):
            sc, res, size, off = (float(lt["scale"][lv]), int(lt["res"][lv]),
                                  int(lt["size"][lv]), int(lt["offset"][lv]))
            acc = 0.0
            for c in range(8):
                w = 1.0
                cc = []
                for d in range(3):
                    pos = sc * float(x[p, d]) + 0.5
                    fl = math.floor(pos)
                    t = pos - fl
                    bit = (c >> d) & 1
                    w *= t if bit else 1 - t
                    cc.append(fl + bit)
                if res ** 3 <= size:
                    e = (cc[0] + cc[1] * res + cc[2] * res * res) % size
                else:
                    e = _hash_py(*cc, size)
                acc += w * float(params[off + e])
            assert abs(out[p, lv] - acc) < 1e-12


def test_continuity_across_cell_faces():
    lt = _lt(19, 4)
    params = torch.from_numpy(np.random.default_rng(5).normal(size=lt["n_params"]))
    sc = float(lt["scale"][3])
    # a face at pos = 20 in x at level 3: x = (20 - 0.5)/scale
    xf = (20 - 0.5) / sc
    xs = np.array([[np.nextafter(np.float32(xf), np.float32(0)), 0.4, 0.6],
                   [np.nextafter(np.float32(xf), np.float32(1)), 0.4, 0.6]], np.float32)
    out = hg.encode(xs, params, lt, dtype=torch.float64).numpy()[:, 12:16]  # level 3
    cells, _, _ = hg.cell_coords(xs[:, :1], lt["scale"][3])
    assert cells[0, 0] + 1 == cells[1, 0]  # the two points straddle the face
    assert np.abs(out[0] - out[1]).max() < 1e-5


def test_autograd_matches_finite_differences():
    lt = hg.level_table(3, 2, 8, 4, 1.7)
    rng = np.random.default_rng(6)
    params = torch.tensor(rng.normal(size=lt["n_params"]), requires_grad=True)
    x = rng.uniform(0, 1, (7, 3)).astype(np.float32)
    wout = torch.from_numpy(rng.normal(size=(7, 6)))
    loss = (hg.encode(x, params, lt, dtype=torch.float64) * wout).sum()
    (g,) = torch.autograd.grad(loss, params)
    eps = 1e-6
    for k in rng.choice(lt["n_params"], 20, replace=False):
        p1 = params.detach().clone(); p1[k] += eps
        p2 = params.detach().clone(); p2[k] -= eps
        fd = ((hg.encode(x, p1, lt, dtype=torch.float64) * wout).sum() -
              (hg.encode(x, p2, lt, dtype=torch.float64) * wout).sum()) / (2 * eps)
        assert abs(fd.item() - g[k].item()) < 1e-7
