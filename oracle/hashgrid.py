"""Oracle: multiresolution hash-grid encoding (CPU, numpy + torch).  TEST INFRASTRUCTURE.

Restates the encoding that projects/neuralangelo/utils/modules.py:33-58 builds
through tcnn.Encoding and calls at :76-95.  The arithmetic lives in the
un-vendored third-party package tiny-cuda-nn (installed from git HEAD, no pin:
requirements.txt:3); it is restated here from the published Instant-NGP
algorithm (Mueller et al., SIGGRAPH 2022, sec. 3 and appendix A):

  level table  log2s = log2f(s); scale_l = exp2f(l*log2s)*base - 1; res_l = ceil(scale_l)+1
               size_l = min(next_mult_8(res_l^3), 2^log2_T); dense iff res_l^3 <= size_l
  cell         pos = fmaf(scale_l, x, 0.5); c = floor(pos); t = pos - c
  index        dense: (cx + cy*res + cz*res^2) mod size_l   (uint32 arithmetic)
               hashed: (cx ^ cy*2654435761 ^ cz*805459861) mod 2^log2_T
  value        sum over 8 corners of prod_d(t_d | 1-t_d) * table[(off_l + idx)*F + f]

Pinned by (tests/test_oracle_hashgrid.py): the level-table figures computed
independently in float64 (resolutions 32..2049, offsets, 45,724,048 entries at
2^22); hand-computed index known-answer tests at levels 0, 5, 6, 10, 15; fmaf
against exact rational arithmetic (fractions) incl. constructed ties; the
multilinear-reproduction invariant (a table holding a linear function of the
lattice coordinates returns that function at dense levels); brute force on a
tiny grid; continuity across cell faces; torch autograd of encode() against
central finite differences.  Parity with tcnn's own binary output is
unpinned (tcnn is absent and unpinned).
"""
import numpy as np
import torch

PRIME_Y = np.uint32(2654435761)
PRIME_Z = np.uint32(805459861)


def level_table(n_levels, n_features, log2_hashmap_size, base_resolution, per_level_scale):
    """Level table; float32 steps rounded one at a time (correctly rounded exp2/log2)."""
    s32 = np.float32(per_level_scale)
    log2s = np.float32(np.log2(np.float64(s32)))
    scale = np.zeros(n_levels, np.float32)
    res = np.zeros(n_levels, np.int64)
    size = np.zeros(n_levels, np.int64)
    off = np.zeros(n_levels + 1, np.int64)
    T = 1 << log2_hashmap_size
    o = 0
    for lv in range(n_levels):
        a = np.float32(np.float32(lv) * log2s)
        e = np.float32(np.exp2(np.float64(a)))
        b = np.float32(e * np.float32(base_resolution))
        sc = np.float32(b - np.float32(1.0))
        r = int(np.ceil(sc)) + 1
        p = min(r ** 3, 2 ** 31 - 1)
        p = (p + 7) // 8 * 8
        p = min(p, T)
        scale[lv], res[lv], size[lv], off[lv] = sc, r, p, o
        o += p
    off[n_levels] = o
    dense = res ** 3 <= size
    return dict(n_levels=n_levels, n_features=n_features, log2_T=log2_hashmap_size,
                scale=scale, res=res, size=size, offset=off, dense=dense,
                n_entries=int(o), n_params=int(o) * n_features, log2s=log2s)


def fmaf32(a, b, c):
    """Correctly rounded float32 fused multiply-add, vectorised (exact emulation).

    a*b is exact in float64; s = fl64(a*b + c) carries error e (TwoSum).  fl32(s)
    equals the correctly rounded result unless s is exactly a float32 midpoint,
    in which case the sign of e decides.
    """
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    c = np.asarray(c, np.float32)
    p = a.astype(np.float64) * b.astype(np.float64)
    cd = np.broadcast_to(c.astype(np.float64), p.shape)
    s = p + cd
    bb = s - p
    err = (p - (s - bb)) + (cd - bb)
    r = s.astype(np.float32)
    rd = r.astype(np.float64)
    toward = np.where(s > rd, np.float32(np.inf), np.float32(-np.inf)).astype(np.float32)
    nb = np.nextafter(r, toward)
    delta = nb.astype(np.float64) - rd
    is_mid = (s != rd) & (2.0 * (s - rd) == delta)
    take_nb = is_mid & (err != 0) & (np.sign(err) == np.sign(delta))
    return np.where(take_nb, nb, r).astype(np.float32)


def cell_coords(x, scale):
    """(uint32 cell, float32 frac) per component for one level's float32 scale."""
    pos = fmaf32(np.float32(scale), x, np.float32(0.5))
    fl = np.floor(pos)
    # (uint32)(int)floor(pos): saturating float->int32 conversion, then reinterpret.
    cell = np.clip(fl, -2.0 ** 31, 2.0 ** 31 - 1).astype(np.int64).astype(np.int32).view(np.uint32)
    t = (pos - fl).astype(np.float32)
    return cell, t, pos


def corner_index(cx, cy, cz, res, size, dense):
    cx = cx.astype(np.uint32)
    cy = cy.astype(np.uint32)
    cz = cz.astype(np.uint32)
    with np.errstate(over="ignore"):
        if dense:
            r = np.uint32(res)
            lin = (cx + cy * r + cz * (r * r)).astype(np.uint32)
            return (lin % np.uint32(size)).astype(np.int64)
        h = (cx ^ (cy * PRIME_Y) ^ (cz * PRIME_Z)).astype(np.uint32)
        return (h & np.uint32(size - 1)).astype(np.int64)


def corner_weights32(t):
    """[..., 8] float32 weights, product in dim order 0,1,2."""
    one = np.float32(1.0)
    ws = []
    for c in range(8):
        w = t[..., 0] if (c & 1) else (one - t[..., 0])
        w = (w * (t[..., 1] if (c & 2) else (one - t[..., 1]))).astype(np.float32)
        w = (w * (t[..., 2] if (c & 4) else (one - t[..., 2]))).astype(np.float32)
        ws.append(w.astype(np.float32))
    return np.stack(ws, axis=-1)


def corners(x, lt):
    """Entry indices (offset included) int64 [P,L,8] and float32 weights [P,L,8]."""
    x = np.asarray(x, np.float32)
    P = x.shape[0]
    L = lt["n_levels"]
    idx = np.zeros((P, L, 8), np.int64)
    w = np.zeros((P, L, 8), np.float32)
    for lv in range(L):
        cell, t, _ = cell_coords(x, lt["scale"][lv])
        for c in range(8):
            cc = [cell[:, d] + np.uint32((c >> d) & 1) for d in range(3)]
            idx[:, lv, c] = lt["offset"][lv] + corner_index(*cc, lt["res"][lv], lt["size"][lv],
                                                            bool(lt["dense"][lv]))
        w[:, lv, :] = corner_weights32(t)
    return idx, w


def corner_weights64(x, lt, x64=None):
    """FP64 weights from float64 positions, using the float32-derived cells.

    x64: the unrounded float64 coordinates when available (else float64(x))."""
    x = np.asarray(x, np.float32)
    xd = x.astype(np.float64) if x64 is None else np.asarray(x64, np.float64)
    P, L = x.shape[0], lt["n_levels"]
    w = np.zeros((P, L, 8), np.float64)
    for lv in range(L):
        cell, _, _ = cell_coords(x, lt["scale"][lv])
        pos = np.float64(lt["scale"][lv]) * xd + 0.5
        t = pos - cell.view(np.int32).astype(np.float64)
        for c in range(8):
            wc = np.ones(P)
            for d in range(3):
                wc = wc * (t[:, d] if (c >> d) & 1 else 1.0 - t[:, d])
            w[:, lv, c] = wc
    return w


def encode(x, params, lt, active_levels=None, dtype=torch.float32, table_dtype=None,
           weights32=False):
    """Encoder output [P, L*F] (torch, CPU), differentiable w.r.t. ``params``.

    x: numpy/torch [P,3] float32 in the unit cube.  params: torch [n_params].
    dtype float32: float32 weights; float64: float64 weights (same cells).
    table_dtype: optional storage rounding of the table (e.g. torch.bfloat16).
    weights32: with dtype float64, use the float32 weights of the kernel algorithm
      (accumulated in float64) instead of float64 weights.
    """
    xo = x.detach().cpu().numpy() if torch.is_tensor(x) else np.asarray(x)
    xn = xo.astype(np.float32)
    L, F = lt["n_levels"], lt["n_features"]
    active = L if active_levels is None else int(active_levels)
    idx, w32 = corners(xn, lt)
    use64 = dtype == torch.float64 and not weights32
    w = torch.from_numpy(corner_weights64(xn, lt, xo if xo.dtype == np.float64 else None)
                         if use64 else w32)
    tab = params
    if table_dtype is not None:
        tab = tab.to(table_dtype)
    tab = tab.to(dtype).view(-1, F)
    vals = tab[torch.from_numpy(idx).view(-1)].view(xn.shape[0], L, 8, F)
    out = (w.to(dtype)[..., None] * vals).sum(dim=2)  # [P,L,F]
    if active < L:
        mask = torch.zeros(L, 1, dtype=dtype)
        mask[:active] = 1
        out = out * mask
    return out.reshape(xn.shape[0], L * F)


def abs_bound(x, params, lt):
    """sum_c |w_c| |v_c| per output element: the scale for floating-point tolerances."""
    xn = x.detach().cpu().numpy() if torch.is_tensor(x) else np.asarray(x, np.float32)
    L, F = lt["n_levels"], lt["n_features"]
    idx, w = corners(xn.astype(np.float32), lt)
    vals = params.detach().abs().view(-1, F)[torch.from_numpy(idx).view(-1)].view(-1, L, 8, F)
    return (torch.from_numpy(w).abs()[..., None] * vals).sum(dim=2).reshape(-1, L * F)
