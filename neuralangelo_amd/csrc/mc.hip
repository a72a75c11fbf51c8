// Marching cubes on the GPU over a block-partitioned lattice:
// classify -> tile scan -> emit.
//
// Replaces the per-block loop of projects/neuralangelo/utils/mesh.py:24-49
// (mcubes.marching_cubes(sdf, 0.) on each (block_res+1)^3 lattice block,
// mesh.py:119-121, then V*intv + xyz[0,0,0] and concatenation).  All blocks
// of the lattice are processed by one launch chain; the output is the
// concatenation, in block order, of the per-block meshes.  Per block the
// vertex/face order is that of the sequential algorithm: cells visited i-major,
// per cell the "far" edges 6, 5, 10 always own a new vertex and edges
// 0,1,2,3,4,7,8,9,11 own one only on the block boundary (0: j=k=0; 1,2: k=0;
// 3: i=k=0; 4: j=0; 7: i=0; 8: i=j=0; 9: j=0; 11: i=0), in that order.
//
// Work unit: a tile of 1024 consecutive cells of one block (blocks never share
// a tile).  classify reads the volume (the only HBM-heavy pass: 4 B per
// lattice point), counts owned vertices and triangles per tile, and writes the
// per-cell cube index and in-tile vertex prefix only for tiles that touch the
// surface.  A shared edge's owner cell owns a vertex, so its tile is always
// one of those.  The tile scan turns tile counts into global offsets; emit
// revisits only surface tiles.
#include <algorithm>

#include "common.h"
#include "mc_tables.h"

#pragma clang fp contract(off)

namespace na {

constexpr int kMcBlock = 256;
constexpr int kMcItems = 4;
constexpr int kMcTile = kMcBlock * kMcItems;  // cells per tile
constexpr int kScanGroup = kMcBlock * kMcItems;  // tiles per scan group

__constant__ int8_t dMcTri[256][16] = NA_MC_TRI_TABLE_INIT;
__constant__ uint16_t dMcEdge[256] = NA_MC_EDGE_TABLE_INIT;

__device__ const int kCx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ const int kCy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ const int kCz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__device__ const int kEa[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
__device__ const int kEb[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};
__device__ const int kOrder[12] = {6, 5, 10, 0, 1, 2, 3, 4, 7, 8, 9, 11};
// position of edge e in kOrder
__device__ const int kRankOf[12] = {3, 4, 5, 6, 7, 1, 0, 8, 9, 10, 2, 11};

struct McLattice {
  const float* v;
  int64_t nx, ny, nz;   // lattice points per axis
  double iso;
  int64_t br[3];        // cells per full block per axis
  int64_t nb[3];        // blocks per axis
  int64_t tpb;          // tiles per block
};

// One block of the lattice: origin (in points) and cell counts.
struct McBlock {
  int64_t ox, oy, oz;
  int64_t cx, cy, cz;
  int64_t ncell;
};

__device__ inline McBlock block_of(const McLattice& g, int64_t b) {
  McBlock k;
  const int64_t bz = b % g.nb[2];
  const int64_t t = b / g.nb[2];
  const int64_t by = t % g.nb[1];
  const int64_t bx = t / g.nb[1];
  k.ox = bx * g.br[0];
  k.oy = by * g.br[1];
  k.oz = bz * g.br[2];
  k.cx = min(g.br[0], g.nx - 1 - k.ox);
  k.cy = min(g.br[1], g.ny - 1 - k.oy);
  k.cz = min(g.br[2], g.nz - 1 - k.oz);
  k.ncell = k.cx * k.cy * k.cz;
  return k;
}

__device__ inline float vox(const McLattice& g, int64_t x, int64_t y, int64_t z) {
  return g.v[(x * g.ny + y) * g.nz + z];
}

__device__ inline int cube_of(const McLattice& g, int64_t x, int64_t y, int64_t z) {
  int c = 0;
#pragma unroll
  for (int m = 0; m < 8; ++m)
    if ((double)vox(g, x + kCx[m], y + kCy[m], z + kCz[m]) <= g.iso) c |= 1 << m;
  return c;
}

// bit e set if edge e is owned by block-local cell (i,j,k)
__device__ inline int owned_mask(int64_t i, int64_t j, int64_t k) {
  int m = (1 << 6) | (1 << 5) | (1 << 10);
  if (j == 0 && k == 0) m |= 1 << 0;
  if (k == 0) m |= (1 << 1) | (1 << 2);
  if (i == 0 && k == 0) m |= 1 << 3;
  if (j == 0) m |= (1 << 4) | (1 << 9);
  if (i == 0) m |= (1 << 7) | (1 << 11);
  if (i == 0 && j == 0) m |= 1 << 8;
  return m;
}

__device__ inline int tri_count(int cube) {
  int n = 0;
  while (n < 5 && dMcTri[cube][3 * n] >= 0) ++n;
  return n;
}

__device__ inline void unravel(int64_t c, int64_t cy, int64_t cz, int64_t& i, int64_t& j,
                               int64_t& k) {
  k = c % cz;
  const int64_t t = c / cz;
  j = t % cy;
  i = t / cy;
}

// Inclusive scan across the 64 lanes of a wavefront.
__device__ inline int wave_incl_scan(int x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x;
}

// Exclusive scan of one value per thread in thread order across the
// workgroup; *total gets the workgroup sum.  sh: kMcBlock/64 ints of LDS.
__device__ inline int block_excl_scan(int x, int* sh, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int inc = wave_incl_scan(x);
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  int before = 0, sum = 0;
#pragma unroll
  for (int q = 0; q < kMcBlock / 64; ++q) {
    const int s = sh[q];
    before += q < w ? s : 0;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return before + inc - x;
}

__device__ inline int64_t block_excl_scan64(int64_t x, int64_t* sh, int64_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t inc = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  int64_t before = 0, sum = 0;
#pragma unroll
  for (int q = 0; q < kMcBlock / 64; ++q) {
    const int64_t s = sh[q];
    before += q < w ? s : 0;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return before + inc - x;
}

// ---------------------------------------------------------------------------
// 1. classify: one workgroup per tile.  Cells of a tile are visited as
//    local index base + it*256 + thread (coalesced volume reads); the vertex
//    prefix is accumulated in that order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kMcBlock)
mc_classify_kernel(McLattice g, uint8_t* __restrict__ cidx, uint16_t* __restrict__ vpre,
                   int* __restrict__ tile_v, int* __restrict__ tile_f) {
  __shared__ int sh[kMcBlock / 64];
  __shared__ int shf[kMcBlock / 64];
  const int64_t gt = blockIdx.x;
  const int64_t b = gt / g.tpb;
  const int64_t base = (gt % g.tpb) * kMcTile;
  const McBlock k = block_of(g, b);
  if (base >= k.ncell) {  // trailing tile of a ragged block (uniform branch)
    if (threadIdx.x == 0) {
      tile_v[gt] = 0;
      tile_f[gt] = 0;
    }
    return;
  }
  int cube[kMcItems], pre[kMcItems];
  int carry = 0, nf_sum = 0;
#pragma unroll
  for (int it = 0; it < kMcItems; ++it) {
    const int64_t lc = base + it * kMcBlock + threadIdx.x;
    int c = 0, nv = 0;
    if (lc < k.ncell) {
      int64_t i, j, l;
      unravel(lc, k.cy, k.cz, i, j, l);
      c = cube_of(g, k.ox + i, k.oy + j, k.oz + l);
      nv = __popc(dMcEdge[c] & owned_mask(i, j, l));
      nf_sum += tri_count(c);
    }
    cube[it] = c;
    int tot;
    pre[it] = carry + block_excl_scan(nv, sh, &tot);
    carry += tot;
  }
  // triangle total of the tile
  int f = nf_sum;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) f += __shfl_xor(f, o);
  if ((threadIdx.x & 63) == 0) shf[threadIdx.x >> 6] = f;
  __syncthreads();
  int tf = 0;
#pragma unroll
  for (int q = 0; q < kMcBlock / 64; ++q) tf += shf[q];
  const int tv = carry;
  if (tv != 0 || tf != 0) {
    const int64_t slot0 = gt * kMcTile;
#pragma unroll
    for (int it = 0; it < kMcItems; ++it) {
      const int64_t s = slot0 + it * kMcBlock + threadIdx.x;
      cidx[s] = (uint8_t)cube[it];
      vpre[s] = (uint16_t)pre[it];
    }
  }
  if (threadIdx.x == 0) {
    tile_v[gt] = tv;
    tile_f[gt] = tf;
  }
}

// ---------------------------------------------------------------------------
// 2. tile scan.  (a) each workgroup scans kScanGroup tile counts (exclusive,
//    within the group) and writes the group sums; (b) one workgroup scans the
//    group sums (exclusive) and writes the totals.  Global offset of tile t =
//    toff[t] + gsum[t / kScanGroup].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kMcBlock)
mc_scan_tiles_kernel(int64_t ntiles, const int* __restrict__ tile_v, const int* __restrict__ tile_f,
                     int64_t* __restrict__ toff_v, int64_t* __restrict__ toff_f,
                     int64_t* __restrict__ gsum_v, int64_t* __restrict__ gsum_f) {
  __shared__ int64_t sh[kMcBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanGroup + threadIdx.x * kMcItems;
  int64_t lv[kMcItems], lf[kMcItems], sv = 0, sf = 0;
#pragma unroll
  for (int it = 0; it < kMcItems; ++it) {
    const int64_t t = base + it;
    lv[it] = t < ntiles ? tile_v[t] : 0;
    lf[it] = t < ntiles ? tile_f[t] : 0;
    sv += lv[it];
    sf += lf[it];
  }
  int64_t tot_v, tot_f;
  int64_t ev = block_excl_scan64(sv, sh, &tot_v);
  int64_t ef = block_excl_scan64(sf, sh, &tot_f);
#pragma unroll
  for (int it = 0; it < kMcItems; ++it) {
    const int64_t t = base + it;
    if (t < ntiles) {
      toff_v[t] = ev;
      toff_f[t] = ef;
    }
    ev += lv[it];
    ef += lf[it];
  }
  if (threadIdx.x == 0) {
    gsum_v[blockIdx.x] = tot_v;
    gsum_f[blockIdx.x] = tot_f;
  }
}

__global__ void __launch_bounds__(kMcBlock)
mc_scan_groups_kernel(int64_t ngroups, int64_t* __restrict__ gsum_v, int64_t* __restrict__ gsum_f,
                      int64_t* __restrict__ totals) {
  __shared__ int64_t sh[kMcBlock / 64];
  int64_t carry_v = 0, carry_f = 0;
  for (int64_t c0 = 0; c0 < ngroups; c0 += kMcBlock) {
    const int64_t c = c0 + threadIdx.x;
    const int64_t xv = c < ngroups ? gsum_v[c] : 0;
    const int64_t xf = c < ngroups ? gsum_f[c] : 0;
    int64_t tv, tf;
    const int64_t ev = block_excl_scan64(xv, sh, &tv);
    const int64_t ef = block_excl_scan64(xf, sh, &tf);
    if (c < ngroups) {
      gsum_v[c] = carry_v + ev;
      gsum_f[c] = carry_f + ef;
    }
    carry_v += tv;
    carry_f += tf;
  }
  if (threadIdx.x == 0) {
    totals[0] = carry_v;
    totals[1] = carry_f;
  }
}

// ---------------------------------------------------------------------------
// 3. emit: one workgroup per surface tile
// ---------------------------------------------------------------------------
__device__ inline double interp(double x1, double x2, double f1, double f2, double iso) {
  if (f2 == f1) return (x1 + x2) / 2;
  const double d = x2 - x1;
  const double n = iso - f1;
  const double q = f2 - f1;
  const double t = d * n;
  const double r = t / q;
  return r + x1;
}

// Owner cell and owner-side edge of edge e of cell (i,j,k) (which the cell
// does not own); block-local coordinates.
__device__ inline void edge_owner(int e, int64_t i, int64_t j, int64_t k, int64_t& oi,
                                  int64_t& oj, int64_t& ok, int& oe) {
  const int a = kEa[e], b = kEb[e];
  const int64_t px = i + min(kCx[a], kCx[b]), py = j + min(kCy[a], kCy[b]),
                pz = k + min(kCz[a], kCz[b]);
  const int axis = e < 8 ? (e & 1) : 2;
  if (axis == 0) {  // x edge at (px, py, pz)
    oi = px;
    if (py >= 1 && pz >= 1) { oj = py - 1; ok = pz - 1; oe = 6; }
    else if (py == 0 && pz == 0) { oj = 0; ok = 0; oe = 0; }
    else if (pz == 0) { oj = py - 1; ok = 0; oe = 2; }
    else { oj = 0; ok = pz - 1; oe = 4; }
  } else if (axis == 1) {  // y edge
    oj = py;
    if (px >= 1 && pz >= 1) { oi = px - 1; ok = pz - 1; oe = 5; }
    else if (px == 0 && pz == 0) { oi = 0; ok = 0; oe = 3; }
    else if (pz == 0) { oi = px - 1; ok = 0; oe = 1; }
    else { oi = 0; ok = pz - 1; oe = 7; }
  } else {  // z edge
    ok = pz;
    if (px >= 1 && py >= 1) { oi = px - 1; oj = py - 1; oe = 10; }
    else if (px == 0 && py == 0) { oi = 0; oj = 0; oe = 8; }
    else if (py == 0) { oi = px - 1; oj = 0; oe = 9; }
    else { oi = 0; oj = py - 1; oe = 11; }
  }
}

// rank of edge e among the owned crossed edges of a cell, in kOrder order
__device__ inline int owned_rank(int e, int crossed_owned) {
  int r = 0;
  const int pos = kRankOf[e];
  for (int q = 0; q < pos; ++q)
    if (crossed_owned & (1 << kOrder[q])) ++r;
  return r;
}

struct McWorld {  // optional index -> world transform (V * intv + axis[block origin])
  const float* ax[3];
  double intv;
};

__global__ void __launch_bounds__(kMcBlock)
mc_emit_kernel(McLattice g, McWorld wt, const uint8_t* __restrict__ cidx,
               const uint16_t* __restrict__ vpre, const int* __restrict__ tile_v,
               const int* __restrict__ tile_f, const int64_t* __restrict__ toff_v,
               const int64_t* __restrict__ toff_f, const int64_t* __restrict__ gsum_v,
               const int64_t* __restrict__ gsum_f, double* __restrict__ V,
               int64_t* __restrict__ F) {
  __shared__ int sh[kMcBlock / 64];
  const int64_t gt = blockIdx.x;
  if (tile_v[gt] == 0 && tile_f[gt] == 0) return;  // uniform
  const int64_t b = gt / g.tpb;
  const int64_t base = (gt % g.tpb) * kMcTile;
  const McBlock k = block_of(g, b);
  const int64_t vbase = toff_v[gt] + gsum_v[gt / kScanGroup];
  int64_t fcarry = toff_f[gt] + gsum_f[gt / kScanGroup];
  const int64_t tile0 = b * g.tpb;  // first tile of this block
  double org[3] = {0.0, 0.0, 0.0};
  if (wt.ax[0]) {
    org[0] = (double)wt.ax[0][k.ox];
    org[1] = (double)wt.ax[1][k.oy];
    org[2] = (double)wt.ax[2][k.oz];
  }
  for (int it = 0; it < kMcItems; ++it) {
    const int64_t lc = base + it * kMcBlock + threadIdx.x;
    const int64_t slot = gt * kMcTile + it * kMcBlock + threadIdx.x;
    int cube = 0;
    if (lc < k.ncell) cube = cidx[slot];
    const int nf = tri_count(cube);
    int ftot;
    const int64_t fb = fcarry + block_excl_scan(nf, sh, &ftot);
    fcarry += ftot;
    const int edges = dMcEdge[cube];
    if (!edges) continue;
    int64_t i, j, l;
    unravel(lc, k.cy, k.cz, i, j, l);
    const int own = edges & owned_mask(i, j, l);
    int64_t idx[12];
    const int64_t vb = vbase + vpre[slot];
    int rank = 0;
    for (int q = 0; q < 12; ++q) {
      const int e = kOrder[q];
      if (!(edges & (1 << e))) continue;
      if (own & (1 << e)) {
        const int a = kEa[e], c = kEb[e];
        const int64_t p[3] = {i + kCx[a], j + kCy[a], l + kCz[a]};
        const int64_t pb[3] = {i + kCx[c], j + kCy[c], l + kCz[c]};
        const int ax = e < 8 ? (e & 1) : 2;
        const double fa = (double)vox(g, k.ox + p[0], k.oy + p[1], k.oz + p[2]);
        const double fc = (double)vox(g, k.ox + pb[0], k.oy + pb[1], k.oz + pb[2]);
        double o[3] = {(double)p[0], (double)p[1], (double)p[2]};
        o[ax] = interp((double)p[ax], (double)pb[ax], fa, fc, g.iso);
        if (wt.ax[0]) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double s = o[d] * wt.intv;
            o[d] = s + org[d];
          }
        }
        double* out = V + 3 * (vb + rank);
        out[0] = o[0];
        out[1] = o[1];
        out[2] = o[2];
        idx[e] = vb + rank;
        ++rank;
      } else {
        int64_t oi, oj, ok;
        int oe;
        edge_owner(e, i, j, l, oi, oj, ok, oe);
        const int64_t olc = (oi * k.cy + oj) * k.cz + ok;
        const int64_t ot = tile0 + olc / kMcTile;
        const int64_t oslot = tile0 * kMcTile + olc;
        const int ocube = cidx[oslot];
        const int oown = dMcEdge[ocube] & owned_mask(oi, oj, ok);
        idx[e] = toff_v[ot] + gsum_v[ot / kScanGroup] + vpre[oslot] + owned_rank(oe, oown);
      }
    }
    int64_t fo = fb;
    for (int q = 0; q < 15 && dMcTri[cube][q] >= 0; q += 3) {
      int64_t* out = F + 3 * fo;
      out[0] = idx[dMcTri[cube][q]];
      out[1] = idx[dMcTri[cube][q + 1]];
      out[2] = idx[dMcTri[cube][q + 2]];
      ++fo;
    }
  }
}

__global__ void mc_zero_totals_kernel(int64_t* totals) {
  if (threadIdx.x < 2) totals[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct McPlan {
  McLattice g;
  int64_t ntiles, ngroups;
  bool empty;
};

static McPlan mc_plan(const float* v, int64_t nx, int64_t ny, int64_t nz, int64_t block_res,
                      double iso) {
  McPlan p{};
  p.empty = nx < 2 || ny < 2 || nz < 2;
  if (p.empty) return p;
  const int64_t n[3] = {nx, ny, nz};
  p.g.v = v;
  p.g.nx = nx;
  p.g.ny = ny;
  p.g.nz = nz;
  p.g.iso = iso;
  for (int a = 0; a < 3; ++a) {
    p.g.br[a] = block_res > 0 ? std::min<int64_t>(block_res, n[a] - 1) : n[a] - 1;
    p.g.nb[a] = ceil_div(n[a] - 1, p.g.br[a]);
  }
  p.g.tpb = ceil_div(p.g.br[0] * p.g.br[1] * p.g.br[2], (int64_t)kMcTile);
  p.ntiles = p.g.nb[0] * p.g.nb[1] * p.g.nb[2] * p.g.tpb;
  p.ngroups = ceil_div(p.ntiles, (int64_t)kScanGroup);
  return p;
}

struct McWorkspace {
  uint8_t* cidx;
  uint16_t* vpre;
  int* tile_v;
  int* tile_f;
  int64_t* toff_v;
  int64_t* toff_f;
  int64_t* gsum_v;
  int64_t* gsum_f;
};

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

static size_t mc_ws_layout(const McPlan& p, void* base, McWorkspace* ws) {
  const size_t nslots = (size_t)p.ntiles * kMcTile;
  uint8_t* b = (uint8_t*)base;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    uint8_t* q = b ? b + off : nullptr;
    off = align256(off + bytes);
    return q;
  };
  McWorkspace w;
  w.cidx = take(nslots);
  w.vpre = (uint16_t*)take(nslots * 2);
  w.tile_v = (int*)take((size_t)p.ntiles * 4);
  w.tile_f = (int*)take((size_t)p.ntiles * 4);
  w.toff_v = (int64_t*)take((size_t)p.ntiles * 8);
  w.toff_f = (int64_t*)take((size_t)p.ntiles * 8);
  w.gsum_v = (int64_t*)take((size_t)p.ngroups * 8);
  w.gsum_f = (int64_t*)take((size_t)p.ngroups * 8);
  if (ws) *ws = w;
  return off;
}

}  // namespace na

using namespace na;

extern "C" int64_t na_mc_workspace_size(int64_t nx, int64_t ny, int64_t nz, int64_t block_res) {
  NA_CHECK_ARG(nx >= 0 && ny >= 0 && nz >= 0, "negative size");
  const McPlan p = mc_plan(nullptr, nx, ny, nz, block_res, 0.0);
  if (p.empty) return 256;
  NA_CHECK_ARG(p.ntiles < (int64_t)INT32_MAX, "lattice too large");
  return (int64_t)mc_ws_layout(p, nullptr, nullptr);
}

extern "C" int na_mc_count(const float* volume, int64_t nx, int64_t ny, int64_t nz,
                           int64_t block_res, double iso, void* workspace, int64_t ws_bytes,
                           int64_t* totals, na_stream_t stream) {
  NA_CHECK_ARG(nx >= 0 && ny >= 0 && nz >= 0, "negative size");
  NA_CHECK_ARG(totals && workspace, "null pointer");
  hipStream_t s = (hipStream_t)stream;
  const McPlan p = mc_plan(volume, nx, ny, nz, block_res, iso);
  if (p.empty) {
    hipLaunchKernelGGL(mc_zero_totals_kernel, dim3(1), dim3(64), 0, s, totals);
    NA_CHECK_LAUNCH();
    return NA_OK;
  }
  NA_CHECK_ARG(volume, "null volume");
  NA_CHECK_ARG(p.ntiles < (int64_t)INT32_MAX, "lattice too large");
  NA_CHECK_ARG((int64_t)mc_ws_layout(p, nullptr, nullptr) <= ws_bytes, "workspace too small");
  McWorkspace w;
  mc_ws_layout(p, workspace, &w);
  hipLaunchKernelGGL(mc_classify_kernel, dim3((unsigned)p.ntiles), dim3(kMcBlock), 0, s, p.g,
                     w.cidx, w.vpre, w.tile_v, w.tile_f);
  NA_CHECK_LAUNCH();
  hipLaunchKernelGGL(mc_scan_tiles_kernel, dim3((unsigned)p.ngroups), dim3(kMcBlock), 0, s,
                     p.ntiles, w.tile_v, w.tile_f, w.toff_v, w.toff_f, w.gsum_v, w.gsum_f);
  NA_CHECK_LAUNCH();
  hipLaunchKernelGGL(mc_scan_groups_kernel, dim3(1), dim3(kMcBlock), 0, s, p.ngroups, w.gsum_v,
                     w.gsum_f, totals);
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int na_mc_emit(const float* volume, int64_t nx, int64_t ny, int64_t nz,
                          int64_t block_res, double iso, void* workspace, int64_t ws_bytes,
                          const float* axis_x, const float* axis_y, const float* axis_z,
                          double intv, double* vertices, int64_t* faces, na_stream_t stream) {
  NA_CHECK_ARG(nx >= 0 && ny >= 0 && nz >= 0, "negative size");
  const McPlan p = mc_plan(volume, nx, ny, nz, block_res, iso);
  if (p.empty) return NA_OK;
  NA_CHECK_ARG(volume && workspace && vertices && faces, "null pointer");
  NA_CHECK_ARG((axis_x && axis_y && axis_z) || (!axis_x && !axis_y && !axis_z),
               "axis arrays must be given together");
  NA_CHECK_ARG((int64_t)mc_ws_layout(p, nullptr, nullptr) <= ws_bytes, "workspace too small");
  McWorkspace w;
  mc_ws_layout(p, workspace, &w);
  McWorld wt{{axis_x, axis_y, axis_z}, intv};
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(mc_emit_kernel, dim3((unsigned)p.ntiles), dim3(kMcBlock), 0, s, p.g, wt,
                     w.cidx, w.vpre, w.tile_v, w.tile_f, w.toff_v, w.toff_f, w.gsum_v, w.gsum_f,
                     vertices, faces);
  NA_CHECK_LAUNCH();
  return NA_OK;
}
