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
// Work unit: a tile of at most 1024 consecutive cells of one block (blocks never
// share a tile), cut along whole z-rows: with cz <= 1024 cells per row a tile is
// R = 1024 / cz rows of one x-plane (the last tile of a plane may hold fewer),
// with cz > 1024 a 1024-cell segment of one row.  Tiles are numbered plane by
// plane, so their concatenation is the block's i-major cell order, and a tile's
// cell c sits at workspace slot tile * 1024 + c.  classify reads the volume (the
// only HBM-heavy pass: 4 B per lattice point), counts owned vertices and
// triangles per tile, and writes the per-cell cube index and in-tile vertex
// prefix only for tiles that touch the surface.  A shared edge's owner cell owns
// a vertex, so its tile is always one of those.  The tile scan turns tile counts
// into global offsets; emit revisits only surface tiles.
#include <algorithm>

#include "common.h"
#include "mc_tables.h"

#pragma clang fp contract(off)

namespace na {

constexpr int kMcBlock = 256;
constexpr int kMcItems = 4;
constexpr int kMcWaves = kMcBlock / 64;
constexpr int kMcTile = kMcBlock * kMcItems;  // cells per tile
static_assert(kMcItems == 4, "the classify's wave vote names four items");
constexpr int kScanGroup = kMcBlock * kMcItems;  // tiles per scan group
constexpr int kMcEmitBlock = 64;                 // emit: one wave per workgroup

// Bourke corner offsets and edge endpoints; kOrder = per-cell vertex order.
constexpr int kCx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
constexpr int kCy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
constexpr int kCz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
constexpr int kEa[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
constexpr int kEb[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};
constexpr int kOrder[12] = {6, 5, 10, 0, 1, 2, 3, 4, 7, 8, 9, 11};

// Edges that precede edge e in kOrder, as a bit mask.
constexpr int before_mask(int e) {
  int m = 0;
  for (int q = 0; q < 12 && kOrder[q] != e; ++q) m |= 1 << kOrder[q];
  return m;
}

// Case word: crossed-edge mask (bits 0..11) | triangle count << 12.
struct McCaseTable {
  uint16_t w[256];
};
constexpr McCaseTable make_case_table() {
  McCaseTable t{};
  constexpr uint16_t edge[256] = NA_MC_EDGE_TABLE_INIT;
  constexpr int8_t tri[256][16] = NA_MC_TRI_TABLE_INIT;
  for (int c = 0; c < 256; ++c) {
    int n = 0;
    while (n < 5 && tri[c][3 * n] >= 0) ++n;
    t.w[c] = (uint16_t)(edge[c] | (n << 12));
  }
  return t;
}
__constant__ McCaseTable dMcCase = make_case_table();
__constant__ __attribute__((aligned(16))) int8_t dMcTri[256][16] = NA_MC_TRI_TABLE_INIT;

struct McLattice {
  const float* v;
  int64_t nx, ny, nz;   // lattice points per axis
  double iso;
  float isof;           // iso as a float, when that is exact (iso_is_f): then
  int iso_is_f;         // v <= iso compares in float with the same result
  int64_t br[3];        // cells per full block per axis
  int64_t nb[3];        // blocks per axis
  int64_t tpb;          // tiles per block
};

// One block of the lattice: origin (in points) and cell counts.
struct McBlock {
  int64_t ox, oy, oz;
  int64_t cx, cy, cz;
  int64_t ncell;
  // tiles: rpt rows of one plane (cz <= 1024; seg = 1) or one 1024-cell segment
  // of a row (cz > 1024; rpt = 1, seg segments per row); tpp tiles per x-plane
  int64_t rpt, seg, tpp;
};

// The cells of tile q of an x-plane: first row j0, rows, first column k0, columns.
struct McTileRows {
  int64_t j0, rows, k0, cols;
};

__host__ __device__ inline void tile_geometry(int64_t cy, int64_t cz, int64_t& rpt, int64_t& seg,
                                              int64_t& tpp) {
  if (cz <= 1024) {
    rpt = cz > 0 ? 1024 / cz : 1;
    seg = 1;
    tpp = (cy + rpt - 1) / rpt;
  } else {
    rpt = 1;
    seg = (cz + 1023) / 1024;
    tpp = cy * seg;
  }
}

__device__ inline McTileRows tile_rows(const McBlock& k, int64_t q) {
  McTileRows t;
  if (k.seg == 1) {
    t.j0 = q * k.rpt;
    t.rows = min(k.rpt, k.cy - t.j0);
    t.k0 = 0;
    t.cols = k.cz;
  } else {
    t.j0 = q / k.seg;
    t.rows = 1;
    t.k0 = (q - t.j0 * k.seg) * 1024;
    t.cols = min((int64_t)1024, k.cz - t.k0);
  }
  return t;
}

__device__ inline McBlock block_of(const McLattice& g, uint32_t b) {
  // block counts per axis are < 2^31 (the host checks the tile count)
  McBlock k;
  const uint32_t nbz = (uint32_t)g.nb[2], nby = (uint32_t)g.nb[1];
  const uint32_t bz = b % nbz;
  const uint32_t t = b / nbz;
  const uint32_t by = t % nby;
  const uint32_t bx = t / nby;
  k.ox = (int64_t)bx * g.br[0];
  k.oy = (int64_t)by * g.br[1];
  k.oz = (int64_t)bz * g.br[2];
  k.cx = min(g.br[0], g.nx - 1 - k.ox);
  k.cy = min(g.br[1], g.ny - 1 - k.oy);
  k.cz = min(g.br[2], g.nz - 1 - k.oz);
  k.ncell = k.cx * k.cy * k.cz;
  tile_geometry(k.cy, k.cz, k.rpt, k.seg, k.tpp);
  return k;
}

// Tile (within the block) and in-tile cell of block-local cell (i, j, k), in
// 32 bits: tile indices are below 2^31 (the host checks the tile count).
__device__ inline void tile_of_cell(const McBlock& b, uint32_t i, uint32_t j, uint32_t k,
                                    uint32_t& tile, uint32_t& c) {
  if (b.seg == 1) {
    const uint32_t jt = j / (uint32_t)b.rpt;
    tile = i * (uint32_t)b.tpp + jt;
    c = (j - jt * (uint32_t)b.rpt) * (uint32_t)b.cz + k;
  } else {
    const uint32_t ks = k >> 10;
    tile = i * (uint32_t)b.tpp + j * (uint32_t)b.seg + ks;
    c = k & 1023u;
  }
}

// bit e set if edge e is owned by block-local cell (i,j,k)
__device__ inline int owned_mask(bool i0, bool j0, bool k0) {
  int m = (1 << 6) | (1 << 5) | (1 << 10);
  if (j0 && k0) m |= 1 << 0;
  if (k0) m |= (1 << 1) | (1 << 2);
  if (i0 && k0) m |= 1 << 3;
  if (j0) m |= (1 << 4) | (1 << 9);
  if (i0) m |= (1 << 7) | (1 << 11);
  if (i0 && j0) m |= 1 << 8;
  return m;
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

__device__ inline int wave_sum(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

// Tile-order exclusive prefix of per-item counts.  Tile order is item-major,
// then wave, then lane.  inc[it] holds the wave-inclusive scan of item it;
// wsum is LDS [kMcItems][kMcWaves] of wave totals (written by lane 63).
// Returns the tile total; pre[it] gets the exclusive prefix.
__device__ inline int tile_prefix(const int (&inc)[kMcItems], const int (&cnt)[kMcItems],
                                  int (&wsum)[kMcItems][kMcWaves], int (&pre)[kMcItems]) {
  const int w = threadIdx.x >> 6;
  int carry = 0;
#pragma unroll
  for (int it = 0; it < kMcItems; ++it) {
    int before = 0, sum = 0;
#pragma unroll
    for (int q = 0; q < kMcWaves; ++q) {
      const int s = wsum[it][q];
      before += q < w ? s : 0;
      sum += s;
    }
    pre[it] = carry + before + inc[it] - cnt[it];
    carry += sum;
  }
  return carry;
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
  for (int q = 0; q < kMcWaves; ++q) {
    const int64_t s = sh[q];
    before += q < w ? s : 0;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return before + inc - x;
}

// ---------------------------------------------------------------------------
// 1. classify, marching along x.  Workgroup (block b, tile column q) visits the
//    block's x-planes in order and classifies tile (i, q) of each.  The lattice
//    points under the tile column in one plane (rows j0 .. j0 + rows, columns
//    k0 .. k0 + cols) are staged in LDS by buffer-to-LDS loads (no registers),
//    into a ring of lead + 2 plane slots: planes i + 2 .. i + 1 + lead are in
//    flight while the cells between planes i and i + 1 are classified, so each
//    workgroup keeps several planes of requests outstanding.  Wave w's u-th
//    load of a plane fills staged elements (w U + u) 64 .. + 63 (U loads per
//    wave per plane; element e is row e / pw, column e % pw).  Each lattice
//    point is read from memory once per tile column (the row after a tile's
//    last row once more).  Tiles without a crossed edge write only their counts
//    (the rest of the count arrays is zero-filled first).
// ---------------------------------------------------------------------------
constexpr int kMcPref = 9;  // loads per wave per plane at most: planes up to 2304 points
constexpr uint32_t kMcOob = 0x80000000u;  // a buffer offset past every plane (reads zero)
typedef __attribute__((address_space(3))) void mc_lds_t;

// vmcnt(n) for n = 0 .. 15
__device__ inline void mc_wait(int n) {
  switch (n) {
    case 0: __builtin_amdgcn_s_waitcnt(0xF70); break;
    case 1: __builtin_amdgcn_s_waitcnt(0xF71); break;
    case 2: __builtin_amdgcn_s_waitcnt(0xF72); break;
    case 3: __builtin_amdgcn_s_waitcnt(0xF73); break;
    case 4: __builtin_amdgcn_s_waitcnt(0xF74); break;
    case 5: __builtin_amdgcn_s_waitcnt(0xF75); break;
    case 6: __builtin_amdgcn_s_waitcnt(0xF76); break;
    case 7: __builtin_amdgcn_s_waitcnt(0xF77); break;
    case 8: __builtin_amdgcn_s_waitcnt(0xF78); break;
    case 9: __builtin_amdgcn_s_waitcnt(0xF79); break;
    case 10: __builtin_amdgcn_s_waitcnt(0xF7A); break;
    case 11: __builtin_amdgcn_s_waitcnt(0xF7B); break;
    case 12: __builtin_amdgcn_s_waitcnt(0xF7C); break;
    case 13: __builtin_amdgcn_s_waitcnt(0xF7D); break;
    case 14: __builtin_amdgcn_s_waitcnt(0xF7E); break;
    default: __builtin_amdgcn_s_waitcnt(0xF7F); break;
  }
}

// Workgroup barrier for the plane pipeline: this wave's plane loads but the
// newest `pending` are done (pending < 0: no wait on them), its LDS accesses are
// done, then s_barrier.  (A __syncthreads would drain every plane load in
// flight: its release fence counts the buffer-to-LDS loads as outstanding
// stores.)
__device__ inline void mc_sync(int pending) {
  if (pending >= 0) mc_wait(pending);
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// Loads per wave per staged plane (U) and planes in flight beyond the two being
// classified (lead), so that lead * U <= 15 loads are outstanding at the wait.
__host__ __device__ inline void mc_pipeline(int64_t np_max, int& U, int& lead) {
  U = (int)((np_max + 4 * 64 - 1) / (4 * 64));
  lead = U > 0 ? (15 / U < 3 ? 15 / U : 3) : 1;
  if (lead < 1) lead = 1;
}

__global__ void __launch_bounds__(256)
mc_zero_counts_kernel(int64_t ntiles, int* __restrict__ tile_v, int* __restrict__ tile_f,
                      int* __restrict__ nlive) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) *nlive = 0;
  if (t < ntiles) {
    tile_v[t] = 0;
    tile_f[t] = 0;
  }
}

__global__ void __launch_bounds__(kMcBlock)
mc_classify_kernel(McLattice g, int64_t tpp_max, int64_t np_max, uint8_t* __restrict__ cidx,
                   uint16_t* __restrict__ vpre, int* __restrict__ tile_v,
                   int* __restrict__ tile_f, int* __restrict__ live, int* __restrict__ nlive) {
  extern __shared__ float ring[];  // lead + 2 slots of 256 U floats (a plane each)
  __shared__ int wv[kMcItems][kMcWaves];
  __shared__ int wf[kMcWaves];
  // per-wave flags of a staged plane: bit 0 a point at or below iso, bit 1 one
  // above (planes p >= 1 alternate between rows 0 and 1, plane 0 uses row 2)
  __shared__ int wfl[3][kMcWaves];
  // the case words in LDS: a vector-memory lookup in the loop would wait for
  // every plane load issued before it (the load counter drains in order)
  __shared__ uint16_t case_w[256];
  const int64_t b = (int64_t)blockIdx.x / tpp_max;
  const int64_t q = (int64_t)blockIdx.x - b * tpp_max;
  const McBlock k = block_of(g, (uint32_t)b);
  if (q >= k.tpp || k.ncell == 0) return;
  const McTileRows tr = tile_rows(k, q);
  const int64_t pw = tr.cols + 1;            // points per staged row
  const int64_t np = (tr.rows + 1) * pw;     // points per staged plane
  const int64_t ncells = tr.rows * tr.cols;  // cells per tile
  const int64_t sy = g.nz, sx = g.ny * g.nz;
  const float* pl0 = g.v + (k.ox * g.ny + (k.oy + tr.j0)) * g.nz + (k.oz + tr.k0);
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int U, lead;
  mc_pipeline(np_max, U, lead);
  const int slots = lead + 2;
  // a slot holds every element the wave loads write, np .. 256 U - 1 included
  // (those lanes read zeros through the out-of-range offset but still write)
  const int64_t stride = 256 * (int64_t)U;
  // this lane's byte offsets within a plane for the wave's U loads (the host
  // checks (rows + 1) * nz * 4 < 2^31)
  uint32_t voff[kMcPref];
#pragma unroll
  for (int u = 0; u < kMcPref; ++u) {
    const int64_t e = ((int64_t)w * U + u) * 64 + lane;
    const int64_t r = e / pw;
    voff[u] = (u < U && e < np) ? (uint32_t)((r * sy + (e - r * pw)) * 4) : kMcOob;
  }
  const int plane_bytes = (int)((tr.rows + 1) * sy * 4);
  auto issue_plane = [&](int64_t p) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void*)(pl0 + p * sx), 0, plane_bytes, 0x00020000);
    float* dst = ring + (p % slots) * stride + (int64_t)w * U * 64;
#pragma unroll
    for (int u = 0; u < kMcPref; ++u)
      if (u < U) __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (mc_lds_t*)(dst + u * 64), 4, voff[u],
                                                         0, 0, 0);
  };
  // this thread's cells c = it * 256 + thread: staged offset of corner (0,0,0)
  // and the block-boundary flags of their owned edges
  int64_t coff[kMcItems];
  int bmask[kMcItems];
  bool cvalid[kMcItems];
#pragma unroll
  for (int it = 0; it < kMcItems; ++it) {
    const int64_t c = (int64_t)it * kMcBlock + threadIdx.x;
    cvalid[it] = c < ncells;
    const int64_t jj = cvalid[it] ? c / tr.cols : 0;
    const int64_t kk = cvalid[it] ? c - jj * tr.cols : 0;
    coff[it] = jj * pw + kk;
    bmask[it] = ((tr.j0 + jj) == 0 ? 1 : 0) | ((tr.k0 + kk) == 0 ? 2 : 0);
  }
  case_w[threadIdx.x] = dMcCase.w[threadIdx.x];  // (published by the loop's first barrier)
  // planes 0 .. lead, then one more at the top of each iteration (points 0 .. cx)
  for (int64_t p = 0; p <= lead && p <= k.cx; ++p) issue_plane(p);
  const double iso = g.iso;
  const float isof = g.isof;
  const bool iso_f = g.iso_is_f != 0;
  const int64_t tile_base = b * g.tpb + q;
  // which of this lane's staged elements are lattice points (the rest read zeros)
  uint32_t own_pts = 0;
#pragma unroll
  for (int u = 0; u < kMcPref; ++u) own_pts |= voff[u] != kMcOob ? 1u << u : 0u;
  // flags of this wave's own staged points of plane p (its loads have landed).
  // All reads are issued before any is used (one LDS round trip, not one per
  // load: u >= U reads the slot's first element and is masked out), and the
  // compare runs in one precision, chosen per launch.
  auto wave_flags = [&](int64_t p) -> int {
    const float* src = ring + (p % slots) * stride + (int64_t)w * U * 64 + lane;
    float v[kMcPref];
#pragma unroll
    for (int u = 0; u < kMcPref; ++u) v[u] = src[(u < U ? u : 0) * 64];
    bool below = false, above = false;
    if (iso_f) {
#pragma unroll
      for (int u = 0; u < kMcPref; ++u) {
        const bool in = (own_pts >> u) & 1u, le = v[u] <= isof;
        below = below || (in && le);
        above = above || (in && !le);
      }
    } else {
#pragma unroll
      for (int u = 0; u < kMcPref; ++u) {
        const bool in = (own_pts >> u) & 1u, le = (double)v[u] <= iso;
        below = below || (in && le);
        above = above || (in && !le);
      }
    }
    return (__any(below) ? 1 : 0) | (__any(above) ? 2 : 0);
  };
  int fl_lo = 0;  // the workgroup's flags of plane i
  for (int64_t i = 0; i < k.cx; ++i) {
    if (i + 1 + lead <= k.cx) issue_plane(i + 1 + lead);
    // planes i and i + 1 have landed once at most the loads of the planes
    // issued after i + 1 (i + 2 .. i + 1 + lead, those that exist) are pending
    const int64_t after = (i + 1 + lead <= k.cx ? i + 1 + lead : k.cx) - (i + 1);
    mc_wait((int)(after > 0 ? after : 0) * U);
    asm volatile("" ::: "memory");
    const int fw1 = wave_flags(i + 1);
    const int fw0 = i == 0 ? wave_flags(0) : 0;
    if (lane == 0) {
      wfl[(i + 1) & 1][w] = fw1;
      if (i == 0) wfl[2][w] = fw0;
    }
    mc_sync(-1);  // both planes and every wave's flags are visible
    int fl_hi = 0;
#pragma unroll
    for (int qq = 0; qq < kMcWaves; ++qq) {
      fl_hi |= wfl[(i + 1) & 1][qq];
      if (i == 0) fl_lo |= wfl[2][qq];
    }
    const bool crossed = (fl_lo | fl_hi) == 3;
    fl_lo = fl_hi;
    // Planes i and i + 1 all at or below iso, or all above: every cell of the
    // slab has case 0 or 255, so no vertex, no face and nothing to write.  No
    // end barrier either: the slab's slots were read only by their own loading
    // waves (the flags above), before this iteration's barrier.
    if (!crossed) continue;
    const float* s0 = ring + (i % slots) * stride;
    const float* s1 = ring + ((i + 1) % slots) * stride;
    int cube[kMcItems], nv[kMcItems], inc[kMcItems];
    int nf = 0;
#pragma unroll
    for (int it = 0; it < kMcItems; ++it) {
      int c = 0, n = 0;
      if (cvalid[it]) {
        const float* p0 = s0 + coff[it];
        const float* p1 = s1 + coff[it];
        const float a00 = p0[0], a01 = p0[pw], a10 = p1[0], a11 = p1[pw];
        const float b00 = p0[1], b01 = p0[pw + 1], b10 = p1[1], b11 = p1[pw + 1];
        if (iso_f) {
          c = (a00 <= isof) | ((a10 <= isof) << 1) | ((a11 <= isof) << 2) | ((a01 <= isof) << 3) |
              ((b00 <= isof) << 4) | ((b10 <= isof) << 5) | ((b11 <= isof) << 6) |
              ((b01 <= isof) << 7);
        } else {
          c = ((double)a00 <= iso) | (((double)a10 <= iso) << 1) |
              (((double)a11 <= iso) << 2) | (((double)a01 <= iso) << 3) |
              (((double)b00 <= iso) << 4) | (((double)b10 <= iso) << 5) |
              (((double)b11 <= iso) << 6) | (((double)b01 <= iso) << 7);
        }
        const int word = case_w[c];
        n = __popc(word & 0xfff & owned_mask(i == 0, bmask[it] & 1, bmask[it] & 2));
        nf += word >> 12;
      }
      cube[it] = c;
      nv[it] = n;
    }
    // the scans only for waves with a crossed edge (most of a lattice has none)
    if (__any((nv[0] | nv[1] | nv[2] | nv[3] | nf) != 0)) {
#pragma unroll
      for (int it = 0; it < kMcItems; ++it) {
        inc[it] = wave_incl_scan(nv[it]);
        if (lane == 63) wv[it][w] = inc[it];
      }
      nf = wave_sum(nf);
    } else {
#pragma unroll
      for (int it = 0; it < kMcItems; ++it) {
        inc[it] = 0;
        if (lane == 63) wv[it][w] = 0;
      }
      nf = 0;
    }
    if (lane == 0) wf[w] = nf;
    mc_sync(-1);  // the wave totals (LDS); plane loads stay in flight
    int pre_v[kMcItems];
    const int tv = tile_prefix(inc, nv, wv, pre_v);
    int tf = 0;
#pragma unroll
    for (int qq = 0; qq < kMcWaves; ++qq) tf += wf[qq];
    const int64_t gt = tile_base + i * k.tpp;
    if (tv != 0 || tf != 0) {
      const int64_t slot0 = gt * kMcTile + threadIdx.x;
#pragma unroll
      for (int it = 0; it < kMcItems; ++it) {
        cidx[slot0 + it * kMcBlock] = (uint8_t)cube[it];
        vpre[slot0 + it * kMcBlock] = (uint16_t)pre_v[it];
      }
      if (threadIdx.x == 0) {
        tile_v[gt] = tv;
        tile_f[gt] = tf;
        live[atomicAdd(nlive, 1)] = (int)gt;  // emit's work list (any order)
      }
    }
    // every wave has read planes i and i + 1 (and the counts) before a slot or
    // the counts are written again
    mc_sync(-1);
  }
  mc_wait(0);  // no buffer-to-LDS load outlives the workgroup
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
  __shared__ int64_t sh[kMcWaves];
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
  __shared__ int64_t sh[kMcWaves];
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

struct McWorld {  // optional index -> world transform (V * intv + axis[block origin])
  const float* ax[3];
  double intv;
};

struct McOffsets {
  const uint8_t* cidx;
  const uint16_t* vpre;
  const int64_t* toff_v;
  const int64_t* gsum_v;
};

// A cell's crossed edges in two phases, so that the loads of all its edges are
// in flight together: emit_load issues them (the two lattice values of an owned
// edge; the owner's cube index, in-tile prefix and tile offsets of a shared
// one), emit_store consumes them (writes the owned vertex, or forms the shared
// vertex's global index) and records the cell's vertex index of each edge.
struct McEdgeData {
  float fa[12], fb[12];
  int64_t otoff[12], ogsum[12];
  int ocx[12], ovp[12], omask[12];
};

// Owner of the vertex on edge E of block-local cell (i,j,k), which that cell
// does not own: the owner cell's workspace slot and tile, and the mask of the
// owner's edges that precede its own edge for this vertex and that it owns.
template <int E>
__device__ inline void shared_owner(uint32_t tile0, uint32_t i, uint32_t j, uint32_t k,
                                    const McBlock& blk, int64_t& oslot, uint32_t& ot, int& omask) {
  constexpr int a = kEa[E], b = kEb[E];
  constexpr int lx = kCx[a] < kCx[b] ? kCx[a] : kCx[b];
  constexpr int ly = kCy[a] < kCy[b] ? kCy[a] : kCy[b];
  constexpr int lz = kCz[a] < kCz[b] ? kCz[a] : kCz[b];
  constexpr int axis = E < 8 ? (E & 1) : 2;
  const uint32_t px = i + lx, py = j + ly, pz = k + lz;  // low end of the lattice edge
  uint32_t oi, oj, ok;
  int bm;  // edges before the owner's edge in kOrder
  if (axis == 0) {
    oi = px;
    if (py >= 1 && pz >= 1) { oj = py - 1; ok = pz - 1; bm = before_mask(6); }
    else if (py == 0 && pz == 0) { oj = 0; ok = 0; bm = before_mask(0); }
    else if (pz == 0) { oj = py - 1; ok = 0; bm = before_mask(2); }
    else { oj = 0; ok = pz - 1; bm = before_mask(4); }
  } else if (axis == 1) {
    oj = py;
    if (px >= 1 && pz >= 1) { oi = px - 1; ok = pz - 1; bm = before_mask(5); }
    else if (px == 0 && pz == 0) { oi = 0; ok = 0; bm = before_mask(3); }
    else if (pz == 0) { oi = px - 1; ok = 0; bm = before_mask(1); }
    else { oi = 0; ok = pz - 1; bm = before_mask(7); }
  } else {
    ok = pz;
    if (px >= 1 && py >= 1) { oi = px - 1; oj = py - 1; bm = before_mask(10); }
    else if (px == 0 && py == 0) { oi = 0; oj = 0; bm = before_mask(8); }
    else if (py == 0) { oi = px - 1; oj = 0; bm = before_mask(9); }
    else { oi = 0; oj = py - 1; bm = before_mask(11); }
  }
  uint32_t otl, oc;
  tile_of_cell(blk, oi, oj, ok, otl, oc);
  ot = tile0 + otl;
  oslot = (int64_t)ot * kMcTile + oc;
  omask = owned_mask(oi == 0, oj == 0, ok == 0) & bm;
}

template <int Q>
__device__ inline void emit_load(const McOffsets& o, const float* p, int64_t sx, int64_t sy,
                                 uint32_t tile0, uint32_t i, uint32_t j, uint32_t k,
                                 const McBlock& blk, int edges, int own, McEdgeData& d) {
  constexpr int e = kOrder[Q];
  if (!(edges & (1 << e))) return;
  if (own & (1 << e)) {
    constexpr int a = kEa[e], b = kEb[e];
    d.fa[e] = p[kCx[a] * sx + kCy[a] * sy + kCz[a]];
    d.fb[e] = p[kCx[b] * sx + kCy[b] * sy + kCz[b]];
  } else {
    int64_t oslot;
    uint32_t ot;
    shared_owner<e>(tile0, i, j, k, blk, oslot, ot, d.omask[e]);
    d.ocx[e] = o.cidx[oslot];
    d.ovp[e] = o.vpre[oslot];
    d.otoff[e] = o.toff_v[ot];
    d.ogsum[e] = o.gsum_v[ot / kScanGroup];
  }
}

template <int Q>
__device__ inline void emit_store(const McLattice& g, const McWorld& wt, const double (&org)[3],
                                  const uint16_t* case_w, uint32_t i, uint32_t j, uint32_t k,
                                  int edges, int own, int64_t vb, const McEdgeData& d,
                                  double* __restrict__ V, int64_t* sidx) {
  constexpr int e = kOrder[Q];
  if (!(edges & (1 << e))) return;
  if (own & (1 << e)) {
    constexpr int a = kEa[e], b = kEb[e];
    constexpr int axis = e < 8 ? (e & 1) : 2;
    double q[3] = {(double)(i + kCx[a]), (double)(j + kCy[a]), (double)(k + kCz[a])};
    const double xb = axis == 0 ? (double)(i + kCx[b]) : axis == 1 ? (double)(j + kCy[b])
                                                                   : (double)(k + kCz[b]);
    q[axis] = interp(q[axis], xb, (double)d.fa[e], (double)d.fb[e], g.iso);
    if (wt.ax[0]) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double s = q[c] * wt.intv;
        q[c] = s + org[c];
      }
    }
    const int64_t vi = vb + __popc(own & edges & before_mask(e));
    double* out = V + 3 * vi;
    out[0] = q[0];
    out[1] = q[1];
    out[2] = q[2];
    sidx[e * kMcEmitBlock] = vi;
  } else {
    sidx[e * kMcEmitBlock] =
        d.otoff[e] + d.ogsum[e] + d.ovp[e] + __popc(case_w[d.ocx[e]] & d.omask[e]);
  }
}

// Persistent, one wave per workgroup: workgroup b takes entries b, b + grid, ...
// of the classify's list of surface tiles (the only tiles with output); each
// tile's output positions come from the tile scan, so the list's order does not
// matter.  Lane l reads the cube indices of cells 16 l .. 16 l + 15 (one 16-byte
// load), scans the face counts and lists the surface cells in LDS; the wave
// then takes the listed cells 64 at a time.  Waves work on their own tiles, with
// no workgroup barrier between them.
__global__ void __launch_bounds__(kMcEmitBlock)
mc_emit_kernel(McLattice g, McWorld wt, McOffsets o, const int* __restrict__ live,
               const int* __restrict__ nlive, const int64_t* __restrict__ toff_f,
               const int64_t* __restrict__ gsum_f, double* __restrict__ V,
               int64_t* __restrict__ F) {
  __shared__ int64_t sidx_all[12 * kMcEmitBlock];
  __shared__ uint16_t case_w[256];
  // the tile's cells with a crossed edge: cell | cube << 10 | face prefix << 18
  __shared__ uint32_t surf[kMcTile];
  const int lane = threadIdx.x;
  const int64_t sy = g.nz, sx = g.ny * g.nz;
  int64_t* sidx = sidx_all + lane;
  const uint32_t tpb = (uint32_t)g.tpb;
#pragma unroll
  for (int q = 0; q < 256 / kMcEmitBlock; ++q)
    case_w[q * kMcEmitBlock + lane] = dMcCase.w[q * kMcEmitBlock + lane];
  __syncthreads();
  const bool cidx16 = ((uintptr_t)o.cidx & 15) == 0;
  const int n_live = *nlive;
  for (int li = blockIdx.x; li < n_live; li += gridDim.x) {
    const uint32_t gt = (uint32_t)live[li];
    const uint32_t b = gt / tpb;
    const uint32_t tl = gt - b * tpb;  // tile within the block
    const McBlock k = block_of(g, b);
    const int64_t vbase = o.toff_v[gt] + o.gsum_v[gt / kScanGroup];
    const int64_t fbase = toff_f[gt] + gsum_f[gt / kScanGroup];
    const uint32_t tile0 = b * tpb;
    double org[3] = {0.0, 0.0, 0.0};
    if (wt.ax[0]) {
      org[0] = (double)wt.ax[0][k.ox];
      org[1] = (double)wt.ax[1][k.oy];
      org[2] = (double)wt.ax[2][k.oz];
    }
    // the tile's place in the block: x-plane ti and tile column tl - ti * tpp
    const uint32_t ti = tl / (uint32_t)k.tpp;
    const McTileRows tr = tile_rows(k, (int64_t)(tl - ti * (uint32_t)k.tpp));
    const uint32_t ncells = (uint32_t)(tr.rows * tr.cols);
    const uint32_t c0 = (uint32_t)lane * 16;
    const uint8_t* cp = o.cidx + (int64_t)gt * kMcTile + c0;
    uint32_t cw[4];
    if (cidx16) {
      const uint4 v = *reinterpret_cast<const uint4*>(cp);
      cw[0] = v.x; cw[1] = v.y; cw[2] = v.z; cw[3] = v.w;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        cw[q] = (uint32_t)cp[4 * q] | ((uint32_t)cp[4 * q + 1] << 8) |
                ((uint32_t)cp[4 * q + 2] << 16) | ((uint32_t)cp[4 * q + 3] << 24);
    }
    // cells past the tile's end hold no case (their slots were never written)
#pragma unroll
    for (int t = 0; t < 16; ++t)
      if (c0 + t >= ncells) cw[t >> 2] &= ~(0xffu << ((t & 3) * 8));
    int nf_l = 0, ns_l = 0;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int wd = case_w[(cw[t >> 2] >> ((t & 3) * 8)) & 255u];
      nf_l += wd >> 12;
      ns_l += (wd & 0xfff) != 0 ? 1 : 0;
    }
    const int nf_inc = wave_incl_scan(nf_l), ns_inc = wave_incl_scan(ns_l);
    const int n_surf = __shfl(ns_inc, 63);
    int fp = nf_inc - nf_l, sb = ns_inc - ns_l;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const uint32_t cb = (cw[t >> 2] >> ((t & 3) * 8)) & 255u;
      const int wd = case_w[cb];
      if ((wd & 0xfff) != 0) surf[sb++] = (c0 + t) | (cb << 10) | ((uint32_t)fp << 18);
      fp += wd >> 12;
    }
    __syncthreads();
#pragma unroll 1
    for (int si = lane; si < n_surf; si += kMcEmitBlock) {
      const uint32_t e0 = surf[si];
      const uint32_t c = e0 & 1023u;
      const int cb = (int)((e0 >> 10) & 255u);
      const int edges = case_w[cb] & 0xfff;
      // triangle list of the case (16 bytes, -1 terminated), loaded with the edges
      const uint4 trow = *reinterpret_cast<const uint4*>(&dMcTri[cb][0]);
      const uint32_t jj = c / (uint32_t)tr.cols;
      const uint32_t i = ti, j = (uint32_t)tr.j0 + jj, kk = (uint32_t)tr.k0 + (c - jj * (uint32_t)tr.cols);
      const int own = edges & owned_mask(i == 0, j == 0, kk == 0);
      const float* p = g.v + ((k.ox + (int64_t)i) * g.ny + (k.oy + (int64_t)j)) * g.nz +
                       (k.oz + (int64_t)kk);
      const int64_t vb = vbase + o.vpre[(int64_t)gt * kMcTile + c];
      // two batches of six edges (the registers of twelve in flight would halve
      // the waves per SIMD)
      McEdgeData d;
#define NA_MC_LOAD(Q) emit_load<Q>(o, p, sx, sy, tile0, i, j, kk, k, edges, own, d)
#define NA_MC_STORE(Q) emit_store<Q>(g, wt, org, case_w, i, j, kk, edges, own, vb, d, V, sidx)
      NA_MC_LOAD(0); NA_MC_LOAD(1); NA_MC_LOAD(2); NA_MC_LOAD(3); NA_MC_LOAD(4); NA_MC_LOAD(5);
      NA_MC_STORE(0); NA_MC_STORE(1); NA_MC_STORE(2); NA_MC_STORE(3); NA_MC_STORE(4); NA_MC_STORE(5);
      asm volatile("" ::: "memory");  // the second batch's loads stay after the first's use
      NA_MC_LOAD(6); NA_MC_LOAD(7); NA_MC_LOAD(8); NA_MC_LOAD(9); NA_MC_LOAD(10); NA_MC_LOAD(11);
      NA_MC_STORE(6); NA_MC_STORE(7); NA_MC_STORE(8); NA_MC_STORE(9); NA_MC_STORE(10);
      NA_MC_STORE(11);
#undef NA_MC_LOAD
#undef NA_MC_STORE
      const uint32_t tw[4] = {trow.x, trow.y, trow.z, trow.w};
      int64_t fo = fbase + (int64_t)(e0 >> 18);
#pragma unroll
      for (int q3 = 0; q3 < 15; q3 += 3) {
        const int t0 = (int8_t)(tw[q3 >> 2] >> ((q3 & 3) * 8));
        if (t0 < 0) break;
        const int t1 = (int8_t)(tw[(q3 + 1) >> 2] >> (((q3 + 1) & 3) * 8));
        const int t2 = (int8_t)(tw[(q3 + 2) >> 2] >> (((q3 + 2) & 3) * 8));
        int64_t* out = F + 3 * fo;
        out[0] = sidx[t0 * kMcEmitBlock];
        out[1] = sidx[t1 * kMcEmitBlock];
        out[2] = sidx[t2 * kMcEmitBlock];
        ++fo;
      }
    }
    __syncthreads();  // surf is read before the next tile writes it
  }
}

__global__ void mc_zero_totals_kernel(int64_t* totals) {
  if (threadIdx.x < 2) totals[threadIdx.x] = 0;
}

// Offsets of every block's vertices and faces in the concatenated output:
// block b starts at tile b * tpb (tiles are block-major, tpb per block), and
// entry nblocks holds the totals.
__global__ void __launch_bounds__(256)
mc_block_offsets_kernel(int64_t nblocks, int64_t tpb, int64_t ntiles,
                        const int* __restrict__ tile_v, const int* __restrict__ tile_f,
                        const int64_t* __restrict__ toff_v, const int64_t* __restrict__ toff_f,
                        const int64_t* __restrict__ gsum_v, const int64_t* __restrict__ gsum_f,
                        int64_t* __restrict__ block_v, int64_t* __restrict__ block_f) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nblocks) return;
  if (b < nblocks) {
    const int64_t t = b * tpb;
    block_v[b] = toff_v[t] + gsum_v[t / kScanGroup];
    block_f[b] = toff_f[t] + gsum_f[t / kScanGroup];
  } else {
    const int64_t t = ntiles - 1;
    block_v[b] = toff_v[t] + gsum_v[t / kScanGroup] + tile_v[t];
    block_f[b] = toff_f[t] + gsum_f[t / kScanGroup] + tile_f[t];
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct McPlan {
  McLattice g;
  int64_t ntiles, ngroups;
  int64_t tpp;     // tiles per x-plane of a full block (the classify grid's tile columns)
  int64_t np_max;  // lattice points of the largest staged plane over the block shapes
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
  p.g.isof = (float)iso;
  p.g.iso_is_f = (double)p.g.isof == iso;
  for (int a = 0; a < 3; ++a) {
    p.g.br[a] = block_res > 0 ? std::min<int64_t>(block_res, n[a] - 1) : n[a] - 1;
    p.g.nb[a] = ceil_div(n[a] - 1, p.g.br[a]);
  }
  // a smaller (last) block along y or z has no more tiles per plane than a full
  // one (fewer rows, and at least as many rows per tile), so the full block's
  // count bounds every block's
  int64_t rpt, seg;
  tile_geometry(p.g.br[1], p.g.br[2], rpt, seg, p.tpp);
  p.g.tpb = p.g.br[0] * p.tpp;
  p.ntiles = p.g.nb[0] * p.g.nb[1] * p.g.nb[2] * p.g.tpb;
  p.ngroups = ceil_div(p.ntiles, (int64_t)kScanGroup);
  p.np_max = 0;
  const int64_t ys[2] = {p.g.br[1], (ny - 1) - (p.g.nb[1] - 1) * p.g.br[1]};
  const int64_t zs[2] = {p.g.br[2], (nz - 1) - (p.g.nb[2] - 1) * p.g.br[2]};
  for (int64_t cy : ys)
    for (int64_t cz : zs) {
      int64_t r, sg, tp;
      tile_geometry(cy, cz, r, sg, tp);
      const int64_t rows = std::min(r, cy), cols = std::min(cz, (int64_t)1024);
      p.np_max = std::max(p.np_max, (rows + 1) * (cols + 1));
    }
  return p;
}

struct McWorkspace {
  uint8_t* cidx;
  uint16_t* vpre;
  int* tile_v;
  int* tile_f;
  int* live;   // surface tiles (classify appends, emit walks)
  int* nlive;  // their count
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
  w.live = (int*)take((size_t)p.ntiles * 4);
  w.nlive = (int*)take(4);
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
  NA_CHECK_ARG(p.np_max <= (int64_t)kMcPref * kMcBlock, "staged plane too large");
  NA_CHECK_ARG(1025 * nz * 4 < (int64_t)INT32_MAX, "lattice rows too long (nz)");
  hipLaunchKernelGGL(mc_zero_counts_kernel, dim3((unsigned)ceil_div(p.ntiles, (int64_t)256)),
                     dim3(256), 0, s, p.ntiles, w.tile_v, w.tile_f, w.nlive);
  NA_CHECK_LAUNCH();
  const int64_t nblocks = p.g.nb[0] * p.g.nb[1] * p.g.nb[2];
  int U, lead;
  mc_pipeline(p.np_max, U, lead);
  const size_t lds = (size_t)(lead + 2) * 256 * U * sizeof(float);  // slots of 256 U floats
  hipLaunchKernelGGL(mc_classify_kernel, dim3((unsigned)(nblocks * p.tpp)), dim3(kMcBlock), lds, s,
                     p.g, p.tpp, p.np_max, w.cidx, w.vpre, w.tile_v, w.tile_f, w.live, w.nlive);
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
  McOffsets o{w.cidx, w.vpre, w.toff_v, w.gsum_v};
  // persistent grid of as many workgroups as are resident at once (the surface
  // tiles are few next to the lattice's tiles)
  int dev = 0, ncu = 256, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    ncu = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mc_emit_kernel, kMcEmitBlock, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const dim3 grid((unsigned)std::min<int64_t>(p.ntiles, (int64_t)ncu * per_cu));
  hipLaunchKernelGGL(mc_emit_kernel, grid, dim3(kMcEmitBlock), 0, s, p.g, wt, o, w.live, w.nlive,
                     w.toff_f, w.gsum_f, vertices, faces);
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int64_t na_mc_num_blocks(int64_t nx, int64_t ny, int64_t nz, int64_t block_res) {
  NA_CHECK_ARG(nx >= 0 && ny >= 0 && nz >= 0, "negative size");
  const McPlan p = mc_plan(nullptr, nx, ny, nz, block_res, 0.0);
  if (p.empty) return 0;
  return p.g.nb[0] * p.g.nb[1] * p.g.nb[2];
}

extern "C" int na_mc_block_offsets(int64_t nx, int64_t ny, int64_t nz, int64_t block_res,
                                   void* workspace, int64_t ws_bytes, int64_t* block_v,
                                   int64_t* block_f, na_stream_t stream) {
  NA_CHECK_ARG(nx >= 0 && ny >= 0 && nz >= 0, "negative size");
  const McPlan p = mc_plan(nullptr, nx, ny, nz, block_res, 0.0);
  if (p.empty) return NA_OK;
  NA_CHECK_ARG(workspace && block_v && block_f, "null pointer");
  NA_CHECK_ARG((int64_t)mc_ws_layout(p, nullptr, nullptr) <= ws_bytes, "workspace too small");
  McWorkspace w;
  mc_ws_layout(p, workspace, &w);
  const int64_t nblocks = p.g.nb[0] * p.g.nb[1] * p.g.nb[2];
  hipLaunchKernelGGL(mc_block_offsets_kernel, dim3((unsigned)ceil_div(nblocks + 1, (int64_t)256)),
                     dim3(256), 0, (hipStream_t)stream, nblocks, p.g.tpb, p.ntiles, w.tile_v,
                     w.tile_f, w.toff_v, w.toff_f, w.gsum_v, w.gsum_f, block_v, block_f);
  NA_CHECK_LAUNCH();
  return NA_OK;
}
