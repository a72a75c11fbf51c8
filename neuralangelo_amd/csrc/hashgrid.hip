// Multiresolution hash-grid encoding: forward gather and backward scatter.
//
// Replaces tcnn.Encoding(3, {"otype": "HashGrid", ...}) as constructed at
// projects/neuralangelo/utils/modules.py:42-50 and called at :84-86.
// Algorithm (Instant-NGP, Mueller et al. 2022, as implemented by tiny-cuda-nn):
//   per level l: pos = fmaf(scale_l, x_d, 0.5); cell = floor(pos); t = pos - cell
//   dense level  (res^3 <= size_l): idx = (cx + cy*res + cz*res^2) mod size_l
//   hashed level (res^3 >  size_l): idx = (cx ^ cy*2654435761 ^ cz*805459861) mod T
//   out[l*F+f] = sum_{corner c} (prod_d t_d or 1-t_d) * table[(off_l+idx_c)*F+f]
// corners are enumerated with bit d of c selecting +1 along dim d, weights are
// multiplied in dim order 0,1,2 and accumulated with fmaf in corner order.
//
// Layout: table is AoS per entry ([entries][F]), levels concatenated; the
// output is [n, L*F] row-major (level-major within a row), which is what the
// SDF MLP consumes.  Forward: one thread per (point, level), level-minor, so
// the L lanes of a point write its output row in one wave instruction (whole
// cache lines; a level-major grid writes each row 16 bytes at a time from
// workgroups launched far apart).  Backward: one thread per (point, feature) so the F
// atomics of one corner are adjacent lanes of one instruction and coalesce
// into a single L2 atomic request.
#include <cstring>

#include "common.h"

namespace na {

struct LevelParams {
  float scale;
  uint32_t res;
  uint32_t off;
  uint32_t size;
  bool hashed;
};

__device__ inline LevelParams level_params(const na_grid_desc& d, int l) {
  LevelParams p;
  p.scale = d.scale[l];
  p.res = d.resolution[l];
  p.off = d.offset[l];
  p.size = d.offset[l + 1] - p.off;
  uint64_t r = p.res;
  p.hashed = r * r * r > (uint64_t)p.size;
  return p;
}

__device__ inline void pos_fract(float x, float scale, uint32_t& cell, float& t) {
  float pos = fmaf(scale, x, 0.5f);
  float fl = floorf(pos);
  cell = (uint32_t)(int)fl;
  t = pos - fl;
}

__device__ inline uint32_t corner_index(uint32_t x, uint32_t y, uint32_t z,
                                        const LevelParams& p) {
  if (p.hashed) {
    return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & (p.size - 1u);
  }
  return (x + y * p.res + z * p.res * p.res) % p.size;
}

__device__ inline float corner_weight(int c, const float t[3]) {
  float w = (c & 1) ? t[0] : 1.0f - t[0];
  w *= (c & 2) ? t[1] : 1.0f - t[1];
  w *= (c & 4) ? t[2] : 1.0f - t[2];
  return w;
}

// The eight corner indices of one cell, as corner_index computes them.  The
// corners differ from corner 0 by fixed per-axis terms (uint32 arithmetic
// wraps the same way), so a hashed level costs two xors and a mask per corner
// and a dense level one add; both are formed and one is selected, since the
// lanes of a wave hold different levels.  The dense wrap (index % size) only
// changes corners past the grid's upper face or of points outside the unit
// cube, so it runs only when some corner of the cell needs it.
__device__ inline void cell_corners(const uint32_t (&cell)[3], const LevelParams& p,
                                    uint32_t (&id)[8]) {
  constexpr uint32_t kPy = 2654435761u, kPz = 805459861u;
  const uint32_t hx[2] = {cell[0], cell[0] + 1u};
  const uint32_t hy[2] = {cell[1] * kPy, cell[1] * kPy + kPy};
  const uint32_t hz[2] = {cell[2] * kPz, cell[2] * kPz + kPz};
  const uint32_t r1 = p.res, r2 = p.res * p.res;
  const uint32_t d0 = cell[0] + cell[1] * r1 + cell[2] * r2;
  const uint32_t dd[4] = {d0, d0 + r1, d0 + r2, d0 + r2 + r1};
  const uint32_t mask = p.size - 1u;
  bool wrap = false;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t h = (hx[c & 1] ^ hy[(c >> 1) & 1] ^ hz[c >> 2]) & mask;
    const uint32_t d = dd[c >> 1] + (uint32_t)(c & 1);
    wrap = wrap || d >= p.size;
    id[c] = p.hashed ? h : d;
  }
  if (__builtin_expect(!p.hashed && wrap, 0)) {
#pragma unroll
    for (int c = 0; c < 8; ++c) id[c] = (dd[c >> 1] + (uint32_t)(c & 1)) % p.size;
  }
}

// Trilinear weights of the eight corners, multiplied in corner_weight's order.
__device__ inline void cell_weights(const float (&t)[3], float (&w)[8]) {
  const float wx[2] = {1.0f - t[0], t[0]}, wy[2] = {1.0f - t[1], t[1]};
  const float wz[2] = {1.0f - t[2], t[2]};
  float wxy[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) wxy[a] = wx[a & 1] * wy[a >> 1];
#pragma unroll
  for (int c = 0; c < 8; ++c) w[c] = wxy[c & 3] * wz[c >> 2];
}

// Load F table values of one entry as floats (vectorised).
template <int F, typename TT> struct EntryLoad;
template <int F> struct EntryLoad<F, float> {
  __device__ static inline void load(const float* __restrict__ p, float v[F]) {
    if constexpr (F == 8) {
      float4 a = *reinterpret_cast<const float4*>(p);
      float4 b = *reinterpret_cast<const float4*>(p + 4);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else if constexpr (F == 4) {
      float4 a = *reinterpret_cast<const float4*>(p);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    } else if constexpr (F == 2) {
      float2 a = *reinterpret_cast<const float2*>(p);
      v[0] = a.x; v[1] = a.y;
    } else {
      v[0] = p[0];
    }
  }
};
template <int F> struct EntryLoad<F, bf16_t> {
  __device__ static inline void load(const bf16_t* __restrict__ p, float v[F]) {
    if constexpr (F == 8) {
      uint4 a = *reinterpret_cast<const uint4*>(p);
      uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        v[2 * k] = bf16_to_f32((bf16_t)(w[k] & 0xffffu));
        v[2 * k + 1] = bf16_to_f32((bf16_t)(w[k] >> 16));
      }
    } else if constexpr (F == 4) {
      uint2 a = *reinterpret_cast<const uint2*>(p);
      v[0] = bf16_to_f32((bf16_t)(a.x & 0xffffu));
      v[1] = bf16_to_f32((bf16_t)(a.x >> 16));
      v[2] = bf16_to_f32((bf16_t)(a.y & 0xffffu));
      v[3] = bf16_to_f32((bf16_t)(a.y >> 16));
    } else if constexpr (F == 2) {
      uint32_t a = *reinterpret_cast<const uint32_t*>(p);
      v[0] = bf16_to_f32((bf16_t)(a & 0xffffu));
      v[1] = bf16_to_f32((bf16_t)(a >> 16));
    } else {
      v[0] = bf16_to_f32(p[0]);
    }
  }
};

template <int F, typename TO> struct RowStore;
template <int F> struct RowStore<F, float> {
  __device__ static inline void store(float* __restrict__ p, const float v[F]) {
    if constexpr (F == 8) {
      *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
      *reinterpret_cast<float4*>(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
    } else if constexpr (F == 4) {
      *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (F == 2) {
      *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
    } else {
      p[0] = v[0];
    }
  }
};
template <int F> struct RowStore<F, bf16_t> {
  __device__ static inline void store(bf16_t* __restrict__ p, const float v[F]) {
    uint32_t w[(F + 1) / 2];
#pragma unroll
    for (int k = 0; k < F / 2; ++k)
      w[k] = (uint32_t)f32_to_bf16(v[2 * k]) | ((uint32_t)f32_to_bf16(v[2 * k + 1]) << 16);
    if constexpr (F == 8) {
      *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
    } else if constexpr (F == 4) {
      *reinterpret_cast<uint2*>(p) = make_uint2(w[0], w[1]);
    } else if constexpr (F == 2) {
      *reinterpret_cast<uint32_t*>(p) = w[0];
    } else {
      p[0] = f32_to_bf16(v[0]);
    }
  }
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;

// Backward launches cover (chunk of points, level) pairs.  Each level reads its
// own 4F or 2F bytes of every dL/dy row, so the rows' cache lines are shared by
// all levels: the levels of one chunk run back to back on one XCD (workgroups
// are dispatched to the eight XCDs round-robin by id), and the rows come from
// HBM about once instead of once per level.  Block b: XCD b mod 8, slot b / 8;
// level slot mod levels, chunk (slot / levels) * 8 + XCD.  The grid holds
// ceil(chunks / 8) * 8 * levels blocks; chunks past the end return at once.
constexpr int kXcds = 8;

__device__ inline void chunk_level(int levels, int64_t& chunk, int& level) {
  const uint32_t b = blockIdx.x, slot = b / kXcds;
  level = (int)(slot % (uint32_t)levels);
  chunk = (int64_t)(slot / (uint32_t)levels) * kXcds + (b % kXcds);
}

static inline int chunk_level_grid(int64_t chunks, int levels, dim3* grid) {
  const int64_t blocks = ceil_div(chunks, (int64_t)kXcds) * kXcds * levels;
  NA_CHECK_ARG(blocks * kBlock <= (int64_t)0xffffffffu, "n too large for one launch");
  *grid = dim3((unsigned)blocks);
  return NA_OK;
}

__global__ void __launch_bounds__(kBlock)
hashgrid_corners_kernel(na_grid_desc d, const float* __restrict__ x, int64_t n,
                        int32_t* __restrict__ idx_out, float* __restrict__ w_out) {
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int l = blockIdx.y;
  if (i >= n) return;
  LevelParams p = level_params(d, l);
  uint32_t cell[3];
  float t[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) pos_fract(x[i * 3 + k], p.scale, cell[k], t[k]);
  int64_t base = (i * d.n_levels + l) * 8;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    uint32_t id = corner_index(cell[0] + (c & 1), cell[1] + ((c >> 1) & 1),
                               cell[2] + ((c >> 2) & 1), p);
    idx_out[base + c] = (int32_t)(p.off + id);
    w_out[base + c] = corner_weight(c, t);
  }
}

// acc = sum over the eight corners of the cell of weight * entry (corner order
// 0..7, as the per-corner loops accumulate).
template <int F, typename TT>
__device__ inline void gather_cell(const TT* __restrict__ tab, const LevelParams& p,
                                   const uint32_t (&cell)[3], const float (&t)[3], float (&acc)[F]) {
  uint32_t id[8];
  float w[8];
  cell_corners(cell, p, id);
  cell_weights(t, w);
#pragma unroll
  for (int f = 0; f < F; ++f) acc[f] = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    float v[F];
    EntryLoad<F, TT>::load(tab + (size_t)id[c] * F, v);
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = fmaf(w[c], v[f], acc[f]);
  }
}

// The eight raw corner entries of a cell (F = 8 bf16: 16 bytes each) and their
// weighted sum, split so that a tap in the centre's cell can reuse the centre's
// entries instead of fetching them again.
__device__ inline void load_cell8(const bf16_t* __restrict__ tab, const LevelParams& p,
                                  const uint32_t (&cell)[3], uint4 (&v)[8]) {
  uint32_t id[8];
  cell_corners(cell, p, id);
#pragma unroll
  for (int c = 0; c < 8; ++c) v[c] = *reinterpret_cast<const uint4*>(tab + (size_t)id[c] * 8);
}
__device__ inline void sum_cell8(const uint4 (&v)[8], const float (&t)[3], float (&acc)[8]) {
  float w[8];
  cell_weights(t, w);
#pragma unroll
  for (int f = 0; f < 8; ++f) acc[f] = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t u[4] = {v[c].x, v[c].y, v[c].z, v[c].w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc[2 * k] = fmaf(w[c], __uint_as_float(u[k] << 16), acc[2 * k]);
      acc[2 * k + 1] = fmaf(w[c], __uint_as_float(u[k] & 0xffff0000u), acc[2 * k + 1]);
    }
  }
}

// Input point -> unit-cube coordinate: u = (x - vmin) / vrange
// (projects/neuralangelo/utils/modules.py:82-83; vmin = 0, vrange = 1 is the
// tcnn convention of inputs already in [0,1]).
__device__ inline float normalize_coord(float x, float vmin, float vrange) {
  return (x - vmin) / vrange;
}

// The same coordinate for the gather kernels, where the division is a visible
// share of the work: a positive normal power-of-two range divides exactly by
// an exponent shift.
__device__ inline float normalize_coord_gather(float x, float vmin, float vrange) {
  const uint32_t b = __float_as_uint(vrange);
  if ((b & 0x807fffffu) == 0u && (b >> 23) != 0u) return ldexpf(x - vmin, 127 - (int)(b >> 23));
  return (x - vmin) / vrange;
}

template <int F, typename TT, typename TO>
__global__ void __launch_bounds__(kBlock)
hashgrid_fwd_kernel(na_grid_desc d, const float* __restrict__ x, int64_t n, float vmin,
                    float vrange, const TT* __restrict__ table, TO* __restrict__ out, int64_t ld,
                    int write_extras, int active_levels) {
  // level-minor mapping: the L lanes of one point write one output row together
  const int L = d.n_levels;
  const int64_t tid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t i = tid / L;
  const int l = (int)(tid % L);
  if (i >= n) return;
  TO* o = out + i * ld + l * F;
  if (write_extras && l == 0) {
    // fp32: [enc(L*F) | x(3) | zero pad up to ld]
    // bf16: [enc(L*F) | hi(3) | lo(3) | hi(3) | zero pad], hi = bf16(x),
    //       lo = bf16(x - hi): x is carried to 16 mantissa bits, and with the
    //       weight columns [W_hi | W_hi | W_lo] the product keeps
    //       x_hi W_hi + x_lo W_hi + x_hi W_lo
    TO* e = out + i * ld + L * F;
    int64_t c0 = L * F + 3;
    if constexpr (sizeof(TO) == 2) {
      for (int k = 0; k < 3; ++k) {
        const float xv = x[i * 3 + k];
        const bf16_t hi = f32_to_bf16(xv);
        const float lo = xv - bf16_to_f32(hi);
        e[k] = hi;
        e[3 + k] = f32_to_bf16(lo);
        e[6 + k] = hi;
      }
      c0 = L * F + 9;
    } else {
      for (int k = 0; k < 3; ++k) scalar_io<TO>::store(e + k, x[i * 3 + k]);
    }
    for (int64_t c = c0; c < ld; ++c) scalar_io<TO>::store(out + i * ld + c, 0.0f);
  }
  float acc[F];
#pragma unroll
  for (int f = 0; f < F; ++f) acc[f] = 0.0f;
  if (l < active_levels) {
    LevelParams p = level_params(d, l);
    uint32_t cell[3];
    float t[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      pos_fract(normalize_coord_gather(x[i * 3 + k], vmin, vrange), p.scale, cell[k], t[k]);
    gather_cell<F, TT>(table + (size_t)p.off * F, p, cell, t, acc);
  }
  RowStore<F, TO>::store(o, acc);
}

template <int F, typename TD>
__global__ void __launch_bounds__(kBlock)
hashgrid_bwd_kernel(na_grid_desc d, const float* __restrict__ x, int64_t n, float vmin,
                    float vrange, const TD* __restrict__ dy, int64_t ld,
                    float* __restrict__ grad, int active) {
  int64_t chunk;
  int l;
  chunk_level(active, chunk, l);
  int64_t tid = chunk * kBlock + threadIdx.x;
  int64_t i = tid / F;
  int f = (int)(tid % F);
  if (i >= n) return;
  const int L = d.n_levels;
  (void)L;
  float g = scalar_io<TD>::load(dy + i * ld + l * F + f);
  if (g == 0.0f) return;  // adding +0 leaves the accumulator unchanged
  LevelParams p = level_params(d, l);
  uint32_t cell[3];
  float t[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    pos_fract(normalize_coord(x[i * 3 + k], vmin, vrange), p.scale, cell[k], t[k]);
  float* gtab = grad + (size_t)p.off * F + f;
  uint32_t id[8];
  float w[8];
  cell_corners(cell, p, id);
  cell_weights(t, w);
#pragma unroll
  for (int c = 0; c < 8; ++c) atomic_add_f32(gtab + (size_t)id[c] * F, w[c] * g);
}

constexpr float kTapDir[4][3] = {{1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}};
// Coordinate class per evaluation (0 the centre, then the four taps) and axis:
// 0 the centre's coordinate, 1 x + eps, 2 x - eps.  The tap scatter groups the
// taps by their x class: taps 0 and 3 on +eps, taps 1 and 2 on -eps.
constexpr int kEvalClass[5][3] = {{0, 0, 0}, {1, 2, 2}, {2, 2, 1}, {2, 1, 2}, {1, 1, 1}};
static_assert(kTapDir[0][0] > 0 && kTapDir[3][0] > 0 && kTapDir[1][0] < 0 && kTapDir[2][0] < 0,
              "tap grouping by x class");

// Centre input and tap increments of the SDF network for the increment
// formulation of the 4-tap gradient (bf16 compute).  Per point: X0c (the centre
// row) and, for each tap t, the point-major row 4 i + t of D = X(x + eps k_t) - X(x),
// evaluated in fp32 and stored as two bf16 parts (hi = bf16(D), lo = bf16(D - hi)),
// so the increments keep about 16 mantissa bits relative to their own (small)
// size and the four taps' second-order sum survives the rounding.  Row layouts:
//   X0c [ld]:   enc (L*F) | x as [hi | lo | hi] (9) | zeros
//   D   [ld_d]: enc hi (L*F) | (x_t - x) as [hi | lo | hi] (9) | zeros to ld |
//               enc lo (L*F) | zeros
// so the first ld columns of D match X0c's layout (the first-layer weights apply
// to both, and the weight gradient reads them) and the lo block pairs with a
// second copy of the encoder weights.  S [ld] holds the taps' sum of D (fp32 sum,
// one bf16 rounding: second order, so it keeps bf16 precision of its own size).
// One thread per (point, level).
__device__ inline void split_store(bf16_t* e, const float v[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const bf16_t hi = f32_to_bf16(v[k]);
    e[k] = hi;
    e[3 + k] = f32_to_bf16(v[k] - bf16_to_f32(hi));
    e[6 + k] = hi;
  }
}

template <int F, typename TT>
__global__ void __launch_bounds__(kBlock)
sdf_taps_input_kernel(na_grid_desc d, const float* __restrict__ x, int64_t n, float vmin,
                      float vrange, float eps, const TT* __restrict__ table,
                      bf16_t* __restrict__ X0c, int64_t ld, bf16_t* __restrict__ D, int64_t ld_d,
                      bf16_t* __restrict__ S, int active_levels) {
  // level-minor mapping: the L lanes of one point write its rows' level segments
  // together, so each output row is written whole by one wave instruction
  const int L = d.n_levels;
  const int64_t tid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t i = tid / L;
  const int l = (int)(tid % L);
  if (i >= n) return;
  const int64_t LF = (int64_t)L * F;
  float xc[3], xt[4][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    xc[k] = x[i * 3 + k];
#pragma unroll
    for (int t = 0; t < 4; ++t) xt[t][k] = xc[k] + kTapDir[t][k] * eps;
  }
  // the six rows' coordinate columns and zero padding, a row per level lane (the
  // point's lanes run together, so one lane writing all six would hold the rest)
  for (int r = l; r < 6; r += L) {  // 0: X0c, 1 .. 4: tap r - 1 of D, 5: S
    float v[3];
    bf16_t* row;
    if (r == 0) {
      row = X0c + i * ld;
#pragma unroll
      for (int k = 0; k < 3; ++k) v[k] = xc[k];
    } else if (r <= 4) {
      row = D + (4 * i + (r - 1)) * ld_d;
#pragma unroll
      for (int k = 0; k < 3; ++k) v[k] = xt[r - 1][k] - xc[k];  // exact (Sterbenz)
    } else {
      row = S + i * ld;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float dsum = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) dsum += xt[t][k] - xc[k];
        v[k] = dsum;
      }
    }
    split_store(row + LF, v);
    for (int64_t c = LF + 9; c < ld; ++c) row[c] = 0;
    if (r >= 1 && r <= 4)
      for (int64_t c = ld + LF; c < ld_d; ++c) row[c] = 0;
  }
  if (l >= active_levels) {  // coarse-to-fine: inactive levels are zero in all outputs
    float z[F];
#pragma unroll
    for (int f = 0; f < F; ++f) z[f] = 0.0f;
    RowStore<F, bf16_t>::store(X0c + i * ld + l * F, z);
    RowStore<F, bf16_t>::store(S + i * ld + l * F, z);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      bf16_t* row = D + (4 * i + t) * ld_d;
      RowStore<F, bf16_t>::store(row + l * F, z);
      RowStore<F, bf16_t>::store(row + ld + l * F, z);
    }
    return;
  }
  const LevelParams p = level_params(d, l);
  const TT* tab = table + (size_t)p.off * F;
  // centre first, then one tap at a time (two F-vectors live at once); with a
  // bf16 table of 8 features the centre's raw corner entries stay in registers
  // and a tap in the same cell reuses them (the fetch runs only for the lanes
  // whose tap crossed a cell boundary)
  constexpr bool kReuse = F == 8 && sizeof(TT) == 2;
  float ac[F];
  uint32_t cc[3];
  uint4 vc[kReuse ? 8 : 1];
  {
    float fr[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      pos_fract(normalize_coord_gather(xc[k], vmin, vrange), p.scale, cc[k], fr[k]);
    if constexpr (kReuse) {
      load_cell8((const bf16_t*)tab, p, cc, vc);
      sum_cell8(vc, fr, ac);
    } else {
      gather_cell<F, TT>(tab, p, cc, fr, ac);
    }
    RowStore<F, bf16_t>::store(X0c + i * ld + l * F, ac);
  }
  float ssum[F];
#pragma unroll
  for (int f = 0; f < F; ++f) ssum[f] = 0.0f;
#pragma unroll 1
  for (int t = 0; t < 4; ++t) {
    uint32_t cell[3];
    float fr[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      pos_fract(normalize_coord_gather(xt[t][k], vmin, vrange), p.scale, cell[k], fr[k]);
    float at[F];
    if constexpr (kReuse) {
      uint4 vt[8];
      if (cell[0] == cc[0] && cell[1] == cc[1] && cell[2] == cc[2]) {
#pragma unroll
        for (int c = 0; c < 8; ++c) vt[c] = vc[c];
      } else {
        load_cell8((const bf16_t*)tab, p, cell, vt);
      }
      sum_cell8(vt, fr, at);
    } else {
      gather_cell<F, TT>(tab, p, cell, fr, at);
    }
    float hi[F], lo[F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const float dv = at[f] - ac[f];
      ssum[f] += dv;
      hi[f] = bf16_to_f32(f32_to_bf16(dv));
      lo[f] = dv - hi[f];  // exact
    }
    bf16_t* row = D + (4 * i + t) * ld_d;
    RowStore<F, bf16_t>::store(row + l * F, hi);
    RowStore<F, bf16_t>::store(row + ld + l * F, lo);
  }
  RowStore<F, bf16_t>::store(S + i * ld + l * F, ssum);
}

// Backward of the five encoder evaluations of one numerical-gradient point
// (the centre and its four tetrahedral taps, modules.py:84-86 and :157-175)
// in one pass.  The taps lie less than one cell from the centre at every
// active level, so the union of the five points' corners fits a 3x3x3 box of
// lattice points; the five contributions to each corner are summed in
// registers.  A thread owns one feature of one level for `run` consecutive
// points (consecutive samples of a ray; the launch tunable taps_bwd.run,
// default 4): while their boxes share one anchor
// the sums keep accumulating, and each touched corner gets one atomic when the
// anchor changes or the run ends.  At coarse levels, where neighbouring
// samples share cells and atomics to one address would otherwise queue, this
// removes most of the atomics.  When a level's cells spread wider than one cell
// (an epsilon larger than the finest active cell), that point falls back to
// one atomic per point and corner.

template <int F>
__device__ __attribute__((always_inline)) inline void taps_flush(float* gtab, const LevelParams& p,
                                                                 const int (&blo)[3],
                                                                 const float (&acc)[27],
                                                                 uint32_t mask) {
  // corner_index of the 27 box corners from per-axis terms: a hashed level
  // xors one term per axis; a dense level adds them, and its wrap (index %
  // size) is the identity while the box lies inside the grid (0 <= blo and
  // blo + 2 < res on every axis, so the index is below res^3 <= size; blo below
  // zero arrives as a large unsigned value and takes the wrap)
  constexpr uint32_t kPy = 2654435761u, kPz = 805459861u;
  const uint32_t x0 = (uint32_t)blo[0], y0 = (uint32_t)blo[1], z0 = (uint32_t)blo[2];
  uint32_t id[27];
  if (p.hashed) {
    const uint32_t m = p.size - 1u;
#pragma unroll
    for (int o = 0; o < 27; ++o)
      id[o] = ((x0 + (uint32_t)(o % 3)) ^ ((y0 + (uint32_t)((o / 3) % 3)) * kPy) ^
               ((z0 + (uint32_t)(o / 9)) * kPz)) & m;
  } else {
    const uint32_t r1 = p.res, r2 = p.res * p.res;
    const uint32_t d0 = x0 + y0 * r1 + z0 * r2;
#pragma unroll
    for (int o = 0; o < 27; ++o)
      id[o] = d0 + (uint32_t)(o % 3) + (uint32_t)((o / 3) % 3) * r1 + (uint32_t)(o / 9) * r2;
    const bool inside =
        p.res > 2u && x0 < p.res - 2u && y0 < p.res - 2u && z0 < p.res - 2u;
    if (!inside) {
#pragma unroll
      for (int o = 0; o < 27; ++o) id[o] %= p.size;
    }
  }
#pragma unroll
  for (int o = 0; o < 27; ++o)
    if (mask & (1u << o)) atomic_add_f32(gtab + (size_t)id[o] * F, acc[o]);
}

// P2: the volume range is a positive normal power of two, 2^(-vshift) of the
// unit cube, and coordinates normalise by an exact exponent shift (no division
// code in that instantiation).
template <int F, typename TD, bool P2>
__global__ void __launch_bounds__(kBlock)
hashgrid_taps_bwd_kernel(na_grid_desc d, const float* __restrict__ x, int64_t n, float vmin,
                         float vrange, int vshift, float eps, const TD* __restrict__ dyc,
                         const TD* __restrict__ dyt, const TD* __restrict__ dyq, int64_t ld,
                         float* __restrict__ grad, int inc, int run, int active) {
  auto unit = [&](float v) -> float {
    return P2 ? ldexpf(v - vmin, vshift) : normalize_coord(v, vmin, vrange);
  };
  int64_t chunk;
  int l;
  chunk_level(active, chunk, l);
  const int64_t tid = chunk * kBlock + threadIdx.x;
  const int64_t i0 = (tid / F) * run;
  const int f = (int)(tid % F);
  if (i0 >= n) return;
  const int64_t col = (int64_t)l * F + f;
  const LevelParams p = level_params(d, l);
  float* gtab = grad + (size_t)p.off * F + f;
  float acc[27];
#pragma unroll
  for (int o = 0; o < 27; ++o) acc[o] = 0.0f;
  uint32_t mask = 0;
  int blo[3] = {0, 0, 0};
  const int64_t iend = i0 + run < n ? i0 + run : n;
  // the rows and coordinates of point i + 1 are loaded while point i is processed
  struct Rows {
    float c, t[4], q, x[3];
  };
  auto fetch = [&](int64_t i, Rows& r) {
    if (inc) {
      r.q = dyq ? scalar_io<TD>::load(dyq + i * ld + col) : 0.0f;
#pragma unroll
      for (int t = 0; t < 4; ++t) r.t[t] = scalar_io<TD>::load(dyt + (4 * i + t) * ld + col);
    } else {
      r.q = 0.0f;
#pragma unroll
      for (int t = 0; t < 4; ++t) r.t[t] = scalar_io<TD>::load(dyt + ((int64_t)t * n + i) * ld + col);
    }
    r.c = scalar_io<TD>::load(dyc + i * ld + col);
#pragma unroll
    for (int k = 0; k < 3; ++k) r.x[k] = x[i * 3 + k];
  };
  Rows nxt;
  fetch(i0, nxt);
  for (int64_t i = i0; i < iend; ++i) {
    const Rows cur = nxt;
    if (i + 1 < iend) fetch(i + 1, nxt);
    float g[5];
    if (inc) {
      // increments: tap rows point-major, gradients of the inputs D_t = X(x + eps k_t)
      // - X(x) (dyt, plus the point row dyq, the gradient of sum_t D_t, shared by the
      // four taps); the centre row holds the gradient of the centre input alone, so
      // the centre's corners receive g_c - sum_t g_t (formed here in fp32)
#pragma unroll
      for (int t = 0; t < 4; ++t) g[1 + t] = cur.t[t] + cur.q;
      g[0] = cur.c - ((g[1] + g[2]) + (g[3] + g[4]));
    } else {
      g[0] = cur.c;
#pragma unroll
      for (int t = 0; t < 4; ++t) g[1 + t] = cur.t[t];
    }
    if (g[0] == 0.0f && g[1] == 0.0f && g[2] == 0.0f && g[3] == 0.0f && g[4] == 0.0f) continue;
    // Per axis the five evaluations take three coordinates: the centre's and
    // x +- eps (each tap moves every axis by +eps or -eps, as the forward built
    // it: x + k_t eps).  Class 0 is the centre, 1 the +eps side, 2 the -eps side;
    // the cells are ordered cl[2] <= cl[0] <= cl[1] (the lattice position is
    // monotone in x).
    int cl[3][3];
    float fl[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float xc = cur.x[k];
      uint32_t c;
      pos_fract(unit(xc), p.scale, c, fl[0][k]);
      cl[0][k] = (int)c;
      pos_fract(unit(xc + eps), p.scale, c, fl[1][k]);
      cl[1][k] = (int)c;
      pos_fract(unit(xc + -eps), p.scale, c, fl[2][k]);
      cl[2][k] = (int)c;
    }
    bool box = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) box = box && (cl[1][k] - cl[2][k] <= 1);
    if (!box) {
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        if (g[q] == 0.0f) continue;
        int cq[3];
        float fq[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          cq[k] = cl[kEvalClass[q][k]][k];
          fq[k] = fl[kEvalClass[q][k]][k];
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const uint32_t id = corner_index((uint32_t)cq[0] + (c & 1), (uint32_t)cq[1] + ((c >> 1) & 1),
                                           (uint32_t)cq[2] + ((c >> 2) & 1), p);
          atomic_add_f32(gtab + (size_t)id * F, corner_weight(c, fq) * g[q]);
        }
      }
      continue;
    }
    // the point joins the running box when all its cells lie in [blo, blo + 1]
    bool fits = mask != 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) fits = fits && cl[2][k] >= blo[k] && cl[1][k] <= blo[k] + 1;
    if (!fits) {
      if (mask) taps_flush<F>(gtab, p, blo, acc, mask);
#pragma unroll
      for (int o = 0; o < 27; ++o) acc[o] = 0.0f;
      mask = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) blo[k] = cl[2][k];
    }
    // The weights are separable: evaluation q adds g_q wx[ox] wy[oy] wz[oz] to
    // box corner (ox, oy, oz), each factor the linear weight of that coordinate
    // in the evaluation's cell (offset 0 or 1 in the box) and zero outside it.
    // Per axis only the three classes' weight vectors exist; the evaluations are
    // grouped by their x class (the centre; taps 0 and 3 on +eps; taps 1 and 2 on
    // -eps) so that each group's yz products are formed once.
    float wv[3][3][3];  // [class][axis][offset]
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const bool o0 = cl[c][k] == blo[k];
        const float f = fl[c][k], f1 = 1.0f - f;
        wv[c][k][0] = o0 ? f1 : 0.0f;
        wv[c][k][1] = o0 ? f : f1;
        wv[c][k][2] = o0 ? 0.0f : f;
      }
    float ayz[3][3][3];  // [x class][oy][oz]
#pragma unroll
    for (int oy = 0; oy < 3; ++oy) {
      const float c0 = g[0] * wv[0][1][oy];
      const float t0 = g[1] * wv[kEvalClass[1][1]][1][oy], t3 = g[4] * wv[kEvalClass[4][1]][1][oy];
      const float t1 = g[2] * wv[kEvalClass[2][1]][1][oy], t2 = g[3] * wv[kEvalClass[3][1]][1][oy];
#pragma unroll
      for (int oz = 0; oz < 3; ++oz) {
        ayz[0][oy][oz] = c0 * wv[0][2][oz];
        ayz[1][oy][oz] = fmaf(t3, wv[kEvalClass[4][2]][2][oz], t0 * wv[kEvalClass[1][2]][2][oz]);
        ayz[2][oy][oz] = fmaf(t2, wv[kEvalClass[3][2]][2][oz], t1 * wv[kEvalClass[2][2]][2][oz]);
      }
    }
#pragma unroll
    for (int oz = 0; oz < 3; ++oz)
#pragma unroll
      for (int oy = 0; oy < 3; ++oy)
#pragma unroll
        for (int ox = 0; ox < 3; ++ox) {
          const int o = (oz * 3 + oy) * 3 + ox;
          float a = fmaf(wv[0][0][ox], ayz[0][oy][oz], acc[o]);
          a = fmaf(wv[1][0][ox], ayz[1][oy][oz], a);
          acc[o] = fmaf(wv[2][0][ox], ayz[2][oy][oz], a);
        }
    // corners touched: the eight-corner pattern of each evaluation with a
    // nonzero gradient, shifted by its cell's offset in the box
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const int shift = (cl[kEvalClass[q][0]][0] - blo[0]) + 3 * (cl[kEvalClass[q][1]][1] - blo[1]) +
                        9 * (cl[kEvalClass[q][2]][2] - blo[2]);
      if (g[q] != 0.0f) mask |= 0x361Bu << shift;  // corners {0,1}^3 of a 3x3x3 box
    }
  }
  if (mask) taps_flush<F>(gtab, p, blo, acc, mask);
}

// ---------------------------------------------------------------------------
// Deterministic (sorted) table-gradient reduction.
//
// The atomic kernels above add every contribution into the gradient as it
// arrives, so the float sum order of a table entry follows the scheduling and
// the last bits vary between runs.  The sorted variant fixes the order: per
// level, contribution id = (i * E + e) * 8 + c (point i, evaluation e of E,
// corner c) gets the key "entry index within the level"; a stable radix sort
// by key lists each entry's contributions in id order, and the first thread of
// each run of equal keys sums the run in that order (fp32 fma, weight times
// gradient) and adds the sum into the entry's F gradient values.  The result
// depends only on the inputs.  Evaluations: E = 1 (one input row per point,
// na_hashgrid_bwd / na_sdf_input_bwd) or E = 5 (the centre and its four taps,
// gradients formed as hashgrid_taps_bwd_kernel forms them).  A contribution
// whose F gradient values are all zero takes key = level size: it sorts after
// every entry and is skipped, as the atomic kernels skip it.
// ---------------------------------------------------------------------------
struct SortedSrc {
  const void* dyc;  // centre (or only) rows [n, ld]
  const void* dyt;  // tap rows: point-major [4n, ld] (inc) or tap-major [4][n, ld]
  const void* dyq;  // shared row of the increment layout [n, ld], or null
  int64_t ld;
  int evals;  // 1 or 5
  int inc;
  float eps;
};

// Gradient of evaluation e of point i at columns col .. col + F - 1.
template <int F, typename TD>
__device__ inline void sorted_eval_grad(const SortedSrc& s, int64_t n, int64_t i, int e,
                                        int64_t col, float (&g)[F]) {
  const TD* dyc = (const TD*)s.dyc;
  const TD* dyt = (const TD*)s.dyt;
  const TD* dyq = (const TD*)s.dyq;
  if (s.evals == 1 || (e == 0 && !s.inc)) {
#pragma unroll
    for (int f = 0; f < F; ++f) g[f] = scalar_io<TD>::load(dyc + i * s.ld + col + f);
    return;
  }
  if (!s.inc) {
#pragma unroll
    for (int f = 0; f < F; ++f)
      g[f] = scalar_io<TD>::load(dyt + ((int64_t)(e - 1) * n + i) * s.ld + col + f);
    return;
  }
#pragma unroll
  for (int f = 0; f < F; ++f) {
    const float q = dyq ? scalar_io<TD>::load(dyq + i * s.ld + col + f) : 0.0f;
    float gt[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) gt[t] = scalar_io<TD>::load(dyt + (4 * i + t) * s.ld + col + f) + q;
    g[f] = e == 0 ? scalar_io<TD>::load(dyc + i * s.ld + col + f) - ((gt[0] + gt[1]) + (gt[2] + gt[3]))
                  : gt[e - 1];
  }
}

// Unit-cube cell and fractions of evaluation e of point i at level p.
__device__ inline void sorted_eval_cell(const float* __restrict__ x, int64_t i, int e, float eps,
                                        float vmin, float vrange, const LevelParams& p,
                                        uint32_t (&cell)[3], float (&fr)[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float xc = x[i * 3 + k];
    const float xe = e == 0 ? xc : xc + kTapDir[e - 1][k] * eps;  // as the forward built it
    pos_fract(normalize_coord_gather(xe, vmin, vrange), p.scale, cell[k], fr[k]);
  }
}

// One thread per (point, evaluation): the eight keys and ids of its corners.
template <int F, typename TD>
__global__ void __launch_bounds__(kBlock)
sorted_keys_kernel(na_grid_desc d, int l, const float* __restrict__ x, int64_t n, float vmin,
                   float vrange, SortedSrc s, uint32_t* __restrict__ keys,
                   uint32_t* __restrict__ ids) {
  const int64_t tid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (tid >= n * s.evals) return;
  const int64_t i = tid / s.evals;
  const int e = (int)(tid % s.evals);
  const LevelParams p = level_params(d, l);
  float g[F];
  sorted_eval_grad<F, TD>(s, n, i, e, (int64_t)l * F, g);
  bool any = false;
#pragma unroll
  for (int f = 0; f < F; ++f) any = any || g[f] != 0.0f;
  uint32_t id[8];
  if (any) {
    uint32_t cell[3];
    float fr[3];
    sorted_eval_cell(x, i, e, s.eps, vmin, vrange, p, cell, fr);
    cell_corners(cell, p, id);
  } else {
#pragma unroll
    for (int c = 0; c < 8; ++c) id[c] = p.size;
  }
  const uint32_t base = (uint32_t)(tid * 8);
  uint4* kp = reinterpret_cast<uint4*>(keys + tid * 8);
  uint4* ip = reinterpret_cast<uint4*>(ids + tid * 8);
  kp[0] = make_uint4(id[0], id[1], id[2], id[3]);
  kp[1] = make_uint4(id[4], id[5], id[6], id[7]);
  ip[0] = make_uint4(base, base + 1, base + 2, base + 3);
  ip[1] = make_uint4(base + 4, base + 5, base + 6, base + 7);
}

// One thread per sorted position; the first position of each run of equal keys
// sums the run in id order and adds the sums into the entry's gradient row.
template <int F, typename TD>
__global__ void __launch_bounds__(kBlock)
sorted_reduce_kernel(na_grid_desc d, int l, const float* __restrict__ x, int64_t n, float vmin,
                     float vrange, SortedSrc s, const uint32_t* __restrict__ keys,
                     const uint32_t* __restrict__ ids, int64_t count, float* __restrict__ grad) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= count) return;
  const LevelParams p = level_params(d, l);
  const uint32_t key = keys[j];
  if (key >= p.size) return;                // zero contributions sort last
  if (j > 0 && keys[j - 1] == key) return;  // not the first of its run
  float acc[F];
#pragma unroll
  for (int f = 0; f < F; ++f) acc[f] = 0.0f;
  for (int64_t m = j; m < count && keys[m] == key; ++m) {
    const uint32_t id = ids[m];
    const int c = (int)(id & 7u);
    const int64_t ie = (int64_t)(id >> 3);
    const int64_t i = ie / s.evals;
    const int e = (int)(ie % s.evals);
    float g[F];
    sorted_eval_grad<F, TD>(s, n, i, e, (int64_t)l * F, g);
    uint32_t cell[3];
    float fr[3];
    sorted_eval_cell(x, i, e, s.eps, vmin, vrange, p, cell, fr);
    const float w = corner_weight(c, fr);
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = fmaf(w, g[f], acc[f]);
  }
  float* gp = grad + ((size_t)p.off + key) * F;
#pragma unroll
  for (int f = 0; f < F; ++f) gp[f] += acc[f];
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
static int check_desc(const na_grid_desc* d) {
  NA_CHECK_ARG(d != nullptr, "desc is null");
  NA_CHECK_ARG(d->n_levels >= 1 && d->n_levels <= NA_MAX_LEVELS, "n_levels out of range");
  NA_CHECK_ARG(d->n_features == 1 || d->n_features == 2 || d->n_features == 4 ||
                   d->n_features == 8,
               "n_features must be 1, 2, 4 or 8");
  return NA_OK;
}

struct FwdLaunch {
  const na_grid_desc* d;
  const float* x;
  int64_t n;
  float vmin, vrange;
  const void* table;
  void* out;
  int64_t ld;
  int extras;
  int active;
  hipStream_t s;
};

template <int F, typename TT, typename TO>
static int launch_fwd(const FwdLaunch& a) {
  NA_CHECK_ARG(a.n <= (int64_t)0x7fffffff * kBlock / a.d->n_levels, "n too large");
  dim3 grid((unsigned)ceil_div(a.n * a.d->n_levels, (int64_t)kBlock));
  hipLaunchKernelGGL((hashgrid_fwd_kernel<F, TT, TO>), grid, dim3(kBlock), 0, a.s, *a.d, a.x, a.n,
                     a.vmin, a.vrange, (const TT*)a.table, (TO*)a.out, a.ld, a.extras, a.active);
  NA_CHECK_LAUNCH();
  return NA_OK;
}

template <int F>
static int dispatch_fwd(const FwdLaunch& a, int tdt, int odt) {
  if (tdt == NA_F32 && odt == NA_F32) return launch_fwd<F, float, float>(a);
  if (tdt == NA_F32 && odt == NA_BF16) return launch_fwd<F, float, bf16_t>(a);
  if (tdt == NA_BF16 && odt == NA_F32) return launch_fwd<F, bf16_t, float>(a);
  if (tdt == NA_BF16 && odt == NA_BF16) return launch_fwd<F, bf16_t, bf16_t>(a);
  set_error("hash-grid forward: unsupported dtype combination");
  return NA_ERR_UNSUPPORTED;
}

static int run_fwd(const FwdLaunch& a, int tdt, int odt) {
  switch (a.d->n_features) {
    case 1: return dispatch_fwd<1>(a, tdt, odt);
    case 2: return dispatch_fwd<2>(a, tdt, odt);
    case 4: return dispatch_fwd<4>(a, tdt, odt);
    default: return dispatch_fwd<8>(a, tdt, odt);
  }
}

struct BwdLaunch {
  const na_grid_desc* d;
  const float* x;
  int64_t n;
  float vmin, vrange;
  const void* dy;
  int64_t ld;
  float* grad;
  int active;
  hipStream_t s;
};

template <int F, typename TD>
static int launch_bwd(const BwdLaunch& a) {
  if (a.active <= 0) return NA_OK;
  dim3 grid;
  if (int st = chunk_level_grid(ceil_div(a.n * F, kBlock), a.active, &grid)) return st;
  hipLaunchKernelGGL((hashgrid_bwd_kernel<F, TD>), grid, dim3(kBlock), 0, a.s, *a.d, a.x, a.n,
                     a.vmin, a.vrange, (const TD*)a.dy, a.ld, a.grad, a.active);
  NA_CHECK_LAUNCH();
  return NA_OK;
}

template <int F>
static int dispatch_bwd(const BwdLaunch& a, int ddt) {
  if (ddt == NA_F32) return launch_bwd<F, float>(a);
  if (ddt == NA_BF16) return launch_bwd<F, bf16_t>(a);
  set_error("hash-grid backward: unsupported dy dtype");
  return NA_ERR_UNSUPPORTED;
}

static int run_bwd(const BwdLaunch& a, int ddt) {
  switch (a.d->n_features) {
    case 1: return dispatch_bwd<1>(a, ddt);
    case 2: return dispatch_bwd<2>(a, ddt);
    case 4: return dispatch_bwd<4>(a, ddt);
    default: return dispatch_bwd<8>(a, ddt);
  }
}

template <int F, typename TD>
static int launch_taps_bwd(const na_grid_desc* d, const float* x, int64_t n, float vmin,
                           float vrange, float eps, const void* dyc, const void* dyt,
                           const void* dyq, int64_t ld, float* grad, int active, int inc,
                           hipStream_t s) {
  // the scatter orders the cells of x - eps, x, x + eps along each axis
  NA_CHECK_ARG(eps >= 0.0f, "tap epsilon must be non-negative");
  const int run = (int)tune_value(TUNE_TAPS_BWD_RUN);
  dim3 grid;
  if (int st = chunk_level_grid(ceil_div(ceil_div(n, (int64_t)run) * F, kBlock), active, &grid))
    return st;
  uint32_t vb;
  memcpy(&vb, &vrange, 4);
  const bool p2 = (vb & 0x807fffffu) == 0u && (vb >> 23) != 0u;  // positive normal power of two
  const int vshift = p2 ? 127 - (int)(vb >> 23) : 0;
  if (p2)
    hipLaunchKernelGGL((hashgrid_taps_bwd_kernel<F, TD, true>), grid, dim3(kBlock), 0, s, *d, x, n,
                       vmin, vrange, vshift, eps, (const TD*)dyc, (const TD*)dyt, (const TD*)dyq,
                       ld, grad, inc, run, active);
  else
    hipLaunchKernelGGL((hashgrid_taps_bwd_kernel<F, TD, false>), grid, dim3(kBlock), 0, s, *d, x, n,
                       vmin, vrange, vshift, eps, (const TD*)dyc, (const TD*)dyt, (const TD*)dyq,
                       ld, grad, inc, run, active);
  NA_CHECK_LAUNCH();
  return NA_OK;
}

template <int F>
static int dispatch_taps_bwd(const na_grid_desc* d, const float* x, int64_t n, float vmin,
                             float vrange, float eps, const void* dyc, const void* dyt,
                             const void* dyq, int ddt, int64_t ld, float* grad, int active,
                             int inc, hipStream_t s) {
  if (ddt == NA_F32)
    return launch_taps_bwd<F, float>(d, x, n, vmin, vrange, eps, dyc, dyt, dyq, ld, grad, active,
                                     inc, s);
  if (ddt == NA_BF16)
    return launch_taps_bwd<F, bf16_t>(d, x, n, vmin, vrange, eps, dyc, dyt, dyq, ld, grad, active,
                                      inc, s);
  set_error("hash-grid taps backward: unsupported dy dtype");
  return NA_ERR_UNSUPPORTED;
}

// sorted reduction (sort.hip): stable radix sort of uint32 pairs on bits [0, end_bit)
int radix_sort_temp_bytes(int64_t count, size_t* bytes);
int radix_sort_pairs_u32(void* temp, size_t temp_bytes, const uint32_t* keys_in,
                         uint32_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                         int64_t count, int end_bit, hipStream_t s);

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Workspace of the sorted reduction: keys and ids twice (sort input and
// output) for n * evals * 8 contributions of one level, then the sort's own
// temporary storage.
static int sorted_workspace(int64_t n, int evals, size_t* total, size_t* arr, size_t* temp) {
  NA_CHECK_ARG(n >= 0, "n < 0");
  NA_CHECK_ARG(evals == 1 || evals == 5, "evals must be 1 or 5");
  const int64_t count = n * evals * 8;
  NA_CHECK_ARG(count <= 0x7fffffff, "n too large for the sorted reduction (n * evals * 8 >= 2^31)");
  int st = radix_sort_temp_bytes(count, temp);
  if (st) return st;
  *arr = align256((size_t)count * 4);
  *total = 4 * *arr + align256(*temp);
  return NA_OK;
}

template <int F, typename TD>
static int launch_sorted(const na_grid_desc* d, const float* x, int64_t n, float vmin,
                         float vrange, const SortedSrc& src, float* grad, int active, void* ws,
                         int64_t ws_bytes, hipStream_t s) {
  NA_CHECK_ARG(ws != nullptr, "workspace is null");
  NA_CHECK_ARG(((uintptr_t)ws & 255u) == 0, "workspace must be 256-byte aligned");
  size_t total = 0, arr = 0, temp = 0;
  int st = sorted_workspace(n, src.evals, &total, &arr, &temp);
  if (st) return st;
  NA_CHECK_ARG(ws_bytes >= (int64_t)total, "workspace smaller than na_hashgrid_sorted_workspace");
  const int64_t count = n * src.evals * 8;
  char* b = (char*)ws;
  uint32_t* k0 = (uint32_t*)b;
  uint32_t* v0 = (uint32_t*)(b + arr);
  uint32_t* k1 = (uint32_t*)(b + 2 * arr);
  uint32_t* v1 = (uint32_t*)(b + 3 * arr);
  void* tmp = b + 4 * arr;
  const dim3 gk((unsigned)ceil_div(n * src.evals, (int64_t)kBlock));
  const dim3 gr((unsigned)ceil_div(count, (int64_t)kBlock));
  for (int l = 0; l < active; ++l) {
    const uint32_t size = d->offset[l + 1] - d->offset[l];
    const int end_bit = 32 - __builtin_clz(size);  // keys run to size (the skip key)
    hipLaunchKernelGGL((sorted_keys_kernel<F, TD>), gk, dim3(kBlock), 0, s, *d, l, x, n, vmin,
                       vrange, src, k0, v0);
    NA_CHECK_LAUNCH();
    st = radix_sort_pairs_u32(tmp, temp, k0, k1, v0, v1, count, end_bit, s);
    if (st) return st;
    hipLaunchKernelGGL((sorted_reduce_kernel<F, TD>), gr, dim3(kBlock), 0, s, *d, l, x, n, vmin,
                       vrange, src, k1, v1, count, grad);
    NA_CHECK_LAUNCH();
  }
  return NA_OK;
}

static int run_sorted(const na_grid_desc* d, const float* x, int64_t n, float vmin, float vrange,
                      const SortedSrc& src, int ddt, float* grad, int active, void* ws,
                      int64_t ws_bytes, hipStream_t s) {
  if (ddt != NA_F32 && ddt != NA_BF16) {
    set_error("sorted table gradient: unsupported dy dtype");
    return NA_ERR_UNSUPPORTED;
  }
  const bool bf = ddt == NA_BF16;
#define NA_SORTED(FF)                                                                          \
  return bf ? launch_sorted<FF, bf16_t>(d, x, n, vmin, vrange, src, grad, active, ws, ws_bytes, s) \
            : launch_sorted<FF, float>(d, x, n, vmin, vrange, src, grad, active, ws, ws_bytes, s)
  switch (d->n_features) {
    case 1: NA_SORTED(1);
    case 2: NA_SORTED(2);
    case 4: NA_SORTED(4);
    default: NA_SORTED(8);
  }
#undef NA_SORTED
}

}  // namespace na

using namespace na;

extern "C" int na_hashgrid_corners(const na_grid_desc* desc, const float* x, int64_t n,
                                   int32_t* idx, float* w, na_stream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  NA_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(x && idx && w, "null pointer");
  dim3 grid((unsigned)ceil_div(n, kBlock), desc->n_levels);
  hipLaunchKernelGGL(hashgrid_corners_kernel, grid, dim3(kBlock), 0, (hipStream_t)stream,
                     *desc, x, n, idx, w);
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int na_hashgrid_fwd(const na_grid_desc* desc, const float* x, int64_t n,
                               const void* table, int32_t table_dtype, void* out,
                               int32_t out_dtype, int32_t active_levels, na_stream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  NA_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(x && table && out, "null pointer");
  NA_CHECK_ARG(n <= (int64_t)0x7fffffff * kBlock / 8, "n too large");
  FwdLaunch a{desc, x, n, 0.0f, 1.0f, table, out,
              (int64_t)desc->n_levels * desc->n_features, 0, active_levels, (hipStream_t)stream};
  return run_fwd(a, table_dtype, out_dtype);
}

extern "C" int na_hashgrid_bwd(const na_grid_desc* desc, const float* x, int64_t n,
                               const void* dy, int32_t dy_dtype, float* grad_table,
                               int32_t active_levels, na_stream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  NA_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(x && dy && grad_table, "null pointer");
  NA_CHECK_ARG(n <= (int64_t)0x7fffffff * kBlock / 8, "n too large");
  int active = active_levels < desc->n_levels ? active_levels : desc->n_levels;
  BwdLaunch a{desc, x, n, 0.0f, 1.0f, dy, (int64_t)desc->n_levels * desc->n_features,
              grad_table, active, (hipStream_t)stream};
  return run_bwd(a, dy_dtype);
}

extern "C" int na_sdf_input_fwd(const na_grid_desc* desc, const float* points, int64_t n,
                                float vmin, float vmax, const void* table, int32_t table_dtype,
                                int32_t active_levels, void* out, int32_t out_dtype, int64_t ld,
                                na_stream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  NA_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(points && table && out, "null pointer");
  NA_CHECK_ARG(vmax > vmin, "vmax must exceed vmin");
  const int64_t LF = (int64_t)desc->n_levels * desc->n_features;
  NA_CHECK_ARG(ld >= LF + (out_dtype == NA_BF16 ? 9 : 3),
               "ld must be >= L*F + 3 (fp32) or L*F + 9 (bf16)");
  NA_CHECK_ARG(ld % (out_dtype == NA_BF16 ? 8 : 4) == 0, "ld must keep rows 16-byte aligned");
  NA_CHECK_ARG(n <= (int64_t)0x7fffffff * kBlock / 8, "n too large");
  FwdLaunch a{desc, points, n, vmin, vmax - vmin, table, out, ld, 1, active_levels,
              (hipStream_t)stream};
  return run_fwd(a, table_dtype, out_dtype);
}

extern "C" int na_sdf_input_bwd(const na_grid_desc* desc, const float* points, int64_t n,
                                float vmin, float vmax, const void* dy, int32_t dy_dtype,
                                int64_t ld, float* grad_table, int32_t active_levels,
                                na_stream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  NA_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(points && dy && grad_table, "null pointer");
  NA_CHECK_ARG(vmax > vmin, "vmax must exceed vmin");
  NA_CHECK_ARG(ld >= (int64_t)desc->n_levels * desc->n_features, "ld too small");
  NA_CHECK_ARG(n <= (int64_t)0x7fffffff * kBlock / 8, "n too large");
  int active = active_levels < desc->n_levels ? active_levels : desc->n_levels;
  BwdLaunch a{desc, points, n, vmin, vmax - vmin, dy, ld, grad_table, active,
              (hipStream_t)stream};
  return run_bwd(a, dy_dtype);
}

extern "C" int na_sdf_taps_bwd(const na_grid_desc* desc, const float* points, int64_t n,
                               float vmin, float vmax, float eps, const void* dy_centre,
                               const void* dy_taps, const void* dy_shared, int32_t dy_dtype,
                               int64_t ld, float* grad_table, int32_t active_levels,
                               int32_t increments, na_stream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  NA_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(points && dy_centre && dy_taps && grad_table, "null pointer");
  NA_CHECK_ARG(increments || !dy_shared, "dy_shared belongs to the increment layout");
  NA_CHECK_ARG(vmax > vmin, "vmax must exceed vmin");
  NA_CHECK_ARG(ld >= (int64_t)desc->n_levels * desc->n_features, "ld must be >= L*F");
  NA_CHECK_ARG(n <= (int64_t)0x7fffffff * kBlock / 8, "n too large");
  const int active = active_levels < desc->n_levels ? active_levels : desc->n_levels;
  if (active <= 0) return NA_OK;
  hipStream_t s = (hipStream_t)stream;
  const float vr = vmax - vmin;
  switch (desc->n_features) {
    case 1: return dispatch_taps_bwd<1>(desc, points, n, vmin, vr, eps, dy_centre, dy_taps,
                                        dy_shared, dy_dtype, ld, grad_table, active,
                                        increments != 0, s);
    case 2: return dispatch_taps_bwd<2>(desc, points, n, vmin, vr, eps, dy_centre, dy_taps,
                                        dy_shared, dy_dtype, ld, grad_table, active,
                                        increments != 0, s);
    case 4: return dispatch_taps_bwd<4>(desc, points, n, vmin, vr, eps, dy_centre, dy_taps,
                                        dy_shared, dy_dtype, ld, grad_table, active,
                                        increments != 0, s);
    default: return dispatch_taps_bwd<8>(desc, points, n, vmin, vr, eps, dy_centre, dy_taps,
                                         dy_shared, dy_dtype, ld, grad_table, active,
                                         increments != 0, s);
  }
}

extern "C" int na_sdf_taps_increment_input(const na_grid_desc* desc, const float* points,
                                           int64_t n, float vmin, float vmax, float eps,
                                           const void* table, int32_t table_dtype,
                                           int32_t active_levels, void* x0_centre, int64_t ld,
                                           void* d_taps, int64_t ld_d, void* d_sum,
                                           na_stream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  NA_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(points && table && x0_centre && d_taps && d_sum, "null pointer");
  NA_CHECK_ARG(vmax > vmin, "vmax must exceed vmin");
  const int64_t LF = (int64_t)desc->n_levels * desc->n_features;
  NA_CHECK_ARG(ld >= LF + 9 && ld % 8 == 0, "ld must be >= L*F + 9 and a multiple of 8");
  NA_CHECK_ARG(ld_d >= ld + LF && ld_d % 8 == 0, "ld_d must be >= ld + L*F and a multiple of 8");
  NA_CHECK_ARG(n <= (int64_t)0x7fffffff * kBlock / 8, "n too large");
  NA_CHECK_ARG(n <= (int64_t)0x7fffffff * kBlock / desc->n_levels, "n too large");
  dim3 grid((unsigned)ceil_div(n * desc->n_levels, (int64_t)kBlock));
  hipStream_t s = (hipStream_t)stream;
  const float vr = vmax - vmin;
#define NA_TAPS_IN(FF, TT)                                                                   \
  hipLaunchKernelGGL((sdf_taps_input_kernel<FF, TT>), grid, dim3(kBlock), 0, s, *desc, points, \
                     n, vmin, vr, eps, (const TT*)table, (bf16_t*)x0_centre, ld,               \
                     (bf16_t*)d_taps, ld_d, (bf16_t*)d_sum, active_levels)
  if (table_dtype != NA_F32 && table_dtype != NA_BF16) {
    set_error("taps input: table must be f32 or bf16");
    return NA_ERR_UNSUPPORTED;
  }
  const bool tb = table_dtype == NA_BF16;
  switch (desc->n_features) {
    case 1: if (tb) NA_TAPS_IN(1, bf16_t); else NA_TAPS_IN(1, float); break;
    case 2: if (tb) NA_TAPS_IN(2, bf16_t); else NA_TAPS_IN(2, float); break;
    case 4: if (tb) NA_TAPS_IN(4, bf16_t); else NA_TAPS_IN(4, float); break;
    default: if (tb) NA_TAPS_IN(8, bf16_t); else NA_TAPS_IN(8, float); break;
  }
#undef NA_TAPS_IN
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int64_t na_hashgrid_sorted_workspace(const na_grid_desc* desc, int64_t n,
                                                int32_t evals) {
  int st = check_desc(desc);
  if (st) return st;
  size_t total = 0, arr = 0, temp = 0;
  st = sorted_workspace(n, evals, &total, &arr, &temp);
  if (st) return st;
  return (int64_t)total;
}

extern "C" int na_hashgrid_bwd_sorted(const na_grid_desc* desc, const float* x, int64_t n,
                                      const void* dy, int32_t dy_dtype, float* grad_table,
                                      int32_t active_levels, void* workspace,
                                      int64_t workspace_bytes, na_stream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  NA_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(x && dy && grad_table, "null pointer");
  const int active = active_levels < desc->n_levels ? active_levels : desc->n_levels;
  if (active <= 0) return NA_OK;
  SortedSrc src{dy, nullptr, nullptr, (int64_t)desc->n_levels * desc->n_features, 1, 0, 0.0f};
  return run_sorted(desc, x, n, 0.0f, 1.0f, src, dy_dtype, grad_table, active, workspace,
                    workspace_bytes, (hipStream_t)stream);
}

extern "C" int na_sdf_input_bwd_sorted(const na_grid_desc* desc, const float* points, int64_t n,
                                       float vmin, float vmax, const void* dy, int32_t dy_dtype,
                                       int64_t ld, float* grad_table, int32_t active_levels,
                                       void* workspace, int64_t workspace_bytes,
                                       na_stream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  NA_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(points && dy && grad_table, "null pointer");
  NA_CHECK_ARG(vmax > vmin, "vmax must exceed vmin");
  NA_CHECK_ARG(ld >= (int64_t)desc->n_levels * desc->n_features, "ld too small");
  const int active = active_levels < desc->n_levels ? active_levels : desc->n_levels;
  if (active <= 0) return NA_OK;
  SortedSrc src{dy, nullptr, nullptr, ld, 1, 0, 0.0f};
  return run_sorted(desc, points, n, vmin, vmax - vmin, src, dy_dtype, grad_table, active,
                    workspace, workspace_bytes, (hipStream_t)stream);
}

extern "C" int na_sdf_taps_bwd_sorted(const na_grid_desc* desc, const float* points, int64_t n,
                                      float vmin, float vmax, float eps, const void* dy_centre,
                                      const void* dy_taps, const void* dy_shared,
                                      int32_t dy_dtype, int64_t ld, float* grad_table,
                                      int32_t active_levels, int32_t increments, void* workspace,
                                      int64_t workspace_bytes, na_stream_t stream) {
  int st = check_desc(desc);
  if (st) return st;
  NA_CHECK_ARG(n >= 0, "n < 0");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(points && dy_centre && dy_taps && grad_table, "null pointer");
  NA_CHECK_ARG(increments || !dy_shared, "dy_shared belongs to the increment layout");
  NA_CHECK_ARG(vmax > vmin, "vmax must exceed vmin");
  NA_CHECK_ARG(ld >= (int64_t)desc->n_levels * desc->n_features, "ld must be >= L*F");
  const int active = active_levels < desc->n_levels ? active_levels : desc->n_levels;
  if (active <= 0) return NA_OK;
  SortedSrc src{dy_centre, dy_taps, dy_shared, ld, 5, increments != 0, eps};
  return run_sorted(desc, points, n, vmin, vmax - vmin, src, dy_dtype, grad_table, active,
                    workspace, workspace_bytes, (hipStream_t)stream);
}
