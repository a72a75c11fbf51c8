// Dense layers of the SDF / colour MLPs on CDNA4 matrix cores.
//
// One GEMM core, C[m,n] = sum_k A(m,k) B(n,k), serves the three products of a
// layer (forward Y = act(X W^T + b), backward-data dX = dPre W, backward-weight
// dW = dPre^T X) through operand-layout flags and a fused epilogue:
//   bias + activation (forward), rank-1 add + activation derivative (backward
//   chaining into the previous layer), column sums (bias gradients) and fp32
//   split-K atomics (weight gradients, reduced over millions of points).
// Tiles: 128x128x32 per 256-thread workgroup, each of the 4 waves owns a 64x64
// sub-tile computed as 2x2 MFMA 32x32 blocks (bf16: v_mfma_f32_32x32x16_bf16,
// fp32: v_mfma_f32_32x32x2_f32), operands staged through LDS.
//
// Narrow layers (N <= 4: SDF head, colour output) use a wave-per-row kernel,
// and weight normalisation (torch.nn.utils.weight_norm, dim=0) has its own
// small kernels.
#include "common.h"

#include <mutex>

namespace na {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(16))) float f32x16_t;

constexpr int GBM = 128, GBN = 128, GNT = 256;

struct EpiParams {
  int epilogue;
  void* C;
  int64_t ldc;
  int c_dtype;
  const float* bias;
  int act;
  float beta;
  const void* dact_y;
  int64_t ld_dact;
  int dact_dtype;
  int dact;
  const float* r1_a;
  const float* r1_b;
  float* colsum;
  float* z_out;
  int64_t ld_z;
  const float* zc;
  int64_t ld_zc;
  int64_t zc_rows;
  const float* head_w;
  float* head_out;
  // tap-increment backward (PK_TAPS_BWD): dL/d(tap SDF increment) per row, the
  // centre rows (read-modify-write), the shared rows and the head-weight sums
  const float* tap_dds;
  void* tap_centre;
  int64_t tap_ld_centre;
  void* tap_q;
  int64_t tap_ld_q;
  float* colsum2;
};

struct GemmParams {
  int64_t M, N, K;
  const void* A;
  int64_t lda;
  const void* B;
  int64_t ldb;
  int64_t k_chunk;
  EpiParams e;
};

__device__ inline float act_forward(int act, float v, float beta) {
  if (act == NA_ACT_RELU) return v > 0.0f ? v : 0.0f;
  if (act == NA_ACT_SOFTPLUS) {
    // torch semantics: linear above the threshold (beta v > 20), else
    // log1p(exp(beta v)) / beta, evaluated as max(v, 0) + log(1 + e^{-beta|v|}) / beta
    // with the hardware exp/log (1 + t lies in (1, 2]: no cancellation)
    const float z = v * beta;
    if (z > 20.0f) return v;
    const float t = __expf(-fabsf(z));
    return fmaxf(v, 0.0f) + __logf(1.0f + t) * (1.0f / beta);
  }
  if (act == NA_ACT_SIGMOID) return 1.0f / (1.0f + expf(-v));
  return v;
}

// Derivative of the activation expressed through its output y.
//   relu: [y > 0];  softplus: sigmoid(beta*x) = -expm1(-beta*y) (1 in the linear
//   region beta*x > 20);  sigmoid: y(1-y).
__device__ inline float act_deriv_from_y(int act, float y, float beta) {
  if (act == NA_ACT_RELU) return y > 0.0f ? 1.0f : 0.0f;
  if (act == NA_ACT_SOFTPLUS) {
    const float z = y * beta;
    return z > 20.0f ? 1.0f : 1.0f - __expf(-z);
  }
  if (act == NA_ACT_SIGMOID) return y * (1.0f - y);
  return 1.0f;
}

// softplus_beta(z0 + dz) - softplus_beta(z0) without cancellation: below the
// linear threshold it is log1p(sigmoid(beta z0) * expm1(beta dz)) / beta.  The
// small-argument branches of expm1 and log1p use their Taylor series (relative
// error < 1e-9 for |x| < 1/32), the rest the hardware exp/log.
__device__ inline float expm1_fast(float x) {
  if (fabsf(x) < 0.03125f) return x * (1.0f + x * (0.5f + x * (1.0f / 6.0f + x * (1.0f / 24.0f))));
  return __expf(x) - 1.0f;
}
__device__ inline float log1p_fast(float y) {
  if (fabsf(y) < 0.03125f)
    return y * (1.0f - y * (0.5f - y * (1.0f / 3.0f - y * (0.25f - y * 0.2f))));
  return __logf(1.0f + y);
}
__device__ inline float softplus_delta(float z0, float dz, float beta) {
  const float z1 = z0 + dz;
  const float a0 = z0 * beta, a1 = z1 * beta;
  if (a0 > 20.0f && a1 > 20.0f) return z1 - z0;
  if (a0 > 20.0f || a1 > 20.0f)
    return act_forward(NA_ACT_SOFTPLUS, z1, beta) - act_forward(NA_ACT_SOFTPLUS, z0, beta);
  const float sig = __frcp_rn(1.0f + __expf(-a0));
  return log1p_fast(sig * expm1_fast(beta * dz)) * (1.0f / beta);
}

__device__ inline float load_as_f32(const void* p, int dtype, int64_t i) {
  return dtype == NA_BF16 ? bf16_to_f32(((const bf16_t*)p)[i]) : ((const float*)p)[i];
}

// Everything of the epilogue after the bias: act, rank-1 term, derivative factor.
__device__ inline float epilogue_value_nobias(const EpiParams& e, int64_t m, int64_t n, float v) {
  if (e.act) v = act_forward(e.act, v, e.beta);
  if (e.r1_a) v += e.r1_a[m] * e.r1_b[n];
  if (e.dact) v *= act_deriv_from_y(e.dact, load_as_f32(e.dact_y, e.dact_dtype, m * e.ld_dact + n), e.beta);
  return v;
}

__device__ inline float epilogue_value(const EpiParams& e, int64_t m, int64_t n, float v) {
  if (e.bias) v += e.bias[n];
  if (e.act) v = act_forward(e.act, v, e.beta);
  if (e.r1_a) v += e.r1_a[m] * e.r1_b[n];
  if (e.dact) v *= act_deriv_from_y(e.dact, load_as_f32(e.dact_y, e.dact_dtype, m * e.ld_dact + n), e.beta);
  return v;
}

__device__ inline void epilogue_store(const EpiParams& e, int64_t m, int64_t n, float v) {
  int64_t off = m * e.ldc + n;
  if (e.epilogue == NA_EPI_ATOMIC) {
    atomic_add_f32((float*)e.C + off, v);
  } else if (e.c_dtype == NA_BF16) {
    ((bf16_t*)e.C)[off] = f32_to_bf16(v);
  } else {
    ((float*)e.C)[off] = v;
  }
}

// Per element type: K-tile depth, LDS row pitch (elements) and MFMA step.
//   bf16: BK = 32, pitch 40 (80 B rows keep 16-byte MFMA operand reads aligned)
//   fp32: BK = 16, pitch 17 (odd pitch: conflict-free single-float operand reads)
template <typename T> struct GemmCfg;
template <> struct GemmCfg<bf16_t> {
  static constexpr int BK = 32, LDK = 40;
};
template <> struct GemmCfg<float> {
  static constexpr int BK = 16, LDK = 17;
};

// Global -> register stage of one [128 x BK] operand tile, then register -> LDS
// ([row][k] layout, k contiguous).  Consecutive threads read consecutive
// 16-byte chunks of the contiguous dimension.
//   TRANS = false: element (r,k) at X[(r0+r)*ld + k0+k]   (k contiguous)
//   TRANS = true : element (r,k) at X[(k0+k)*ld + r0+r]   (r contiguous)
template <typename T, bool TRANS>
struct TileLoader {
  static constexpr int BK = GemmCfg<T>::BK;
  static constexpr int LDK = GemmCfg<T>::LDK;
  static constexpr int VEC = 16 / (int)sizeof(T);
  static constexpr int NVEC = GBM * BK / VEC / GNT;
  union Vec {
    uint4 q;
    T e[VEC];
  };
  static_assert(sizeof(T) != 2 || NVEC == 2, "bf16 transposed loads pair two k rows");
  Vec reg[NVEC];

  __device__ inline void load(const T* __restrict__ X, int64_t ld, int64_t r0, int64_t rmax,
                              int64_t k0, int64_t kend, int tid) {
#pragma unroll
    for (int j = 0; j < NVEC; ++j) {
      const int v = j * GNT + tid;
      if (!TRANS) {
        const int r = v / (BK / VEC), k = (v % (BK / VEC)) * VEC;
        const int64_t gr = r0 + r, gk = k0 + k;
        if (gr < rmax && gk < kend) {
          reg[j].q = *reinterpret_cast<const uint4*>(X + gr * ld + gk);
        } else {
          reg[j].q = make_uint4(0, 0, 0, 0);
        }
      } else if constexpr (sizeof(T) == 2) {
        // bf16: thread t owns rows 8*(t%16)..+7 at k = 2*(t/16) + j
        const int r = (tid % (GBM / VEC)) * VEC, k = 2 * (tid / (GBM / VEC)) + j;
        const int64_t gr = r0 + r, gk = k0 + k;
        if (gk < kend && gr + VEC <= rmax) {
          reg[j].q = *reinterpret_cast<const uint4*>(X + gk * ld + gr);
        } else {
#pragma unroll
          for (int e = 0; e < VEC; ++e)
            reg[j].e[e] = (gk < kend && gr + e < rmax) ? X[gk * ld + gr + e] : T(0);
        }
      } else {
        const int k = v / (GBM / VEC), r = (v % (GBM / VEC)) * VEC;
        const int64_t gr = r0 + r, gk = k0 + k;
        if (gk < kend && gr + VEC <= rmax) {
          reg[j].q = *reinterpret_cast<const uint4*>(X + gk * ld + gr);
        } else {
#pragma unroll
          for (int e = 0; e < VEC; ++e)
            reg[j].e[e] = (gk < kend && gr + e < rmax) ? X[gk * ld + gr + e] : T(0);
        }
      }
    }
  }

  __device__ inline void store(T* __restrict__ s, int tid) const {
#pragma unroll
    for (int j = 0; j < NVEC; ++j) {
      const int v = j * GNT + tid;
      if (!TRANS) {
        const int r = v / (BK / VEC), k = (v % (BK / VEC)) * VEC;
        if constexpr (sizeof(T) == 2) {
          *reinterpret_cast<uint4*>(s + r * LDK + k) = reg[j].q;
        } else {
#pragma unroll
          for (int e = 0; e < VEC; ++e) s[r * LDK + k + e] = reg[j].e[e];
        }
      } else if constexpr (sizeof(T) == 2) {
        // transpose the thread's 2 (k) x 8 (r) block in registers: one 32-bit
        // LDS store per row holds the (k, k+1) pair
        if (j == 0) {
          const int r = (tid % (GBM / VEC)) * VEC, k = 2 * (tid / (GBM / VEC));
          const uint32_t a[4] = {reg[0].q.x, reg[0].q.y, reg[0].q.z, reg[0].q.w};
          const uint32_t b[4] = {reg[1].q.x, reg[1].q.y, reg[1].q.z, reg[1].q.w};
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const uint32_t lo = (a[e] & 0xffffu) | (b[e] << 16);
            const uint32_t hi = (a[e] >> 16) | (b[e] & 0xffff0000u);
            *reinterpret_cast<uint32_t*>(s + (r + 2 * e) * LDK + k) = lo;
            *reinterpret_cast<uint32_t*>(s + (r + 2 * e + 1) * LDK + k) = hi;
          }
        }
      } else {
        const int k = v / (GBM / VEC), r = (v % (GBM / VEC)) * VEC;
#pragma unroll
        for (int e = 0; e < VEC; ++e) s[(r + e) * LDK + k] = reg[j].e[e];
      }
    }
  }
};

template <typename T> struct MfmaStep;
template <> struct MfmaStep<bf16_t> {
  static constexpr int LDK = GemmCfg<bf16_t>::LDK;
  __device__ static inline void run(const bf16_t* As, const bf16_t* Bs, int wm, int wn, int lane,
                                    f32x16_t acc[2][2]) {
#pragma unroll
    for (int kk = 0; kk < GemmCfg<bf16_t>::BK; kk += 16) {
      bf16x8_t a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = *reinterpret_cast<const bf16x8_t*>(
            As + (wm * 64 + i * 32 + (lane & 31)) * LDK + kk + (lane >> 5) * 8);
        b[i] = *reinterpret_cast<const bf16x8_t*>(
            Bs + (wn * 64 + i * 32 + (lane & 31)) * LDK + kk + (lane >> 5) * 8);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
};
template <> struct MfmaStep<float> {
  static constexpr int LDK = GemmCfg<float>::LDK;
  __device__ static inline void run(const float* As, const float* Bs, int wm, int wn, int lane,
                                    f32x16_t acc[2][2]) {
#pragma unroll
    for (int kk = 0; kk < GemmCfg<float>::BK; kk += 2) {
      float a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = As[(wm * 64 + i * 32 + (lane & 31)) * LDK + kk + (lane >> 5)];
        b[i] = Bs[(wn * 64 + i * 32 + (lane & 31)) * LDK + kk + (lane >> 5)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
};

// Epilogue shared by the GEMM kernels: per half (i), each wave stages a 32 x 64
// accumulator block through LDS (`smem`, at least 4*32*65 floats, free to
// reuse) so that every accumulator register is addressed statically and the
// global stores run along rows (64 consecutive columns per wave instruction).
// Epilogue shared by the GEMM kernels: per half (i), each wave stages a 32 x 64
// accumulator block through LDS (`smem`, at least 4*32*65 floats, free to
// reuse) so that every accumulator register is addressed statically and the
// global stores run along rows (64 consecutive columns per wave instruction).
// Rows go in batches of RB; the per-row operand of the next batch (the stored
// activation for dact, or the stored pre-activation in delta mode) is loaded
// while the current batch is processed.
__device__ __attribute__((always_inline)) inline float epi_row_operand(const EpiParams& e,
                                                                       int64_t m, int64_t n,
                                                                       bool ok) {
  if (!ok) return 0.0f;
  if (e.zc) return e.zc[(m % e.zc_rows) * e.ld_zc + n];
  if (e.dact) return load_as_f32(e.dact_y, e.dact_dtype, m * e.ld_dact + n);
  return 0.0f;
}

template <int RB, bool PIPE>
__device__ __attribute__((always_inline)) inline void gemm_epilogue(
    const GemmParams& p, unsigned char* smem, f32x16_t (&acc)[2][2], int64_t m0, int64_t n0,
    int wm, int wn, int wave, int lane) {
  constexpr int CP = 65;
  float* Cw = reinterpret_cast<float*>(smem) + wave * 32 * CP;
  const EpiParams& e = p.e;
  const int64_t n = n0 + wn * 64 + lane;
  const bool ncol = n < p.N;
  const bool head = e.head_w != nullptr;
  const bool rowop = PIPE && (e.zc || e.dact);
  const int64_t part = 2 * (2 * (int64_t)blockIdx.y + wn);  // first of the wave's two 32-column parts
  // per-column operands, loaded once
  const float hw = (head && ncol) ? e.head_w[n] : 0.0f;
  const float bias_n = (e.bias && ncol) ? e.bias[n] : 0.0f;
  const float r1b_n = (e.r1_a && ncol) ? e.r1_b[n] : 0.0f;
  float csum = 0.0f;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int64_t mb = m0 + wm * 64 + i * 32;
    float nxt[RB];
#pragma unroll
    for (int u = 0; u < RB; ++u) nxt[u] = rowop ? epi_row_operand(e, mb + u, n, ncol && mb + u < p.M) : 0.0f;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int col = j * 32 + (lane & 31);
        Cw[row * CP + col] = acc[i][j][r];
      }
    __syncthreads();
    for (int rb = 0; rb < 32; rb += RB) {
      float cur[RB], v[RB], hv[RB];
#pragma unroll
      for (int u = 0; u < RB; ++u) {
        cur[u] = nxt[u];
        v[u] = Cw[(rb + u) * CP + lane];
      }
      if (rowop && rb + RB < 32) {
#pragma unroll
        for (int u = 0; u < RB; ++u) {
          const int64_t m = mb + rb + RB + u;
          nxt[u] = epi_row_operand(e, m, n, ncol && m < p.M);
        }
      }
#pragma unroll
      for (int u = 0; u < RB; ++u) {
        const int64_t m = mb + rb + u;
        const bool ok = ncol && m < p.M;
        float x = v[u], h = 0.0f;
        if (ok) {
          if (!PIPE) cur[u] = epi_row_operand(e, m, n, true);
          if (e.zc) {  // delta mode: x is an increment of the stored pre-activation
            const float z0 = cur[u];
            h = softplus_delta(z0, x, e.beta);
            x = act_forward(NA_ACT_SOFTPLUS, z0 + x, e.beta);
          } else {
            x += bias_n;
            if (e.z_out) e.z_out[m * e.ld_z + n] = x;
            if (e.act) x = act_forward(e.act, x, e.beta);
            if (e.r1_a) x += e.r1_a[m] * r1b_n;
            if (e.dact) x *= act_deriv_from_y(e.dact, cur[u], e.beta);
            h = x;
          }
          csum += x;
          epilogue_store(e, m, n, x);
        }
        hv[u] = hw * h;
      }
      if (head) {
#pragma unroll
        for (int u = 0; u < RB; ++u) Cw[(rb + u) * CP + lane] = hv[u];
      }
    }
    if (head) {
      __syncthreads();
      // lane l reduces row l % 32 over the 32 columns of half l / 32 (fixed order)
      const int rr = lane & 31, half = lane >> 5;
      if (mb + rr < p.M) {
        float sum = 0.0f;
        for (int c = 0; c < 32; ++c) sum += Cw[rr * CP + 32 * half + c];
        e.head_out[(part + half) * p.M + mb + rr] = sum;
      }
    }
    __syncthreads();
  }
  if (e.colsum && ncol) atomic_add_f32(e.colsum + n, csum);
}

// Two LDS stages: the next K-tile is fetched into registers while the MFMAs
// consume the current stage, then written to the other stage; one barrier per
// K-tile.  The epilogue stages each wave's accumulators through LDS (reusing
// the operand buffers) so that every accumulator register is addressed
// statically and the global stores run along rows (64 consecutive columns per
// wave instruction).
template <typename T> struct GemmSmem {
  static constexpr int LDK = GemmCfg<T>::LDK;
  static constexpr int OPERAND_BYTES = 2 * (GBM + GBN) * LDK * (int)sizeof(T);
  static constexpr int CPITCH = 65;                       // floats per staged row
  static constexpr int C_BYTES = 4 * 32 * CPITCH * 4;    // 4 waves x 32 rows
  static constexpr int BYTES = OPERAND_BYTES > C_BYTES ? OPERAND_BYTES : C_BYTES;
};

template <typename T, bool TA, bool TB>
__global__ void __launch_bounds__(GNT, 3) gemm_kernel(GemmParams p) {
  constexpr int BK = GemmCfg<T>::BK;
  constexpr int LDK = GemmCfg<T>::LDK;
  __shared__ __attribute__((aligned(16))) unsigned char smem[GemmSmem<T>::BYTES];
  T* As0 = reinterpret_cast<T*>(smem);
  T* Bs0 = As0 + 2 * GBM * LDK;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t m0 = (int64_t)blockIdx.x * GBM;
  const int64_t n0 = (int64_t)blockIdx.y * GBN;
  const int64_t kbeg = (int64_t)blockIdx.z * p.k_chunk;
  const int64_t kend = kbeg + p.k_chunk < p.K ? kbeg + p.k_chunk : p.K;
  const T* A = (const T*)p.A;
  const T* B = (const T*)p.B;

  f32x16_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  TileLoader<T, TA> la;
  TileLoader<T, TB> lb;
  if (kbeg < kend) {
    la.load(A, p.lda, m0, p.M, kbeg, kend, tid);
    lb.load(B, p.ldb, n0, p.N, kbeg, kend, tid);
    la.store(As0, tid);
    lb.store(Bs0, tid);
  }
  __syncthreads();
  int buf = 0;
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    const bool more = k0 + BK < kend;
    if (more) {
      la.load(A, p.lda, m0, p.M, k0 + BK, kend, tid);
      lb.load(B, p.ldb, n0, p.N, k0 + BK, kend, tid);
    }
    MfmaStep<T>::run(As0 + buf * GBM * LDK, Bs0 + buf * GBN * LDK, wm, wn, lane, acc);
    if (more) {
      la.store(As0 + (buf ^ 1) * GBM * LDK, tid);
      lb.store(Bs0 + (buf ^ 1) * GBN * LDK, tid);
    }
    __syncthreads();
    buf ^= 1;
  }

  gemm_epilogue<4, false>(p, smem, acc, m0, n0, wm, wn, wave, lane);
}

// out[n] += sum_m X[m*ld + n] (fp32 accumulation): column sums for bias gradients.
// One thread per column, workgroups over (column block, row slab); each thread
// reduces its slab with 8 independent rows in flight, then adds one atomic, so
// every output sees M / kColSlab atomics instead of one per GEMM tile.
constexpr int kColSlab = 1024;

template <typename TX>
__global__ void __launch_bounds__(256)
colsum_kernel(const TX* __restrict__ X, int64_t ld, int64_t M, int64_t N, float* __restrict__ out) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.y * kColSlab;
  if (n >= N) return;
  const int64_t r1 = r0 + kColSlab < M ? r0 + kColSlab : M;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int64_t m = r0;
  for (; m + 8 <= r1; m += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += scalar_io<TX>::load(X + (m + u) * ld + n);
  }
  for (; m < r1; ++m) acc[0] += scalar_io<TX>::load(X + m * ld + n);
  const float s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  atomic_add_f32(out + n, s);
}

// out[m] = (bias ? bias[0] : 0) + sum_p parts[p*M + m], parts summed in order.
__global__ void __launch_bounds__(256)
head_sum_kernel(const float* __restrict__ parts, int nparts, int64_t M,
                const float* __restrict__ bias, float* __restrict__ out) {
  const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (m >= M) return;
  float v = bias ? bias[0] : 0.0f;
  int q = 0;
  for (; q + 8 <= nparts; q += 8) {  // eight loads in flight, summed in order
    float t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = parts[(q + u) * M + m];
#pragma unroll
    for (int u = 0; u < 8; ++u) v += t[u];
  }
  for (; q < nparts; ++q) v += parts[q * M + m];
  out[m] = v;
}

// ---------------------------------------------------------------------------
// narrow layers (N <= 4)
// ---------------------------------------------------------------------------
// 8 consecutive values to memory in one 16-byte (bf16) or two 16-byte (fp32) stores
__device__ inline void store8v(float* p, const float (&v)[8]) {
  *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
  *reinterpret_cast<float4*>(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
}
__device__ inline void store8v(bf16_t* p, const float (&v)[8]) {
  typedef __attribute__((ext_vector_type(8))) float f8_t;
  f8_t x;
#pragma unroll
  for (int c = 0; c < 8; ++c) x[c] = v[c];
  *reinterpret_cast<bf16x8_t*>(p) = __builtin_convertvector(x, bf16x8_t);
}

constexpr int kNarrowMax = 4;

// Vector load of 8 consecutive elements as floats (16 B of bf16 / 32 B of fp32).
__device__ inline void load8(const bf16_t* __restrict__ p, float v[8]) {
  uint4 q = *reinterpret_cast<const uint4*>(p);
  uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[2 * i] = bf16_to_f32((bf16_t)(w[i] & 0xffffu));
    v[2 * i + 1] = bf16_to_f32((bf16_t)(w[i] >> 16));
  }
}
__device__ inline void load8(const float* __restrict__ p, float v[8]) {
  float4 a = *reinterpret_cast<const float4*>(p);
  float4 b = *reinterpret_cast<const float4*>(p + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

// Eight consecutive elements of a row held as raw 16-byte chunks (one for
// bf16, two for fp32) until they are used.
// The rows stream through once, so loads and stores carry the non-temporal hint.
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4_t;
template <typename T> struct Chunk8 { uint4 q[sizeof(T) / 2]; };
template <typename T> __device__ inline void chunk_load(const T* p, Chunk8<T>& c) {
  const u32x4_t* s = reinterpret_cast<const u32x4_t*>(p);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 2); ++i) {
    const u32x4_t r = __builtin_nontemporal_load(s + i);
    c.q[i] = make_uint4(r.x, r.y, r.z, r.w);
  }
}
__device__ inline void store8v_nt(float* p, const float (&v)[8]) {
  typedef __attribute__((ext_vector_type(4))) float f4_t;
  const f4_t a = {v[0], v[1], v[2], v[3]}, b = {v[4], v[5], v[6], v[7]};
  __builtin_nontemporal_store(a, reinterpret_cast<f4_t*>(p));
  __builtin_nontemporal_store(b, reinterpret_cast<f4_t*>(p + 4));
}
__device__ inline void store8v_nt(bf16_t* p, const float (&v)[8]) {
  typedef __attribute__((ext_vector_type(8))) float f8_t;
  f8_t x;
#pragma unroll
  for (int c = 0; c < 8; ++c) x[c] = v[c];
  const bf16x8_t b = __builtin_convertvector(x, bf16x8_t);
  __builtin_nontemporal_store(__builtin_bit_cast(u32x4_t, b), reinterpret_cast<u32x4_t*>(p));
}

// Row-parallel narrow layer: a half-wave (32 lanes) owns one row; lane l covers
// k = 8*l + 256*c (c < KC) with 16/32-byte vector loads; each wave takes two
// rows per step and issues the loads of U steps before their arithmetic.
template <typename TX, int KC>
__global__ void __launch_bounds__(256)
narrow_fwd_vec_kernel(const TX* __restrict__ X, int64_t ldx, int64_t M, int64_t K,
                      const float* __restrict__ W, const float* __restrict__ b, int N, int act,
                      float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int hl = lane & 31, half = lane >> 5;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  float w[kNarrowMax][KC][8];
#pragma unroll
  for (int n = 0; n < kNarrowMax; ++n)
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t k = 8 * hl + 256 * c + e;
        w[n][c][e] = (n < N && k < K) ? W[n * K + k] : 0.0f;
      }
  constexpr int U = KC == 1 ? 4 : 2;
  const int64_t pairs = (M + 1) / 2;
  for (int64_t pr = wave; pr < pairs; pr += U * nwaves) {
    Chunk8<TX> xr[U][KC];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t m = 2 * (pr + u * nwaves) + half;
      if (m < M) {
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int64_t k = 8 * hl + 256 * c;
          if (k < K) chunk_load(X + m * ldx + k, xr[u][c]);
        }
      }
    }
    float s[U][kNarrowMax];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t m = 2 * (pr + u * nwaves) + half;
#pragma unroll
      for (int n = 0; n < kNarrowMax; ++n) s[u][n] = 0.0f;
      if (m < M) {
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int64_t k = 8 * hl + 256 * c;
          if (k < K) {
            float x[8];
            chunk_f32(xr[u][c], x);
#pragma unroll
            for (int n = 0; n < kNarrowMax; ++n)
#pragma unroll
              for (int e = 0; e < 8; ++e) s[u][n] = fmaf(x[e], w[n][c][e], s[u][n]);
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int n = 0; n < kNarrowMax; ++n)
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) s[u][n] += __shfl_xor(s[u][n], o);
    if (hl == 0) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t m = 2 * (pr + u * nwaves) + half;
        if (m < M)
          for (int n = 0; n < N; ++n) out[m * N + n] = act_forward(act, s[u][n] + b[n], 1.0f);
      }
    }
  }
}

// Workgroup size of the narrow backward (waves).
template <int NM> struct NarrowBwdCfg { static constexpr int WPG = 8; };

__device__ inline void chunk_f32(const Chunk8<bf16_t>& c, float v[8]) {
  const uint32_t w[4] = {c.q[0].x, c.q[0].y, c.q[0].z, c.q[0].w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[2 * i] = bf16_to_f32((bf16_t)(w[i] & 0xffffu));
    v[2 * i + 1] = bf16_to_f32((bf16_t)(w[i] >> 16));
  }
}
__device__ inline void chunk_f32(const Chunk8<float>& c, float v[8]) {
  const uint32_t w[8] = {c.q[0].x, c.q[0].y, c.q[0].z, c.q[0].w,
                         c.q[1].x, c.q[1].y, c.q[1].z, c.q[1].w};
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = __uint_as_float(w[i]);
}

// X_ONLY: the only row operand is X (no dX accumulation, and the activation
// derivative, if any, is taken from X itself), so a batch holds just the raw
// X chunks and more rows fit in flight.
template <typename TX, typename TDX, int KC, int NM, bool X_ONLY>
__global__ void __launch_bounds__(64 * NarrowBwdCfg<NM>::WPG)
narrow_bwd_vec_kernel(const TX* __restrict__ X, int64_t ldx, int64_t M, int64_t K,
                      const float* __restrict__ W, int N, const float* __restrict__ dpre,
                      float* __restrict__ dW, float* __restrict__ db, TDX* __restrict__ dX,
                      int64_t lddx, int add_dx, const void* Y, int64_t ldy, int y_dtype,
                      int dact, float beta, float* __restrict__ dx_colsum) {
  const int lane = threadIdx.x & 63;
  const int hl = lane & 31, half = lane >> 5;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  float w[NM][KC][8], gw[NM][KC][8], cs[KC][8];
  float gb[NM];
#pragma unroll
  for (int n = 0; n < NM; ++n) gb[n] = 0.0f;
#pragma unroll
  for (int n = 0; n < NM; ++n)
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t k = 8 * hl + 256 * c + e;
        w[n][c][e] = (n < N && k < K) ? W[n * K + k] : 0.0f;
        gw[n][c][e] = 0.0f;
        if (n == 0) cs[c][e] = 0.0f;
      }
  // Rows go in batches of NB per half-wave: all loads of a batch are issued
  // before its arithmetic.  When Y is X (the layer's own stored activation) the
  // derivative reads the loaded X values.
  // (rows per batch: what the registers of each variant hold)
  constexpr int NB = X_ONLY ? (KC == 1 ? 4 : (NM == 1 ? 4 : 1))
                            : ((KC > 1 && NM > 1) ? 1 : 2);
  const int64_t stride = 2 * nwaves;
  const bool y_is_x = dact && (const void*)X == Y && ldy == ldx &&
                      y_dtype == (sizeof(TX) == 2 ? NA_BF16 : NA_F32);
  if constexpr (X_ONLY) {
    for (int64_t mb = 2 * wave + half; mb < M; mb += stride * NB) {
      float d[NB][NM];
      Chunk8<TX> xr[NB][KC];
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int64_t m = mb + r * stride;
        const bool ok = m < M;
#pragma unroll
        for (int n = 0; n < NM; ++n) d[r][n] = (ok && n < N) ? dpre[m * N + n] : 0.0f;
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int64_t k = 8 * hl + 256 * c;
          if (ok && k < K) chunk_load(X + m * ldx + k, xr[r][c]);
        }
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        const int64_t m = mb + r * stride;
        if (m >= M) continue;
#pragma unroll
        for (int n = 0; n < NM; ++n) gb[n] += d[r][n];
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          const int64_t k = 8 * hl + 256 * c;
          if (k >= K) continue;
          float xv[8], dx[8];
          chunk_f32(xr[r][c], xv);
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            dx[e] = 0.0f;
#pragma unroll
            for (int n = 0; n < NM; ++n) {
              gw[n][c][e] = fmaf(d[r][n], xv[e], gw[n][c][e]);
              dx[e] = fmaf(d[r][n], w[n][c][e], dx[e]);
            }
          }
          if (dX) {
            if (dact) {
#pragma unroll
              for (int e = 0; e < 8; ++e) dx[e] *= act_deriv_from_y(dact, xv[e], beta);
            }
            store8v_nt(dX + m * lddx + k, dx);
#pragma unroll
            for (int e = 0; e < 8; ++e) cs[c][e] += dx[e];
          }
        }
      }
    }
  } else {
  for (int64_t mb = 2 * wave + half; mb < M; mb += stride * NB) {
    float d[NB][NM], xv[NB][KC][8], yv[NB][KC][8], ov[NB][KC][8];
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      const int64_t m = mb + r * stride;
      const bool ok = m < M;
#pragma unroll
      for (int n = 0; n < NM; ++n) d[r][n] = (ok && n < N) ? dpre[m * N + n] : 0.0f;
#pragma unroll
      for (int c = 0; c < KC; ++c) {
        const int64_t k = 8 * hl + 256 * c;
        const bool kok = ok && k < K;
#pragma unroll
        for (int e = 0; e < 8; ++e) xv[r][c][e] = yv[r][c][e] = ov[r][c][e] = 0.0f;
        if (kok) {
          load8(X + m * ldx + k, xv[r][c]);
          if (dX && add_dx) load8(dX + m * lddx + k, ov[r][c]);
          if (dX && dact && !y_is_x) {
            if (y_dtype == NA_BF16) load8((const bf16_t*)Y + m * ldy + k, yv[r][c]);
            else load8((const float*)Y + m * ldy + k, yv[r][c]);
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      const int64_t m = mb + r * stride;
      if (m >= M) continue;
#pragma unroll
      for (int n = 0; n < NM; ++n) gb[n] += d[r][n];
#pragma unroll
      for (int c = 0; c < KC; ++c) {
        const int64_t k = 8 * hl + 256 * c;
        if (k >= K) continue;
        float dx[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          dx[e] = ov[r][c][e];
#pragma unroll
          for (int n = 0; n < NM; ++n) {
            gw[n][c][e] = fmaf(d[r][n], xv[r][c][e], gw[n][c][e]);
            dx[e] = fmaf(d[r][n], w[n][c][e], dx[e]);
          }
        }
        if (dX) {
          if (dact) {
#pragma unroll
            for (int e = 0; e < 8; ++e)
              dx[e] *= act_deriv_from_y(dact, y_is_x ? xv[r][c][e] : yv[r][c][e], beta);
          }
          store8v(dX + m * lddx + k, dx);
#pragma unroll
          for (int e = 0; e < 8; ++e) cs[c][e] += dx[e];
        }
      }
    }
  }
  }
  // combine the two half-waves (same k ranges), then the waves of the
  // workgroup through LDS; one atomic per element leaves the workgroup
#pragma unroll
  for (int n = 0; n < NM; ++n) {
    gb[n] += __shfl_xor(gb[n], 32);
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
      for (int e = 0; e < 8; ++e) gw[n][c][e] += __shfl_xor(gw[n][c][e], 32);
  }
#pragma unroll
  for (int c = 0; c < KC; ++c)
#pragma unroll
    for (int e = 0; e < 8; ++e) cs[c][e] += __shfl_xor(cs[c][e], 32);
  constexpr int PER = (NM + 1) * KC * 256 + NM;  // partials of one wave
  constexpr int WPG = NarrowBwdCfg<NM>::WPG;
  __shared__ float red[(WPG - 1) * PER];
  const int wv = threadIdx.x >> 6;
  if (wv > 0 && half == 0) {
    float* r = red + (wv - 1) * PER;
#pragma unroll
    for (int n = 0; n < NM; ++n)
#pragma unroll
      for (int c = 0; c < KC; ++c)
#pragma unroll
        for (int e = 0; e < 8; ++e) r[(n * KC + c) * 256 + 8 * hl + e] = gw[n][c][e];
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
      for (int e = 0; e < 8; ++e) r[(NM * KC + c) * 256 + 8 * hl + e] = cs[c][e];
    if (hl == 0)
#pragma unroll
      for (int n = 0; n < NM; ++n) r[(NM + 1) * KC * 256 + n] = gb[n];
  }
  __syncthreads();
  if (wv == 0 && half == 0) {
#pragma unroll
    for (int q = 0; q < WPG - 1; ++q) {
      const float* r = red + q * PER;
#pragma unroll
      for (int n = 0; n < NM; ++n)
#pragma unroll
        for (int c = 0; c < KC; ++c)
#pragma unroll
          for (int e = 0; e < 8; ++e) gw[n][c][e] += r[(n * KC + c) * 256 + 8 * hl + e];
#pragma unroll
      for (int c = 0; c < KC; ++c)
#pragma unroll
        for (int e = 0; e < 8; ++e) cs[c][e] += r[(NM * KC + c) * 256 + 8 * hl + e];
#pragma unroll
      for (int n = 0; n < NM; ++n) gb[n] += r[(NM + 1) * KC * 256 + n];
    }
    for (int n = 0; n < N; ++n) {
#pragma unroll
      for (int c = 0; c < KC; ++c)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int64_t k = 8 * hl + 256 * c + e;
          if (k < K) atomic_add_f32(dW + n * K + k, gw[n][c][e]);
        }
    }
    if (dX && dx_colsum) {
#pragma unroll
      for (int c = 0; c < KC; ++c)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int64_t k = 8 * hl + 256 * c + e;
          if (k < K) atomic_add_f32(dx_colsum + k, cs[c][e]);
        }
    }
    if (hl == 0)
      for (int n = 0; n < N; ++n) atomic_add_f32(db + n, gb[n]);
  }
}

// out = dy * act'(y) with optional column sums (bias gradient of the layer)
template <typename TD, typename TY, typename TO>
__global__ void __launch_bounds__(256)
act_grad_kernel(const TD* __restrict__ dy, int64_t ld_dy, const TY* __restrict__ y, int64_t ld_y,
                int64_t M, int64_t N, int act, float beta, TO* __restrict__ out, int64_t ld_out,
                float* __restrict__ colsum) {
  // block: 256 threads = 64 columns x 4 row lanes; grid.x over row chunks, grid.y over columns
  const int64_t n = (int64_t)blockIdx.y * 64 + (threadIdx.x & 63);
  const int rl = threadIdx.x >> 6;
  const int64_t rows_per_block = 256;
  const int64_t m0 = (int64_t)blockIdx.x * rows_per_block;
  float cs = 0.0f;
  if (n < N) {
    for (int64_t m = m0 + rl; m < M && m < m0 + rows_per_block; m += 4) {
      float v = scalar_io<TD>::load(dy + m * ld_dy + n) *
                act_deriv_from_y(act, scalar_io<TY>::load(y + m * ld_y + n), beta);
      scalar_io<TO>::store(out + m * ld_out + n, v);
      cs += v;
    }
  }
  if (colsum) {
    __shared__ float red[256];
    red[threadIdx.x] = cs;
    __syncthreads();
    if (rl == 0 && n < N) {
      float s = red[threadIdx.x] + red[threadIdx.x + 64] + red[threadIdx.x + 128] + red[threadIdx.x + 192];
      atomic_add_f32(colsum + n, s);
    }
  }
}

// ---------------------------------------------------------------------------
// weight normalisation (one block per output row)
// ---------------------------------------------------------------------------
__device__ inline float block_sum_256(float v, float* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  float s = 0.0f;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  return s;
}

__global__ void __launch_bounds__(256)
weight_norm_fwd_kernel(const float* __restrict__ v, const float* __restrict__ g, int64_t K,
                       float* __restrict__ W, void* Wlow, int low_dtype, int64_t ld_low) {
  __shared__ float red[4];
  const int64_t n = blockIdx.x;
  float ss = 0.0f;
  for (int64_t k = threadIdx.x; k < K; k += blockDim.x) {
    float x = v[n * K + k];
    ss = fmaf(x, x, ss);
  }
  float norm = sqrtf(block_sum_256(ss, red));
  float scale = g[n] / norm;
  for (int64_t k = threadIdx.x; k < K; k += blockDim.x) {
    float w = v[n * K + k] * scale;
    W[n * K + k] = w;
    if (Wlow) {
      if (low_dtype == NA_BF16) ((bf16_t*)Wlow)[n * ld_low + k] = f32_to_bf16(w);
      else ((float*)Wlow)[n * ld_low + k] = w;
    }
  }
  if (Wlow) {  // zero the padding columns of the low-precision copy
    for (int64_t k = K + threadIdx.x; k < ld_low; k += blockDim.x) {
      if (low_dtype == NA_BF16) ((bf16_t*)Wlow)[n * ld_low + k] = 0;
      else ((float*)Wlow)[n * ld_low + k] = 0.0f;
    }
  }
}

__global__ void __launch_bounds__(256)
weight_norm_bwd_kernel(const float* __restrict__ v, const float* __restrict__ g,
                       const float* __restrict__ dW, int64_t K, float* __restrict__ dv,
                       float* __restrict__ dg) {
  __shared__ float red[4];
  const int64_t n = blockIdx.x;
  float ss = 0.0f, dot = 0.0f;
  for (int64_t k = threadIdx.x; k < K; k += blockDim.x) {
    float x = v[n * K + k];
    ss = fmaf(x, x, ss);
    dot = fmaf(x, dW[n * K + k], dot);
  }
  float norm = sqrtf(block_sum_256(ss, red));
  float d = block_sum_256(dot, red);
  float dgn = d / norm;
  if (threadIdx.x == 0) dg[n] = dgn;
  float s = g[n] / norm;
  for (int64_t k = threadIdx.x; k < K; k += blockDim.x) {
    dv[n * K + k] = s * (dW[n * K + k] - dgn * v[n * K + k] / norm);
  }
}

// ---------------------------------------------------------------------------
// host dispatch
// ---------------------------------------------------------------------------
// bf16 GEMM, both operands k-contiguous (A[m*lda + k], B[n*ldb + k]): a
// four-stage pipeline of direct global -> LDS loads (global_load_lds_dwordx4,
// no register staging).  Stage t+3 is in flight while the MFMAs consume
// stage t, which hides HBM latency behind three K-tiles of MFMA work.
//
// Stage layout: [128 rows][32 k] bf16 per operand, 64-byte rows, with the four
// 16-byte chunks of row r stored at chunk slot c ^ ((r >> 2) & 3).  One wave
// load instruction fills 1 KB (16 rows); lane l lands at slot l & 3 of row
// l >> 2 and fetches the global chunk that belongs there.  The MFMA operand
// reads (16 B per lane, 32 rows per half-wave) then touch 16 distinct bank
// groups per 16 lanes.  Chunks past M/N or K read a zero block instead.
// ---------------------------------------------------------------------------
constexpr int kNNStages = 4;
constexpr int kNNBK = 32;
constexpr int kNNStageElems = GBM * kNNBK;  // per operand

__device__ uint4 gemm_zero_chunks[64];  // zero source for out-of-range chunks

typedef __attribute__((address_space(3))) void lds_void_t;

__device__ inline void nn_issue_stage(const bf16_t* __restrict__ X, int64_t ld, int64_t r0,
                                      int64_t rmax, int64_t k0, int64_t kend, bf16_t* stage,
                                      int wave, int lane) {
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int q = wave * 2 + u;            // 16-row block of the stage
    const int row = q * 16 + (lane >> 2);
    const int chunk = (lane & 3) ^ ((row >> 2) & 3);
    const int64_t gr = r0 + row, gk = k0 + chunk * 8;
    const void* src = (gr < rmax && gk < kend) ? (const void*)(X + gr * ld + gk)
                                               : (const void*)(gemm_zero_chunks + lane);
    __builtin_amdgcn_global_load_lds(src, (lds_void_t*)(stage + q * 16 * kNNBK), 16, 0, 0);
  }
}

// Wait until at most `pending` of this wave's stage loads are outstanding
// (4 load instructions per stage).  vmcnt is an immediate operand.
__device__ inline void nn_wait_stages(int pending) {
  // gfx9 s_waitcnt immediate: vmcnt[3:0], expcnt[6:4] = 7, lgkmcnt[11:8] = 15
  if (pending >= 2) __builtin_amdgcn_s_waitcnt(0xF70 | 8);
  else if (pending == 1) __builtin_amdgcn_s_waitcnt(0xF70 | 4);
  else __builtin_amdgcn_s_waitcnt(0xF70);
}

__global__ void __launch_bounds__(GNT)
gemm_bf16_nn_kernel(GemmParams p) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[kNNStages * 2 * kNNStageElems * 2];
  bf16_t* st0 = reinterpret_cast<bf16_t*>(smem);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t m0 = (int64_t)blockIdx.x * GBM;
  const int64_t n0 = (int64_t)blockIdx.y * GBN;
  const int64_t kbeg = (int64_t)blockIdx.z * p.k_chunk;
  const int64_t kend = kbeg + p.k_chunk < p.K ? kbeg + p.k_chunk : p.K;
  const bf16_t* A = (const bf16_t*)p.A;
  const bf16_t* B = (const bf16_t*)p.B;
  const int nk = kbeg < kend ? (int)((kend - kbeg + kNNBK - 1) / kNNBK) : 0;

  f32x16_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  auto issue = [&](int t) {
    bf16_t* sa = st0 + (t % kNNStages) * 2 * kNNStageElems;
    const int64_t k0 = kbeg + (int64_t)t * kNNBK;
    nn_issue_stage(A, p.lda, m0, p.M, k0, kend, sa, wave, lane);
    nn_issue_stage(B, p.ldb, n0, p.N, k0, kend, sa + kNNStageElems, wave, lane);
  };
#pragma unroll
  for (int t = 0; t < kNNStages - 1; ++t)
    if (t < nk) issue(t);

  for (int t = 0; t < nk; ++t) {
    const int ahead = nk - 1 - t;  // stages issued after t, capped by the pipeline depth
    nn_wait_stages(ahead < kNNStages - 2 ? ahead : kNNStages - 2);
    __builtin_amdgcn_s_barrier();   // stage t complete in every wave; stage t-1 consumed
    if (t + kNNStages - 1 < nk) issue(t + kNNStages - 1);
    const bf16_t* sa = st0 + (t % kNNStages) * 2 * kNNStageElems;
    const bf16_t* sb = sa + kNNStageElems;
#pragma unroll
    for (int kk = 0; kk < kNNBK; kk += 16) {
      bf16x8_t a[2], b[2];
      const int ck = (kk >> 3) + (lane >> 5);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int ra = wm * 64 + i * 32 + (lane & 31);
        const int rb = wn * 64 + i * 32 + (lane & 31);
        a[i] = *reinterpret_cast<const bf16x8_t*>(sa + ra * kNNBK + ((ck ^ ((ra >> 2) & 3)) << 3));
        b[i] = *reinterpret_cast<const bf16x8_t*>(sb + rb * kNNBK + ((ck ^ ((rb >> 2) & 3)) << 3));
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  __builtin_amdgcn_s_waitcnt(0xF70);
  __syncthreads();  // all operand reads done before the epilogue reuses the stages
  gemm_epilogue<8, true>(p, smem, acc, m0, n0, wm, wn, wave, lane);
}

// ---------------------------------------------------------------------------
// Persistent BF16 NN GEMM for tall-skinny products (M >> N, K of a few
// hundred): C = epi(A[M,K] . B[N,K]^T) with the store epilogue.  Each
// workgroup walks the tiles blockIdx.x, blockIdx.x + gridDim.x, ... and its
// direct-to-LDS load pipeline (NS stages of 32-wide K-slices, same swizzled
// layout as above) runs across tile boundaries: the first K-slices of the next
// tile are in flight while the current tile's epilogue runs.  A tile is 128
// rows x BN = 64*WN columns (BN = 256 covers a 256-wide layer, so A is read
// from HBM once); 2 x WN waves each own 64 x 64 outputs in 32x32x16 MFMA
// accumulators.
//
// Epilogue: each wave stages one 32 x 32 accumulator block at a time in a
// private LDS region (pitch 36 floats = 4 x odd, so the 16 rows read by a
// 16-lane group fall in distinct banks) and reads it back as 8 consecutive
// columns of one row per lane; every global load / store of the epilogue then
// moves 16 bytes, and bf16 results use the hardware pack-convert (RNE).
// ---------------------------------------------------------------------------
constexpr int kPEP = 36;
constexpr int kPEpiFloats = 32 * kPEP;

template <int WN> struct PersistTile {
  static constexpr int NW = 2 * WN, BN = 64 * WN, ROWS = GBM + BN;
  static constexpr int NS = WN == 4 ? 5 : 3;                 // 1 resp. 2 workgroups per CU
  static constexpr int STAGE_EL = ROWS * kNNBK;
  static constexpr int LOADS = ROWS / 16 / NW;               // 1 KB loads per wave per stage
  static constexpr int LDS_BYTES = NS * STAGE_EL * 2 + NW * kPEpiFloats * 4;
};

typedef __attribute__((ext_vector_type(8))) float f32x8_t;

// Stage loads: one 1 KB buffer-to-LDS load per 16 rows.  Lane l lands at slot
// l & 3 of row l >> 2 and fetches the chunk that belongs there; its byte offset
// within the tile (row * ld * 2 + chunk * 16) is fixed for the whole kernel.
// The buffer's record count ends at the last row of the matrix, so rows past
// M (or N) read zeros; K-tail chunks get an offset past the record count.
constexpr uint32_t kOobOffset = 0x80000000u;

// Buffer of one stage: the tile's first row starts at `base` and `ext` bytes of
// the matrix follow it; the stage starts kb bytes further along the rows.
__device__ inline __amdgpu_buffer_rsrc_t stage_rsrc(const char* base, int64_t ext, int64_t kb) {
  const int64_t bytes = ext - kb;
  const int n = bytes > 0x7fffffffLL ? 0x7fffffff : (bytes > 0 ? (int)bytes : 0);
  return __builtin_amdgcn_make_buffer_rsrc((void*)(base + kb), 0, n, 0x00020000);
}

template <int WN>
__device__ inline void persist_issue(const __amdgpu_buffer_rsrc_t ra, const __amdgpu_buffer_rsrc_t rb,
                                     const uint32_t (&voff)[PersistTile<WN>::LOADS],
                                     const int (&kofs)[PersistTile<WN>::LOADS], int krem,
                                     bf16_t* stage, int wave) {
  using PT = PersistTile<WN>;
#pragma unroll
  for (int u = 0; u < PT::LOADS; ++u) {
    const int q = wave * PT::LOADS + u;  // 16-row block of the stage: A rows, then B rows
    const uint32_t off = kofs[u] < krem ? voff[u] : kOobOffset;
    __builtin_amdgcn_raw_ptr_buffer_load_lds(q < GBM / 16 ? ra : rb,
                                             (lds_void_t*)(stage + q * 16 * kNNBK), 16, off, 0, 0, 0);
  }
}

// Wait until at most `ahead` stages (LOADS instructions each) issued after the
// one needed now are outstanding.  vmcnt is an immediate operand.
template <int LOADS>
__device__ inline void persist_wait(int ahead) {
  switch (ahead) {
    case 0: __builtin_amdgcn_s_waitcnt(0xF70); break;
    case 1: __builtin_amdgcn_s_waitcnt(0xF70 | (1 * LOADS)); break;
    case 2: __builtin_amdgcn_s_waitcnt(0xF70 | (2 * LOADS)); break;
    default: __builtin_amdgcn_s_waitcnt(0xF70 | (3 * LOADS)); break;
  }
}

__device__ inline void vload8(const float* p, float (&v)[8]) {
  const float4 a = *reinterpret_cast<const float4*>(p), b = *reinterpret_cast<const float4*>(p + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ inline void vload8(const bf16_t* p, float (&v)[8]) {
  const uint4 a = *reinterpret_cast<const uint4*>(p);
  const uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    v[2 * c] = __uint_as_float(w[c] << 16);
    v[2 * c + 1] = __uint_as_float(w[c] & 0xffff0000u);
  }
}
__device__ inline void vstore8(float* p, const float (&v)[8]) {
  *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
  *reinterpret_cast<float4*>(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
}
__device__ inline void vstore8(bf16_t* p, const float (&v)[8]) {
  f32x8_t x;
#pragma unroll
  for (int c = 0; c < 8; ++c) x[c] = v[c];
  *reinterpret_cast<bf16x8_t*>(p) = __builtin_convertvector(x, bf16x8_t);
}
// 8 columns of a row: vector access when FULL, else only the first `nc`
template <bool FULL, typename T>
__device__ inline void load_cols(const T* p, int nc, float (&v)[8]) {
  if (FULL) { vload8(p, v); return; }
#pragma unroll
  for (int c = 0; c < 8; ++c) v[c] = c < nc ? scalar_io<T>::load(p + c) : 0.0f;
}
template <bool FULL, typename T>
__device__ inline void store_cols(T* p, int nc, const float (&v)[8]) {
  if (FULL) { vstore8(p, v); return; }
#pragma unroll
  for (int c = 0; c < 8; ++c)
    if (c < nc) scalar_io<T>::store(p + c, v[c]);
}

// Epilogue kinds of the persistent kernel (the products of the SDF and colour
// MLPs); other combinations run on the general kernel.
enum PersistKind : int {
  PK_PLAIN = 0,         // C = acc
  PK_BIAS_RELU = 1,     // C = relu(acc + b)
  PK_BIAS_SOFTPLUS = 2, // C = softplus(acc + b) [z_out = acc + b] [head]
  PK_DACT = 3,          // C = (acc [+ r1_a r1_b]) * act'(y), relu or softplus  [colsum]
  PK_DELTA = 4,         // C = softplus(zc + acc), head over softplus increments
  // Tap increments of the 4-tap numerical gradient, rows point-major (row m is tap
  // m % 4 of point m / 4, whose stored centre pre-activation is zc row m / 4):
  PK_TAPS_FWD = 5,      // head over softplus(z + dz) - softplus(z) only (no C)
  PK_TAPS_BWD = 6,      // backward of the increments: C = tap rows, centre +=, shared rows, sums
};

// softplus_beta with a select for the linear region (no branch); ib = 1 / beta
__device__ inline float softplus_sel(float v, float beta, float ib) {
  const float z = v * beta;
  const float r = fmaxf(v, 0.0f) + __logf(1.0f + __expf(-fabsf(z))) * ib;
  return z > 20.0f ? v : r;
}

// Increment h = softplus(z0 + dz) - softplus(z0) and value y1 = softplus(z0 + dz)
// for softplus(z) = ln(1 + exp(b z)) / b.  With a0 = b z0, u = b dz and the two
// factors t0 = exp(-|a0|), e = exp(-|u|) in (0, 1], taking the larger exponent
// out of ln(1 + exp(a0 + u)) in each sign case gives
//   b y1 = max(a0, 0) + max(u, 0) + ln P,   P = 1 + t0 e  (z0, dz of one sign)
//                                           P = t0 + e    (opposite signs)
//   b h  = max(u, 0) + ln P - ln(1 + t0).
// P lies in [max(t0, e), 2], so nothing overflows, and the difference of the
// two logarithms (both of numbers near 1 when dz is small) leaves h an absolute
// error of a few float32 ulps of 1 / b at any step size: four transcendentals
// and no data-dependent branch.  P can underflow only when both |a0| and |u|
// exceed 64; those elements take the direct difference of the two values.  The
// smooth softplus differs from the thresholded one (linear for b z > 20) by at
// most log1p(exp(-20)) / b, far below the float32 spacing of those values.
__device__ inline void softplus_step(float z0, float dz, float beta, float ib, float& h, float& y1) {
  constexpr float kLog2e = 1.4426950408889634f, kLn2 = 0.6931471805599453f;
  const float a0 = z0 * beta, u = dz * beta;
  const float t0 = __builtin_amdgcn_exp2f(-fabsf(a0) * kLog2e);
  const float e = __builtin_amdgcn_exp2f(-fabsf(u) * kLog2e);
  const bool same = (z0 >= 0.0f) == (dz >= 0.0f);
  const float P = same ? fmaf(t0, e, 1.0f) : t0 + e;
  const float lP = __builtin_amdgcn_logf(P), l0 = __builtin_amdgcn_logf(1.0f + t0);  // log2
  const float s = kLn2 * ib, mu = fmaxf(dz, 0.0f);
  h = fmaf(lP - l0, s, mu);
  y1 = fmaf(lP, s, fmaxf(z0, 0.0f) + mu);
  if (__builtin_expect(fminf(fabsf(a0), fabsf(u)) > 64.0f, 0)) {
    const float z1 = z0 + dz;
    const float t1 = __builtin_amdgcn_exp2f(-fabsf(z1 * beta) * kLog2e);
    y1 = fmaf(__builtin_amdgcn_logf(1.0f + t1), s, fmaxf(z1, 0.0f));
    h = y1 - fmaf(l0, s, fmaxf(z0, 0.0f));
  }
}

// 1 - exp(-u) without cancellation: a Taylor series below |u| = 1/16 (relative
// error under 1e-9), the exponential above (relative error a few ulps of 1/16).
__device__ inline float one_minus_exp_neg(float u) {
  if (fabsf(u) < 0.0625f)
    return u * (1.0f - 0.5f * u * (1.0f - (1.0f / 3.0f) * u * (1.0f - 0.25f * u * (1.0f - 0.2f * u))));
  return 1.0f - __expf(-u);
}

// Derivatives of softplus_beta: s1 = sp'(z0 + dz) = sigmoid(b (z0 + dz)), s0 = sp'(z0)
// and the increment ds = sp'(z0 + dz) - sp'(z0) = sigmoid(a + u) sigmoid(-a) (1 - e^-u)
// (a = b z0, u = b dz): a product of factors, so it keeps float32 precision relative
// to its own size however small dz is.
__device__ inline void softplus_deriv_step(float z0, float dz, float beta, float& s1, float& ds,
                                           float& s0) {
  const float a = z0 * beta, u = dz * beta;
  s1 = 1.0f / (1.0f + __expf(-(a + u)));
  const float sna = 1.0f / (1.0f + __expf(a));
  ds = s1 * sna * one_minus_exp_neg(u);
  s0 = 1.0f - sna;  // sp'(z0)
}

// Branch-free forms for the tap backward epilogue (sixteen elements per lane in
// flight; a data-dependent branch per element splits its schedule).
// Increment h = softplus(z0 + dz) - softplus(z0) as softplus_step forms it, with
// ln P = ln(1 + e^{-(|a0| + |u|)}) (same signs) or -m + ln(1 + e^{-(M - m)})
// (opposite signs; m, M the smaller and larger of |a0|, |u|): the larger
// exponent is taken out first, so P cannot underflow and no slow path is needed.
__device__ inline float softplus_incr_nb(float z0, float dz, float beta, float ib) {
  constexpr float kLog2e = 1.4426950408889634f, kLn2 = 0.6931471805599453f;
  const float a0 = fabsf(z0 * beta), u = fabsf(dz * beta);
  const bool same = (z0 >= 0.0f) == (dz >= 0.0f);
  const float m = fminf(a0, u);
  const float sp = same ? a0 + u : fmaxf(a0, u) - m;
  const float off = same ? 0.0f : m * kLog2e;
  const float lP = __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(-sp * kLog2e)) - off;
  const float l0 = __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(-a0 * kLog2e));
  return fmaf(lP - l0, kLn2 * ib, fmaxf(dz, 0.0f));
}

// 1 - exp(-u) as one_minus_exp_neg, both branches evaluated and selected.
__device__ inline float one_minus_exp_neg_nb(float u) {
  const float t = u * (1.0f - 0.5f * u * (1.0f - (1.0f / 3.0f) * u * (1.0f - 0.25f * u * (1.0f - 0.2f * u))));
  const float x = 1.0f - __expf(-u);
  return fabsf(u) < 0.0625f ? t : x;
}

__device__ inline void softplus_deriv_step_nb(float z0, float dz, float beta, float& s1, float& ds,
                                              float& s0) {
  const float a = z0 * beta, u = dz * beta;
  s1 = 1.0f / (1.0f + __expf(-(a + u)));
  const float sna = 1.0f / (1.0f + __expf(a));
  ds = s1 * sna * one_minus_exp_neg_nb(u);
  s0 = 1.0f - sna;
}

// Backward terms of one (point, column) of the tap-increment layer, from the
// centre pre-activation z and the four taps' increments dz_t, with the terms of z
// alone formed once for the four taps.  With a = b z, u_t = b dz_t (b = beta),
// t0 = e^{-|a|}, e_t = e^{-|u_t|} and p_t = e^{-|a + u_t|}:
//   sp'(z) = sigmoid(a), sigmoid(-a)            one reciprocal of 1 + t0
//   sp'(z + dz_t) = sigmoid(a + u_t)            one reciprocal of 1 + p_t
//   b h_t = max(u_t, 0) + ln(1 + p_t) - [opposite signs] min(|a|, |u_t|)
//           - ln(1 + t0)                        (softplus_step's P, with the
//                                               larger exponent taken out: no
//                                               underflow, no slow path)
//   dsp'_t = sigmoid(a + u_t) sigmoid(-a) (1 - e^{-u_t}), where 1 - e^{-u} is
//           1 - e_t (a series below |u| = 1/16) for u >= 0 and -(1 - e_t) / e_t
//           below zero: a product of factors, precise relative to its own size.
// Outputs: tap row values o_t = w (b_t sp'(z + dz_t) + am dsp'_t) (b_t = dv_t -
// am), ev = sum_t dv_t w dsp'_t, hs += sum_t dv_t h_t (in tap order), and sp'(z).
__device__ inline void tap_bwd_column(float z, const float (&dz)[4], const float (&dv)[4], float am,
                                      float w, float beta, float ib, float (&o)[4], float& ev,
                                      float& hs, float& s0) {
  constexpr float kLog2e = 1.4426950408889634f, kLn2 = 0.6931471805599453f;
  const float za = z * beta, aa = fabsf(za);
  const float t0 = __builtin_amdgcn_exp2f(-aa * kLog2e);
  const float l0 = __builtin_amdgcn_logf(1.0f + t0);
  const float r0 = __builtin_amdgcn_rcpf(1.0f + t0);
  s0 = za >= 0.0f ? r0 : t0 * r0;
  const float sna = za >= 0.0f ? t0 * r0 : r0;
  const float sc = kLn2 * ib;
  ev = 0.0f;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const float u = dz[t] * beta, au = fabsf(u), v = za + u;
    const float e = __builtin_amdgcn_exp2f(-au * kLog2e);
    const float pp = __builtin_amdgcn_exp2f(-fabsf(v) * kLog2e);
    const bool same = (z >= 0.0f) == (dz[t] >= 0.0f);
    const float off = same ? 0.0f : fminf(aa, au) * kLog2e;
    const float hv = fmaf(__builtin_amdgcn_logf(1.0f + pp) - off - l0, sc, fmaxf(dz[t], 0.0f));
    const float r1 = __builtin_amdgcn_rcpf(1.0f + pp);
    const float s1 = v >= 0.0f ? r1 : pp * r1;
    const float ser = au * (1.0f - 0.5f * au * (1.0f - (1.0f / 3.0f) * au *
                                                  (1.0f - 0.25f * au * (1.0f - 0.2f * au))));
    const float om = au < 0.0625f ? ser : 1.0f - e;
    const float ome = u >= 0.0f ? om : -om * __builtin_amdgcn_rcpf(e);
    const float dsp = s1 * sna * ome;
    o[t] = w * fmaf(dv[t] - am, s1, am * dsp);
    ev = fmaf(dv[t] * w, dsp, ev);
    hs = fmaf(dv[t], hv, hs);
  }
}

// One 32 x 32 block of the wave, staged in eb: lane (rl, cg) handles columns
// nb .. nb+7 of rows mrow0 + rl and mrow0 + 16 + rl.  cs collects column sums,
// hs the head partial of each of the two rows.
template <int KIND, bool FULL>
__device__ __attribute__((always_inline)) inline void persist_rows(
    const GemmParams& p, const float* eb, int64_t mrow0, int64_t nb, int nc, int rl, int cg,
    const float (&cb)[8], const float (&hw8)[8], float (&hs)[2], float (&cs)[8], int64_t zrow0,
    float ib) {
  const EpiParams& e = p.e;
#pragma unroll
  for (int hh = 0; hh < 2; ++hh) {
    const int rr = hh * 16 + rl;
    const int64_t m = mrow0 + rr;
    float x[8];
    {
      const float4 a = *reinterpret_cast<const float4*>(eb + rr * kPEP + cg * 8);
      const float4 b = *reinterpret_cast<const float4*>(eb + rr * kPEP + cg * 8 + 4);
      x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w; x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
    }
    if (!FULL && (m >= p.M || nc <= 0)) continue;
    float hv[8];
    if constexpr (KIND == PK_BIAS_RELU || KIND == PK_BIAS_SOFTPLUS) {
#pragma unroll
      for (int c = 0; c < 8; ++c) x[c] += cb[c];
      if (KIND == PK_BIAS_SOFTPLUS && e.z_out) store_cols<FULL>(e.z_out + m * e.ld_z + nb, nc, x);
#pragma unroll
      for (int c = 0; c < 8; ++c)
        x[c] = KIND == PK_BIAS_RELU ? fmaxf(x[c], 0.0f) : softplus_sel(x[c], e.beta, ib);
#pragma unroll
      for (int c = 0; c < 8; ++c) hv[c] = x[c];
    } else if constexpr (KIND == PK_DACT) {
      float y[8];
      if (e.dact_dtype == NA_BF16)
        load_cols<FULL>((const bf16_t*)e.dact_y + m * e.ld_dact + nb, nc, y);
      else
        load_cols<FULL>((const float*)e.dact_y + m * e.ld_dact + nb, nc, y);
      if (e.r1_a) {
        const float ra = e.r1_a[m];
#pragma unroll
        for (int c = 0; c < 8; ++c) x[c] = fmaf(ra, cb[c], x[c]);  // cb holds r1_b here
      }
      if (e.dact == NA_ACT_RELU) {
#pragma unroll
        for (int c = 0; c < 8; ++c) x[c] = y[c] > 0.0f ? x[c] : 0.0f;
      } else {
#pragma unroll
        for (int c = 0; c < 8; ++c) x[c] *= act_deriv_from_y(NA_ACT_SOFTPLUS, y[c], e.beta);
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) cs[c] += x[c];  // columns past N hold 0
    } else if constexpr (KIND == PK_DELTA) {
      // row m reads stored row m mod zc_rows; zrow0 + rr is that index before the reduction
      const int64_t off = zrow0 + rr;
      const int64_t mz = e.zc_rows >= GBM ? (off >= e.zc_rows ? off - e.zc_rows : off)
                                          : (int64_t)((uint32_t)m % (uint32_t)e.zc_rows);
      float z0[8];
      load_cols<FULL>(e.zc + mz * e.ld_zc + nb, nc, z0);
#pragma unroll
      for (int c = 0; c < 8; ++c) softplus_step(z0[c], x[c], e.beta, ib, hv[c], x[c]);
    }
    if (e.c_dtype == NA_BF16) store_cols<FULL>((bf16_t*)e.C + m * e.ldc + nb, nc, x);
    else store_cols<FULL>((float*)e.C + m * e.ldc + nb, nc, x);
    if constexpr (KIND == PK_BIAS_SOFTPLUS || KIND == PK_DELTA) {
      if (e.head_w) {
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) s += hw8[c] * hv[c];
        hs[hh] += s;
      }
    }
  }
}

__device__ __attribute__((always_inline)) inline void persist_stage(float* eb, const f32x16_t& a, int lane) {
#pragma unroll
  for (int r = 0; r < 16; ++r) eb[(8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * kPEP + (lane & 31)] = a[r];
}

// Epilogue of a wave tile that lies wholly inside C (64 x 64).  Phase 1 asks
// for every per-element operand of the four 32 x 32 blocks (zc rows, bf16 or
// fp32 activations, r1_a) and the column vectors, then stages, reads and
// transforms the blocks into registers; phase 2 issues all stores.  Vector
// memory operations complete in issue order, so a staging read placed after one
// of this wave's stores would wait for that store: here none is.  Lane (rl, cg)
// owns rows mw + 32 i + 16 hh + rl and columns nw + 32 j + 8 cg .. + 7.
template <int WN, int KIND>
__device__ __attribute__((always_inline)) inline void persist_epilogue_full(
    const GemmParams& p, float* eb, f32x16_t (&acc)[2][2], int64_t mw, int64_t nw, int lane,
    float (&csum)[2][8], int64_t zr0, float ib, int part) {
  const EpiParams& e = p.e;
  const int rl = lane & 15, cg = lane >> 4;
  constexpr bool kHead = KIND == PK_BIAS_SOFTPLUS || KIND == PK_DELTA;
  constexpr bool kOpd = KIND == PK_DELTA || KIND == PK_DACT;
  float opd[kOpd ? 4 : 1][2][8];
  float ra[2][2];
  if constexpr (KIND == PK_DELTA) {
#pragma unroll
    for (int blk = 0; blk < 4; ++blk)
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const int64_t off = zr0 + (blk >> 1) * 32 + hh * 16 + rl;
        const int64_t m = mw + (blk >> 1) * 32 + hh * 16 + rl;
        const int64_t mz = e.zc_rows >= GBM ? (off >= e.zc_rows ? off - e.zc_rows : off)
                                            : (int64_t)((uint32_t)m % (uint32_t)e.zc_rows);
        vload8(e.zc + mz * e.ld_zc + nw + (blk & 1) * 32 + cg * 8, opd[blk][hh]);
      }
  } else if constexpr (KIND == PK_DACT) {
#pragma unroll
    for (int blk = 0; blk < 4; ++blk)
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const int64_t m = mw + (blk >> 1) * 32 + hh * 16 + rl;
        const int64_t o = m * e.ld_dact + nw + (blk & 1) * 32 + cg * 8;
        if (e.dact_dtype == NA_BF16) vload8((const bf16_t*)e.dact_y + o, opd[blk][hh]);
        else vload8((const float*)e.dact_y + o, opd[blk][hh]);
      }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) ra[i][hh] = e.r1_a ? e.r1_a[mw + i * 32 + hh * 16 + rl] : 0.0f;
  }
  float cb[2][8], hw[2][8];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int64_t nb = nw + j * 32 + cg * 8;
#pragma unroll
    for (int c = 0; c < 8; ++c) cb[j][c] = hw[j][c] = 0.0f;
    if constexpr (KIND == PK_BIAS_RELU || KIND == PK_BIAS_SOFTPLUS) vload8(e.bias + nb, cb[j]);
    if constexpr (KIND == PK_DACT) { if (e.r1_a) vload8(e.r1_b + nb, cb[j]); }
    if constexpr (kHead) { if (e.head_w) vload8(e.head_w + nb, hw[j]); }
  }
  float oc[4][2][8];
  float oz[KIND == PK_BIAS_SOFTPLUS ? 4 : 1][2][8];
  float hs[2][2][2] = {};  // [row block][column block][half row] head partials
#pragma unroll
  for (int blk = 0; blk < 4; ++blk) {
    const int i = blk >> 1, j = blk & 1;
    persist_stage(eb, acc[i][j], lane);
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const int rr = hh * 16 + rl;
      float x[8];
      {
        const float4 a = *reinterpret_cast<const float4*>(eb + rr * kPEP + cg * 8);
        const float4 b = *reinterpret_cast<const float4*>(eb + rr * kPEP + cg * 8 + 4);
        x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w; x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
      }
      float hv[8];
      if constexpr (KIND == PK_BIAS_RELU || KIND == PK_BIAS_SOFTPLUS) {
#pragma unroll
        for (int c = 0; c < 8; ++c) x[c] += cb[j][c];
        if constexpr (KIND == PK_BIAS_SOFTPLUS) {
#pragma unroll
          for (int c = 0; c < 8; ++c) oz[blk][hh][c] = x[c];
        }
#pragma unroll
        for (int c = 0; c < 8; ++c)
          x[c] = KIND == PK_BIAS_RELU ? fmaxf(x[c], 0.0f) : softplus_sel(x[c], e.beta, ib);
#pragma unroll
        for (int c = 0; c < 8; ++c) hv[c] = x[c];
      } else if constexpr (KIND == PK_DACT) {
        if (e.r1_a) {
#pragma unroll
          for (int c = 0; c < 8; ++c) x[c] = fmaf(ra[i][hh], cb[j][c], x[c]);
        }
        if (e.dact == NA_ACT_RELU) {
#pragma unroll
          for (int c = 0; c < 8; ++c) x[c] = opd[blk][hh][c] > 0.0f ? x[c] : 0.0f;
        } else {
#pragma unroll
          for (int c = 0; c < 8; ++c) x[c] *= act_deriv_from_y(NA_ACT_SOFTPLUS, opd[blk][hh][c], e.beta);
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) csum[j][c] += x[c];
      } else if constexpr (KIND == PK_DELTA) {
#pragma unroll
        for (int c = 0; c < 8; ++c) softplus_step(opd[blk][hh][c], x[c], e.beta, ib, hv[c], x[c]);
      }
      if constexpr (kHead) {
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) s += hw[j][c] * hv[c];
        hs[i][j][hh] += s;
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) oc[blk][hh][c] = x[c];
    }
  }
  // phase 2: stores
#pragma unroll
  for (int blk = 0; blk < 4; ++blk)
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const int64_t m = mw + (blk >> 1) * 32 + hh * 16 + rl;
      const int64_t nb = nw + (blk & 1) * 32 + cg * 8;
      if (e.c_dtype == NA_BF16) vstore8((bf16_t*)e.C + m * e.ldc + nb, oc[blk][hh]);
      else vstore8((float*)e.C + m * e.ldc + nb, oc[blk][hh]);
      if constexpr (KIND == PK_BIAS_SOFTPLUS) {
        if (e.z_out) vstore8(e.z_out + m * e.ld_z + nb, oz[blk][hh]);
      }
    }
  if constexpr (kHead) {
    if (e.head_w) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int hh = 0; hh < 2; ++hh) {
            float s = hs[i][j][hh];
            s += __shfl_xor(s, 16);
            s += __shfl_xor(s, 32);
            if (cg == 0) e.head_out[(int64_t)(part + j) * p.M + mw + i * 32 + hh * 16 + rl] = s;
          }
    }
  }
}

// Edge tiles: the four 32 x 32 blocks go through one (rolled) loop so that each kind
// carries a single copy of its element math per bounds mode.
template <int WN, int KIND>
__device__ __attribute__((always_inline)) inline void persist_epilogue(
    const GemmParams& p, float* eb, f32x16_t (&acc)[2][2], int64_t m0, int64_t n0, int wm, int wn,
    int lane, float (&csum)[2][8], int64_t zbase, float ib) {
  // zbase = m0 mod zc_rows (increment mode)
  const EpiParams& e = p.e;
  const int rl = lane & 15, cg = lane >> 4;
  const int64_t mw = m0 + wm * 64, nw = n0 + wn * 64;
  const bool full = mw + 64 <= p.M && nw + 64 <= p.N;  // wave-uniform
  constexpr bool kHead = KIND == PK_BIAS_SOFTPLUS || KIND == PK_DELTA;
  if (full) {
    persist_epilogue_full<WN, KIND>(p, eb, acc, mw, nw, lane, csum, zbase + wm * 64, ib,
                                    (int)(nw / 32));
    return;
  }
  float hs[2] = {0.0f, 0.0f};
#pragma unroll 1
  for (int blk = 0; blk < 4; ++blk) {
    const int i = blk >> 1, j = blk & 1;
    if (blk == 0) persist_stage(eb, acc[0][0], lane);
    else if (blk == 1) persist_stage(eb, acc[0][1], lane);
    else if (blk == 2) persist_stage(eb, acc[1][0], lane);
    else persist_stage(eb, acc[1][1], lane);
    const int64_t nb = nw + j * 32 + cg * 8;
    const int64_t ncl = p.N - nb;
    const int nc = ncl >= 8 ? 8 : (ncl > 0 ? (int)ncl : 0);
    float cb[8], hw8[8], cs[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) cb[c] = hw8[c] = cs[c] = 0.0f;
    const int64_t zr = zbase + wm * 64 + i * 32;
    if (nc > 0) {
      if constexpr (KIND == PK_BIAS_RELU || KIND == PK_BIAS_SOFTPLUS) load_cols<false>(e.bias + nb, nc, cb);
      if constexpr (KIND == PK_DACT) { if (e.r1_a) load_cols<false>(e.r1_b + nb, nc, cb); }
      if constexpr (kHead) { if (e.head_w) load_cols<false>(e.head_w + nb, nc, hw8); }
    }
    persist_rows<KIND, false>(p, eb, mw + i * 32, nb, nc, rl, cg, cb, hw8, hs, cs, zr, ib);
    if constexpr (KIND == PK_DACT) {
      if (j == 0) {
#pragma unroll
        for (int c = 0; c < 8; ++c) csum[0][c] += cs[c];
      } else {
#pragma unroll
        for (int c = 0; c < 8; ++c) csum[1][c] += cs[c];
      }
    }
    if constexpr (kHead) {
      if (e.head_w) {
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
          float s = hs[hh];
          s += __shfl_xor(s, 16);
          s += __shfl_xor(s, 32);
          // parts = one per 32 columns, 4 * ceil(N / 128) of them (zero where past N)
          const int64_t part = (nw + 32 * j) / 32, m = mw + i * 32 + hh * 16 + rl;
          if (cg == 0 && m < p.M && part < 4 * ((p.N + 127) / 128)) e.head_out[part * p.M + m] = s;
          hs[hh] = 0.0f;
        }
      }
    }
  }
}

// Column sums of the workgroup: each wave reduces its 64 columns over the 16
// lanes that share a column group, the two waves of a column block combine
// through LDS, and one atomic per column leaves the workgroup.  Called by every
// wave of the workgroup (it holds a barrier).
template <int WN>
__device__ inline void persist_flush_colsum(float* colsum, int64_t N, int64_t ncol0, int lane,
                                            float (&csum)[2][8], float* epi_base, int wm, int wn) {
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      float s = csum[j][c];
      s += __shfl_xor(s, 1);
      s += __shfl_xor(s, 2);
      s += __shfl_xor(s, 4);
      s += __shfl_xor(s, 8);
      csum[j][c] = s;
    }
  float* partner = epi_base + (WN + wn) * kPEpiFloats;  // region of the wm = 1 wave
  if (wm == 1 && (lane & 15) == 0) {
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int c = 0; c < 8; ++c) partner[(lane >> 4) * 16 + j * 8 + c] = csum[j][c];
  }
  __syncthreads();
  if (wm == 0 && (lane & 15) == 0) {
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int64_t n = ncol0 + j * 32 + (lane >> 4) * 8 + c;
        const float s = csum[j][c] + partner[(lane >> 4) * 16 + j * 8 + c];
        if (n < N) atomic_add_f32(colsum + n, s);
      }
  }
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int c = 0; c < 8; ++c) csum[j][c] = 0.0f;
}

template <int WN, int KIND>
__global__ void __launch_bounds__(128 * WN, WN == 2 ? 2 : 1)  // WN = 2: two workgroups per CU
gemm_bf16_nn_persist_kernel(GemmParams p) {
  using PT = PersistTile<WN>;
  constexpr int NS = PT::NS;
  __shared__ __attribute__((aligned(16))) unsigned char smem[PT::LDS_BYTES];
  // the wave index is uniform: held in a scalar register, the stage buffers and
  // LDS destinations chosen from it stay scalar too
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave / WN, wn = wave % WN;
  bf16_t* st0 = reinterpret_cast<bf16_t*>(smem);
  float* eb = reinterpret_cast<float*>(smem + NS * PT::STAGE_EL * 2) + wave * kPEpiFloats;
  const bf16_t* A = (const bf16_t*)p.A;
  const bf16_t* B = (const bf16_t*)p.B;
  // tile indices fit 32 bits (M < 2^31 is checked on the host)
  const uint32_t n_mt = (uint32_t)((p.M + GBM - 1) / GBM);
  const uint32_t n_tiles = n_mt * (uint32_t)((p.N + PT::BN - 1) / PT::BN);
  const int nk = (int)((p.K + kNNBK - 1) / kNNBK);
  const int64_t my_tiles = blockIdx.x < n_tiles ? (n_tiles - 1 - blockIdx.x) / gridDim.x + 1 : 0;
  const int64_t total = my_tiles * nk;

  // per-lane byte offsets of this wave's stage loads within a tile, and the K
  // offset of the lane's chunk (a chunk is loaded when kofs < K - k0)
  uint32_t voff[PT::LOADS];
  int kofs[PT::LOADS];
#pragma unroll
  for (int u = 0; u < PT::LOADS; ++u) {
    const int q = wave * PT::LOADS + u;
    const int row = q * 16 + (lane >> 2);
    const int chunk = (lane & 3) ^ ((row >> 2) & 3);
    const int64_t ld = q < GBM / 16 ? p.lda : p.ldb;
    const int lrow = q < GBM / 16 ? row : row - GBM;
    voff[u] = (uint32_t)(lrow * ld * 2 + chunk * 16);
    kofs[u] = chunk * 8;
  }

  // issue cursor: next stage to load
  int64_t issued = 0;
  uint32_t is_tile = blockIdx.x;
  int is_t = 0, is_slot = 0;
  // per-tile buffer bases (first row of the tile) and the bytes that follow them
  const char *is_a = nullptr, *is_b = nullptr;
  int64_t is_ae = 0, is_be = 0;
  auto set_issue_tile = [&]() {
    const int64_t m0 = (int64_t)(is_tile % n_mt) * GBM, n0 = (int64_t)(is_tile / n_mt) * PT::BN;
    is_a = reinterpret_cast<const char*>(A + m0 * p.lda);
    is_b = reinterpret_cast<const char*>(B + n0 * p.ldb);
    is_ae = (p.M - m0) * p.lda * 2;
    is_be = (p.N - n0) * p.ldb * 2;
  };
  set_issue_tile();
  auto issue_next = [&]() {
    const int64_t k0 = (int64_t)is_t * kNNBK;
    const int64_t kr = p.K - k0;
    persist_issue<WN>(stage_rsrc(is_a, is_ae, k0 * 2), stage_rsrc(is_b, is_be, k0 * 2), voff, kofs,
                      kr < kNNBK ? (int)kr : kNNBK, st0 + is_slot * PT::STAGE_EL, wave);
    ++issued;
    is_slot = is_slot + 1 == NS ? 0 : is_slot + 1;
    if (++is_t == nk) {
      is_t = 0;
      is_tile += gridDim.x;
      set_issue_tile();
    }
  };
#pragma unroll
  for (int s = 0; s < NS - 1; ++s)
    if (issued < total) issue_next();

  float csum[2][8];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int c = 0; c < 8; ++c) csum[j][c] = 0.0f;
  const float ib = 1.0f / p.e.beta;
  int slot = 0;
  int64_t g = 0;
  for (int64_t jt = 0; jt < my_tiles; ++jt) {
    const uint32_t tile = blockIdx.x + (uint32_t)jt * gridDim.x;
    const int64_t m0 = (int64_t)(tile % n_mt) * GBM, n0 = (int64_t)(tile / n_mt) * PT::BN;
    f32x16_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    for (int t = 0; t < nk; ++t, ++g) {
      persist_wait<PT::LOADS>((int)(issued - g - 1));
      __builtin_amdgcn_s_barrier();  // stage g landed in every wave; stage g-1 consumed
      if (issued < total) issue_next();
      const bf16_t* sa = st0 + slot * PT::STAGE_EL;
      const bf16_t* sb = sa + GBM * kNNBK;
      slot = slot + 1 == NS ? 0 : slot + 1;
#pragma unroll
      for (int kk = 0; kk < kNNBK; kk += 16) {
        bf16x8_t a[2], b[2];
        const int ck = (kk >> 3) + (lane >> 5);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int ra = wm * 64 + i * 32 + (lane & 31);
          const int rb = wn * 64 + i * 32 + (lane & 31);
          a[i] = *reinterpret_cast<const bf16x8_t*>(sa + ra * kNNBK + ((ck ^ ((ra >> 2) & 3)) << 3));
          b[i] = *reinterpret_cast<const bf16x8_t*>(sb + rb * kNNBK + ((ck ^ ((rb >> 2) & 3)) << 3));
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    }
    const int64_t zbase = KIND == PK_DELTA ? (int64_t)((uint32_t)m0 % (uint32_t)p.e.zc_rows) : 0;
    persist_epilogue<WN, KIND>(p, eb, acc, m0, n0, wm, wn, lane, csum, zbase, ib);
    if constexpr (KIND == PK_DACT) {
      if (p.e.colsum) {
        const uint32_t nxt = tile + gridDim.x;
        if (jt + 1 == my_tiles || nxt / n_mt != tile / n_mt)
          persist_flush_colsum<WN>(p.e.colsum, p.N, n0 + wn * 64, lane, csum,
                                   reinterpret_cast<float*>(smem + NS * PT::STAGE_EL * 2), wm, wn);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Weights-stationary GEMM for the 256-wide MLP layers: C = epi(A . B^T) with
// N = 256, M a multiple of 128 and NK 32-wide K stages.  Eight waves each own
// 32 output columns and hold those columns' weights over the whole K range in
// registers (2 NK fragments of 8 bf16 per lane), so only activation rows
// stream: 128 x 32 bf16 stages (8 KB, one buffer-to-LDS load per wave) in
// kWsNS LDS slots, up to kWsNS - 1 of them in flight, and no weight tile is
// re-read per row tile.  The products take the weights as the first MFMA
// operand, which leaves one output row per lane (four 4-column groups); a
// permlane32 swap between the two halves of the wave turns them into two
// 8-column groups per lane, so the epilogue stores 16-byte row segments from
// registers with no LDS staging.
constexpr int kWsStage = GBM * kNNBK;
// NWV waves own 32 columns each (N = 32 NWV); eight waves: one workgroup per CU
// with 15 stages; four waves: two workgroups per CU with 9 stages each.
template <int NWV> struct WsCfg {
  static constexpr int NS = NWV == 8 ? 15 : 9;
  static constexpr int ROWS = GBM / NWV;  // stage rows loaded per wave
};

// vmcnt(n) for n = 0 .. 15 (larger counts wait for 15, which is also safe)
__device__ inline void ws_wait(int ahead) {
  switch (ahead) {
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

__device__ inline void ws_issue(__amdgpu_buffer_rsrc_t r, uint32_t off, bf16_t* dst) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void_t*)dst, 16, off, 0, 0, 0);
}

// Four-element row chunks (16 B of fp32, 8 B of bf16).
typedef __attribute__((ext_vector_type(4))) float f32x4_t;
typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4_t;

__device__ inline void vload4(const float* p, float (&v)[4]) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ inline void vload4(const bf16_t* p, float (&v)[4]) {
  const uint2 a = *reinterpret_cast<const uint2*>(p);
  v[0] = __uint_as_float(a.x << 16);
  v[1] = __uint_as_float(a.x & 0xffff0000u);
  v[2] = __uint_as_float(a.y << 16);
  v[3] = __uint_as_float(a.y & 0xffff0000u);
}
__device__ inline void vstore4(bf16_t* p, const float (&v)[4]) {
  const f32x4_t x = {v[0], v[1], v[2], v[3]};
  *reinterpret_cast<bf16x4_t*>(p) = __builtin_convertvector(x, bf16x4_t);
}

// Lane K of this lane's quad, and the quad's sum (the same value in all four
// lanes: (v0 + v1) + (v2 + v3)), through DPP quad permutes.
template <int K>
__device__ inline float quad_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), K * 0x55, 0xf, 0xf, false));
}
__device__ inline float quad_sum(float v) {
  const float s = v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xf, 0xf, false));
  return s + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s), 0x4E, 0xf, 0xf, false));
}

// Sums over the 32 rows held by the lanes of each wave half (the lanes that share
// h); every lane of the half ends with the sums of its 16 columns.
__device__ inline void ws_reduce_rows(float (&cs)[2][8]) {
#pragma unroll
  for (int gq = 0; gq < 2; ++gq)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      float s = cs[gq][c];
      s += __shfl_xor(s, 1);
      s += __shfl_xor(s, 2);
      s += __shfl_xor(s, 4);
      s += __shfl_xor(s, 8);
      s += __shfl_xor(s, 16);
      cs[gq][c] = s;
    }
}

// Column sums of a weights-stationary wave: sums over the 32 rows of the lanes that
// share h, then one atomic per column (dst + 16 g + c, the lane's columns).
__device__ inline void ws_flush_colsum(float* dst, float (&cs)[2][8], int lane) {
#pragma unroll
  for (int gq = 0; gq < 2; ++gq)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      float s = cs[gq][c];
      s += __shfl_xor(s, 1);
      s += __shfl_xor(s, 2);
      s += __shfl_xor(s, 4);
      s += __shfl_xor(s, 8);
      s += __shfl_xor(s, 16);
      cs[gq][c] = s;
    }
  if ((lane & 31) == 0) {
#pragma unroll
    for (int gq = 0; gq < 2; ++gq)
#pragma unroll
      for (int c = 0; c < 8; ++c) atomic_add_f32(dst + 16 * gq + c, cs[gq][c]);
  }
}

template <int NWV, int NK, int KIND>
__global__ void __launch_bounds__(64 * NWV, NWV == 4 ? 2 : 1) gemm_bf16_ws_kernel(GemmParams p) {
  constexpr int kWsNS = WsCfg<NWV>::NS;
  constexpr int LOADS = WsCfg<NWV>::ROWS / 16;
  __shared__ __attribute__((aligned(16))) unsigned char smem[kWsNS * kWsStage * 2];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int h = lane >> 5;
  bf16_t* st0 = reinterpret_cast<bf16_t*>(smem);
  const bf16_t* A = (const bf16_t*)p.A;
  const EpiParams& e = p.e;
  const uint32_t n_mt = (uint32_t)(p.M / GBM);
  const int64_t my_tiles = blockIdx.x < n_mt ? (n_mt - 1 - blockIdx.x) / gridDim.x + 1 : 0;
  const int64_t total = my_tiles * NK;
  const int nw = wave * 32;  // first column of this wave

  // weights of row nw + lane % 32 of B: 8 k-values (offset 8 h) of each 16-wide step
  bf16x8_t bw[2 * NK];
  {
    const bf16_t* brow = (const bf16_t*)p.B + (int64_t)(nw + (lane & 31)) * p.ldb;
#pragma unroll
    for (int s = 0; s < 2 * NK; ++s) {
      // unconditional load (the row's last chunk stands in past K) and a select,
      // so the loads carry no branch and the compiler counts them exactly
      const int64_t k = 16 * s + 8 * h;
      uint4 u = *reinterpret_cast<const uint4*>(brow + (k < p.K ? k : p.K - 8));
      if (k >= p.K) u = make_uint4(0u, 0u, 0u, 0u);
      bw[s] = __builtin_bit_cast(bf16x8_t, u);
    }
  }
  // bias of the lane's columns nw + 16 g + 8 h + c
  float bb[2][8];
#pragma unroll
  for (int gq = 0; gq < 2; ++gq)
#pragma unroll
    for (int c = 0; c < 8; ++c) bb[gq][c] = 0.0f;
  if constexpr (KIND == PK_BIAS_RELU || KIND == PK_BIAS_SOFTPLUS) {
#pragma unroll
    for (int gq = 0; gq < 2; ++gq) vload8(e.bias + nw + 16 * gq + 8 * h, bb[gq]);
  }

  // stage loads: wave w fills the 16-row blocks LOADS w .. LOADS w + LOADS - 1;
  // lane l lands at slot l & 3 of row l >> 2 of a block and fetches the chunk
  // that belongs there (the chunk index is the same for every block)
  const int arow = wave * 16 * LOADS + (lane >> 2);
  const int achunk = (lane & 3) ^ ((arow >> 2) & 3);
  const uint32_t avoff = (uint32_t)(arow * p.lda * 2 + achunk * 16);
  const uint32_t ablk = (uint32_t)(16 * p.lda * 2);  // byte step between a wave's blocks
  const int akofs = achunk * 8;
  int64_t issued = 0;
  uint32_t is_tile = blockIdx.x;
  int is_t = 0, is_slot = 0;
  const char* is_a = reinterpret_cast<const char*>(A + (int64_t)is_tile * GBM * p.lda);
  int64_t is_ae = (p.M - (int64_t)is_tile * GBM) * p.lda * 2;
  auto issue_next = [&]() {
    const int64_t k0 = (int64_t)is_t * kNNBK;
    const __amdgpu_buffer_rsrc_t rs = stage_rsrc(is_a, is_ae, k0 * 2);
    const bool in_k = akofs < p.K - k0;
#pragma unroll
    for (int u = 0; u < LOADS; ++u)
      ws_issue(rs, in_k ? avoff + u * ablk : kOobOffset,
               st0 + is_slot * kWsStage + (wave * LOADS + u) * 16 * kNNBK);
    ++issued;
    is_slot = is_slot + 1 == kWsNS ? 0 : is_slot + 1;
    if (++is_t == NK) {
      is_t = 0;
      is_tile += gridDim.x;
      is_a = reinterpret_cast<const char*>(A + (int64_t)is_tile * GBM * p.lda);
      is_ae = (p.M - (int64_t)is_tile * GBM) * p.lda * 2;
    }
  };
  // Stages are issued unconditionally: those past the last tile read nothing
  // (their buffer covers no bytes) and land in slots no wave reads, and the
  // workgroup waits for them before it ends.  With a fixed count of loads per
  // iteration the compiler's waits for the per-tile operand registers below
  // count exactly the stage loads issued after them.
  (void)total;
#pragma unroll
  for (int s = 0; s < kWsNS - 1; ++s) issue_next();

  const float ib = 1.0f / e.beta;
  // derivative kind: rank-1 coefficients of the lane's columns and column sums
  constexpr bool kHead = KIND == PK_BIAS_SOFTPLUS || KIND == PK_DELTA || KIND == PK_TAPS_FWD ||
                        KIND == PK_TAPS_BWD;
  constexpr bool kTB = KIND == PK_TAPS_BWD;
  float rb[2][8], cs[2][8], hw[2][8];
#pragma unroll
  for (int gq = 0; gq < 2; ++gq)
#pragma unroll
    for (int c = 0; c < 8; ++c) rb[gq][c] = cs[gq][c] = hw[gq][c] = 0.0f;
  if constexpr (kHead && !kTB) {
    if (e.head_w) {
#pragma unroll
      for (int gq = 0; gq < 2; ++gq) vload8(e.head_w + nw + 16 * gq + 8 * h, hw[gq]);
    }
  }
  if constexpr (KIND == PK_DACT) {
    if (e.r1_a) {
#pragma unroll
      for (int gq = 0; gq < 2; ++gq) vload8(e.r1_b + nw + 16 * gq + 8 * h, rb[gq]);
    }
  }
  // tap backward: lane t of each lane quad (the four taps of one point, rows
  // 4 p .. 4 p + 3) owns columns 4 t .. 4 t + 3 of the quad's 16 in the centre and
  // shared rows; hs2 collects the head-weight gradient of the lane's columns
  // shared rows; the column sums of a tile are reduced over the wave's rows and
  // kept in LDS (one region per wave: [sum][h][16 columns]), so no sums stay in
  // registers across the product loop
  const int tq = lane & 3;
  const int64_t tb_col = nw + 16 * (tq >> 1) + 8 * h + 4 * (tq & 1);
  __shared__ float tb_sums[kTB ? NWV * 64 : 1];
  if constexpr (kTB) tb_sums[wave * 64 + lane] = 0.0f;
  int slot = 0;
  int64_t g = 0;
  for (int64_t jt = 0; jt < my_tiles; ++jt) {
    const int64_t m0 = (int64_t)(blockIdx.x + (uint32_t)jt * gridDim.x) * GBM;
    // derivative kind: the tile's stored activations (bf16) and rank-1 row terms,
    // requested before this tile's stage loads
    uint4 yv[KIND == PK_DACT ? 4 : 1][2];
    float r1v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // increment kind: the stored pre-activation rows (m mod zc_rows; zc_rows is a
    // multiple of 128, so a tile's rows do not wrap)
    float zv[KIND == PK_DELTA || KIND == PK_TAPS_FWD ? 4 : 1][2][8];
    if constexpr (KIND == PK_DELTA) {
      const int64_t z0 = m0 % e.zc_rows;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int gq = 0; gq < 2; ++gq)
          vload8(e.zc + (z0 + i * 32 + (lane & 31)) * e.ld_zc + nw + 16 * gq + 8 * h, zv[i][gq]);
    }
    // point-major taps: row m belongs to point m / 4
    auto load_zc = [&](int i, float (&z)[2][8]) {
      const int64_t zr = (m0 + i * 32 + (lane & 31)) >> 2;
#pragma unroll
      for (int gq = 0; gq < 2; ++gq) vload8(e.zc + zr * e.ld_zc + nw + 16 * gq + 8 * h, z[gq]);
    };
    if constexpr (KIND == PK_TAPS_FWD) {
#pragma unroll
      for (int i = 0; i < 4; ++i) load_zc(i, zv[i]);
    }
    // tap backward: the first row group's centre pre-activations now, the others'
    // at the start of their group
    if constexpr (kTB) load_zc(0, zv[0]);
    // tap backward: each row's dL/d(increment) (dds is point-major [M / 4, 4], so
    // row m's entry is dds[m]) and the centre chunk this lane updates
    float dvr[kTB ? 4 : 1];
    if constexpr (kTB) {
#pragma unroll
      for (int i = 0; i < 4; ++i) dvr[i] = e.tap_dds[m0 + i * 32 + (lane & 31)];
    }
    if constexpr (KIND == PK_DACT) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t m = m0 + i * 32 + (lane & 31);
#pragma unroll
        for (int gq = 0; gq < 2; ++gq)
          yv[i][gq] = *reinterpret_cast<const uint4*>((const bf16_t*)e.dact_y + m * e.ld_dact + nw + 16 * gq + 8 * h);
        if (e.r1_a) r1v[i] = e.r1_a[m];
      }
    }
    f32x16_t acc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][r] = 0.0f;
#pragma unroll
    for (int t = 0; t < NK; ++t, ++g) {
      ws_wait((int)(issued - g - 1) * LOADS);
      __builtin_amdgcn_s_barrier();  // stage g landed in every wave; stage g-1 consumed
      issue_next();
      const bf16_t* sa = st0 + slot * kWsStage;
      slot = slot + 1 == kWsNS ? 0 : slot + 1;
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        const int ck = 2 * kk + h;
        bf16x8_t a[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int ra = i * 32 + (lane & 31);
          a[i] = *reinterpret_cast<const bf16x8_t*>(sa + ra * kNNBK + ((ck ^ ((ra >> 2) & 3)) << 3));
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
          acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bw[2 * t + kk], a[i], acc[i], 0, 0, 0);
      }
    }
    // epilogue: lane holds row m0 + 32 i + lane % 32; register 8 g + 4 k2 + q holds
    // column nw + 16 g + 8 k2 + 4 h + q; after the swap, o[g][0..7] holds the
    // columns nw + 16 g + 8 h + 0 .. 7
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t m = m0 + i * 32 + (lane & 31);
      float o[2][8];
      float hsum = 0.0f;
#pragma unroll
      for (int gq = 0; gq < 2; ++gq)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc[i][8 * gq + q]),
                                                          __float_as_uint(acc[i][8 * gq + 4 + q]), false, false);
          o[gq][q] = __uint_as_float(r[0]);
          o[gq][4 + q] = __uint_as_float(r[1]);
        }
      if constexpr (kTB) {
        if (i == 0) {
#pragma unroll
          for (int gq = 0; gq < 2; ++gq)
#pragma unroll
            for (int c = 0; c < 8; ++c) cs[gq][c] = rb[gq][c] = 0.0f;  // this tile's sums
        }
        // row m is tap t = m % 4 of point m / 4; o holds its dz for 16 columns.
        // With dds_t = a + b_t (a the quad's mean): tap row w (b_t sp'(z + dz) +
        // a dsp'), centre += sum_t dds_t w dsp' (a quad sum), shared row a w sp'(z),
        // and the column sums of the centre increments (b0) and of dds_t (sp(z + dz)
        // - sp(z)) (head weights)
        float (&z)[2][8] = zv[0];
        if (i > 0) load_zc(i, z);
        const int64_t pt = m >> 2;
        const uint2 cen = *reinterpret_cast<const uint2*>((const bf16_t*)e.tap_centre +
                                                          pt * e.tap_ld_centre + tb_col);
        float w8[2][8];  // head weights of the lane's columns (cache hits)
#pragma unroll
        for (int gq = 0; gq < 2; ++gq) vload8(e.head_w + nw + 16 * gq + 8 * h, w8[gq]);
        const float dv = dvr[i];
        const float a = 0.25f * ((quad_bcast<0>(dv) + quad_bcast<1>(dv)) +
                                 (quad_bcast<2>(dv) + quad_bcast<3>(dv)));
        const float b = dv - a;
        // the lane keeps the quad sums and shared-row values of its own 4 columns
        // (chunk tq = 2 gq + c / 4 of its 16)
        float ce[4], cq[4];
#pragma unroll
        for (int gq = 0; gq < 2; ++gq)
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            float s1, dsp, s0;
            const float hv = softplus_incr_nb(z[gq][c], o[gq][c], e.beta, ib);
            softplus_deriv_step_nb(z[gq][c], o[gq][c], e.beta, s1, dsp, s0);
            const float w = w8[gq][c];
            const float ev = (dv * w) * dsp;
            cs[gq][c] += ev;
            rb[gq][c] = fmaf(dv, hv, rb[gq][c]);
            const float es = quad_sum(ev), qv = (a * w) * s0;
            const bool mine = tq == 2 * gq + c / 4;
            if (gq == 0 && c < 4) {
              ce[c & 3] = es;
              cq[c & 3] = qv;
            } else {
              ce[c & 3] = mine ? es : ce[c & 3];
              cq[c & 3] = mine ? qv : cq[c & 3];
            }
            o[gq][c] = w * fmaf(b, s1, a * dsp);
          }
#pragma unroll
        for (int gq = 0; gq < 2; ++gq)
          vstore8((bf16_t*)e.C + m * e.ldc + nw + 16 * gq + 8 * h, o[gq]);
        const uint32_t cw[2] = {cen.x, cen.y};
        float cf[4];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          cf[2 * k] = __uint_as_float(cw[k] << 16) + ce[2 * k];
          cf[2 * k + 1] = __uint_as_float(cw[k] & 0xffff0000u) + ce[2 * k + 1];
        }
        vstore4((bf16_t*)e.tap_centre + pt * e.tap_ld_centre + tb_col, cf);
        vstore4((bf16_t*)e.tap_q + pt * e.tap_ld_q + tb_col, cq);
        if (i == 3) {
          // the tile's column sums over the 32 rows of each lane half, into LDS
          ws_reduce_rows(cs);
          ws_reduce_rows(rb);
          if ((lane & 31) == 0) {
#pragma unroll
            for (int gq = 0; gq < 2; ++gq)
#pragma unroll
              for (int c = 0; c < 8; ++c) {
                tb_sums[wave * 64 + 16 * h + 8 * gq + c] += cs[gq][c];
                tb_sums[wave * 64 + 32 + 16 * h + 8 * gq + c] += rb[gq][c];
              }
          }
        }
        continue;
      }
#pragma unroll
      for (int gq = 0; gq < 2; ++gq) {
        if constexpr (KIND == PK_DACT) {
          const uint32_t yw[4] = {yv[i][gq].x, yv[i][gq].y, yv[i][gq].z, yv[i][gq].w};
          float y[8];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            y[2 * c] = __uint_as_float(yw[c] << 16);
            y[2 * c + 1] = __uint_as_float(yw[c] & 0xffff0000u);
          }
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            float x = fmaf(r1v[i], rb[gq][c], o[gq][c]);
            x = e.dact == NA_ACT_RELU ? (y[c] > 0.0f ? x : 0.0f)
                                      : x * act_deriv_from_y(NA_ACT_SOFTPLUS, y[c], e.beta);
            cs[gq][c] += x;
            o[gq][c] = x;
          }
        } else if constexpr (KIND == PK_DELTA || KIND == PK_TAPS_FWD) {
          if constexpr (KIND == PK_TAPS_FWD) {
            // the increments themselves, kept for the backward when asked for
            if (e.z_out) vstore8(e.z_out + m * e.ld_z + nw + 16 * gq + 8 * h, o[gq]);
          }
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            float hv, y1;
            softplus_step(zv[i][gq][c], o[gq][c], e.beta, ib, hv, y1);
            hsum = fmaf(hw[gq][c], hv, hsum);
            o[gq][c] = y1;
          }
        } else {
          const int64_t col = nw + 16 * gq + 8 * h;
#pragma unroll
          for (int c = 0; c < 8; ++c) o[gq][c] += bb[gq][c];
          if constexpr (KIND == PK_BIAS_SOFTPLUS) {
            if (e.z_out) vstore8(e.z_out + m * e.ld_z + col, o[gq]);
          }
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            float x = o[gq][c];
            if constexpr (KIND == PK_BIAS_RELU) x = fmaxf(x, 0.0f);
            if constexpr (KIND == PK_BIAS_SOFTPLUS) {
              x = softplus_sel(x, e.beta, ib);
              hsum = fmaf(hw[gq][c], x, hsum);
            }
            o[gq][c] = x;
          }
        }
        const int64_t col = nw + 16 * gq + 8 * h;
        if constexpr (KIND != PK_TAPS_FWD) {
          if (e.c_dtype == NA_BF16) vstore8((bf16_t*)e.C + m * e.ldc + col, o[gq]);
          else vstore8((float*)e.C + m * e.ldc + col, o[gq]);
        }
      }
      if constexpr (kHead) {
        // row m over the wave's 32 columns: this lane's 16 and its partner's 16
        if (e.head_w) {
          hsum += __shfl_xor(hsum, 32);
          if (h == 0) e.head_out[(int64_t)wave * p.M + m] = hsum;
        }
      }
    }
  }
  if constexpr (KIND == PK_DACT) {
    if (e.colsum) ws_flush_colsum(e.colsum + nw + 8 * h, cs, lane);
  }
  if constexpr (kTB) {
    // lane j adds entry j of the wave's sums: (sum j / 32, h, group, column)
    const int j = lane & 31, hh = j >> 4, gq = (j >> 3) & 1, c = j & 7;
    float* dst = lane < 32 ? e.colsum : e.colsum2;
    atomic_add_f32(dst + nw + 16 * gq + 8 * hh + c, tb_sums[wave * 64 + lane]);
  }
  __builtin_amdgcn_s_waitcnt(0xF70);  // no buffer-to-LDS load outlives the workgroup
}

// ---------------------------------------------------------------------------
// Backward of the tap-increment layer (na_taps_increment_bwd), elementwise from
// the recomputed increments dz = D W^T (fp32 [M = 4 P, N], point-major rows
// 4 i + t).  With dds_t = a + b_t (a = the four taps' mean, the Hessian-trace
// coefficient shared by every tap; sum_t b_t = 0 since sum_t k_t = 0), s1_t =
// sp'(z + dz_t), s0 = sp'(z) and dsp_t = s1_t - s0 (formed as a product, relative
// precision of its own size):
//   tap row   dpre[4 i + t] = w (b_t s1_t + a dsp_t)      (pairs with D_t)
//   point row q[i]          = a w s0                      (pairs with sum_t D_t)
//   centre[i]              += sum_t dds_t w dsp_t         (and db0 += its column sums)
//   dws                    += sum_t dds_t (sp(z + dz_t) - sp(z))
// Each bf16 row holds a term of its natural size: no difference of nearly equal
// terms is rounded to bf16.  A thread owns 4 columns of a point and walks points
// with a grid stride; the workgroup's column sums are combined through LDS and
// added with one atomic per column.
// ---------------------------------------------------------------------------
constexpr int kTibBlock = 256;
__global__ void __launch_bounds__(kTibBlock)
taps_increment_bwd_kernel(const float* __restrict__ dz, int64_t P, int N,
                          const float* __restrict__ zc, int64_t ld_zc,
                          const float* __restrict__ ws, float beta,
                          const float* __restrict__ dds, bf16_t* __restrict__ dpre,
                          int64_t ld_dpre, bf16_t* __restrict__ centre, int64_t ld_centre,
                          bf16_t* __restrict__ q, int64_t ld_q, float* __restrict__ db0,
                          float* __restrict__ dws) {
  __shared__ float red[2][kTibBlock][4];
  const int cpp = N / 4;               // 4-column chunks per point (64 or 32)
  const int ppb = kTibBlock / cpp;     // points per workgroup pass (4 or 8)
  const int chunk = threadIdx.x % cpp, slot = threadIdx.x / cpp;
  const int c0 = chunk * 4;
  const float ib = 1.0f / beta;
  float w[4], cs[4], hs[4];
  vload4(ws + c0, w);
#pragma unroll
  for (int c = 0; c < 4; ++c) cs[c] = hs[c] = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * ppb + slot; i < P; i += (int64_t)gridDim.x * ppb) {
    float z0[4], x[4][4];
    vload4(zc + i * ld_zc + c0, z0);
#pragma unroll
    for (int t = 0; t < 4; ++t) vload4(dz + (4 * i + t) * N + c0, x[t]);
    const float4 d4 = *reinterpret_cast<const float4*>(dds + 4 * i);
    const float dv[4] = {d4.x, d4.y, d4.z, d4.w};
    const float a = 0.25f * ((dv[0] + dv[1]) + (dv[2] + dv[3]));
    float ev[4], s0v[4], o[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float dzc[4] = {x[0][c], x[1][c], x[2][c], x[3][c]};
      float oc[4];
      tap_bwd_column(z0[c], dzc, dv, a, w[c], beta, ib, oc, ev[c], hs[c], s0v[c]);
#pragma unroll
      for (int t = 0; t < 4; ++t) o[t][c] = oc[t];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) vstore4(dpre + (4 * i + t) * ld_dpre + c0, o[t]);
    float qv[4], cv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) qv[c] = (a * w[c]) * s0v[c];
    vstore4(q + i * ld_q + c0, qv);
    vload4(centre + i * ld_centre + c0, cv);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      cv[c] += ev[c];
      cs[c] += ev[c];
    }
    vstore4(centre + i * ld_centre + c0, cv);
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    red[0][threadIdx.x][c] = cs[c];
    red[1][threadIdx.x][c] = hs[c];
  }
  __syncthreads();
  if (slot == 0) {
    for (int k = 1; k < ppb; ++k) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        cs[c] += red[0][k * cpp + chunk][c];
        hs[c] += red[1][k * cpp + chunk][c];
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      atomic_add_f32(db0 + c0 + c, cs[c]);
      atomic_add_f32(dws + c0 + c, hs[c]);
    }
  }
}

// ---------------------------------------------------------------------------
// Weight-gradient GEMM: C[n, k] += sum_m A[m, n] B[m, k] (bf16 in, fp32 out),
// A = [M, lda] and B = [M, ldb] row-major, M (the points) much larger than N
// and K (the layer widths).  Each workgroup owns a C tile of 256 x 64*KB in
// MFMA accumulators (8 waves: 4 along n x 2 along k, 2 x KB 32x32 blocks per
// wave) and reduces one contiguous range of rows; one atomic add per output
// element leaves the workgroup at the end.
//
// Rows stream through NS LDS stages of 32 rows (buffer-to-LDS loads, 16 B per
// lane, rows past the range or columns past N / K read zeros).  A stage keeps
// the row-major layout of memory with a row pitch of PA = 512 B (A) and PB
// bytes (B), and 16-byte chunk c of row r stored at slot c ^ 4*(r % 4).  The
// MFMA operands need 8 consecutive rows of one column per lane, which
// ds_read_tr16_b64 delivers from this layout: in each 16-lane group, source
// lane j addresses row r0 + j/4 and the 4 columns c0 + 4*(j%4) .. +3, and lane i
// receives column c0 + i of rows r0 .. r0 + 3.  The swizzle puts the four rows
// of a read on distinct banks.
// ---------------------------------------------------------------------------
typedef short v4s_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v4s_t lds_v4s_t;

constexpr int kWgPA = 512;

template <int KB> struct WgTile {
  static constexpr int BK = 64 * KB;
  static constexpr int PB = ((2 * BK + 255) / 256) * 256;
  static constexpr int STAGE_BYTES = 32 * (kWgPA + PB);
  static constexpr int NS = 150000 / STAGE_BYTES < 4 ? 150000 / STAGE_BYTES : 4;
  static constexpr int LOADS = STAGE_BYTES / 1024 / 8;  // 1 KB loads per wave per stage
  static_assert(STAGE_BYTES % 8192 == 0, "stage must split into 8 waves of 1 KB loads");
};

__device__ inline bf16x8_t tr_pair(unsigned char* p0, unsigned char* p1) {
  const v4s_t lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s_t*)p0);
  const v4s_t hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s_t*)p1);
  typedef short v8s_t __attribute__((ext_vector_type(8)));
  const v8s_t v = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  return __builtin_bit_cast(bf16x8_t, v);
}

// Cache policy of a buffer-to-LDS load whose rows are read once (the
// non-temporal bit of the instruction's cache-policy field).
constexpr int kStreamAux = 2;

// One 32-row slice of A and B into a stage; rows from `me` on read zeros.
template <int KB>
__device__ inline void wgrad_issue(const bf16_t* A, int64_t lda, const bf16_t* B, int64_t ldb,
                                   int64_t m, int64_t me, const uint32_t (&voff)[WgTile<KB>::LOADS],
                                   const bool (&isa)[WgTile<KB>::LOADS], unsigned char* st, int wave) {
  using WT = WgTile<KB>;
  const int64_t rem = me - m;
  const int64_t ab = rem * lda * 2, bb = rem * ldb * 2;
  const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(A + m * lda), 0, ab > 0x7fffffffLL ? 0x7fffffff : (int)ab, 0x00020000);
  const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(B + m * ldb), 0, bb > 0x7fffffffLL ? 0x7fffffff : (int)bb, 0x00020000);
#pragma unroll
  for (int u = 0; u < WT::LOADS; ++u)
    __builtin_amdgcn_raw_ptr_buffer_load_lds(isa[u] ? ra : rb,
                                             (lds_void_t*)(st + (wave * WT::LOADS + u) * 1024), 16,
                                             voff[u], 0, 0, kStreamAux);
}

template <int KB>
__global__ void __launch_bounds__(512)
gemm_bf16_wgrad_kernel(const bf16_t* __restrict__ A, int64_t lda, const bf16_t* __restrict__ B,
                       int64_t ldb, int64_t M, int64_t N, int64_t K, int64_t rows_per,
                       float* __restrict__ C, int64_t ldc, float* __restrict__ part) {
  using WT = WgTile<KB>;
  constexpr int NS = WT::NS, PB = WT::PB;
  __shared__ __attribute__((aligned(16))) unsigned char smem[NS * WT::STAGE_BYTES];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wn = wave >> 1, wk = wave & 1;
  const int64_t n0 = (int64_t)blockIdx.y * 256, k0 = (int64_t)blockIdx.z * WT::BK;
  const int64_t mb = (int64_t)blockIdx.x * rows_per;
  const int64_t me = mb + rows_per < M ? mb + rows_per : M;
  const int nsl = mb < me ? (int)((me - mb + 31) / 32) : 0;

  // per-lane load offsets (bytes within a 32-row slice) and validity
  uint32_t voff[WT::LOADS];
  bool isa[WT::LOADS];
#pragma unroll
  for (int u = 0; u < WT::LOADS; ++u) {
    const int off = (wave * WT::LOADS + u) * 1024 + lane * 16;  // byte offset in the stage
    const bool a = off < 32 * kWgPA;
    const int loc = a ? off : off - 32 * kWgPA;
    const int pitch = a ? kWgPA : PB;
    const int row = loc / pitch, slot = (loc % pitch) / 16;
    const int chunk = slot ^ (4 * (row & 3));
    const int64_t col = (a ? n0 : k0) + chunk * 8;
    const bool ok = col < (a ? N : K) && chunk * 8 < (a ? 256 : WT::BK);
    voff[u] = ok ? (uint32_t)(row * (a ? lda : ldb) * 2 + col * 2) : kOobOffset;
    isa[u] = (wave * WT::LOADS + u) * 1024 < 32 * kWgPA;  // wave-uniform per u
  }
  auto issue = [&](int sl) {
    wgrad_issue<KB>(A, lda, B, ldb, mb + (int64_t)sl * 32, me, voff, isa,
                    smem + (sl % NS) * WT::STAGE_BYTES, wave);
  };

  // per-lane LDS read offsets: row (L>>5)*8 + ((L>>2)&3), column group 16*((L>>4)&1) + 4*(L&3)
  const int rrow = (lane >> 5) * 8 + ((lane >> 2) & 3);
  const int swz = 4 * ((lane >> 2) & 3);
  const int csub = 16 * ((lane >> 4) & 1) + 4 * (lane & 3);  // column within a 32-block
  int aoff[2], boff[KB];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int c = wn * 64 + i * 32 + csub;
    aoff[i] = rrow * kWgPA + (((c >> 3) ^ swz) << 4) + (c & 7) * 2;
  }
#pragma unroll
  for (int j = 0; j < KB; ++j) {
    const int c = wk * 32 * KB + j * 32 + csub;
    boff[j] = 32 * kWgPA + rrow * PB + (((c >> 3) ^ swz) << 4) + (c & 7) * 2;
  }

  f32x16_t acc[2][KB];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < KB; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

#pragma unroll
  for (int s = 0; s < NS - 1; ++s)
    if (s < nsl) issue(s);
  for (int sl = 0; sl < nsl; ++sl) {
    const int ahead = nsl - 1 - sl < NS - 2 ? nsl - 1 - sl : NS - 2;
    persist_wait<WT::LOADS>(ahead);
    __builtin_amdgcn_s_barrier();  // slice sl landed in every wave; slice sl-1 consumed
    if (sl + NS - 1 < nsl) issue(sl + NS - 1);
    unsigned char* st = smem + (sl % NS) * WT::STAGE_BYTES;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {  // 16 rows per MFMA step
      bf16x8_t a[2], b[KB];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        a[i] = tr_pair(st + aoff[i] + kk * 16 * kWgPA, st + aoff[i] + (kk * 16 + 4) * kWgPA);
#pragma unroll
      for (int j = 0; j < KB; ++j)
        b[j] = tr_pair(st + boff[j] + kk * 16 * PB, st + boff[j] + (kk * 16 + 4) * PB);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < KB; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  // acc[i][j][r]: n = n0 + wn*64 + i*32 + 8*(r>>2) + 4*(lane>>5) + (r&3), k = k0 + wk*32*KB + j*32 + (lane&31)
  if (part) {
    // this split's partial product, [split][n][k] (zeros for a split without rows)
    float* pp = part + (int64_t)blockIdx.x * N * K;
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      const int64_t k = k0 + wk * 32 * KB + j * 32 + (lane & 31);
      if (k >= K) continue;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int64_t n = n0 + wn * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          if (n < N) __builtin_nontemporal_store(acc[i][j][r], pp + n * K + k);
        }
    }
    return;
  }
  if (nsl == 0) return;
#pragma unroll
  for (int j = 0; j < KB; ++j) {
    const int64_t k = k0 + wk * 32 * KB + j * 32 + (lane & 31);
    if (k >= K) continue;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t n = n0 + wn * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (n < N) atomic_add_f32(C + n * ldc + k, acc[i][j][r]);
      }
  }
}

// C[n, k] += sum over splits s (four interleaved chains, each in split order) of
// part[s][n][k].  A workgroup of 4 waves takes 64 consecutive outputs; wave w
// sums splits w, w + 4, ...; the four sums combine as (0 + 1) + (2 + 3).
__global__ void __launch_bounds__(256)
wgrad_reduce_kernel(const float* __restrict__ part, int64_t splits, int64_t N, int64_t K,
                    float* __restrict__ C, int64_t ldc) {
  __shared__ float red[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t total = N * K;
  const int64_t o = (int64_t)blockIdx.x * 64 + lane;
  float s = 0.0f;
  if (o < total) {
    const float* p = part + o;
    int64_t sp = w;
    for (; sp + 28 < splits; sp += 32) {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = __builtin_nontemporal_load(p + (sp + 4 * u) * total);
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; sp < splits; sp += 4) s += __builtin_nontemporal_load(p + sp * total);
  }
  red[w][lane] = s;
  __syncthreads();
  if (w == 0 && o < total) {
    const float v = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    const int64_t n = o / K, k = o - n * K;
    C[n * ldc + k] += v;
  }
}

// Workgroups per CU and CU count for the persistent launch, per device.
struct PersistOcc { int cus = 0, wg2 = 0, wg4 = 0; };
static PersistOcc g_persist_occ[64];

static int persist_grid(int WN, int64_t n_tiles) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  PersistOcc& o = g_persist_occ[dev];
  if (o.cus == 0) {
    int cus = 0, b2 = 0, b4 = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b2, gemm_bf16_nn_persist_kernel<2, PK_PLAIN>, 256, 0) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b4, gemm_bf16_nn_persist_kernel<4, PK_PLAIN>, 512, 0) != hipSuccess) return 0;
    o.wg2 = b2; o.wg4 = b4; o.cus = cus;
  }
  const int64_t per = (int64_t)o.cus * (WN == 4 ? o.wg4 : o.wg2);
  if (per <= 0) return 0;
  return (int)(n_tiles < per ? n_tiles : per);
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// Epilogue kind of the persistent kernel for these arguments, or -1 when the
// general kernel runs them (other operand types, split-K, atomic epilogue,
// unaligned rows, or epilogue terms outside the kinds).
static int persist_kind(const na_gemm_args* a, int split) {
  if (a->in_dtype != NA_BF16 || a->trans_a || a->trans_b) return -1;
  if (a->epilogue != NA_EPI_STORE || split != 1 || a->K < 1) return -1;
  if (a->M >= 0x7fffffffLL || a->lda * 2 * GBM >= 0x7fffffffLL || a->ldb * 2 * 256 >= 0x7fffffffLL) return -1;
  const int cb = a->c_dtype == NA_BF16 ? 2 : 4;
  if (!aligned16(a->C) || (a->ldc * cb) % 16) return -1;
  if (a->r1_a && (!a->dact || !aligned16(a->r1_b))) return -1;
  if (a->bias && !aligned16(a->bias)) return -1;
  if (a->head_w && !aligned16(a->head_w)) return -1;
  if (a->z_out && (!aligned16(a->z_out) || (a->ld_z * 4) % 16)) return -1;
  if (a->zc) {
    if (!aligned16(a->zc) || (a->ld_zc * 4) % 16 || a->colsum || a->zc_rows >= 0x7fffffffLL) return -1;
    return PK_DELTA;  // softplus, no bias / dact / z_out: checked by the caller
  }
  if (a->dact) {
    const int db = a->dact_dtype == NA_BF16 ? 2 : 4;
    if ((a->dact != NA_ACT_RELU && a->dact != NA_ACT_SOFTPLUS) || a->act || a->bias || a->z_out ||
        a->head_w)
      return -1;
    if (!aligned16(a->dact_y) || (a->ld_dact * db) % 16) return -1;
    return PK_DACT;
  }
  if (a->colsum) return -1;
  if (!a->bias && !a->act && !a->z_out && !a->head_w) return PK_PLAIN;
  if (a->bias && a->act == NA_ACT_RELU && !a->z_out && !a->head_w) return PK_BIAS_RELU;
  if (a->bias && a->act == NA_ACT_SOFTPLUS) return PK_BIAS_SOFTPLUS;
  return -1;
}

template <int WN>
static void launch_persist(int kind, int grid, const GemmParams& p, hipStream_t s) {
  const dim3 g(grid), b(128 * WN);
  switch (kind) {
    case PK_PLAIN: hipLaunchKernelGGL((gemm_bf16_nn_persist_kernel<WN, PK_PLAIN>), g, b, 0, s, p); break;
    case PK_BIAS_RELU: hipLaunchKernelGGL((gemm_bf16_nn_persist_kernel<WN, PK_BIAS_RELU>), g, b, 0, s, p); break;
    case PK_BIAS_SOFTPLUS: hipLaunchKernelGGL((gemm_bf16_nn_persist_kernel<WN, PK_BIAS_SOFTPLUS>), g, b, 0, s, p); break;
    case PK_DACT: hipLaunchKernelGGL((gemm_bf16_nn_persist_kernel<WN, PK_DACT>), g, b, 0, s, p); break;
    default: hipLaunchKernelGGL((gemm_bf16_nn_persist_kernel<WN, PK_DELTA>), g, b, 0, s, p); break;
  }
}

// Weights-stationary path: N = 256 or 128, whole 128-row tiles, K of 5, 8 or 9
// stages; bf16 activations for the derivative kind, stored rows in whole tiles
// for the increment kind.
static int ws_stages(const na_gemm_args* a, int kind) {
  if (kind == PK_DACT && (a->dact_dtype != NA_BF16 || a->ld_dact % 8 || !aligned16(a->dact_y))) return 0;
  if (kind == PK_DELTA && (a->zc_rows % GBM || a->ld_zc % 4)) return 0;
  if ((a->z_out || a->head_w) && kind != PK_BIAS_SOFTPLUS && kind != PK_DELTA) return 0;
  if (a->z_out && a->ld_z % 4) return 0;
  if ((a->N != 256 && a->N != 128) || a->M % GBM != 0 || a->K % 8 != 0 || a->lda % 8 || a->ldb % 8)
    return 0;
  if (!aligned16(a->A) || !aligned16(a->B)) return 0;
  const int64_t nk = ceil_div(a->K, (int64_t)kNNBK);
  if (kind == PK_DELTA && nk != 5) return 0;  // the increment layer's width (K <= 160)
  return (nk == 5 || nk == 8 || nk == 9) ? (int)nk : 0;
}

// one workgroup per CU (eight waves) or two (four waves)
static int ws_grid(int64_t n_mt, int nwv) {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (cus[dev] == 0 && hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  const int64_t per = (int64_t)cus[dev] * (nwv == 8 ? 1 : 2);
  return (int)(n_mt < per ? n_mt : per);
}

template <int NWV>
static void launch_ws(int nk, int kind, int grid, const GemmParams& p, hipStream_t s,
                      int halves = 1) {
  const dim3 g(grid, halves), b(64 * NWV);
#define NA_WS(NK)                                                                                             \
  switch (kind) {                                                                                             \
    case PK_PLAIN: hipLaunchKernelGGL((gemm_bf16_ws_kernel<NWV, NK, PK_PLAIN>), g, b, 0, s, p); break;         \
    case PK_BIAS_RELU: hipLaunchKernelGGL((gemm_bf16_ws_kernel<NWV, NK, PK_BIAS_RELU>), g, b, 0, s, p); break; \
    case PK_DACT: hipLaunchKernelGGL((gemm_bf16_ws_kernel<NWV, NK, PK_DACT>), g, b, 0, s, p); break;           \
    case PK_DELTA: hipLaunchKernelGGL((gemm_bf16_ws_kernel<NWV, NK, PK_DELTA>), g, b, 0, s, p); break;         \
    case PK_TAPS_FWD: hipLaunchKernelGGL((gemm_bf16_ws_kernel<NWV, NK, PK_TAPS_FWD>), g, b, 0, s, p); break;   \
    case PK_TAPS_BWD: hipLaunchKernelGGL((gemm_bf16_ws_kernel<NWV, NK, PK_TAPS_BWD>), g, b, 0, s, p); break;   \
    default: hipLaunchKernelGGL((gemm_bf16_ws_kernel<NWV, NK, PK_BIAS_SOFTPLUS>), g, b, 0, s, p); break;       \
  }
  if (nk == 5) { NA_WS(5) } else if (nk == 8) { NA_WS(8) } else { NA_WS(9) }
#undef NA_WS
}

// Weight-gradient products (both operands transposed, bf16, atomic fp32
// output, no epilogue terms) run on the row-range reduction kernel.
static bool wgrad_ok(const na_gemm_args* a) {
  return a->in_dtype == NA_BF16 && a->trans_a && a->trans_b && a->epilogue == NA_EPI_ATOMIC &&
         a->c_dtype == NA_F32 && !a->bias && !a->act && !a->dact && !a->r1_a && !a->colsum &&
         !a->z_out && !a->zc && !a->head_w && a->K >= 1 && a->lda * 2 * 32 < 0x7fffffffLL &&
         a->ldb * 2 * 32 < 0x7fffffffLL;
}

// Launch plan of the weight-gradient kernel: output [a->M, a->N] = sum over
// a->K rows, split into `splits` row ranges of rows_per rows (kernel names:
// n = output row, k = output column).
struct WgradPlan { int64_t No, Ko, Mr, splits, rows_per; int kb; };

static WgradPlan wgrad_plan(const na_gemm_args* a) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  WgradPlan w;
  w.No = a->M; w.Ko = a->N; w.Mr = a->K;
  w.kb = w.Ko <= 128 ? 2 : (w.Ko <= 192 ? 3 : (w.Ko <= 256 ? 4 : 5));
  const int64_t tiles = ceil_div(w.No, (int64_t)256) * ceil_div(w.Ko, (int64_t)(64 * w.kb));
  int64_t splits = cus / tiles;
  if (splits < 1) splits = 1;
  int64_t rows_per = ceil_div(ceil_div(w.Mr, splits), (int64_t)32) * 32;
  if (rows_per < 128) rows_per = 128;
  w.rows_per = rows_per;
  w.splits = ceil_div(w.Mr, rows_per);
  return w;
}

static int launch_wgrad(const na_gemm_args* a, hipStream_t s) {
  const WgradPlan w = wgrad_plan(a);
  const int64_t bk = 64 * w.kb;
  NA_CHECK_ARG(w.splits < 0x7fffffffLL && ceil_div(w.No, (int64_t)256) < 65536 &&
                   ceil_div(w.Ko, bk) < 65536,
               "weight-gradient grid too large");
  const bool use_part = a->wg_partials && a->wg_partials_len >= w.splits * w.No * w.Ko;
  float* part = use_part ? a->wg_partials : nullptr;
  const dim3 grid((unsigned)w.splits, (unsigned)ceil_div(w.No, (int64_t)256), (unsigned)ceil_div(w.Ko, bk));
  const bf16_t* A = (const bf16_t*)a->A;
  const bf16_t* B = (const bf16_t*)a->B;
  float* C = (float*)a->C;
#define NA_WG(KB) hipLaunchKernelGGL((gemm_bf16_wgrad_kernel<KB>), grid, dim3(512), 0, s, A, a->lda, B, \
                                     a->ldb, w.Mr, w.No, w.Ko, w.rows_per, C, a->ldc, part)
  if (w.kb == 2) NA_WG(2);
  else if (w.kb == 3) NA_WG(3);
  else if (w.kb == 4) NA_WG(4);
  else NA_WG(5);
#undef NA_WG
  NA_CHECK_LAUNCH();
  if (use_part) {
    const int64_t blocks = ceil_div(w.No * w.Ko, (int64_t)64);
    NA_CHECK_ARG(blocks < 0x7fffffffLL, "weight-gradient output too large");
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, s, part, w.splits,
                       w.No, w.Ko, C, a->ldc);
    NA_CHECK_LAUNCH();
  }
  return NA_OK;
}

// ---------------------------------------------------------------------------
template <typename T>
static int launch_gemm(const na_gemm_args* a, const GemmParams& p, dim3 grid, hipStream_t s) {
  if (!a->trans_a && !a->trans_b) {
    if constexpr (sizeof(T) == 2)
      hipLaunchKernelGGL(gemm_bf16_nn_kernel, grid, dim3(GNT), 0, s, p);
    else
      hipLaunchKernelGGL((gemm_kernel<T, false, false>), grid, dim3(GNT), 0, s, p);
  }
  else if (!a->trans_a && a->trans_b)
    hipLaunchKernelGGL((gemm_kernel<T, false, true>), grid, dim3(GNT), 0, s, p);
  else if (a->trans_a && a->trans_b)
    hipLaunchKernelGGL((gemm_kernel<T, true, true>), grid, dim3(GNT), 0, s, p);
  else
    hipLaunchKernelGGL((gemm_kernel<T, true, false>), grid, dim3(GNT), 0, s, p);
  NA_CHECK_LAUNCH();
  return NA_OK;
}


// Vectorised act_grad for rows of whole 16-byte chunks (N a multiple of 8 with
// N / 8 dividing 256): thread (r, c) of a block handles columns 8 c .. 8 c + 7
// of rows r, r + R, ... (R = 256 / (N / 8) row lanes) over the block's row
// range; column sums are reduced over the row lanes in LDS and leave the block
// as one atomic per column.
template <typename TD, typename TY, typename TO>
__global__ void __launch_bounds__(256)
act_grad_vec_kernel(const TD* __restrict__ dy, int64_t ld_dy, const TY* __restrict__ y, int64_t ld_y,
                    int64_t M, int64_t N, int act, float beta, TO* __restrict__ out, int64_t ld_out,
                    float* __restrict__ colsum, int64_t rows_per_block) {
  __shared__ float red[256 * 8];
  const int cpr = (int)(N / 8);
  const int c = threadIdx.x % cpr, r = threadIdx.x / cpr, lanes = 256 / cpr;
  const int64_t m0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t m1 = m0 + rows_per_block < M ? m0 + rows_per_block : M;
  float cs[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) cs[k] = 0.0f;
  // four rows per thread in flight (raw 16-byte chunks, non-temporal: every
  // row streams through once); each thread still visits its rows in order
  constexpr int U = 4;
  for (int64_t mb = m0 + r; mb < m1; mb += U * lanes) {
    Chunk8<TD> gr[U];
    Chunk8<TY> vr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t m = mb + u * lanes;
      if (m < m1) {
        chunk_load(dy + m * ld_dy + 8 * c, gr[u]);
        chunk_load(y + m * ld_y + 8 * c, vr[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t m = mb + u * lanes;
      if (m >= m1) continue;
      float g[8], v[8];
      chunk_f32(gr[u], g);
      chunk_f32(vr[u], v);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        g[k] *= act_deriv_from_y(act, v[k], beta);
        cs[k] += g[k];
      }
      store8v_nt(out + m * ld_out + 8 * c, g);
    }
  }
  if (colsum) {
#pragma unroll
    for (int k = 0; k < 8; ++k) red[k * 256 + threadIdx.x] = cs[k];
    __syncthreads();
    if (r == 0) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        float s = 0.0f;
        for (int q = 0; q < lanes; ++q) s += red[k * 256 + q * cpr + c];
        atomic_add_f32(colsum + 8 * c + k, s);
      }
    }
  }
}

}  // namespace na

using namespace na;

// Tap-increment products of the BF16 4-tap gradient on the weights-stationary
// kernel: rows point-major (row m = tap m % 4 of point m / 4), inputs the increments
// D of the first layer's input, weights W [N, K], the centre pre-activations zc
// (fp32, point rows) and the SDF head ws [N].
static int taps_increment(bool backward, const void* D, int64_t ld_d, int64_t M, int64_t K,
                          const void* W, int64_t ldw, int64_t N, const float* zc, int64_t ld_zc,
                          const float* ws, float beta, float* head_out, const float* dds,
                          void* dpre, int64_t ld_dpre, void* centre, int64_t ld_centre,
                          void* q, int64_t ld_q, float* db0, float* dws, float* work,
                          int work_holds_dz, hipStream_t s) {
  NA_CHECK_ARG(M >= 0 && M % GBM == 0, "rows must be a multiple of 128 (pad the points to 32)");
  NA_CHECK_ARG(N == 256 || N == 128, "layer width must be 128 or 256");
  const int64_t nk = ceil_div(K, (int64_t)kNNBK);
  NA_CHECK_ARG(K >= 1 && K % 8 == 0 && (nk == 5 || nk == 8 || nk == 9),
               "K must be a multiple of 8 in 129..160, 225..288 (5, 8 or 9 stages of 32)");
  NA_CHECK_ARG(ld_d >= K && ld_d % 8 == 0 && ldw >= K && ldw % 8 == 0 && ld_zc >= N &&
                   ld_zc % 4 == 0,
               "bad leading dimensions");
  NA_CHECK_ARG(aligned16(D) && aligned16(W) && aligned16(zc) && aligned16(ws),
               "operands must be 16-byte aligned");
  NA_CHECK_ARG(M / GBM < 0x7fffffffLL, "too many rows");
  if (!backward) {
    NA_CHECK_ARG(head_out != nullptr, "head_out is null");
    NA_CHECK_ARG(aligned16(work), "dz_out must be 16-byte aligned");
  } else {
    const bool fused = tune_value(TUNE_TIB_FUSED) != 0;
    NA_CHECK_ARG(fused || (work && aligned16(work)), "work is null (the split backward needs it)");
    NA_CHECK_ARG(dds && dpre && centre && q && db0 && dws && aligned16(dpre) &&
                     aligned16(centre) && aligned16(q) && aligned16(dds) &&
                     ld_dpre >= N && ld_dpre % 8 == 0 &&
                     ld_centre >= N && ld_centre % 8 == 0 && ld_q >= N && ld_q % 8 == 0,
                 "bad backward operands");
  }
  if (M == 0) return NA_OK;
  GemmParams p{};
  p.M = M; p.N = N; p.K = K;
  p.A = D; p.lda = ld_d; p.B = W; p.ldb = ldw;
  p.k_chunk = nk * kNNBK;
  p.e.epilogue = NA_EPI_STORE;
  p.e.beta = beta;
  const int nwv = N == 256 ? 8 : 4;
  const int wg = ws_grid(M / GBM, nwv);
  NA_CHECK_ARG(wg > 0, "could not size the launch");
  if (!backward) {
    p.e.act = NA_ACT_SOFTPLUS;
    p.e.zc = zc; p.e.ld_zc = ld_zc; p.e.zc_rows = M / 4;
    p.e.head_w = ws; p.e.head_out = head_out;
    p.e.z_out = work; p.e.ld_z = N;  // dz_out (optional)
    if (nwv == 8) launch_ws<8>((int)nk, PK_TAPS_FWD, wg, p, s);
    else launch_ws<4>((int)nk, PK_TAPS_FWD, wg, p, s);
    NA_CHECK_LAUNCH();
    return NA_OK;
  }
  if (tune_value(TUNE_TIB_FUSED) != 0) {
    // backward in the weights-stationary product's epilogue: dz never leaves
    // the registers (no workspace)
    p.e.zc = zc; p.e.ld_zc = ld_zc; p.e.zc_rows = M / 4;
    p.e.head_w = ws;
    p.e.C = dpre; p.e.ldc = ld_dpre; p.e.c_dtype = NA_BF16;
    p.e.tap_dds = dds;
    p.e.tap_centre = centre; p.e.tap_ld_centre = ld_centre;
    p.e.tap_q = q; p.e.tap_ld_q = ld_q;
    p.e.colsum = db0; p.e.colsum2 = dws;
    if (nwv == 8) launch_ws<8>((int)nk, PK_TAPS_BWD, wg, p, s);
    else launch_ws<4>((int)nk, PK_TAPS_BWD, wg, p, s);
    NA_CHECK_LAUNCH();
    return NA_OK;
  }
  // backward: recompute dz = D W^T into the fp32 workspace (unless the forward
  // kept it there), then the elementwise pass
  if (!work_holds_dz) {
    p.e.C = work; p.e.ldc = N; p.e.c_dtype = NA_F32;
    if (nwv == 8) launch_ws<8>((int)nk, PK_PLAIN, wg, p, s);
    else launch_ws<4>((int)nk, PK_PLAIN, wg, p, s);
    NA_CHECK_LAUNCH();
  }
  const int64_t P = M / 4;
  const int ppb = kTibBlock / (int)(N / 4);
  const int64_t groups = ceil_div(P, (int64_t)ppb);
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    ncu = 256;
  const int64_t grid =
      std::max<int64_t>(1, std::min<int64_t>(groups, (int64_t)ncu * tune_value(TUNE_TIB_WG_PER_CU)));
  hipLaunchKernelGGL(taps_increment_bwd_kernel, dim3((unsigned)grid), dim3(kTibBlock), 0, s,
                     (const float*)work, P, (int)N, zc, ld_zc, ws, beta, dds, (bf16_t*)dpre,
                     ld_dpre, (bf16_t*)centre, ld_centre, (bf16_t*)q, ld_q, db0, dws);
  NA_CHECK_LAUNCH();
  return NA_OK;
}

// Partial products of na_gemm_wgrad_static: a file-scope device array (one copy
// per device, held until the process ends).  Its name is a host-side stand-in;
// kernels receive the device address from hipGetSymbolAddress, looked up once per
// device with that device current.
static __device__ __align__(256) float scratch[NA_WGRAD_STATIC_MAX_FLOATS];
static float* g_scratch_ptr[64];
static std::mutex g_scratch_mu;

static float* scratch_address(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    float* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error(std::string("na_gemm_wgrad_static: hipGetSymbolAddress failed: ") +
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

extern "C" int na_gemm_wgrad_static(const na_gemm_args* a, na_stream_t stream) {
  NA_CHECK_ARG(a != nullptr, "args is null");
  NA_CHECK_ARG(wgrad_ok(a), "weight-gradient products only (bf16, trans_a = trans_b = 1, atomic "
                            "fp32 epilogue, no epilogue terms, K >= 1, rows addressable)");
  if (a->M == 0 || a->N == 0) return NA_OK;
  const WgradPlan w = wgrad_plan(a);
  const int64_t need = w.splits * w.No * w.Ko;
  NA_CHECK_ARG(need <= NA_WGRAD_STATIC_MAX_FLOATS,
               "needs " + std::to_string(need) + " floats of partial products; the static "
               "scratch holds " + std::to_string((int64_t)NA_WGRAD_STATIC_MAX_FLOATS));
  int dev = 0;
  NA_CHECK_HIP(hipGetDevice(&dev));
  NA_CHECK_ARG(dev >= 0 && dev < 64, "device index out of range");
  float* ptr = scratch_address(dev);
  if (!ptr) return NA_ERR_HIP;
  na_gemm_args b = *a;
  b.wg_partials = ptr;
  b.wg_partials_len = NA_WGRAD_STATIC_MAX_FLOATS;
  return na_gemm(&b, stream);
}

extern "C" int na_taps_increment_fwd(const void* D, int64_t ld_d, int64_t M, int64_t K,
                                     const void* W, int64_t ldw, int64_t N, const float* zc,
                                     int64_t ld_zc, const float* ws, float beta, float* head_out,
                                     float* dz_out, na_stream_t stream) {
  return taps_increment(false, D, ld_d, M, K, W, ldw, N, zc, ld_zc, ws, beta, head_out,
                        nullptr, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, nullptr, dz_out, 0,
                        (hipStream_t)stream);
}

extern "C" int na_taps_increment_bwd(const void* D, int64_t ld_d, int64_t M, int64_t K,
                                     const void* W, int64_t ldw, int64_t N, const float* zc,
                                     int64_t ld_zc, const float* ws, float beta, const float* dds,
                                     void* dpre, int64_t ld_dpre, void* centre, int64_t ld_centre,
                                     void* q, int64_t ld_q, float* db0, float* dws, float* work,
                                     int32_t work_holds_dz, na_stream_t stream) {
  return taps_increment(true, D, ld_d, M, K, W, ldw, N, zc, ld_zc, ws, beta, nullptr, dds,
                        dpre, ld_dpre, centre, ld_centre, q, ld_q, db0, dws, work,
                        work_holds_dz, (hipStream_t)stream);
}

extern "C" int64_t na_gemm_wgrad_workspace(const na_gemm_args* a) {
  NA_CHECK_ARG(a != nullptr, "args is null");
  if (a->M <= 0 || a->N <= 0 || a->K <= 0 || !wgrad_ok(a)) return 0;
  const WgradPlan w = wgrad_plan(a);
  return w.splits * w.No * w.Ko;
}

extern "C" int na_gemm(const na_gemm_args* a, na_stream_t stream) {
  NA_CHECK_ARG(a != nullptr, "args is null");
  NA_CHECK_ARG(a->M >= 0 && a->N >= 0 && a->K >= 0, "negative size");
  if (a->M == 0 || a->N == 0) return NA_OK;
  NA_CHECK_ARG(a->in_dtype == NA_F32 || a->in_dtype == NA_BF16, "in_dtype must be F32 or BF16");
  NA_CHECK_ARG(a->A && a->B && a->C, "null operand");
  const int vec = a->in_dtype == NA_BF16 ? 8 : 4;
  NA_CHECK_ARG(a->lda % vec == 0 && a->ldb % vec == 0, "leading dimensions must be multiples of 8 (bf16) / 4 (fp32)");
  NA_CHECK_ARG(((uintptr_t)a->A & 15) == 0 && ((uintptr_t)a->B & 15) == 0, "operands must be 16-byte aligned");
  NA_CHECK_ARG(a->K % vec == 0 || wgrad_ok(a), "K must be a multiple of 8 (bf16) / 4 (fp32)");
  NA_CHECK_ARG(a->epilogue == NA_EPI_STORE || a->epilogue == NA_EPI_ATOMIC, "bad epilogue");
  NA_CHECK_ARG(a->epilogue == NA_EPI_ATOMIC || a->c_dtype == NA_F32 || a->c_dtype == NA_BF16, "bad c_dtype");
  NA_CHECK_ARG(a->epilogue == NA_EPI_STORE || a->c_dtype == NA_F32, "atomic epilogue needs fp32 C");
  int split = a->split_k < 1 ? 1 : a->split_k;
  NA_CHECK_ARG(split == 1 || a->epilogue == NA_EPI_ATOMIC, "split_k > 1 needs the atomic epilogue");
  NA_CHECK_ARG(split == 1 || (!a->act && !a->dact && !a->bias && !a->r1_a),
               "split_k > 1 allows no nonlinear epilogue terms");
  NA_CHECK_ARG(!a->dact || a->dact_y, "dact needs dact_y");
  NA_CHECK_ARG(!a->r1_a || a->r1_b, "r1_a needs r1_b");
  const bool extras = a->z_out || a->zc || a->head_w;
  NA_CHECK_ARG(!extras || (split == 1 && a->epilogue == NA_EPI_STORE),
               "z_out / zc / head need split_k = 1 and the store epilogue");
  NA_CHECK_ARG(!a->head_w || a->head_out, "head_w needs head_out");
  NA_CHECK_ARG(!a->zc || (a->act == NA_ACT_SOFTPLUS && !a->bias && !a->dact && !a->r1_a &&
                          !a->z_out && a->zc_rows >= 1),
               "delta mode needs softplus, no bias/dact/r1/z_out, zc_rows >= 1");
  NA_CHECK_ARG((a->M + GBM - 1) / GBM < 0x7fffffffLL, "M too large");
  GemmParams p;
  p.M = a->M; p.N = a->N; p.K = a->K;
  p.A = a->A; p.lda = a->lda; p.B = a->B; p.ldb = a->ldb;
  const int64_t bk = a->in_dtype == NA_BF16 ? GemmCfg<bf16_t>::BK : GemmCfg<float>::BK;
  int64_t kc = ceil_div(ceil_div(a->K, split), bk) * bk;
  if (kc < bk) kc = bk;
  p.k_chunk = kc;
  int64_t nsplit = a->K == 0 ? 1 : ceil_div(a->K, kc);
  p.e.epilogue = a->epilogue; p.e.C = a->C; p.e.ldc = a->ldc; p.e.c_dtype = a->c_dtype;
  p.e.bias = a->bias; p.e.act = a->act; p.e.beta = a->act_beta;
  p.e.dact_y = a->dact_y; p.e.ld_dact = a->ld_dact; p.e.dact_dtype = a->dact_dtype; p.e.dact = a->dact;
  p.e.r1_a = a->r1_a; p.e.r1_b = a->r1_b; p.e.colsum = a->colsum;
  p.e.z_out = a->z_out; p.e.ld_z = a->ld_z; p.e.zc = a->zc; p.e.ld_zc = a->ld_zc;
  p.e.zc_rows = a->zc_rows; p.e.head_w = a->head_w; p.e.head_out = a->head_out;
  dim3 grid((unsigned)ceil_div(a->M, GBM), (unsigned)ceil_div(a->N, GBN), (unsigned)nsplit);
  hipStream_t s = (hipStream_t)stream;
  if (wgrad_ok(a)) return launch_wgrad(a, s);
  const int kind = persist_kind(a, (int)nsplit);
  if (kind >= 0) {
    const int nk = ws_stages(a, kind);
    const int nwv = a->N == 256 ? 8 : 4;
    const int wg = nk ? ws_grid(a->M / GBM, nwv) : 0;
    if (wg > 0) {
      if (nwv == 8) launch_ws<8>(nk, kind, wg, p, s);
      else launch_ws<4>(nk, kind, wg, p, s);
      NA_CHECK_LAUNCH();
      return NA_OK;
    }
  }
  if (kind >= 0) {
    const int wn = a->N > 128 ? 4 : 2;
    const int64_t n_tiles = ceil_div(a->M, GBM) * ceil_div(a->N, (int64_t)(64 * wn));
    const int pg = persist_grid(wn, n_tiles);
    if (pg > 0) {
      if (wn == 4) launch_persist<4>(kind, pg, p, s);
      else launch_persist<2>(kind, pg, p, s);
      NA_CHECK_LAUNCH();
      return NA_OK;
    }
  }
  if (a->in_dtype == NA_BF16) return launch_gemm<bf16_t>(a, p, grid, s);
  return launch_gemm<float>(a, p, grid, s);
}

static bool narrow_vec_ok(int64_t K, int64_t ld, const void* X, int dtype) {
  const int64_t align = dtype == NA_BF16 ? 16 : 32;
  return K % 8 == 0 && K <= 512 && ld % 8 == 0 && ((uintptr_t)X % align) == 0;
}

extern "C" int na_linear_narrow_fwd(const void* X, int64_t ldx, int32_t x_dtype, int64_t M,
                                    int64_t K, const float* W, const float* b, int32_t N,
                                    int32_t act, float* out, na_stream_t stream) {
  NA_CHECK_ARG(N >= 1 && N <= kNarrowMax, "N must be in [1, 4]");
  NA_CHECK_ARG(M >= 0 && K >= 1, "bad sizes");
  if (M == 0) return NA_OK;
  NA_CHECK_ARG(X && W && b && out, "null pointer");
  NA_CHECK_ARG(narrow_vec_ok(K, ldx, X, x_dtype),
               "narrow layer needs K % 8 == 0, K <= 512, ldx % 8 == 0 and aligned rows");
  int64_t waves = ceil_div(ceil_div(M, 2), 2);
  if (waves > 16384) waves = 16384;
  dim3 grid((unsigned)ceil_div(waves, 4));
  hipStream_t s = (hipStream_t)stream;
#define NA_NFWD(TX, KC)                                                                    \
  hipLaunchKernelGGL((narrow_fwd_vec_kernel<TX, KC>), grid, dim3(256), 0, s, (const TX*)X, ldx, \
                     M, K, W, b, N, act, out)
  if (x_dtype == NA_BF16) {
    if (K <= 256) NA_NFWD(bf16_t, 1); else NA_NFWD(bf16_t, 2);
  } else {
    if (K <= 256) NA_NFWD(float, 1); else NA_NFWD(float, 2);
  }
#undef NA_NFWD
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int na_linear_narrow_bwd(const void* X, int64_t ldx, int32_t x_dtype, int64_t M,
                                    int64_t K, const float* W, int32_t N, const float* dpre,
                                    float* dW, float* db, void* dX, int64_t lddx, int32_t dx_dtype,
                                    int32_t add_dx, const void* Y, int64_t ldy, int32_t y_dtype,
                                    int32_t dact, float act_beta, float* dx_colsum,
                                    na_stream_t stream) {
  NA_CHECK_ARG(N >= 1 && N <= kNarrowMax, "N must be in [1, 4]");
  NA_CHECK_ARG(M >= 0 && K >= 1, "bad sizes");
  if (M == 0) return NA_OK;
  NA_CHECK_ARG(X && W && dpre && dW && db, "null pointer");
  NA_CHECK_ARG(!dact || Y, "dact needs Y");
  NA_CHECK_ARG(narrow_vec_ok(K, ldx, X, x_dtype) && (!dX || narrow_vec_ok(K, lddx, dX, dx_dtype)) &&
                   (!dact || narrow_vec_ok(K, ldy, Y, y_dtype)),
               "narrow layer needs K % 8 == 0, K <= 512, leading dims % 8 == 0 and aligned rows");
  // 2 workgroups per CU at most: every workgroup ends with one atomic per
  // weight / column-sum element, so fewer, longer workgroups contend less
  int64_t waves = ceil_div(M, 16);
  if (waves > 4096) waves = 4096;
  dim3 grid8((unsigned)ceil_div(waves, 8));
  hipStream_t s = (hipStream_t)stream;
  // X is the only row operand: no dX accumulation, and the derivative (if
  // any) is read from X itself
  const bool x_only = !add_dx && (!dX || !dact || (Y == X && ldy == ldx && y_dtype == x_dtype));
#define NA_NBWD_L(TX, TDX, KC, NM, XO)                                                         \
  hipLaunchKernelGGL((narrow_bwd_vec_kernel<TX, TDX, KC, NM, XO>), grid8, dim3(512), 0, s,     \
                     (const TX*)X, ldx, M, K, W, N, dpre, dW, db, (TDX*)dX, lddx, add_dx, Y,   \
                     ldy, y_dtype, dact, act_beta, dx_colsum)
#define NA_NBWD(TX, TDX, KC)                                                                   \
  do {                                                                                         \
    if (N == 1) {                                                                              \
      if (x_only) NA_NBWD_L(TX, TDX, KC, 1, true); else NA_NBWD_L(TX, TDX, KC, 1, false);      \
    } else {                                                                                   \
      if (x_only) NA_NBWD_L(TX, TDX, KC, kNarrowMax, true);                                    \
      else NA_NBWD_L(TX, TDX, KC, kNarrowMax, false);                                          \
    }                                                                                          \
  } while (0)
  const bool x16 = x_dtype == NA_BF16, d16 = dx_dtype == NA_BF16;
  if (K <= 256) {
    if (x16 && d16) NA_NBWD(bf16_t, bf16_t, 1);
    else if (x16) NA_NBWD(bf16_t, float, 1);
    else if (d16) NA_NBWD(float, bf16_t, 1);
    else NA_NBWD(float, float, 1);
  } else {
    if (x16 && d16) NA_NBWD(bf16_t, bf16_t, 2);
    else if (x16) NA_NBWD(bf16_t, float, 2);
    else if (d16) NA_NBWD(float, bf16_t, 2);
    else NA_NBWD(float, float, 2);
  }
#undef NA_NBWD
#undef NA_NBWD_L
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int na_act_grad(const void* dy, int64_t ld_dy, int32_t dy_dtype, const void* y,
                           int64_t ld_y, int32_t y_dtype, int64_t M, int64_t N, int32_t act,
                           float act_beta, void* out, int64_t ld_out, int32_t out_dtype,
                           float* colsum, na_stream_t stream) {
  NA_CHECK_ARG(M >= 0 && N >= 0, "bad sizes");
  if (M == 0 || N == 0) return NA_OK;
  NA_CHECK_ARG(dy && y && out, "null pointer");
  hipStream_t s = (hipStream_t)stream;
  const bool d16 = dy_dtype == NA_BF16, y16 = y_dtype == NA_BF16, o16 = out_dtype == NA_BF16;
  // whole 16-byte chunks: the vectorised kernel, about two blocks per CU
  const bool vec = N % 8 == 0 && 256 % (N / 8) == 0 && ld_dy % 8 == 0 && ld_y % 8 == 0 &&
                   ld_out % 8 == 0 && ((uintptr_t)dy | (uintptr_t)y | (uintptr_t)out) % 16 == 0 &&
                   (d16 || ld_dy % 4 == 0);
  if (vec && d16 && y16) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    const int64_t lanes = 256 / (N / 8);
    int64_t blocks = 2 * (int64_t)cus;
    int64_t rpb = ceil_div(ceil_div(M, blocks), lanes) * lanes;
    if (rpb < 64) rpb = 64;
    blocks = ceil_div(M, rpb);
    if (o16)
      hipLaunchKernelGGL((act_grad_vec_kernel<bf16_t, bf16_t, bf16_t>), dim3((unsigned)blocks), dim3(256), 0,
                         s, (const bf16_t*)dy, ld_dy, (const bf16_t*)y, ld_y, M, N, act, act_beta,
                         (bf16_t*)out, ld_out, colsum, rpb);
    else
      hipLaunchKernelGGL((act_grad_vec_kernel<bf16_t, bf16_t, float>), dim3((unsigned)blocks), dim3(256), 0,
                         s, (const bf16_t*)dy, ld_dy, (const bf16_t*)y, ld_y, M, N, act, act_beta,
                         (float*)out, ld_out, colsum, rpb);
    NA_CHECK_LAUNCH();
    return NA_OK;
  }
  dim3 grid((unsigned)ceil_div(M, 256), (unsigned)ceil_div(N, 64));
#define NA_ACTG(TD, TY, TO)                                                                   \
  hipLaunchKernelGGL((act_grad_kernel<TD, TY, TO>), grid, dim3(256), 0, s, (const TD*)dy, ld_dy, \
                     (const TY*)y, ld_y, M, N, act, act_beta, (TO*)out, ld_out, colsum)
  if (!d16 && !y16 && !o16) NA_ACTG(float, float, float);
  else if (!d16 && !y16 && o16) NA_ACTG(float, float, bf16_t);
  else if (!d16 && y16 && !o16) NA_ACTG(float, bf16_t, float);
  else if (!d16 && y16 && o16) NA_ACTG(float, bf16_t, bf16_t);
  else if (d16 && !y16 && !o16) NA_ACTG(bf16_t, float, float);
  else if (d16 && !y16 && o16) NA_ACTG(bf16_t, float, bf16_t);
  else if (d16 && y16 && !o16) NA_ACTG(bf16_t, bf16_t, float);
  else NA_ACTG(bf16_t, bf16_t, bf16_t);
#undef NA_ACTG
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int na_weight_norm_fwd(const float* v, const float* g, int64_t N, int64_t K, float* W,
                                  void* Wlow, int32_t low_dtype, int64_t ld_low,
                                  na_stream_t stream) {
  NA_CHECK_ARG(N >= 1 && K >= 1, "bad sizes");
  NA_CHECK_ARG(v && g && W, "null pointer");
  NA_CHECK_ARG(!Wlow || ld_low >= K, "ld_low < K");
  hipLaunchKernelGGL(weight_norm_fwd_kernel, dim3((unsigned)N), dim3(256), 0, (hipStream_t)stream,
                     v, g, K, W, Wlow, low_dtype, ld_low);
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int na_weight_norm_bwd(const float* v, const float* g, const float* dW, int64_t N,
                                  int64_t K, float* dv, float* dg, na_stream_t stream) {
  NA_CHECK_ARG(N >= 1 && K >= 1, "bad sizes");
  NA_CHECK_ARG(v && g && dW && dv && dg, "null pointer");
  hipLaunchKernelGGL(weight_norm_bwd_kernel, dim3((unsigned)N), dim3(256), 0, (hipStream_t)stream,
                     v, g, dW, K, dv, dg);
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int na_head_sum(const float* parts, int32_t nparts, int64_t M, const float* bias,
                           float* out, na_stream_t stream) {
  NA_CHECK_ARG(M >= 0 && nparts >= 1, "bad sizes");
  if (M == 0) return NA_OK;
  NA_CHECK_ARG(parts && out, "null pointer");
  hipLaunchKernelGGL(head_sum_kernel, dim3((unsigned)ceil_div(M, (int64_t)256)), dim3(256), 0,
                     (hipStream_t)stream, parts, nparts, M, bias, out);
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int na_colsum(const void* X, int64_t ld, int32_t x_dtype, int64_t M, int64_t N,
                         float* out, na_stream_t stream) {
  NA_CHECK_ARG(M >= 0 && N >= 0 && ld >= N, "bad sizes");
  if (M == 0 || N == 0) return NA_OK;
  NA_CHECK_ARG(X && out, "null pointer");
  NA_CHECK_ARG(ceil_div(M, (int64_t)kColSlab) < 65536, "M too large");
  dim3 grid((unsigned)ceil_div(N, (int64_t)256), (unsigned)ceil_div(M, (int64_t)kColSlab));
  hipStream_t s = (hipStream_t)stream;
  if (x_dtype == NA_BF16)
    hipLaunchKernelGGL(colsum_kernel<bf16_t>, grid, dim3(256), 0, s, (const bf16_t*)X, ld, M, N, out);
  else if (x_dtype == NA_F32)
    hipLaunchKernelGGL(colsum_kernel<float>, grid, dim3(256), 0, s, (const float*)X, ld, M, N, out);
  else {
    set_error("colsum: x dtype must be f32 or bf16");
    return NA_ERR_UNSUPPORTED;
  }
  NA_CHECK_LAUNCH();
  return NA_OK;
}
