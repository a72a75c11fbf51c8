// Fused AdamW over a flat parameter tensor (HBM-bound: 7 streams of 4 bytes
// per parameter, plus 2 bytes for the optional bf16 copy).
#include <math.h>

#include "common.h"

namespace na {

struct AdamArgs {
  float lr, beta1, beta2, eps, decay;   // decay = 1 - lr * weight_decay
  float step_size, inv_bc2_sqrt;        // lr / bc1, 1 / sqrt(bc2)
};

__device__ inline float lerp_torch(float a, float b, float w) {
  // torch.lerp: a + w (b - a) for |w| < 0.5, else b - (b - a)(1 - w)
  return fabsf(w) < 0.5f ? a + w * (b - a) : b - (b - a) * (1.0f - w);
}

// Streaming accesses: every byte is touched once per step, so loads and stores
// carry the non-temporal hint.
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef unsigned int nt_u2 __attribute__((ext_vector_type(2)));
__device__ inline float4 ld_nt(const float* p) {
  const nt_f4 r = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
  return make_float4(r.x, r.y, r.z, r.w);
}
__device__ inline void st_nt(float* p, const float4& v) {
  const nt_f4 r = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(r, reinterpret_cast<nt_f4*>(p));
}

// The update of the 4-parameter chunk at i (element-wise tail past n - 4).
template <bool LOWP, bool CLEAR = false>
__device__ inline void adam_chunk(float* __restrict__ p, const float* __restrict__ g,
                                  float* __restrict__ m, float* __restrict__ v, int64_t n,
                                  int64_t i, const AdamArgs& a, bf16_t* __restrict__ lowp) {
  if (i >= n) return;
  if (i + 4 <= n) {
    float4 pp = ld_nt(p + i);
    float4 gg = ld_nt(g + i);
    if (CLEAR) st_nt(const_cast<float*>(g) + i, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    float4 mm = ld_nt(m + i);
    float4 vv = ld_nt(v + i);
    float* P = &pp.x; const float* G = &gg.x; float* Mv = &mm.x; float* V = &vv.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float pe = P[e] * a.decay;
      float me = lerp_torch(Mv[e], G[e], 1.0f - a.beta1);
      float ve = V[e] * a.beta2 + (1.0f - a.beta2) * G[e] * G[e];
      float den = sqrtf(ve) * a.inv_bc2_sqrt + a.eps;
      pe = pe - a.step_size * me / den;
      P[e] = pe; Mv[e] = me; V[e] = ve;
    }
    st_nt(p + i, pp);
    st_nt(m + i, mm);
    st_nt(v + i, vv);
    if (LOWP) {
      const nt_u2 w = {(uint32_t)f32_to_bf16(pp.x) | ((uint32_t)f32_to_bf16(pp.y) << 16),
                       (uint32_t)f32_to_bf16(pp.z) | ((uint32_t)f32_to_bf16(pp.w) << 16)};
      __builtin_nontemporal_store(w, reinterpret_cast<nt_u2*>(lowp + i));
    }
  } else {
    for (int64_t j = i; j < n; ++j) {
      const float gj = g[j];
      if (CLEAR) const_cast<float*>(g)[j] = 0.0f;
      float pe = p[j] * a.decay;
      float me = lerp_torch(m[j], gj, 1.0f - a.beta1);
      float ve = v[j] * a.beta2 + (1.0f - a.beta2) * gj * gj;
      float den = sqrtf(ve) * a.inv_bc2_sqrt + a.eps;
      pe = pe - a.step_size * me / den;
      p[j] = pe; m[j] = me; v[j] = ve;
      if (LOWP) lowp[j] = f32_to_bf16(pe);
    }
  }
}

// One chunk of 4 parameters per thread (no grid-stride loop: the launch holds
// n / 4 threads, which keeps the most loads in flight).
template <bool LOWP, bool CLEAR>
__global__ void __launch_bounds__(256)
adamw_kernel(float* __restrict__ p, float* __restrict__ g, float* __restrict__ m,
             float* __restrict__ v, int64_t n, AdamArgs a, bf16_t* __restrict__ lowp) {
  adam_chunk<LOWP, CLEAR>(p, g, m, v, n, ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4, a,
                          lowp);
}

// Several tensors in one launch: the list travels in the kernel arguments;
// tensor t owns workgroups [first_block[t], first_block[t + 1]).
struct AdamMultiList {
  float* p[NA_ADAMW_MULTI_MAX];
  const float* g[NA_ADAMW_MULTI_MAX];
  float* m[NA_ADAMW_MULTI_MAX];
  float* v[NA_ADAMW_MULTI_MAX];
  int64_t n[NA_ADAMW_MULTI_MAX];
  int64_t first_block[NA_ADAMW_MULTI_MAX + 1];
  int count;
};

__global__ void __launch_bounds__(256) adamw_multi_kernel(AdamMultiList L, AdamArgs a) {
  const int64_t b = blockIdx.x;
  int t = 0;
  while (t + 1 < L.count && L.first_block[t + 1] <= b) ++t;
  const int64_t i = ((b - L.first_block[t]) * blockDim.x + threadIdx.x) * 4;
  adam_chunk<false>(L.p[t], L.g[t], L.m[t], L.v[t], L.n[t], i, a, nullptr);
}

}  // namespace na

using namespace na;

static AdamArgs adam_args(float lr, float beta1, float beta2, float eps, float weight_decay,
                           int64_t step) {
  AdamArgs a;
  a.lr = lr; a.beta1 = beta1; a.beta2 = beta2; a.eps = eps;
  a.decay = (float)(1.0 - (double)lr * (double)weight_decay);
  const double bc1 = 1.0 - pow((double)beta1, (double)step);
  const double bc2 = 1.0 - pow((double)beta2, (double)step);
  a.step_size = (float)((double)lr / bc1);
  a.inv_bc2_sqrt = (float)(1.0 / sqrt(bc2));
  return a;
}

extern "C" int na_adamw_step(float* param, float* grad, float* exp_avg, float* exp_avg_sq,
                             int64_t n, float lr, float beta1, float beta2, float eps,
                             float weight_decay, int64_t step, void* lowp, int32_t lowp_dtype,
                             int32_t clear_grad, na_stream_t stream) {
  NA_CHECK_ARG(n >= 0 && step >= 1, "bad sizes / step");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(param && grad && exp_avg && exp_avg_sq, "null pointer");
  NA_CHECK_ARG(((uintptr_t)param | (uintptr_t)grad | (uintptr_t)exp_avg | (uintptr_t)exp_avg_sq) % 16 == 0,
               "buffers must be 16-byte aligned");
  NA_CHECK_ARG(!lowp || lowp_dtype == NA_BF16, "lowp copy must be bf16");
  NA_CHECK_ARG(!lowp || (uintptr_t)lowp % 8 == 0, "lowp copy must be 8-byte aligned");
  const AdamArgs a = adam_args(lr, beta1, beta2, eps, weight_decay, step);
  const int64_t threads = ceil_div(n, 4);
  const int64_t blocks = ceil_div(threads, 256);
  NA_CHECK_ARG(blocks <= 0x7fffffffLL, "too many parameters for one launch");
  hipStream_t s = (hipStream_t)stream;
#define NA_ADAM(L, C)                                                                          \
  hipLaunchKernelGGL((adamw_kernel<L, C>), dim3((unsigned)blocks), dim3(256), 0, s, param, grad, \
                     exp_avg, exp_avg_sq, n, a, (bf16_t*)lowp)
  if (lowp) {
    if (clear_grad) NA_ADAM(true, true); else NA_ADAM(true, false);
  } else {
    if (clear_grad) NA_ADAM(false, true); else NA_ADAM(false, false);
  }
#undef NA_ADAM
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int na_adamw_step_multi(const na_adamw_tensor* tensors, int32_t count, float lr,
                                   float beta1, float beta2, float eps, float weight_decay,
                                   int64_t step, na_stream_t stream) {
  NA_CHECK_ARG(count >= 0 && count <= NA_ADAMW_MULTI_MAX, "count must be in [0, NA_ADAMW_MULTI_MAX]");
  NA_CHECK_ARG(step >= 1, "bad step");
  NA_CHECK_ARG(count == 0 || tensors, "null tensor list");
  AdamMultiList L;
  L.count = 0;
  int64_t blocks = 0;
  for (int i = 0; i < count; ++i) {
    const na_adamw_tensor& t = tensors[i];
    NA_CHECK_ARG(t.n >= 0, "bad size");
    if (t.n == 0) continue;
    NA_CHECK_ARG(t.param && t.grad && t.exp_avg && t.exp_avg_sq, "null pointer");
    NA_CHECK_ARG(((uintptr_t)t.param | (uintptr_t)t.grad | (uintptr_t)t.exp_avg |
                  (uintptr_t)t.exp_avg_sq) % 16 == 0,
                 "buffers must be 16-byte aligned");
    const int c = L.count++;
    L.p[c] = t.param; L.g[c] = t.grad; L.m[c] = t.exp_avg; L.v[c] = t.exp_avg_sq; L.n[c] = t.n;
    L.first_block[c] = blocks;
    blocks += ceil_div(ceil_div(t.n, (int64_t)4), (int64_t)256);
  }
  if (L.count == 0) return NA_OK;
  L.first_block[L.count] = blocks;
  NA_CHECK_ARG(blocks <= 0x7fffffffLL, "too many parameters for one launch");
  const AdamArgs a = adam_args(lr, beta1, beta2, eps, weight_decay, step);
  hipLaunchKernelGGL(adamw_multi_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     L, a);
  NA_CHECK_LAUNCH();
  return NA_OK;
}
