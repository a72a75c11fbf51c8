// Background NeRF of the reference's object+background renderer
// (projects/neuralangelo/utils/modules.py:232-311, projects/neuralangelo/model.py:192-205,
// projects/nerf/utils/render.py:68-84): the inverted-sphere Fourier input of its MLP
// and the density -> alpha quadrature.  The MLP layers themselves run on the GEMM
// kernels.
#include <math.h>

#include "common.h"

namespace na {

// One thread per point: q = [p / |p|, 1 / |p|] and its Fourier code, written as
// [q(4) | sin(q_d 2^l pi), l < L | cos(q_d 2^l pi), l < L  for d = 0..3 | 0 pad].
// The frequencies are float32(2^l) * float32(pi), the products float32, as torch
// evaluates nerf_util.positional_encoding (nerf_util.py:134-147).
template <typename TO>
__global__ void __launch_bounds__(256)
bg_encode_kernel(const float* __restrict__ pts, int64_t n, int L, TO* __restrict__ out, int64_t ld) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = pts[i * 3], y = pts[i * 3 + 1], z = pts[i * 3 + 2];
  const float r = sqrtf(x * x + y * y + z * z);
  const float q[4] = {x / r, y / r, z / r, 1.0f / r};
  TO* row = out + i * ld;
#pragma unroll
  for (int d = 0; d < 4; ++d) scalar_io<TO>::store(row + d, q[d]);
  const float pi = 3.14159265358979323846f;
  for (int d = 0; d < 4; ++d) {
    float freq = pi;
    for (int l = 0; l < L; ++l) {
      const float s = q[d] * freq;
      scalar_io<TO>::store(row + 4 + d * 2 * L + l, sinf(s));
      scalar_io<TO>::store(row + 4 + d * 2 * L + L + l, cosf(s));
      freq *= 2.0f;
    }
  }
  for (int64_t c = 4 + 8 * L; c < ld; ++c) scalar_io<TO>::store(row + c, 0.0f);
}

// softplus (beta 1, linear above 20, torch semantics) and its derivative
__device__ inline float softplus1(float v) { return v > 20.0f ? v : log1pf(expf(v)); }
__device__ inline float sigmoid1(float v) { return 1.0f / (1.0f + expf(-v)); }

// alpha_i = 1 - exp(-softplus(o_i) (d_{i+1} - d_i)), d_N = 1e10 (render.py:68-84);
// o = column 0 of the MLP output rows (stride ld), samples [R, N] row-major.
__global__ void __launch_bounds__(256)
bg_alpha_fwd_kernel(const float* __restrict__ o, int64_t ld, const float* __restrict__ dists,
                    int64_t R, int N, float* __restrict__ density, float* __restrict__ alpha) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R * N) return;
  const int i = (int)(t % N);
  const float sigma = softplus1(o[t * ld]);
  const float dnext = i + 1 < N ? dists[t + 1] : 1e10f;
  const float sd = sigma * (dnext - dists[t]);
  if (density) density[t] = sigma;
  alpha[t] = 1.0f - expf(-sd);
}

__global__ void __launch_bounds__(256)
bg_alpha_bwd_kernel(const float* __restrict__ o, int64_t ld, const float* __restrict__ dists,
                    int64_t R, int N, const float* __restrict__ dalpha,
                    const float* __restrict__ ddensity, float* __restrict__ dout, int64_t ld_d) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R * N) return;
  const int i = (int)(t % N);
  const float v = o[t * ld];
  const float sigma = softplus1(v);
  const float dnext = i + 1 < N ? dists[t + 1] : 1e10f;
  const float delta = dnext - dists[t];
  float ds = dalpha ? dalpha[t] * expf(-sigma * delta) * delta : 0.0f;
  if (ddensity) ds += ddensity[t];
  dout[t * ld_d] = ds * (v > 20.0f ? 1.0f : sigmoid1(v));
}

}  // namespace na

using namespace na;

extern "C" int na_bg_encode(const float* points, int64_t n, int32_t levels, void* out,
                            int32_t out_dtype, int64_t ld, na_stream_t stream) {
  NA_CHECK_ARG(n >= 0 && levels >= 0 && ld >= 4 + 8 * (int64_t)levels, "bad sizes");
  if (n == 0) return NA_OK;
  NA_CHECK_ARG(points && out, "null pointer");
  const dim3 grid((unsigned)ceil_div(n, (int64_t)256));
  hipStream_t s = (hipStream_t)stream;
  if (out_dtype == NA_BF16)
    hipLaunchKernelGGL(bg_encode_kernel<bf16_t>, grid, dim3(256), 0, s, points, n, levels,
                       (bf16_t*)out, ld);
  else if (out_dtype == NA_F32)
    hipLaunchKernelGGL(bg_encode_kernel<float>, grid, dim3(256), 0, s, points, n, levels,
                       (float*)out, ld);
  else {
    set_error("na_bg_encode: out dtype must be f32 or bf16");
    return NA_ERR_UNSUPPORTED;
  }
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int na_bg_alpha_fwd(const float* mlp_out, int64_t ld, const float* dists, int64_t R,
                               int32_t N, float* density, float* alpha, na_stream_t stream) {
  NA_CHECK_ARG(R >= 0 && N >= 1 && ld >= 1, "bad sizes");
  if (R == 0) return NA_OK;
  NA_CHECK_ARG(mlp_out && dists && alpha, "null pointer");
  hipLaunchKernelGGL(bg_alpha_fwd_kernel, dim3((unsigned)ceil_div(R * N, (int64_t)256)), dim3(256),
                     0, (hipStream_t)stream, mlp_out, ld, dists, R, N, density, alpha);
  NA_CHECK_LAUNCH();
  return NA_OK;
}

extern "C" int na_bg_alpha_bwd(const float* mlp_out, int64_t ld, const float* dists, int64_t R,
                               int32_t N, const float* dalpha, const float* ddensity,
                               float* dmlp_out, int64_t ld_d, na_stream_t stream) {
  NA_CHECK_ARG(R >= 0 && N >= 1 && ld >= 1 && ld_d >= 1, "bad sizes");
  if (R == 0) return NA_OK;
  NA_CHECK_ARG(mlp_out && dists && dmlp_out && (dalpha || ddensity), "null pointer");
  hipLaunchKernelGGL(bg_alpha_bwd_kernel, dim3((unsigned)ceil_div(R * N, (int64_t)256)), dim3(256),
                     0, (hipStream_t)stream, mlp_out, ld, dists, R, N, dalpha, ddensity, dmlp_out,
                     ld_d);
  NA_CHECK_LAUNCH();
  return NA_OK;
}
