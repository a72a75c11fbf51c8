/*
 * neuralangelo_amd — C ABI of the MI355X (gfx950) hot-path library.
 *
 * Conventions (every entry point):
 *   - arguments are raw device pointers, int64 element counts and POD structs;
 *     no C++ types and no exceptions cross this boundary;
 *   - buffers are owned by the caller; the library never allocates device
 *     memory on these paths;
 *   - work is enqueued on `stream` (a hipStream_t, e.g. torch's current stream)
 *     and the call returns without synchronising;
 *   - the return value is NA_OK (0) or a negative status; na_last_error()
 *     returns a thread-local message describing the last failure.
 *
 * Each declaration names the reference interface it replaces
 * (NVlabs/neuralangelo path:line, or the third-party call made there).
 */
#ifndef NEURALANGELO_AMD_H
#define NEURALANGELO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NA_OK 0
#define NA_ERR_ARGUMENT (-1)
#define NA_ERR_HIP (-2)
#define NA_ERR_UNSUPPORTED (-3)

/* element types */
#define NA_F32 0
#define NA_BF16 1
#define NA_F64 3

#define NA_MAX_LEVELS 32

typedef void* na_stream_t; /* hipStream_t */

/* Host-computed multiresolution hash-grid level table.  Both the kernels and
 * the host read this one table, so hash indices never depend on device
 * transcendental functions.  offset[] and the per-level sizes are in table
 * entries (one entry = n_features parameters). */
typedef struct na_grid_desc {
  int32_t n_levels;
  int32_t n_features;        /* 1, 2, 4 or 8 */
  int32_t log2_hashmap_size;
  int32_t base_resolution;
  float per_level_scale;
  float log2_per_level_scale;
  float scale[NA_MAX_LEVELS];           /* exp2(l*log2 s)*base - 1, float32 */
  uint32_t resolution[NA_MAX_LEVELS];   /* ceil(scale)+1 */
  uint32_t offset[NA_MAX_LEVELS + 1];   /* prefix of per-level entry counts */
} na_grid_desc;

/* library */
const char* na_last_error(void);
int na_version(void);

/* Launch tunables (no counterpart in the reference, which has no launches of its
 * own to shape; the project's autotuner, neuralangelo_amd/autotune.py, picks them):
 *   "taps_bwd.run"                   samples per thread of the tap scatter (1..32, 4)
 *   "taps_increment_bwd.wg_per_cu"   workgroups per CU of the tap-increment
 *                                    backward (1..32, 8)
 *   "packed.rays_per_wg"             rays per workgroup of the packed compositing
 *                                    kernels (64..256, multiple of 64; 64)
 *   "taps_increment_bwd.fused"       tap-increment backward in the epilogue of
 *                                    its recomputed product (1, no workspace) or
 *                                    split in two passes (0); 0
 *   "packed.wg_per_cu"               workgroups per CU of the persistent packed
 *                                    launches (1..16; 4)
 * na_tune_set returns the previous value (or an error code for an unknown name or
 * a value out of range); the launchers read the current value at every launch.
 * Set them between launches: a change is not ordered with launches in flight on
 * other host threads.  na_tune_count / na_tune_name enumerate the names. */
int64_t na_tune_set(const char* name, int64_t value);
int64_t na_tune_get(const char* name);
int na_tune_count(void);
const char* na_tune_name(int32_t index);
int na_device_synchronize(void);

/* ---------------------------------------------------------------------------
 * Hash-grid encoding (replaces tcnn.Encoding(3, {"otype": "HashGrid", ...})
 * constructed at projects/neuralangelo/utils/modules.py:42-50 and called at
 * projects/neuralangelo/utils/modules.py:84-86).
 * ------------------------------------------------------------------------- */

/* Fill *desc (host only, no GPU).  Returns the total number of parameters
 * (entries * n_features) or a negative status. */
int64_t na_hashgrid_layout(int32_t n_levels, int32_t n_features,
                           int32_t log2_hashmap_size, int32_t base_resolution,
                           float per_level_scale, na_grid_desc* desc);

/* Debug/parity entry: per point, level and corner, the table entry index
 * (offset included) and the trilinear weight.  idx: int32 [n, L, 8],
 * w: f32 [n, L, 8]. */
int na_hashgrid_corners(const na_grid_desc* desc, const float* x, int64_t n,
                        int32_t* idx, float* w, na_stream_t stream);

/* Forward.  x: f32 [n,3] in the unit cube (tcnn input convention).
 * table: [n_params] of table_dtype (NA_F32 | NA_BF16).
 * out: [n, L*F] of out_dtype (NA_F32 | NA_BF16), level-major.
 * Levels >= active_levels are written as zero (coarse-to-fine mask of
 * projects/neuralangelo/utils/modules.py:91-93,109-113). */
int na_hashgrid_fwd(const na_grid_desc* desc, const float* x, int64_t n,
                    const void* table, int32_t table_dtype, void* out,
                    int32_t out_dtype, int32_t active_levels,
                    na_stream_t stream);

/* Backward w.r.t. the table: grad_table[(offset_l+idx)*F+f] +=
 * w * dy[p, l*F+f] (fp32 atomics).  dy: [n, L*F] of dy_dtype. */
int na_hashgrid_bwd(const na_grid_desc* desc, const float* x, int64_t n,
                    const void* dy, int32_t dy_dtype, float* grad_table,
                    int32_t active_levels, na_stream_t stream);

/* ---------------------------------------------------------------------------
 * Dense layers of the SDF and colour MLPs (replace torch.nn.Linear +
 * activation at projects/neuralangelo/utils/mlp.py:55-69 and
 * projects/nerf/utils/nerf_util.py:189-196; weight-normed weights
 * W = g * v / ||v||_row as torch.nn.utils.weight_norm, mlp.py:42-43).
 * ------------------------------------------------------------------------- */
#define NA_ACT_NONE 0
#define NA_ACT_RELU 1
#define NA_ACT_SOFTPLUS 2 /* log1p(exp(beta*x))/beta, linear where beta*x > 20 */
#define NA_ACT_SIGMOID 3

#define NA_EPI_STORE 0  /* C[m,n] = value (c_dtype) */
#define NA_EPI_ATOMIC 1 /* C[m,n] += value (fp32 atomics; needed when split_k > 1) */

/* C[m,n] = epilogue( sum_k A(m,k) * B(n,k) ),  with
 *   A(m,k) = trans_a ? A[k*lda + m] : A[m*lda + k]
 *   B(n,k) = trans_b ? B[k*ldb + n] : B[n*ldb + k]
 * epilogue, in order: v += bias[n]; v = act(v); v += r1_a[m]*r1_b[n];
 *   v *= act'(dact_y[m*ld_dact+n]) (derivative expressed through the
 *   activation's output y); colsum[n] += v (atomic); store/accumulate v.
 * A and B share in_dtype (NA_F32: fp32 MFMA, NA_BF16: bf16 MFMA); the
 * accumulation is fp32.  K and the leading dimensions must be multiples of 8
 * (bf16) or 4 (fp32) and the pointers 16-byte aligned. */
typedef struct na_gemm_args {
  int64_t M, N, K;
  const void* A;
  int64_t lda;
  int32_t trans_a;
  const void* B;
  int64_t ldb;
  int32_t trans_b;
  int32_t in_dtype;
  int32_t epilogue;
  void* C;
  int64_t ldc;
  int32_t c_dtype;
  const float* bias;
  int32_t act;
  float act_beta;
  const void* dact_y;
  int64_t ld_dact;
  int32_t dact_dtype;
  int32_t dact;
  const float* r1_a;
  const float* r1_b;
  float* colsum;
  int32_t split_k;
  /* optional extras (NULL = off; store epilogue, split_k = 1):
   *   z_out:    fp32 copy of the pre-activation value (after bias) [M, ld_z];
   *   zc:       delta mode -- the accumulator is an increment dz of a stored
   *             fp32 pre-activation z0 = zc[(m % zc_rows) * ld_zc + n]; the
   *             output is act(z0 + dz) (act must be NA_ACT_SOFTPLUS, no bias);
   *   head_w:   a narrow head fused into the epilogue: head_out[part * M + m] =
   *             sum over the tile's columns n of head_w[n] * h, h = the
   *             activated output, or act(z0 + dz) - act(z0) in delta mode
   *             (evaluated without cancellation); part = n / 32, i.e.
   *             4 * ceil(N / 128) partial sums per row (zero past N). */
  float* z_out;
  int64_t ld_z;
  const float* zc;
  int64_t ld_zc;
  int64_t zc_rows;
  const float* head_w;
  float* head_out;
  /* optional workspace of the weight-gradient path (bf16, trans_a = trans_b = 1,
   * atomic epilogue, no epilogue terms): each row split stores its partial
   * product here and one more launch adds the splits into C in split order
   * (no atomics; the result does not depend on scheduling).  wg_partials_len
   * floats, at least na_gemm_wgrad_workspace(args); NULL or shorter = atomics. */
  float* wg_partials;
  int64_t wg_partials_len;
} na_gemm_args;

int na_gemm(const na_gemm_args* args, na_stream_t stream);

/* Tap increments of the BF16 4-tap numerical gradient
 * (projects/neuralangelo/utils/modules.py:157-175 with the SDF MLP of
 * projects/neuralangelo/utils/mlp.py:55-69), on the weights-stationary kernel.
 * Rows are point-major: row m is tap m % 4 of point m / 4 (M = 4 * points, a
 * multiple of 128).  D [M, ld_d] bf16 holds the first layer's input increments
 * X(x + eps k_t) - X(x) (na_sdf_taps_increment_input), W [N, ldw] bf16 the
 * matching first-layer weights, zc [M/4, ld_zc] fp32 the centre pre-activations
 * z = W0 X(x) + b0, ws [N] fp32 the SDF head; N = 128 or 256, K in 129..160 or
 * 225..288.  With dz = D W^T and softplus sp (beta):
 *   fwd: head_out[w * M + m] = sum over wave w's 32 columns n of
 *        ws[n] (sp(z + dz) - sp(z)) (evaluated without cancellation; N / 32 parts);
 *   bwd: with dds [M] the gradients of the four increments of each point, a = their
 *        mean (the part every tap shares) and b_t = dds - a, and dsp'_t =
 *        sp'(z + dz_t) - sp'(z):
 *          dpre[m, n] = ws[n] (b_t sp'(z + dz_t) + a dsp'_t)  (bf16; pairs with D_t),
 *          q[p, n] = a ws[n] sp'(z)                           (bf16; pairs with sum_t D_t),
 *          centre[p, n] += sum_t dds_t ws[n] dsp'_t  (bf16 read-modify-write: centre
 *                          holds the centre's own pre-activation gradient),
 *          db0[n] += column sums of that addition, dws[n] += sum_m dds[m] (sp(z + dz) -
 *          sp(z)).
 *        Together these are the gradient of the four sp(z + dz_t) - sp(z) split into
 *        terms of their natural size.  Increments and their derivatives are formed
 *        in fp32 from z and dz, so the second-order sums (the Hessian trace and its
 *        gradient) keep fp32 precision relative to their own size.  The backward
 *        recomputes dz = D W^T: into work [M, N], a caller-owned 16-byte aligned
 *        workspace of M * N floats, followed by an elementwise pass (the launch
 *        tunable "taps_increment_bwd.fused" = 0, the default), or in the epilogue
 *        of that product, where dz never leaves the registers and no workspace is
 *        needed (= 1; work may be NULL); dds [M] is 16-byte aligned (four taps per
 *        point).  The forward can keep dz: dz_out [M, N] fp32 (16-byte aligned, or
 *        NULL); a backward given that buffer as work with work_holds_dz = 1 skips
 *        the recomputation (split form; the fused form ignores it). */
int na_taps_increment_fwd(const void* D, int64_t ld_d, int64_t M, int64_t K, const void* W,
                          int64_t ldw, int64_t N, const float* zc, int64_t ld_zc,
                          const float* ws, float beta, float* head_out, float* dz_out,
                          na_stream_t stream);
int na_taps_increment_bwd(const void* D, int64_t ld_d, int64_t M, int64_t K, const void* W,
                          int64_t ldw, int64_t N, const float* zc, int64_t ld_zc,
                          const float* ws, float beta, const float* dds, void* dpre,
                          int64_t ld_dpre, void* centre, int64_t ld_centre, void* q,
                          int64_t ld_q, float* db0, float* dws, float* work,
                          int32_t work_holds_dz, na_stream_t stream);

/* Maximum partial-product floats of na_gemm_wgrad_static (80 MiB per device):
 * every weight gradient of the train step fits (the largest, the 280-wide colour
 * input layer over 256 row splits, needs 18.35 M floats). */
#define NA_WGRAD_STATIC_MAX_FLOATS 20971520

/* No counterpart in the reference.  na_gemm's deterministic weight-gradient path
 * (the split-ordered partial-product reduction that a wg_partials workspace
 * selects) without any allocation or workspace argument: the partial products
 * live in a file-scope device array of NA_WGRAD_STATIC_MAX_FLOATS floats, one per
 * device, held until the process ends.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments as na_gemm for a weight-gradient product (bf16, trans_a = trans_b = 1,
 * NA_EPI_ATOMIC, fp32 C, no epilogue terms); args->wg_partials is ignored.
 * Returns NA_ERR_ARGUMENT for other products or when the plan needs more than
 * NA_WGRAD_STATIC_MAX_FLOATS floats.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (concurrent streams or threads) race. */
int na_gemm_wgrad_static(const na_gemm_args* args, na_stream_t stream);

/* Floats of wg_partials that na_gemm's weight-gradient path uses for these
 * arguments (0 when another path runs them; NA_ERR_ARGUMENT for a NULL args). */
int64_t na_gemm_wgrad_workspace(const na_gemm_args* args);

/* out[n] += sum_m X[m*ld + n] (fp32 accumulation, X of x_dtype): the bias gradient
 * of a linear layer from its pre-activation gradient (torch.nn.Linear backward,
 * mlp.py:55-69 / nerf_util.py:189-196).  Accumulates: the caller zeroes out. */
int na_colsum(const void* X, int64_t ld, int32_t x_dtype, int64_t M, int64_t N, float* out,
              na_stream_t stream);

/* out[m] = (bias ? bias[0] : 0) + sum_{p < nparts} parts[p*M + m] (fixed order):
 * completes the fused narrow head of na_gemm (head_out partial sums). */
int na_head_sum(const float* parts, int32_t nparts, int64_t M, const float* bias, float* out,
                na_stream_t stream);

/* Narrow output layer (N <= 4): out[m,n] = act(sum_k X[m*ldx+k] W[n*K+k] + b[n]),
 * X of x_dtype, W/b fp32, out fp32 [M, N].  Used for the SDF head
 * (mlp.py:50-52,66) and the colour output layer + sigmoid (modules.py:219). */
int na_linear_narrow_fwd(const void* X, int64_t ldx, int32_t x_dtype, int64_t M, int64_t K,
                         const float* W, const float* b, int32_t N, int32_t act,
                         float* out, na_stream_t stream);

/* Backward of the narrow layer.  dpre [M,N] fp32 is the gradient w.r.t. the
 * pre-activation.  dW[n,k] += sum_m dpre[m,n] X[m,k]; db[n] += sum_m dpre[m,n];
 * if dX != NULL: dX[m,k] = (add_dx ? dX[m,k] : 0) + sum_n dpre[m,n] W[n,k], then
 * multiplied by act'(Y[m*ldy+k]) when dact != NA_ACT_NONE (Y of y_dtype); dX
 * is written as dx_dtype. */
int na_linear_narrow_bwd(const void* X, int64_t ldx, int32_t x_dtype, int64_t M, int64_t K,
                         const float* W, int32_t N, const float* dpre, float* dW, float* db,
                         void* dX, int64_t lddx, int32_t dx_dtype, int32_t add_dx,
                         const void* Y, int64_t ldy, int32_t y_dtype, int32_t dact,
                         float act_beta, float* dx_colsum, na_stream_t stream);

/* out[m, n] = dy[m*ld_dy + n] * act'(y[m*ld_y + n]) for n < N (derivative through
 * the activation output y, as na_gemm's dact); colsum[n] += out[m,n] when
 * colsum != NULL.  Converts an output gradient into a pre-activation gradient. */
int na_act_grad(const void* dy, int64_t ld_dy, int32_t dy_dtype, const void* y, int64_t ld_y,
                int32_t y_dtype, int64_t M, int64_t N, int32_t act, float act_beta, void* out,
                int64_t ld_out, int32_t out_dtype, float* colsum, na_stream_t stream);

/* Weight normalisation W = g * v / ||v||_row (v [N,K], g [N]); optionally also
 * writes a bf16 copy Wb.  Backward writes (does not accumulate)
 * dg = sum_k dW*v/||v|| and dv = g/||v|| * (dW - dg * v/||v||). */
int na_weight_norm_fwd(const float* v, const float* g, int64_t N, int64_t K, float* W,
                       void* Wlow, int32_t low_dtype, int64_t ld_low, na_stream_t stream);
int na_weight_norm_bwd(const float* v, const float* g, const float* dW, int64_t N, int64_t K,
                       float* dv, float* dg, na_stream_t stream);

/* ---------------------------------------------------------------------------
 * SDF network input per point, row stride ld: [enc(L*F) | x(3) | zero pad] for an
 * fp32 output; [enc(L*F) | hi(3) | lo(3) | hi(3) | zero pad] for a bf16 output,
 * hi = bf16(x), lo = bf16(x - hi), so the point keeps 16 mantissa bits when the
 * first-layer weight columns are laid out [W_hi | W_hi | W_lo].
 * Fuses the range normalisation of projects/neuralangelo/utils/modules.py:82-83,
 * the encoder (modules.py:84-86), the coarse-to-fine mask (:91-93) and the
 * concatenation with the raw points (:94); the encoder block comes first so
 * its gradient slice is aligned (the first-layer weight columns are permuted
 * to match).  Backward reads the encoder block of dY with the same stride.
 * ------------------------------------------------------------------------- */
int na_sdf_input_fwd(const na_grid_desc* desc, const float* points, int64_t n, float vmin,
                     float vmax, const void* table, int32_t table_dtype, int32_t active_levels,
                     void* out, int32_t out_dtype, int64_t ld, na_stream_t stream);
int na_sdf_input_bwd(const na_grid_desc* desc, const float* points, int64_t n, float vmin,
                     float vmax, const void* dy, int32_t dy_dtype, int64_t ld,
                     float* grad_table, int32_t active_levels, na_stream_t stream);

/* Encoder backward of one numerical-gradient point set in a single pass: the
 * centre evaluation (modules.py:84-86, reached from model.py:160) and the four
 * tetrahedral taps x + eps*k_i, k = (1,-1,-1), (-1,-1,1), (-1,1,-1), (1,1,1)
 * (modules.py:157-175) -- five tcnn backward calls in the reference.
 * points: centre points [n,3] (world); dy_centre [n, ld] and dy_taps [4n, ld]
 * hold the encoder gradient in columns 0..L*F-1.  increments = 0: dy_taps is
 * tap-major (row t*n + i) and each point's gradient is scattered with its own
 * trilinear weights (dy_shared must be NULL).  increments = 1 (the increment
 * formulation): dy_taps is point-major (row 4i + t) and is the gradient of the
 * input increment D_t = X(x + eps k_t) - X(x), dy_shared [n, ld] (or NULL) the
 * gradient of sum_t D_t, shared by the four taps, and dy_centre that of the
 * centre input alone: tap t's corners receive g_t = dy_taps + dy_shared and the
 * centre's dy_centre - sum_t g_t (formed in fp32).  Contributions of the five
 * points to a shared corner are summed before one atomic add into grad_table
 * (accumulated; the caller zeroes it). */
int na_sdf_taps_bwd(const na_grid_desc* desc, const float* points, int64_t n, float vmin,
                    float vmax, float eps, const void* dy_centre, const void* dy_taps,
                    const void* dy_shared, int32_t dy_dtype, int64_t ld, float* grad_table,
                    int32_t active_levels, int32_t increments, na_stream_t stream);

/* Deterministic (sorted) variants of the three table-gradient scatters above:
 * na_hashgrid_bwd (tcnn's backward, modules.py:84-86), na_sdf_input_bwd and
 * na_sdf_taps_bwd, same arguments plus a workspace.  The atomic scatters add
 * each contribution as it arrives, so a table entry's float sum order, and its
 * last bits, vary between runs (tcnn's atomic backward behaves the same way).
 * Here, per level, contribution id = (i * E + e) * 8 + c (point i, evaluation e
 * of E, corner c; E = 1, or 5 for the centre and four taps) is sorted stably by
 * table entry, and each entry's contributions are summed in id order in fp32
 * and added into grad_table: the result depends only on the inputs.  The
 * workspace is caller-owned, 256-byte aligned, of na_hashgrid_sorted_workspace
 * (desc, n, E) bytes (16 n E 8 bytes plus the radix sort's temporary storage);
 * n * E * 8 must stay below 2^31.  Returns the size (>= 0) or an error code. */
int64_t na_hashgrid_sorted_workspace(const na_grid_desc* desc, int64_t n, int32_t evals);
int na_hashgrid_bwd_sorted(const na_grid_desc* desc, const float* x, int64_t n, const void* dy,
                           int32_t dy_dtype, float* grad_table, int32_t active_levels,
                           void* workspace, int64_t workspace_bytes, na_stream_t stream);
int na_sdf_input_bwd_sorted(const na_grid_desc* desc, const float* points, int64_t n, float vmin,
                            float vmax, const void* dy, int32_t dy_dtype, int64_t ld,
                            float* grad_table, int32_t active_levels, void* workspace,
                            int64_t workspace_bytes, na_stream_t stream);
int na_sdf_taps_bwd_sorted(const na_grid_desc* desc, const float* points, int64_t n, float vmin,
                           float vmax, float eps, const void* dy_centre, const void* dy_taps,
                           const void* dy_shared, int32_t dy_dtype, int64_t ld,
                           float* grad_table, int32_t active_levels, int32_t increments,
                           void* workspace, int64_t workspace_bytes, na_stream_t stream);

/* SDF network inputs of a numerical-gradient point set for the increment
 * formulation (bf16; replaces the encoder calls of modules.py:84-86 and
 * :157-175 for one point set): x0_centre [n, ld] as na_sdf_input_fwd writes it
 * for x, and d_taps [4n, ld_d] with point-major row 4i + t holding the
 * increment X(x + eps k_t) - X(x) (k_t as na_sdf_taps_bwd), evaluated in fp32
 * and stored as hi = bf16(D) in columns 0..L*F-1 and lo = bf16(D - hi) in
 * columns ld..ld+L*F-1; columns L*F..L*F+8 hold the exact point offset in the
 * split [hi | lo | hi] layout, all other columns zero (ld_d >= ld + L*F).
 * d_sum [n, ld] receives sum_t D_t (fp32 sum, one bf16 rounding) in x0_centre's
 * layout. */
int na_sdf_taps_increment_input(const na_grid_desc* desc, const float* points, int64_t n,
                                float vmin, float vmax, float eps, const void* table,
                                int32_t table_dtype, int32_t active_levels, void* x0_centre,
                                int64_t ld, void* d_taps, int64_t ld_d, void* d_sum,
                                na_stream_t stream);

/* ---------------------------------------------------------------------------
 * Per-ray operators.  R rays, N samples per ray, row-major [R, N(, K)].
 * ------------------------------------------------------------------------- */

/* Unit-sphere bounds and coarse samples (projects/neuralangelo/model.py:207-213,
 * projects/nerf/utils/nerf_util.py:199-205 and :20-38):
 * near/far from |c + t v| = 1, near = relu(near); rays that miss get
 * outside = 1, near = 1, far = 1.2; dists[r,i] = (i + u[r,i]) / N * (far-near) + near
 * with u = rands (stratified) or 0.5 when rands == NULL. */
int na_ray_bounds_sample(const float* center, const float* ray_unit, const float* rands,
                         int64_t R, int32_t N, float* near, float* far, uint8_t* outside,
                         float* dists, na_stream_t stream);

/* Sphere crossings (projects/nerf/utils/nerf_util.py:199-205, intersect_with_sphere):
 * near/far = -<c,v> -/+ sqrt(<c,v>^2 - (<c,c> - radius^2)) per ray, unclamped,
 * NaN where the ray misses.  center, ray_unit [R,3] -> near, far [R]. */
int na_sphere_crossings(const float* center, const float* ray_unit, int64_t R, float radius,
                        float* near, float* far, na_stream_t stream);

/* Pixel-centre camera rays (projects/nerf/utils/camera.py:283-311, get_center_and_ray,
 * with the pixel selection of projects/nerf/utils/nerf_util.py:127-131, slice_by_ray_idx):
 * pose [B,3,4] world-to-camera [R|t], intr [B,3,3]; pixel p = ray_idx[b,r] (int64,
 * in [0, H*W)) or p = r when ray_idx is NULL (then R must be H*W) sits at
 * (p mod W + 0.5, p div W + 0.5).  center [B,R,3] = -R^T t, ray [B,R,3] =
 * R^T K^-1 [x,y,1]^T (depth 1), ray_unit [B,R,3] = ray/max(|ray|,1e-12) (may be NULL). */
int na_camera_rays(const float* pose, const float* intr, int64_t B, int64_t H, int64_t W,
                   const int64_t* ray_idx, int64_t R, float* center, float* ray, float* ray_unit,
                   na_stream_t stream);

/* Stratified distances (projects/nerf/utils/nerf_util.py:20-38, sample_dists):
 * dists[r,i] = (i + u) / N * (hi - lo) + lo with u = rands[r,i] (stratified) or
 * 0.5 when rands == NULL; lo / hi are per-ray [R] arrays, or the constants
 * lo_const / hi_const when the pointer is NULL.  -> dists [R, N]. */
int na_sample_dists(const float* lo, const float* hi, float lo_const, float hi_const,
                    const float* rands, int64_t R, int32_t N, float* dists, na_stream_t stream);

/* Real spherical harmonics of unit directions
 * (projects/neuralangelo/utils/spherical_harmonics.py:47-71, get_spherical_harmonics):
 * dirs [n,3] -> out [n, (levels+1)^2], degree 0..3 (NA_ERR_UNSUPPORTED above 3). */
int na_sh_encode(const float* dirs, int64_t n, int32_t levels, float* out, na_stream_t stream);

/* Inverse-CDF resampling (projects/nerf/utils/nerf_util.py:41-68):
 * pdf = w / max(sum|w|, 1e-12); cdf = [0, serial fp32 prefix of pdf];
 * u_j = linspace midpoints; idx = searchsorted(cdf, u, right=True); lerp.
 * bins [R, N], weights [R, N-1] -> out [R, Nf]. */
int na_sample_pdf(const float* bins, const float* weights, int64_t R, int32_t N, int32_t Nf,
                  float* out, na_stream_t stream);

/* One round of hierarchical sampling (projects/neuralangelo/model.py:239-247,250-267):
 * NeuS alphas of the N-1 intervals from the sorted samples' SDFs (robust cos,
 * inv_s), compositing weights, sample_pdf of Nf new dists, and a stable merge
 * (ties keep coarse samples first) into merged [R, N+Nf].  pos_coarse [R,N] and
 * pos_fine [R,Nf] receive each sample's position in the merged row. */
int na_hier_resample(const float* dists, const float* sdfs, int64_t R, int32_t N, float inv_s,
                     int32_t Nf, float* merged, float* fine, int32_t* pos_coarse,
                     int32_t* pos_fine, na_stream_t stream);

/* out[r, pos[r,i]] = src[r,i] for both lists (SDF gather after the merge). */
int na_scatter_merge(const float* src_a, const int32_t* pos_a, int32_t Na, const float* src_b,
                     const int32_t* pos_b, int32_t Nb, int64_t R, float* out,
                     na_stream_t stream);

/* NeuS SDF -> alpha (projects/neuralangelo/model.py:275-298).  sdfs [R,N]
 * (rays with outside[r] != 0 use outside_val), grads [R,N,3], dists [R,N],
 * far [R] (NULL: 1e10), s_var: device scalar, inv_s = exp(s_var). */
int na_neus_alpha_fwd(const float* ray_unit, const float* sdfs, const float* grads,
                      const float* dists, const float* far, const uint8_t* outside,
                      float outside_val, const float* s_var, float progress, float anneal_end,
                      int64_t R, int32_t N, float* alpha, na_stream_t stream);
/* Backward: dsdf [R,N], dgrad [R,N,3] written; ds_var (device scalar) += . */
int na_neus_alpha_bwd(const float* ray_unit, const float* sdfs, const float* grads,
                      const float* dists, const float* far, const uint8_t* outside,
                      float outside_val, const float* s_var, float progress, float anneal_end,
                      int64_t R, int32_t N, const float* dalpha, float* dsdf, float* dgrad,
                      float* ds_var, na_stream_t stream);

/* Alpha compositing + composite (projects/nerf/utils/render.py:87-112):
 * T_i = exp(sum_{j<i} log1p(-alpha_j)), w_i = alpha_i T_i, out[r,k] = sum_i w_i q[r,i,k].
 * weights [R,N] always written; q [R,N,K] / out [R,K] optional (K may be 0). */
int na_composite_fwd(const float* alpha, const float* q, int64_t R, int32_t N, int32_t K,
                     float* weights, float* out, na_stream_t stream);
/* Backward with upstream dout [R,K] (optional) and dweights [R,N] (optional):
 * dalpha [R,N] written (exact at alpha = 1, no division), dq [R,N,K] written if non-NULL. */
int na_composite_bwd(const float* alpha, const float* q, const float* weights, int64_t R,
                     int32_t N, int32_t K, const float* dout, const float* dweights,
                     float* dalpha, float* dq, na_stream_t stream);

/* composite with given weights (projects/nerf/utils/render.py:102-112):
 * out[r,k] = sum_i q[r,i,k] w[r,i]; backward dq = w dout, dw[r,i] = sum_k q dout. */
int na_weighted_sum_fwd(const float* q, const float* w, int64_t R, int32_t N, int32_t K,
                        float* out, na_stream_t stream);
int na_weighted_sum_bwd(const float* q, const float* w, int64_t R, int32_t N, int32_t K,
                        const float* dout, float* dq, float* dw, na_stream_t stream);

/* Packed (ragged) rays: the same operators over rays with uneven sample counts.
 * Ray r owns samples [offsets[r], offsets[r+1]) of every per-sample array
 * (offsets int32[R+1], non-decreasing, offsets[0] = 0); per-ray arrays are [R, ...].
 * They replace, for packed batches, alpha_compositing_weights / composite
 * (projects/nerf/utils/render.py:87-112) and Model.compute_neus_alphas
 * (projects/neuralangelo/model.py:275-298, the last sample of a ray uses its
 * far distance as at model.py:170-171); the reference itself renders a fixed N
 * per ray, which is the packed layout with equal counts.
 * na_packed_ray_index writes the ray of every sample (int32[S]). */
int na_packed_ray_index(const int32_t* offsets, int64_t R, int32_t* ray_idx, na_stream_t stream);
int na_composite_packed_fwd(const float* alpha, const float* q, const int32_t* offsets, int64_t R,
                            int32_t K, float* weights, float* out, na_stream_t stream);
int na_composite_packed_bwd(const float* alpha, const float* q, const float* weights,
                            const int32_t* offsets, int64_t R, int32_t K, const float* dout,
                            const float* dweights, float* dalpha, float* dq, na_stream_t stream);
int na_weighted_sum_packed_fwd(const float* q, const float* w, const int32_t* offsets, int64_t R,
                               int32_t K, float* out, na_stream_t stream);
int na_weighted_sum_packed_bwd(const float* q, const float* w, const int32_t* ray_idx, int64_t S,
                               int32_t K, const float* dout, float* dq, float* dw,
                               na_stream_t stream);
int na_neus_alpha_packed_fwd(const float* ray_unit, const float* sdfs, const float* grads,
                             const float* dists, const float* far, const uint8_t* outside,
                             float outside_val, const float* s_var, float progress,
                             float anneal_end, const int32_t* offsets, const int32_t* ray_idx,
                             int64_t S, float* alpha, na_stream_t stream);
int na_neus_alpha_packed_bwd(const float* ray_unit, const float* sdfs, const float* grads,
                             const float* dists, const float* far, const uint8_t* outside,
                             float outside_val, const float* s_var, float progress,
                             float anneal_end, const int32_t* offsets, const int32_t* ray_idx,
                             int64_t S, const float* dalpha, float* dsdf, float* dgrad,
                             float* ds_var, na_stream_t stream);

/* The packed NeuS alpha and compositing of all rays in one persistent launch each
 * way (the separate packed entries above, fused; the same per-sample terms, the
 * transmittance and its backward recurrence as wave scans).  A grid of a few
 * workgroups per CU whose waves take chunks of 8 consecutive rays from *queue (an
 * int32 the caller zeroes before the call; the launch leaves it non-zero) until all
 * rays are done, a wave rendering one ray at a time with its lanes over the ray's
 * samples, so uneven sample counts balance across the chip.  Tile shape: the launch tunables "packed.rays_per_wg"
 * (workgroup size) and "packed.wg_per_cu".
 *   fwd: alpha [S] (model.py:275-298), weights [S] and out [R, K] = sum of q [S, K]
 *        times weights (render.py:87-112), K <= 4;
 *   bwd: from dout [R, K] (or NULL), dweights [S] (or NULL) and a direct dalpha [S]
 *        (or NULL): dsdf [S], dgrad [S, 3], dq [S, K] (or NULL), ds_var += (one
 *        atomic per wave, so the last bits follow the schedule). */
int na_neus_composite_packed_fwd(const float* ray_unit, const float* sdfs, const float* grads,
                                 const float* dists, const float* far, const uint8_t* outside,
                                 float outside_val, const float* s_var, float progress,
                                 float anneal_end, const int32_t* offsets, int64_t R,
                                 const float* q, int32_t K, float* alpha, float* weights,
                                 float* out, int32_t* queue, na_stream_t stream);
int na_neus_composite_packed_bwd(const float* ray_unit, const float* sdfs, const float* grads,
                                 const float* dists, const float* far, const uint8_t* outside,
                                 float outside_val, const float* s_var, float progress,
                                 float anneal_end, const int32_t* offsets, int64_t R,
                                 const float* q, int32_t K, const float* alpha,
                                 const float* weights, const float* dout, const float* dweights,
                                 const float* dalpha, float* dsdf, float* dgrad, float* dq,
                                 float* ds_var, int32_t* queue, na_stream_t stream);

/* Background NeRF (projects/neuralangelo/utils/modules.py:232-311).
 * na_bg_encode: points [n,3] -> rows [n, ld] of [q | Fourier(q)] with
 *   q = [p/|p|, 1/|p|] (modules.py:290-293) and nerf_util.positional_encoding
 *   (nerf_util.py:134-147, `levels` frequencies 2^l pi, sin block then cos block
 *   per component), zero-padded to ld (>= 4 + 8 levels); f32 or bf16 out.
 * na_bg_alpha_fwd: alpha = 1 - exp(-softplus(o) (d_{i+1} - d_i)), d_N = 1e10
 *   (render.volume_rendering_alphas_dist, render.py:68-84, with the density
 *   activation softplus of modules.py:270-272); o = column 0 of mlp_out rows
 *   (stride ld), dists [R, N]; density optional output [R, N].
 * na_bg_alpha_bwd: d(o) from dalpha [R,N] and/or ddensity [R,N], written to column 0
 *   of dmlp_out rows (stride ld_d). */
int na_bg_encode(const float* points, int64_t n, int32_t levels, void* out, int32_t out_dtype,
                 int64_t ld, na_stream_t stream);
int na_bg_alpha_fwd(const float* mlp_out, int64_t ld, const float* dists, int64_t R, int32_t N,
                    float* density, float* alpha, na_stream_t stream);
int na_bg_alpha_bwd(const float* mlp_out, int64_t ld, const float* dists, int64_t R, int32_t N,
                    const float* dalpha, const float* ddensity, float* dmlp_out, int64_t ld_d,
                    na_stream_t stream);

/* 4-tap numerical gradient (projects/neuralangelo/utils/modules.py:157-175).
 * taps: writes x + eps*k_i for k = (1,-1,-1), (-1,-1,1), (-1,1,-1), (1,1,1) as
 * [4, n, 3] (eps = float32 of normal_eps/sqrt(3)).  combine: s_taps [4, n],
 * sdf [n] (centre SDF after the outside fill), grad [n,3] = sum k_i s_i / (4 eps),
 * hess [n,3] = ((sum s_i)/2 - 2 sdf)/eps_sq / 3 (only if hess != NULL), with
 * eps_sq = float32 of the float64 square, as the reference evaluates eps ** 2. */
int na_taps_points(const float* x, int64_t n, float eps, float* out, na_stream_t stream);
int na_taps_combine_fwd(const float* s_taps, const float* sdf, int64_t n, float eps,
                        float eps_sq, float* grad, float* hess, na_stream_t stream);
int na_taps_combine_bwd(int64_t n, float eps, float eps_sq, const float* dgrad,
                        const float* dhess, float* ds_taps, float* dsdf, int32_t add_dsdf,
                        na_stream_t stream);

/* Increment form of the 4-tap combination (modules.py:163-175): ds [4, n] holds
 * sdf(x + eps k_i) - sdf(x), sc [n] the centre SDF as evaluated and sdf [n] the
 * value passed to compute_gradients (outside fill included).  grad [n,3] =
 * sum_i k_i ds_i / (4 eps); hess [n,3] = h/3 with h = (sum_i ds_i / 2 + 2 (sc -
 * sdf)) / eps_sq (algebraically the reference's ((sum s_i)/2 - 2 sdf)/eps^2).
 * hess may be NULL (evaluation).  Backward writes dds [4,n], dsc [n], dsdf [n]. */
int na_taps_delta_fwd(const float* ds, const float* sc, const float* sdf, int64_t n, float eps,
                      float eps_sq, float* grad, float* hess, na_stream_t stream);
int na_taps_delta_bwd(int64_t n, float eps, float eps_sq, const float* dgrad, const float* dhess,
                      float* dds, float* dsc, float* dsdf, na_stream_t stream);

/* Colour-MLP input (projects/neuralangelo/utils/modules.py:209-220, SH degree 3
 * from projects/neuralangelo/utils/spherical_harmonics.py:47-71).  Row layout
 * [feat(col0) | x(3) | SH16(ray) | normalize(grad)(3) | zero pad up to ld]; this
 * call writes columns [col0, ld) (the SDF feature layer writes [0, col0) in
 * place).  ray_unit is per ray ([n / n_per_ray, 3]).  normalize != 0: `grad`
 * holds raw SDF gradients and the normal block is F.normalize(grad) (eps 1e-12,
 * projects/neuralangelo/model.py:167); normalize == 0: `grad` already holds the
 * normals.  Backward: d(grad) from the normal block (through the normalisation
 * when normalize != 0), added to dgrad if add != 0. */
int na_rgb_input_fwd(const float* x, const float* ray_unit, int64_t n_per_ray,
                     const float* grad, int32_t normalize, int64_t n, int32_t col0, void* out,
                     int32_t out_dtype, int64_t ld, na_stream_t stream);
int na_rgb_input_bwd(const float* grad, int32_t normalize, int64_t n, int32_t col0,
                     const void* dy, int32_t dy_dtype, int64_t ld, float* dgrad, int32_t add,
                     na_stream_t stream);

/* Losses (projects/neuralangelo/utils/misc.py:74-89, trainer.py:43):
 * eikonal = mean((|g|-1)^2 * !outside), curvature = mean(|sum hess| * !outside)
 * with NaN/Inf -> 0; both over all n = R * n_per_ray samples.  sums[0] += sum of
 * eikonal terms, sums[1] += sum of curvature terms.  Backward writes the
 * per-sample gradients scaled by w[0] / n (eikonal) and w[1] / n (curvature);
 * w is a device array (the upstream gradients of the two means). */
int na_sdf_losses_fwd(const float* grad, const float* hess, const uint8_t* outside,
                      int64_t n, int64_t n_per_ray, float* sums, na_stream_t stream);
int na_sdf_losses_bwd(const float* grad, const float* hess, const uint8_t* outside,
                      int64_t n, int64_t n_per_ray, const float* w, float* dgrad,
                      int32_t add_dgrad, float* dhess, na_stream_t stream);

/* ---------------------------------------------------------------------------
 * AdamW step (replaces torch.optim.AdamW built at imaginaire/trainers/utils/
 * get_trainer.py:106-150 and stepped at imaginaire/trainers/base.py:454-457,
 * lr 1e-3 / wd 1e-2 from projects/neuralangelo/configs/base.yaml:117-121).
 * Same update as torch's: p *= 1 - lr*wd; m = lerp(m, g, 1-beta1);
 * v = beta2 v + (1-beta2) g^2; p -= lr/bc1 * m / (sqrt(v)/sqrt(bc2) + eps),
 * bc_i = 1 - beta_i^step.  If lowp != NULL the bf16 copy of the updated
 * parameter is written in the same pass (the hash table's gather copy).  With
 * clear_grad != 0 the gradient is zeroed in the same pass, after it is read
 * (the optimizer.step + zero_grad pair of imaginaire/trainers/base.py:447-457,
 * with the zeroed buffer kept for the next backward instead of a new fill).
 * ------------------------------------------------------------------------- */
int na_adamw_step(float* param, float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                  float lr, float beta1, float beta2, float eps, float weight_decay,
                  int64_t step, void* lowp, int32_t lowp_dtype, int32_t clear_grad,
                  na_stream_t stream);

/* The same update for up to NA_ADAMW_MULTI_MAX tensors that share the
 * hyper-parameters and the step count, in one launch (the optimizer's loop
 * over its parameter group, get_trainer.py:106-150, for the small tensors).
 * `tensors` is a host array (copied into the launch); entries with n = 0 are
 * skipped; no bf16 copy.  Results equal na_adamw_step's bit for bit. */
#define NA_ADAMW_MULTI_MAX 32
typedef struct na_adamw_tensor {
  float* param;
  const float* grad;
  float* exp_avg;
  float* exp_avg_sq;
  int64_t n;
} na_adamw_tensor;
int na_adamw_step_multi(const na_adamw_tensor* tensors, int32_t count, float lr, float beta1,
                        float beta2, float eps, float weight_decay, int64_t step,
                        na_stream_t stream);

/* ---------------------------------------------------------------------------
 * Marching cubes over a block-partitioned lattice.  Replaces the per-block
 * loop of projects/neuralangelo/utils/mesh.py:24-49 (LatticeGrid blocks of
 * block_res+1 points, :68-100; mcubes.marching_cubes(sdf, 0.) per block,
 * :119-121; V*intv + xyz[0,0,0], :123) and, with block_res <= 0 (one block
 * spanning the volume), mcubes.marching_cubes(volume, iso) itself.
 *
 * volume: f32 [nx, ny, nz] C order, the whole lattice.  Blocks hold block_res
 * cells per axis (the last block of an axis may hold fewer) and are visited
 * x-major, as the reference's block index.  A cell's cube index has bit m set
 * iff value[m] <= iso.  Output = the concatenation, in block order, of each
 * block's mesh: float64 vertices [nv, 3] and int64 faces [nf, 3] indexing the
 * concatenated vertex list.  Within a block the order is that of the
 * sequential algorithm (cells i-major; per cell the owned edges 6, 5, 10, then
 * 0,1,2,3,4,7,8,9,11 on the block boundary; triangles in table order).
 * Vertices are block-local array-index coordinates, or, when the three f32
 * axis arrays (the lattice coordinates, length nx/ny/nz) are given,
 * local * intv + axis[block origin] in float64.
 *
 * Two calls: na_mc_count classifies, scans and writes totals[0] = nv,
 * totals[1] = nf (device int64[2]); after the caller sizes the outputs,
 * na_mc_emit writes them.  The workspace (na_mc_workspace_size bytes, caller
 * owned, 256-byte aligned) carries state from count to emit.
 * ------------------------------------------------------------------------- */
int64_t na_mc_workspace_size(int64_t nx, int64_t ny, int64_t nz, int64_t block_res);
int na_mc_count(const float* volume, int64_t nx, int64_t ny, int64_t nz, int64_t block_res,
                double iso, void* workspace, int64_t ws_bytes, int64_t* totals,
                na_stream_t stream);
int na_mc_emit(const float* volume, int64_t nx, int64_t ny, int64_t nz, int64_t block_res,
               double iso, void* workspace, int64_t ws_bytes, const float* axis_x,
               const float* axis_y, const float* axis_z, double intv, double* vertices,
               int64_t* faces, na_stream_t stream);
/* Per-block ranges of the concatenated output, valid after na_mc_count on the same
 * workspace: block_v / block_f [nblocks + 1] (device int64) receive the offset of each
 * block's first vertex / face (x-major block order, projects/neuralangelo/utils/
 * mesh.py:82-96) and the totals; the reference's per-block post-processing
 * (mesh.py:119-133: sphere filter, largest component) acts on these ranges. */
int64_t na_mc_num_blocks(int64_t nx, int64_t ny, int64_t nz, int64_t block_res);
int na_mc_block_offsets(int64_t nx, int64_t ny, int64_t nz, int64_t block_res, void* workspace,
                        int64_t ws_bytes, int64_t* block_v, int64_t* block_f, na_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* NEURALANGELO_AMD_H */
