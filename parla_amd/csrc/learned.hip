// Learned leverage-score sketch: the MLP surrogate's forward and backward
// kernels, its Adam step, and the row-sampling gather.
//
// Network: 16 -> 64 -> 64 -> 1, ReLU hidden layers, scalar output (the
// predicted log of the normalised leverage score).  The inputs pass through
// asinh first: pre-sketch features of coherent rows are orders of magnitude
// larger than the rest, and asinh keeps their scale logarithmic, which is
// the scale of the target.  Weights are BF16 (the
// optimiser keeps FP32 masters and rounds them to nearest-even after each
// step); every product and sum is FP32.  Parameter vector layout (FP32
// master, and the same order for gradients):
//   W1[16][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3        (5313)
// One thread evaluates one row; the BF16 weights are widened once per
// workgroup into LDS and read there by same-address (broadcast) loads.
// Backward recomputes the forward pass (it is cheaper than storing
// activations), keeps each row's h1, dh2 and dh1 in that row's LDS vectors,
// and reduces the weight gradients of the workgroup's 256 rows from LDS;
// workgroup partials are summed in a fixed order by a second kernel, so the
// gradients are deterministic.
#include "common.h"
#include "rng.h"

namespace {

constexpr int MLP_IN = 16, MLP_H = 64;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_IN * MLP_H, OFF_W2 = OFF_B1 + MLP_H,
              OFF_B2 = OFF_W2 + MLP_H * MLP_H, OFF_W3 = OFF_B2 + MLP_H, OFF_B3 = OFF_W3 + MLP_H,
              MLP_NPARAM = OFF_B3 + 1;
constexpr int MLP_THREADS = 256;

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));  // inf/nan
  const uint32_t lsb = (u >> 16) & 1u;
  return (uint16_t)((u + 0x7fffu + lsb) >> 16);
}

struct MlpLds {
  float w1[MLP_IN][MLP_H];
  float b1[MLP_H];
  float w2[MLP_H][MLP_H];
  float b2[MLP_H];
  float w3[MLP_H];
  float b3;
};

__device__ inline void load_weights(MlpLds& s, const uint16_t* __restrict__ wb, const float* __restrict__ bias) {
  // wb: BF16 weights W1 | W2 | w3 (16*64 + 64*64 + 64); bias: FP32 b1 | b2 | b3
  for (int e = threadIdx.x; e < MLP_IN * MLP_H; e += blockDim.x) (&s.w1[0][0])[e] = bf16_to_f32(wb[e]);
  for (int e = threadIdx.x; e < MLP_H * MLP_H; e += blockDim.x)
    (&s.w2[0][0])[e] = bf16_to_f32(wb[MLP_IN * MLP_H + e]);
  for (int e = threadIdx.x; e < MLP_H; e += blockDim.x) {
    s.w3[e] = bf16_to_f32(wb[MLP_IN * MLP_H + MLP_H * MLP_H + e]);
    s.b1[e] = bias[e];
    s.b2[e] = bias[MLP_H + e];
  }
  if (threadIdx.x == 0) s.b3 = bias[2 * MLP_H];
}

// h1 = relu(W1^T x + b1) -> hrow (this thread's LDS row), h2 = relu(W2^T h1
// + b2) in registers, y = w3 . h2 + b3.  The second layer walks k through
// LDS (a register array may only be indexed with compile-time indices).
__device__ __forceinline__ float mlp_eval(const MlpLds& s, const float (&x)[MLP_IN], float* hrow,
                                          float (&h2)[MLP_H]) {
#pragma unroll
  for (int j = 0; j < MLP_H; ++j) {
    float a = s.b1[j];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) a = fmaf(x[k], s.w1[k][j], a);
    hrow[j] = fmaxf(a, 0.f);
  }
#pragma unroll
  for (int j = 0; j < MLP_H; ++j) h2[j] = s.b2[j];
  for (int k = 0; k < MLP_H; ++k) {
    const float hk = hrow[k];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) h2[j] = fmaf(hk, s.w2[k][j], h2[j]);
  }
  float y = s.b3;
#pragma unroll
  for (int j = 0; j < MLP_H; ++j) {
    h2[j] = fmaxf(h2[j], 0.f);
    y = fmaf(h2[j], s.w3[j], y);
  }
  return y;
}

__global__ __launch_bounds__(MLP_THREADS) void mlp_fwd_kernel(int64_t rows, const float* __restrict__ X, int64_t ldx,
                                                               const uint16_t* __restrict__ wb,
                                                               const float* __restrict__ bias, float* __restrict__ y) {
  __shared__ MlpLds s;
  __shared__ float hrow[MLP_THREADS][MLP_H + 1];
  load_weights(s, wb, bias);
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float x[MLP_IN], h2[MLP_H];
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) x[k] = asinhf(X[r * ldx + k]);
  y[r] = mlp_eval(s, x, hrow[threadIdx.x], h2);
}

// Loss = (1 / rows_total) sum_r (y_r - t_r)^2 over this call's rows; writes
// per-workgroup partial gradients part[blockIdx.x][0 .. MLP_NPARAM) and the
// partial loss in part[blockIdx.x][MLP_NPARAM].  va / vb are per-row LDS
// vectors, reused phase by phase: (h1, dh2) -> dW2, db2; (h1 mask) -> dh1;
// (dy h2, dh1) -> dw3, db1; (x, dh1) -> dW1.
__global__ __launch_bounds__(MLP_THREADS) void mlp_bwd_kernel(int64_t rows, double inv_total,
                                                              const float* __restrict__ X, int64_t ldx,
                                                              const float* __restrict__ t,
                                                              const uint16_t* __restrict__ wb,
                                                              const float* __restrict__ bias,
                                                              float* __restrict__ part) {
  __shared__ MlpLds s;
  __shared__ float va[MLP_THREADS][MLP_H + 1];
  __shared__ float vb[MLP_THREADS][MLP_H + 1];
  __shared__ float red[MLP_THREADS / 64][2];
  load_weights(s, wb, bias);
  __syncthreads();
  const int tid = threadIdx.x;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + tid;
  const bool valid = r < rows;
  float x[MLP_IN], h2[MLP_H];
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) x[k] = valid ? asinhf(X[r * ldx + k]) : 0.f;
  const float yv = mlp_eval(s, x, va[tid], h2);  // h1 -> va[tid]
  const float diff = valid ? yv - t[r] : 0.f;
  const float dy = (float)(2.0 * inv_total) * diff;
  const float lossc = (float)inv_total * diff * diff;
  // dh2 = dy w3 (h2 > 0) -> vb[tid]
#pragma unroll
  for (int j = 0; j < MLP_H; ++j) vb[tid][j] = (h2[j] > 0.f) ? dy * s.w3[j] : 0.f;
  float* out = part + (int64_t)blockIdx.x * (MLP_NPARAM + 1);
  __syncthreads();
  {  // dW2[a][b] = sum_rows h1[a] dh2[b]: 16 (a, b) pairs per thread
    const int a = tid >> 2, b0 = (tid & 3) * 16;
    float acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
    for (int rr = 0; rr < MLP_THREADS; ++rr) {
      const float ha = va[rr][a];
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[q] = fmaf(ha, vb[rr][b0 + q], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) out[OFF_W2 + a * MLP_H + b0 + q] = acc[q];
  }
  if (tid < MLP_H) {  // db2
    float acc = 0.f;
    for (int rr = 0; rr < MLP_THREADS; ++rr) acc += vb[rr][tid];
    out[OFF_B2 + tid] = acc;
  }
  __syncthreads();
  {  // dh1[a] = (h1[a] > 0) sum_b W2[a][b] dh2[b], into vb[tid] (own row)
    float dh2[MLP_H];
#pragma unroll
    for (int b = 0; b < MLP_H; ++b) dh2[b] = vb[tid][b];
    for (int a = 0; a < MLP_H; ++a) {
      float acc = 0.f;
#pragma unroll
      for (int b = 0; b < MLP_H; ++b) acc = fmaf(s.w2[a][b], dh2[b], acc);
      vb[tid][a] = (va[tid][a] > 0.f) ? acc : 0.f;
    }
  }
#pragma unroll
  for (int j = 0; j < MLP_H; ++j) va[tid][j] = dy * h2[j];
  __syncthreads();
  if (tid < MLP_H) {  // db1, dw3
    float acc1 = 0.f, acc3 = 0.f;
    for (int rr = 0; rr < MLP_THREADS; ++rr) {
      acc1 += vb[rr][tid];
      acc3 += va[rr][tid];
    }
    out[OFF_B1 + tid] = acc1;
    out[OFF_W3 + tid] = acc3;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) va[tid][k] = x[k];
  __syncthreads();
  {  // dW1[k][a] = sum_rows x[k] dh1[a]: 4 entries per thread
    const int k = tid >> 4, a0 = (tid & 15) * 4;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int rr = 0; rr < MLP_THREADS; ++rr) {
      const float xk = va[rr][k];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = fmaf(xk, vb[rr][a0 + q], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) out[OFF_W1 + k * MLP_H + a0 + q] = acc[q];
  }
  // db3 and the loss
  const float sdy = wave_sum(dy), sl = wave_sum(lossc);
  if ((tid & 63) == 0) {
    red[tid >> 6][0] = sdy;
    red[tid >> 6][1] = sl;
  }
  __syncthreads();
  if (tid == 0) {
    float a = 0.f, l = 0.f;
    for (int w = 0; w < MLP_THREADS / 64; ++w) {
      a += red[w][0];
      l += red[w][1];
    }
    out[OFF_B3] = a;
    out[MLP_NPARAM] = l;
  }
}

__global__ void mlp_reduce_kernel(int64_t nparts, const float* __restrict__ part, float* __restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e > MLP_NPARAM) return;
  double acc = 0.0;
  for (int64_t g = 0; g < nparts; ++g) acc += (double)part[g * (MLP_NPARAM + 1) + e];
  grad[e] = (float)acc;
}

// Adam on the FP32 masters; refresh the BF16 weights and FP32 biases.
__global__ void mlp_adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                const float* __restrict__ g, float lr, float beta1, float beta2, float eps,
                                float bc1, float bc2, uint16_t* __restrict__ wb, float* __restrict__ bias) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= MLP_NPARAM) return;
  const float ge = g[e];
  const float me = beta1 * m[e] + (1.f - beta1) * ge;
  const float ve = beta2 * v[e] + (1.f - beta2) * ge * ge;
  m[e] = me;
  v[e] = ve;
  const float we = w[e] - lr * (me / bc1) / (sqrtf(ve / bc2) + eps);
  w[e] = we;
  // publish: BF16 weights W1 | W2 | w3, FP32 biases b1 | b2 | b3
  if (e < OFF_B1) {
    wb[e] = f32_to_bf16_rne(we);
  } else if (e < OFF_W2) {
    bias[e - OFF_B1] = we;
  } else if (e < OFF_B2) {
    wb[MLP_IN * MLP_H + (e - OFF_W2)] = f32_to_bf16_rne(we);
  } else if (e < OFF_W3) {
    bias[MLP_H + (e - OFF_B2)] = we;
  } else if (e < OFF_B3) {
    wb[MLP_IN * MLP_H + MLP_H * MLP_H + (e - OFF_W3)] = f32_to_bf16_rne(we);
  } else {
    bias[2 * MLP_H] = we;
  }
}

// out[r][:] = scale[r] * A[idx[r]][:]
__global__ void row_gather_kernel(int64_t d, int64_t n, const int64_t* __restrict__ idx,
                                  const double* __restrict__ scale, const double* __restrict__ A, int64_t lda,
                                  double* __restrict__ out, int64_t ldo) {
  const int64_t r = blockIdx.x;
  const double sc = scale ? scale[r] : 1.0;
  const double* src = A + idx[r] * lda;
  for (int64_t c = threadIdx.x; c < n; c += blockDim.x) out[r * ldo + c] = sc * src[c];
}

// out[idx[r], :] = A[r, :]: one workgroup per source row
__global__ void row_scatter_kernel(int64_t n, const int64_t* __restrict__ idx, const double* __restrict__ A,
                                   int64_t lda, double* __restrict__ out, int64_t ldo) {
  const int64_t r = blockIdx.x;
  double* dst = out + idx[r] * ldo;
  const double* src = A + r * lda;
  for (int64_t c = threadIdx.x; c < n; c += blockDim.x) dst[c] = src[c];
}

}  // namespace

extern "C" int pamd_mlp_nparams(void) { return MLP_NPARAM; }

extern "C" int pamd_mlp_fwd(int64_t rows, const float* X, int64_t ldx, const uint16_t* wb, const float* bias,
                            float* y, void* stream) {
  PAMD_ARG_CHECK(rows >= 0, 1);
  PAMD_ARG_CHECK(X != nullptr || rows == 0, 2);
  PAMD_ARG_CHECK(ldx >= MLP_IN, 3);
  PAMD_ARG_CHECK(wb != nullptr, 4);
  PAMD_ARG_CHECK(bias != nullptr, 5);
  PAMD_ARG_CHECK(y != nullptr || rows == 0, 6);
  if (rows == 0) return PAMD_OK;
  mlp_fwd_kernel<<<(unsigned)pamd_cdiv(rows, MLP_THREADS), MLP_THREADS, 0, (hipStream_t)stream>>>(rows, X, ldx, wb,
                                                                                                   bias, y);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_mlp_bwd_workspace(int64_t rows, int64_t* bytes) {
  PAMD_ARG_CHECK(rows >= 0, 1);
  PAMD_ARG_CHECK(bytes != nullptr, 2);
  *bytes = imax64(1, pamd_cdiv(rows, MLP_THREADS)) * (MLP_NPARAM + 1) * (int64_t)sizeof(float);
  return PAMD_OK;
}

extern "C" int pamd_mlp_bwd(int64_t rows, int64_t rows_total, const float* X, int64_t ldx, const float* t,
                            const uint16_t* wb, const float* bias, float* grad, void* workspace, void* stream) {
  PAMD_ARG_CHECK(rows >= 1, 1);
  PAMD_ARG_CHECK(rows_total >= rows, 2);
  PAMD_ARG_CHECK(X != nullptr, 3);
  PAMD_ARG_CHECK(ldx >= MLP_IN, 4);
  PAMD_ARG_CHECK(t != nullptr, 5);
  PAMD_ARG_CHECK(wb != nullptr, 6);
  PAMD_ARG_CHECK(bias != nullptr, 7);
  PAMD_ARG_CHECK(grad != nullptr, 8);
  PAMD_ARG_CHECK(workspace != nullptr, 9);
  hipStream_t st = (hipStream_t)stream;
  const int64_t nblk = pamd_cdiv(rows, MLP_THREADS);
  float* part = (float*)workspace;
  mlp_bwd_kernel<<<(unsigned)nblk, MLP_THREADS, 0, st>>>(rows, 1.0 / (double)rows_total, X, ldx, t, wb, bias, part);
  PAMD_LAUNCH_CHECK();
  mlp_reduce_kernel<<<(unsigned)pamd_cdiv(MLP_NPARAM + 1, 256), 256, 0, st>>>(nblk, part, grad);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_mlp_adam(float* w, float* m, float* v, const float* grad, float lr, float beta1, float beta2,
                             float eps, int64_t step, uint16_t* wb, float* bias, void* stream) {
  PAMD_ARG_CHECK(w != nullptr, 1);
  PAMD_ARG_CHECK(m != nullptr, 2);
  PAMD_ARG_CHECK(v != nullptr, 3);
  PAMD_ARG_CHECK(grad != nullptr, 4);
  PAMD_ARG_CHECK(lr > 0.f, 5);
  PAMD_ARG_CHECK(step >= 1, 9);
  PAMD_ARG_CHECK(wb != nullptr, 10);
  PAMD_ARG_CHECK(bias != nullptr, 11);
  const float bc1 = 1.f - powf(beta1, (float)step), bc2 = 1.f - powf(beta2, (float)step);
  mlp_adam_kernel<<<(unsigned)pamd_cdiv(MLP_NPARAM, 256), 256, 0, (hipStream_t)stream>>>(w, m, v, grad, lr, beta1,
                                                                                         beta2, eps, bc1, bc2, wb,
                                                                                         bias);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_row_gather_f64(int64_t d, int64_t m, int64_t n, const int64_t* idx, const double* scale,
                                   const double* A, int64_t lda, double* out, int64_t ldo, void* stream) {
  PAMD_ARG_CHECK(d >= 0, 1);
  PAMD_ARG_CHECK(m >= 1, 2);
  PAMD_ARG_CHECK(n >= 0, 3);
  PAMD_ARG_CHECK(idx != nullptr || d == 0, 4);
  PAMD_ARG_CHECK(A != nullptr, 6);
  PAMD_ARG_CHECK(lda >= n, 7);
  PAMD_ARG_CHECK(out != nullptr || d == 0, 8);
  PAMD_ARG_CHECK(ldo >= n, 9);
  if (d == 0 || n == 0) return PAMD_OK;
  row_gather_kernel<<<(unsigned)d, 256, 0, (hipStream_t)stream>>>(d, n, idx, scale, A, lda, out, ldo);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

extern "C" int pamd_row_scatter_f64(int64_t d, int64_t m, int64_t n, const int64_t* idx, const double* A,
                                    int64_t lda, double* out, int64_t ldo, void* stream) {
  PAMD_ARG_CHECK(d >= 0, 1);
  PAMD_ARG_CHECK(m >= d, 2);
  PAMD_ARG_CHECK(n >= 0, 3);
  PAMD_ARG_CHECK(idx != nullptr || d == 0, 4);
  PAMD_ARG_CHECK(A != nullptr || d == 0, 5);
  PAMD_ARG_CHECK(lda >= n, 6);
  PAMD_ARG_CHECK(out != nullptr || d == 0, 7);
  PAMD_ARG_CHECK(ldo >= n, 8);
  if (d == 0 || n == 0) return PAMD_OK;
  row_scatter_kernel<<<(unsigned)d, 256, 0, (hipStream_t)stream>>>(n, idx, A, lda, out, ldo);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

namespace {

// Inclusive prefix sums of the m row weights p in one 1024-thread workgroup:
// thread t sums its contiguous chunk in order, thread 0 scans the 1024 chunk
// totals in order, then each thread writes its chunk's running sums.  A fixed
// summation order, so the CDF is run-to-run identical.
__global__ __launch_bounds__(1024) void weights_cdf_kernel(int64_t m, const double* __restrict__ p,
                                                           double* __restrict__ cdf) {
  __shared__ double tot[1024];
  const int t = threadIdx.x;
  const int64_t chunk = pamd_cdiv(m, 1024);
  const int64_t i0 = imin64(m, (int64_t)t * chunk), i1 = imin64(m, i0 + chunk);
  // the chunk's weights eight at a time (loads issued together, added in order)
  double s = 0.0;
  for (int64_t i = i0; i < i1; i += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = (i + u < i1) ? p[i + u] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (i + u < i1) s += v[u];
  }
  tot[t] = s;
  __syncthreads();
  if (t == 0) {
    double run = 0.0;
    for (int k = 0; k < 1024; k += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = tot[k + u];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        tot[k + u] = run;
        run += v[u];
      }
    }
  }
  __syncthreads();
  double run = tot[t];
  for (int64_t i = i0; i < i1; i += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = (i + u < i1) ? p[i + u] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (i + u < i1) {
        run += v[u];
        cdf[i + u] = run;
      }
  }
}

// Draw i: u in (0, 1] from Philox-4x32-10 at counter i, the row j with
// cdf[j - 1] < u * total <= cdf[j] (binary search; a row of weight 0 is never
// chosen since u > 0), and the scale 1 / sqrt(d p_j / total) of a sampling
// sketch with replacement.
__global__ void sample_rows_kernel(int64_t m, int64_t d, uint64_t seed, const double* __restrict__ p,
                                   const double* __restrict__ cdf, int64_t* __restrict__ idx,
                                   double* __restrict__ scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d) return;
  const U4 r = philox4x32_10(U4{(uint32_t)i, (uint32_t)(i >> 32), 0x3C6EF372u, 0x510E527Fu}, (uint32_t)seed,
                             (uint32_t)(seed >> 32));
  const double total = cdf[m - 1];
  const double target = u53_open0(r.x, r.y) * total;
  int64_t lo = 0, hi = m - 1;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (cdf[mid] >= target)
      hi = mid;
    else
      lo = mid + 1;
  }
  idx[i] = lo;
  scale[i] = 1.0 / sqrt((double)d * p[lo] / total);
}

}  // namespace

extern "C" int pamd_sample_rows_workspace(int64_t m, int64_t* bytes) {
  PAMD_ARG_CHECK(m >= 1, 1);
  PAMD_ARG_CHECK(bytes != nullptr, 2);
  *bytes = m * (int64_t)sizeof(double);
  return PAMD_OK;
}

extern "C" int pamd_sample_rows_f64(int64_t m, const double* p, int64_t d, uint64_t seed, int64_t* idx,
                                    double* scale, void* workspace, void* stream) {
  PAMD_ARG_CHECK(m >= 1, 1);
  PAMD_ARG_CHECK(p != nullptr, 2);
  PAMD_ARG_CHECK(d >= 0, 3);
  PAMD_ARG_CHECK(idx != nullptr || d == 0, 5);
  PAMD_ARG_CHECK(scale != nullptr || d == 0, 6);
  PAMD_ARG_CHECK(workspace != nullptr, 7);
  if (d == 0) return PAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  double* cdf = (double*)workspace;
  weights_cdf_kernel<<<1, 1024, 0, st>>>(m, p, cdf);
  PAMD_LAUNCH_CHECK();
  sample_rows_kernel<<<(unsigned)pamd_cdiv(d, 256), 256, 0, st>>>(m, d, seed, p, cdf, idx, scale);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}
