// Preconditioned LSQR with every recurrence on the device.
//
// Operator: A_pc = [A; sd I] M with M = R^{-1} (sd = sqrt(delta)), or M = I.
// M is applied by backward-stable substitution (R x = v, R^T x = y) inside the
// tail kernel, in LDS -- an explicit inverse would be faster to apply but its
// forward error grows like kappa(R) u and costs an order of magnitude in the
// final backward error.  One LSQR iteration is two kernels:
//
//   lsqr_pass  (multi-workgroup, one read of A): for each row i
//                u_i <- c1 * (A w)_i + c2 * u_i,   y_g += u_i * A[i, :]
//              where w = M v.  The bidiagonalisation needs A_pc v and then
//              A_pc^T applied to the *normalised* new u; since
//              A^T (u / beta) = (A^T u) / beta, both products come out of the
//              same pass: one read of A per iteration instead of two.
//   lsqr_tail  (one 1024-thread workgroup): sums the per-workgroup partials
//              y_g in fixed order (plus the lifted rows sd * I), then beta,
//              v <- R^{-T} y / beta - beta v, alpha, the Givens recurrences
//              and stopping tests of PARLA's lsqr (comps/determiter/lsqr.py:
//              412-526), the x / w updates and w_next = R^{-1} v.  The first call
//              (phase 0) is the set-up of lsqr.py:362-395 instead.
//
// Every kernel returns at once when `done` is set, so the host enqueues
// iterations in batches and reads the state back once per batch.  u is kept
// unnormalised: the 1/beta of lsqr.py:425 is folded into the next pass's
// c2 = -alpha / beta.
#include "common.h"
#include "trsv_dev.h"

namespace {

struct LsqrState {
  double alfa, beta, anorm, acond, ddnorm, res2, xnorm, xxnorm, z, cs2, sn2;
  double rhobar, phibar, rnorm, r1norm, r2norm, arnorm, bnorm;
  double c1, c2;
  int64_t itn, istop, done, early, phase;
};

struct LsqrConst {
  int64_t m, n, lda, ldr, iter_lim;
  double sd, damp, atol, btol, ctol;
  int has_r, calc_var, has_x0, grid;
  int xs_global;  // the substitution vector lives in the workspace (n too large for LDS)
};

// largest n whose substitution vector the tail keeps in LDS next to the solve
// scratch (<= 64 KiB of dynamic LDS)
constexpr int64_t LSQR_LDS_MAX_N = 1900;

// Pass workgroups of lsqr_waves(m, n) wavefronts, one wavefront per row at a
// time: four, or more where that still leaves at least one workgroup per CU
// (wider workgroups leave fewer partials for the tail to sum; the register
// tile of 4 * NPL doubles per lane bounds the width).
__host__ __device__ inline int lsqr_npl(int64_t n) {
  return n <= 64 ? 1 : n <= 128 ? 2 : n <= 256 ? 4 : n <= 512 ? 8 : n <= 1024 ? 16 : 0;
}
// wavefronts over the rows of A: one per 16 rows, at most 4096 (16 per CU)
__host__ __device__ inline int64_t lsqr_total_waves(int64_t m) { return imax64(1, imin64(4096, pamd_cdiv(m, 16))); }
__host__ __device__ inline int lsqr_waves(int64_t m, int64_t n) {
  const int npl = lsqr_npl(n);
  const int wide = (npl >= 1 && npl <= 8) ? 16 : (npl == 16 ? 8 : 4);
  return pamd_cdiv(lsqr_total_waves(m), wide) >= 256 ? wide : 4;
}
__host__ __device__ inline int lsqr_grid(int64_t m, int64_t n) {
  return (int)pamd_cdiv(lsqr_total_waves(m), lsqr_waves(m, n));
}

struct LsqrLayout {
  int64_t st, u, v, w, work, y, pv, pb, xs, dinv, total;
};

__host__ __device__ inline LsqrLayout lsqr_layout(int64_t m, int64_t n) {
  LsqrLayout L;
  const int G = lsqr_grid(m, n);
  L.st = 0;
  L.u = 64;
  L.v = L.u + m + n;
  L.w = L.v + n;
  L.work = L.w + n;
  L.y = L.work + n;
  L.pv = (L.y + n + 31) / 32 * 32;
  L.pb = L.pv + (int64_t)G * n;
  L.xs = (L.pb + (int64_t)G * 2 + 31) / 32 * 32;
  L.dinv = (L.xs + (n > LSQR_LDS_MAX_N ? n : 0) + 31) / 32 * 32;  // inverses of R's 8 x 8 diagonal sub-blocks
  L.total = L.dinv + pamd_cdiv(n, 8) * 64 + 32;
  return L;
}

// ---------------------------------------------------------------------------
// pass over A
// ---------------------------------------------------------------------------
// NPL = columns per lane held in registers (n <= 64 * NPL); NPL = 0: any n,
// the row is read twice (the second read hits cache).  WAVES wavefronts per
// workgroup (lsqr_waves); the workgroup's partial of A^T u' is the sum of its
// wavefronts' in wavefront order.
template <int NPL, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void lsqr_pass_kernel(LsqrConst K, const double* __restrict__ A,
                                                               double* __restrict__ ws) {
  constexpr int THREADS = 64 * WAVES;
  constexpr int RED_ROWS = 4;  // wavefronts whose register tiles share LDS at a time
  LsqrState* S = reinterpret_cast<LsqrState*>(ws);
  if (S->done) return;
  const LsqrLayout L = lsqr_layout(K.m, K.n);
  const double* __restrict__ w = ws + L.work;
  double* __restrict__ u = ws + L.u;
  double* __restrict__ pv = ws + L.pv;
  double* __restrict__ pb = ws + L.pb;
  const double c1 = S->c1, c2 = S->c2;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t n = K.n;
  const int64_t gw = (int64_t)gridDim.x * WAVES;
  double ssq_new = 0.0, ssq_in = 0.0;

  __shared__ double red[RED_ROWS][64 * (NPL > 0 ? NPL : 1)];
  __shared__ double sred[WAVES][2];

  if constexpr (NPL > 0) {
    double wr[NPL], acc[NPL];
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int64_t c = lane + 64 * q;
      wr[q] = (c < n) ? w[c] : 0.0;
      acc[q] = 0.0;
    }
    int64_t row = (int64_t)blockIdx.x * WAVES + wave;
    double a[NPL];
    if (row < K.m) {
#pragma unroll
      for (int q = 0; q < NPL; ++q) {
        const int64_t c = lane + 64 * q;
        a[q] = (c < n) ? A[row * K.lda + c] : 0.0;
      }
    }
    for (; row < K.m; row += gw) {
      // prefetch the next row while this one is reduced
      const int64_t nrow = row + gw;
      double an[NPL];
      if (nrow < K.m) {
#pragma unroll
        for (int q = 0; q < NPL; ++q) {
          const int64_t c = lane + 64 * q;
          an[q] = (c < n) ? A[nrow * K.lda + c] : 0.0;
        }
      }
      const double uin = u[row];
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < NPL; ++q) s += a[q] * wr[q];
      s = wave_sum(s);
      const double un = c1 * s + c2 * uin;
#pragma unroll
      for (int q = 0; q < NPL; ++q) acc[q] += un * a[q];
      ssq_new += un * un;
      ssq_in += uin * uin;
      if (lane == 0) u[row] = un;
      if (nrow < K.m) {
#pragma unroll
        for (int q = 0; q < NPL; ++q) a[q] = an[q];
      }
    }
    if (lane == 0) {
      sred[wave][0] = ssq_new;
      sred[wave][1] = ssq_in;
    }
    // the wavefronts' tiles pass through LDS RED_ROWS at a time; thread
    // columns c = threadIdx.x + THREADS * j add them in wavefront order
    constexpr int CPT = (64 * NPL + THREADS - 1) / THREADS;
    double t[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) t[j] = 0.0;
#pragma unroll
    for (int w0 = 0; w0 < WAVES; w0 += RED_ROWS) {
      if (wave >= w0 && wave < w0 + RED_ROWS) {
#pragma unroll
        for (int q = 0; q < NPL; ++q) red[wave - w0][lane + 64 * q] = acc[q];
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < CPT; ++j) {
        const int c = threadIdx.x + THREADS * j;
        if (c < 64 * NPL) {
#pragma unroll
          for (int wv = 0; wv < RED_ROWS; ++wv)
            if (w0 + wv < WAVES) t[j] += red[wv][c];
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      const int c = threadIdx.x + THREADS * j;
      if (c < n) pv[(int64_t)blockIdx.x * n + c] = t[j];
    }
  } else {
    for (int64_t c = threadIdx.x; c < n; c += THREADS) pv[(int64_t)blockIdx.x * n + c] = 0.0;
    __syncthreads();
    // uniform trip count across the block (the loop body has barriers)
    for (int64_t base = (int64_t)blockIdx.x * WAVES; base < K.m; base += gw) {
      const int64_t row = base + wave;
      const bool valid = row < K.m;
      const double* ar = A + (valid ? row : 0) * K.lda;
      double un = 0.0;
      if (valid) {
        double s = 0.0;
        for (int64_t c = lane; c < n; c += 64) s += ar[c] * w[c];
        s = wave_sum(s);
        const double uin = u[row];
        un = c1 * s + c2 * uin;
        ssq_new += un * un;
        ssq_in += uin * uin;
        if (lane == 0) u[row] = un;
      }
      // accumulate into this block's partial, one wave at a time (fixed order)
      for (int wv = 0; wv < WAVES; ++wv) {
        __syncthreads();
        if (wv == wave && valid)
          for (int64_t c = lane; c < n; c += 64) pv[(int64_t)blockIdx.x * n + c] += un * ar[c];
      }
    }
    if (lane == 0) {
      sred[wave][0] = ssq_new;
      sred[wave][1] = ssq_in;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double a0 = 0.0, a1 = 0.0;
    for (int wv = 0; wv < WAVES; ++wv) {
      a0 += sred[wv][0];
      a1 += sred[wv][1];
    }
    pb[2 * blockIdx.x] = a0;
    pb[2 * blockIdx.x + 1] = a1;
  }
}

__device__ inline double dsign(double a) { return (a > 0.0) ? 1.0 : ((a < 0.0) ? -1.0 : 0.0); }

// Stable Givens rotation of comps/determiter/lsqr.py:63-95.
__device__ inline void sym_ortho(double a, double b, double& c, double& s, double& r) {
  if (b == 0.0) {
    c = dsign(a);
    s = 0.0;
    r = fabs(a);
  } else if (a == 0.0) {
    c = 0.0;
    s = dsign(b);
    r = fabs(b);
  } else if (fabs(b) > fabs(a)) {
    const double tau = a / b;
    s = dsign(b) / sqrt(1.0 + tau * tau);
    c = s * tau;
    r = b / s;
  } else {
    const double tau = b / a;
    c = dsign(a) / sqrt(1.0 + tau * tau);
    s = c * tau;
    r = a / c;
  }
}

// y = sum of pass partials (+ lifted rows) and the ssq totals.  The columns
// go in chunks of 64 (lanes); wpc wavefronts share a chunk, wavefront `sub`
// of it summing partials g = sub, sub + wpc, ... in order with sixteen loads
// in flight, and the wpc sums are added in order.  Every thread gets the
// totals; y is complete on return.
__device__ inline void gather_partials(const LsqrConst& K, double* ws, const LsqrLayout& L, const LsqrState* S,
                                       double* red, double (*cred)[64], double& ssq_new, double& ssq_in) {
  constexpr int GL = 16;
  const int64_t n = K.n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const double* pv = ws + L.pv;
  const int ng = K.grid;
  double* y = ws + L.y;
  double a = 0.0, b = 0.0;
  for (int g = threadIdx.x; g < ng; g += blockDim.x) {
    a += ws[L.pb + 2 * g];
    b += ws[L.pb + 2 * g + 1];
  }
  const double c1 = S->c1, c2 = S->c2;
  const int nch = (int)pamd_cdiv(n, 64);
  const int wpc = max(1, nw / nch);  // wavefronts per column chunk
  const int cpp = nw / wpc;          // chunks per pass
  const int sub = wave % wpc;
  for (int ch0 = 0; ch0 < nch; ch0 += cpp) {
    const int ch = ch0 + wave / wpc;
    const int64_t c = (int64_t)ch * 64 + lane;
    const bool mine = (wave / wpc < cpp) && ch < nch && c < n;
    double t = 0.0;
    if (mine)
      for (int g = sub; g < ng; g += GL * wpc) {
        double pg[GL];
#pragma unroll
        for (int u = 0; u < GL; ++u) pg[u] = (g + u * wpc < ng) ? pv[(int64_t)(g + u * wpc) * n + c] : 0.0;
#pragma unroll
        for (int u = 0; u < GL; ++u)
          if (g + u * wpc < ng) t += pg[u];
      }
    cred[wave][lane] = t;
    __syncthreads();
    if (mine && sub == 0) {
      double tot = 0.0;
      for (int s2 = 0; s2 < wpc; ++s2) tot += cred[wave + s2][lane];
      if (K.sd > 0.0) {
        double* ul = ws + L.u + K.m;
        const double uin = ul[c];
        const double un = c1 * (K.sd * ws[L.work + c]) + c2 * uin;
        ul[c] = un;
        tot += K.sd * un;
        a += un * un;
        b += uin * uin;
      }
      y[c] = tot;
    }
    __syncthreads();
  }
  ssq_new = block_sum(a, red);
  ssq_in = block_sum(b, red);
}

// v_out = R^{-T} (y * inv) + coef * v_in via substitution in LDS; returns
// ||v_out||^2 (block-wide).  v_in may alias v_out.
__device__ inline double apply_adjoint(const LsqrConst& K, const double* __restrict__ R, double* ws,
                                       const LsqrLayout& L, double inv, const double* v_in, double coef,
                                       double* v_out, double* xs, double* scratch, double* red) {
  const int64_t n = K.n;
  const double* y = ws + L.y;
  double ss = 0.0;
  if (K.has_r) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) xs[i] = y[i] * inv;
    __syncthreads();
    trsv_upper_block(1, (int)n, R, K.ldr, xs, scratch, ws + L.dinv);
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      const double t = (coef != 0.0) ? xs[i] + coef * v_in[i] : xs[i];
      v_out[i] = t;
      ss += t * t;
    }
  } else {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      const double t = (coef != 0.0) ? y[i] * inv + coef * v_in[i] : y[i] * inv;
      v_out[i] = t;
      ss += t * t;
    }
  }
  return block_sum(ss, red);
}

// out = R^{-1} v (or v); v complete before the call, out complete after;
// dinv: the inverses of R's 8 x 8 diagonal sub-blocks (trsv_dinv8)
__device__ inline void apply_forward(const LsqrConst& K, const double* __restrict__ R, const double* v,
                                     double* out, double* xs, double* scratch, const double* dinv) {
  if (K.has_r) {
    for (int64_t i = threadIdx.x; i < K.n; i += blockDim.x) xs[i] = v[i];
    __syncthreads();
    trsv_upper_block(0, (int)K.n, R, K.ldr, xs, scratch, dinv);
    for (int64_t i = threadIdx.x; i < K.n; i += blockDim.x) out[i] = xs[i];
  } else {
    for (int64_t i = threadIdx.x; i < K.n; i += blockDim.x) out[i] = v[i];
  }
  __syncthreads();
}

// state, x = x0 or 0, u = b (lifted part 0), work = M x0 with (c1, c2) =
// (-1, 1), or work = 0 with (c1, c2) = (0, 1)
__global__ __launch_bounds__(1024) void lsqr_init_kernel(LsqrConst K, const double* __restrict__ R,
                                                         const double* __restrict__ b, const double* __restrict__ x0,
                                                         double* __restrict__ x, double* __restrict__ var,
                                                         double* __restrict__ ws) {
  extern __shared__ double smem[];
  const LsqrLayout L = lsqr_layout(K.m, K.n);
  LsqrState* S = reinterpret_cast<LsqrState*>(ws);
  double* u = ws + L.u;
  for (int64_t i = threadIdx.x; i < K.m; i += blockDim.x) u[i] = b[i];
  for (int64_t i = threadIdx.x; i < K.n; i += blockDim.x) {
    u[K.m + i] = 0.0;
    x[i] = K.has_x0 ? x0[i] : 0.0;
    if (var) var[i] = 0.0;
  }
  if (K.has_r) trsv_dinv8((int)K.n, R, K.ldr, ws + L.dinv);  // once per solve: R is fixed
  __syncthreads();
  if (K.has_x0) {
    double* xs = K.xs_global ? ws + L.xs : smem;
    apply_forward(K, R, x, ws + L.work, xs, K.xs_global ? smem : smem + K.n, ws + L.dinv);
  } else {
    for (int64_t i = threadIdx.x; i < K.n; i += blockDim.x) ws[L.work + i] = 0.0;
  }
  if (threadIdx.x == 0) {
    LsqrState s = {};
    s.cs2 = -1.0;
    s.c1 = K.has_x0 ? -1.0 : 0.0;
    s.c2 = 1.0;
    *S = s;
  }
}

__global__ __launch_bounds__(1024) void lsqr_tail_kernel(LsqrConst K, const double* __restrict__ R,
                                                         double* __restrict__ x, double* __restrict__ var,
                                                         double* __restrict__ arnorms, double* __restrict__ ws) {
  LsqrState* S = reinterpret_cast<LsqrState*>(ws);
  if (S->done) return;
  extern __shared__ double smem[];
  __shared__ double red[16];
  __shared__ double cred[16][64];
  __shared__ double bc[2];
  const LsqrLayout L = lsqr_layout(K.m, K.n);
  double* xs = K.xs_global ? ws + L.xs : smem;
  double* scratch = K.xs_global ? smem : smem + K.n;
  double* v = ws + L.v;
  double* w = ws + L.w;
  const int64_t n = K.n;
  const double eps = 2.220446049250313e-16;

  double ssq_new, ssq_in;
  gather_partials(K, ws, L, S, red, cred, ssq_new, ssq_in);
  const double beta = sqrt(ssq_new);

  if (S->phase == 0) {
    // set-up after the first pass (lsqr.py:362-395)
    const double bnorm = sqrt(ssq_in);
    double alfa = 0.0;
    if (beta > 0.0) {
      alfa = sqrt(apply_adjoint(K, R, ws, L, 1.0 / beta, v, 0.0, v, xs, scratch, red));
    } else {
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x) v[i] = x[i];
    }
    __syncthreads();
    if (alfa > 0.0) {
      const double ia = 1.0 / alfa;
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x) v[i] = ia * v[i];
    }
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) w[i] = v[i];
    __syncthreads();
    const double arnorm = alfa * beta;
    const bool stop = (arnorm == 0.0) || (K.iter_lim <= 0);
    if (!stop) apply_forward(K, R, v, ws + L.work, xs, scratch, ws + L.dinv);
    __syncthreads();
    if (threadIdx.x == 0) {
      S->bnorm = bnorm;
      S->beta = beta;
      S->alfa = alfa;
      S->rhobar = alfa;
      S->phibar = beta;
      S->rnorm = S->r1norm = S->r2norm = beta;
      S->arnorm = arnorm;
      S->istop = 0;
      S->itn = 0;
      S->early = (arnorm == 0.0) ? 1 : 0;
      S->c1 = 1.0;
      S->c2 = -alfa * (1.0 / beta);
      S->phase = 1;
      S->done = stop ? 1 : 0;
    }
    return;
  }

  double alfa = S->alfa;
  double anorm = S->anorm;
  if (beta > 0.0) {
    anorm = sqrt(anorm * anorm + alfa * alfa + beta * beta + K.damp * K.damp);
    alfa = sqrt(apply_adjoint(K, R, ws, L, 1.0 / beta, v, -beta, v, xs, scratch, red));
    __syncthreads();
    if (alfa > 0.0) {
      const double ia = 1.0 / alfa;
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x) v[i] = ia * v[i];
    }
  }
  // scalar rotations (identical on every thread)
  const double damp = K.damp;
  const double rhobar0 = S->rhobar, phibar0 = S->phibar;
  const double rhobar1 = sqrt(rhobar0 * rhobar0 + damp * damp);
  const double cs1 = rhobar0 / rhobar1;
  const double sn1 = damp / rhobar1;
  const double psi = sn1 * phibar0;
  double phibar = cs1 * phibar0;
  double cs, sn, rho;
  sym_ortho(rhobar1, beta, cs, sn, rho);
  const double theta = sn * alfa;
  const double rhobar = -cs * alfa;
  const double phi = cs * phibar;
  phibar = sn * phibar;
  const double tau = sn * phi;
  const double t1 = phi / rho;
  const double t2 = -theta / rho;
  const double irho = 1.0 / rho;
  __syncthreads();
  double dd = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double wi = w[i];
    const double dk = irho * wi;
    dd += dk * dk;
    if (K.calc_var) var[i] += dk * dk;
    x[i] = x[i] + t1 * wi;
    w[i] = v[i] + t2 * wi;
  }
  const double ddsum = block_sum(dd, red);
  if (threadIdx.x == 0) {
    const int64_t itn_prev = S->itn;
    arnorms[itn_prev] = S->arnorm;  // logged at the top of the iteration (lsqr.py:413)
    const int64_t itn = itn_prev + 1;
    const double ddnorm = S->ddnorm + ddsum;
    const double delta = S->sn2 * rho;
    const double gambar = -S->cs2 * rho;
    const double rhs = phi - delta * S->z;
    const double zbar = rhs / gambar;
    const double xnorm = sqrt(S->xxnorm + zbar * zbar);
    const double gamma = sqrt(gambar * gambar + theta * theta);
    const double cs2 = gambar / gamma;
    const double sn2 = theta / gamma;
    const double z = rhs / gamma;
    const double xxnorm = S->xxnorm + z * z;
    const double acond = anorm * sqrt(ddnorm);
    const double res1 = phibar * phibar;
    const double res2 = S->res2 + psi * psi;
    const double rnorm = sqrt(res1 + res2);
    const double arnorm = alfa * fabs(tau);
    const double r1sq = rnorm * rnorm - damp * damp * xxnorm;
    double r1norm = sqrt(fabs(r1sq));
    if (r1sq < 0) r1norm = -r1norm;
    const double bnorm = S->bnorm;
    const double test1 = rnorm / bnorm;
    const double test2 = arnorm / (anorm * rnorm + eps);
    const double test3 = 1.0 / (acond + eps);
    const double tt1 = test1 / (1.0 + anorm * xnorm / bnorm);
    const double rtol = K.btol + K.atol * anorm * xnorm / bnorm;
    int64_t istop = 0;
    if (itn >= K.iter_lim) istop = 7;
    if (1.0 + test3 <= 1.0) istop = 6;
    if (1.0 + test2 <= 1.0) istop = 5;
    if (1.0 + tt1 <= 1.0) istop = 4;
    if (test3 <= K.ctol) istop = 3;
    if (test2 <= K.atol) istop = 2;
    if (test1 <= rtol) istop = 1;
    S->itn = itn;
    S->istop = istop;
    S->alfa = alfa;
    S->beta = beta;
    S->anorm = anorm;
    S->acond = acond;
    S->ddnorm = ddnorm;
    S->res2 = res2;
    S->xnorm = xnorm;
    S->xxnorm = xxnorm;
    S->z = z;
    S->cs2 = cs2;
    S->sn2 = sn2;
    S->rhobar = rhobar;
    S->phibar = phibar;
    S->rnorm = rnorm;
    S->r1norm = r1norm;
    S->r2norm = rnorm;
    S->arnorm = arnorm;
    S->c1 = 1.0;
    S->c2 = (beta > 0.0) ? -alfa * (1.0 / beta) : -alfa;
    bc[0] = (istop != 0) ? 1.0 : 0.0;
  }
  __syncthreads();
  if (bc[0] == 0.0) apply_forward(K, R, v, ws + L.work, xs, scratch, ws + L.dinv);
  if (threadIdx.x == 0 && bc[0] != 0.0) S->done = 1;
}

// x_orig = R^{-1} x (the M_fwd of comps/determiter/saddle.py:198)
__global__ __launch_bounds__(1024) void lsqr_xorig_kernel(LsqrConst K, const double* __restrict__ R,
                                                          const double* __restrict__ x, double* __restrict__ xo,
                                                          double* __restrict__ ws) {
  extern __shared__ double smem[];
  const LsqrLayout L = lsqr_layout(K.m, K.n);
  double* xs = K.xs_global ? ws + L.xs : smem;
  apply_forward(K, R, x, xo, xs, K.xs_global ? smem : smem + K.n, ws + L.dinv);
}

int launch_pass(const LsqrConst& K, const double* A, double* ws, hipStream_t st) {
  const int npl = lsqr_npl(K.n);
  const bool wide = lsqr_waves(K.m, K.n) > 4;
#define PAMD_PASS(NPL, WAVES) lsqr_pass_kernel<NPL, WAVES><<<K.grid, 64 * WAVES, 0, st>>>(K, A, ws)
  if (npl == 1)
    wide ? PAMD_PASS(1, 16) : PAMD_PASS(1, 4);
  else if (npl == 2)
    wide ? PAMD_PASS(2, 16) : PAMD_PASS(2, 4);
  else if (npl == 4)
    wide ? PAMD_PASS(4, 16) : PAMD_PASS(4, 4);
  else if (npl == 8)
    wide ? PAMD_PASS(8, 16) : PAMD_PASS(8, 4);
  else if (npl == 16)
    wide ? PAMD_PASS(16, 8) : PAMD_PASS(16, 4);
  else
    PAMD_PASS(0, 4);
#undef PAMD_PASS
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

}  // namespace

extern "C" int pamd_lsqr_workspace(int64_t m, int64_t n, int64_t iter_lim, int64_t* bytes) {
  PAMD_ARG_CHECK(m >= 0, 1);
  PAMD_ARG_CHECK(n >= 1, 2);
  PAMD_ARG_CHECK(bytes != nullptr, 4);
  (void)iter_lim;
  *bytes = lsqr_layout(m, n).total * (int64_t)sizeof(double);
  return PAMD_OK;
}

namespace {

// Checks the arguments shared by the blocking and the asynchronous entry
// point, fills K and the dynamic LDS size, and launches the set-up (init,
// first pass over A, tail phase 0).
int lsqr_start(const pamd_lsqr_problem* prob, double* x, double* var, double* arnorms, const void* result,
               void* workspace, hipStream_t st, LsqrConst& K, size_t& shmem) {
  PAMD_ARG_CHECK(prob != nullptr, 1);
  const pamd_lsqr_problem& P = *prob;
  PAMD_ARG_CHECK(P.m >= 0 && P.n >= 1, 1);
  PAMD_ARG_CHECK(P.A != nullptr && P.lda >= P.n, 1);
  PAMD_ARG_CHECK(P.b != nullptr, 1);
  // with R the tail keeps the substitution vector in LDS up to LSQR_LDS_MAX_N, in the workspace beyond
  PAMD_ARG_CHECK(P.R == nullptr || P.ldr >= P.n, 1);
  PAMD_ARG_CHECK(P.n < ((int64_t)1 << 30), 1);
  PAMD_ARG_CHECK(x != nullptr, 2);
  PAMD_ARG_CHECK(var != nullptr || !P.calc_var, 4);
  PAMD_ARG_CHECK(arnorms != nullptr || P.iter_lim <= 0, 5);
  PAMD_ARG_CHECK(result != nullptr, 6);
  PAMD_ARG_CHECK(workspace != nullptr, 7);
  K.m = P.m;
  K.n = P.n;
  K.lda = P.lda;
  K.iter_lim = P.iter_lim;
  K.sd = P.sqrt_delta > 0 ? P.sqrt_delta : 0.0;
  K.damp = P.damp;
  K.atol = P.atol;
  K.btol = P.btol;
  K.ctol = (P.conlim > 0) ? 1.0 / P.conlim : 0.0;
  K.has_r = P.R != nullptr;
  K.calc_var = P.calc_var != 0;
  K.has_x0 = P.x0 != nullptr;
  K.grid = lsqr_grid(P.m, P.n);
  K.ldr = P.ldr;
  K.xs_global = (K.has_r && P.n > LSQR_LDS_MAX_N) ? 1 : 0;
  double* ws = (double*)workspace;
  shmem = (size_t)(K.has_r ? ((K.xs_global ? 0 : P.n) + TRSV_SCRATCH_DOUBLES) : 1) * sizeof(double);
  lsqr_init_kernel<<<1, 1024, shmem, st>>>(K, P.R, P.b, P.x0, x, var, ws);
  PAMD_LAUNCH_CHECK();
  int rc = launch_pass(K, P.A, ws, st);
  if (rc) return rc;
  lsqr_tail_kernel<<<1, 1024, shmem, st>>>(K, P.R, x, var, arnorms, ws);  // phase 0: set-up
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}

// the result fields from the device state (one thread)
__global__ void lsqr_result_kernel(const double* __restrict__ ws, pamd_lsqr_result* __restrict__ r) {
  if (threadIdx.x != 0) return;
  const LsqrState* S = reinterpret_cast<const LsqrState*>(ws);
  r->istop = S->istop;
  r->itn = S->itn;
  r->r1norm = S->r1norm;
  r->r2norm = S->r2norm;
  r->anorm = S->anorm;
  r->acond = S->acond;
  r->arnorm = S->arnorm;
  r->xnorm = S->xnorm;
  r->n_log = S->early ? -1 : S->itn;
}

}  // namespace

extern "C" int pamd_lsqr_run_f64(const pamd_lsqr_problem* prob, double* x, double* x_orig, double* var,
                                 double* arnorms, pamd_lsqr_result* result, void* workspace, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  LsqrConst K;
  size_t shmem = 0;
  int rc = lsqr_start(prob, x, var, arnorms, result, workspace, st, K, shmem);
  if (rc) return rc;
  const pamd_lsqr_problem& P = *prob;
  double* ws = (double*)workspace;

  LsqrState hs;
  int64_t batch = 4;
  int64_t launched = 0;
  for (;;) {
    PAMD_HIP_CHECK(hipMemcpyAsync(&hs, ws, sizeof(LsqrState), hipMemcpyDeviceToHost, st));
    PAMD_HIP_CHECK(hipStreamSynchronize(st));
    if (hs.done || launched >= P.iter_lim) break;
    const int64_t todo = imin64(batch, P.iter_lim - launched);
    for (int64_t it = 0; it < todo; ++it) {
      rc = launch_pass(K, P.A, ws, st);
      if (rc) return rc;
      lsqr_tail_kernel<<<1, 1024, shmem, st>>>(K, P.R, x, var, arnorms, ws);
      PAMD_LAUNCH_CHECK();
    }
    launched += todo;
    batch = imin64(batch * 2, 64);
  }
  if (x_orig) {
    lsqr_xorig_kernel<<<1, 1024, shmem, st>>>(K, P.R, x, x_orig, ws);
    PAMD_LAUNCH_CHECK();
  }
  result->istop = hs.istop;
  result->itn = hs.itn;
  result->r1norm = hs.r1norm;
  result->r2norm = hs.r2norm;
  result->anorm = hs.anorm;
  result->acond = hs.acond;
  result->arnorm = hs.arnorm;
  result->xnorm = hs.xnorm;
  result->n_log = hs.early ? -1 : hs.itn;
  return PAMD_OK;
}

extern "C" int pamd_lsqr_launch_f64(const pamd_lsqr_problem* prob, double* x, double* x_orig, double* var,
                                    double* arnorms, pamd_lsqr_result* result, void* workspace, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  LsqrConst K;
  size_t shmem = 0;
  int rc = lsqr_start(prob, x, var, arnorms, result, workspace, st, K, shmem);
  if (rc) return rc;
  const pamd_lsqr_problem& P = *prob;
  double* ws = (double*)workspace;
  // every iteration is launched; those after the stop return at once
  for (int64_t it = 0; it < P.iter_lim; ++it) {
    rc = launch_pass(K, P.A, ws, st);
    if (rc) return rc;
    lsqr_tail_kernel<<<1, 1024, shmem, st>>>(K, P.R, x, var, arnorms, ws);
    PAMD_LAUNCH_CHECK();
  }
  if (x_orig) {
    lsqr_xorig_kernel<<<1, 1024, shmem, st>>>(K, P.R, x, x_orig, ws);
    PAMD_LAUNCH_CHECK();
  }
  lsqr_result_kernel<<<1, 64, 0, st>>>(ws, result);
  PAMD_LAUNCH_CHECK();
  return PAMD_OK;
}
