// GMRES for (1 - T W) b = T a on the GPU.
//
// Vectors live on the device in the padded layout [n_part][B] (complex FP64);
// slots a particle does not have carry T = 0 and stay zero. Arnoldi uses
// classical Gram-Schmidt applied twice (CGS2) in FP64 with fused multi-dot
// and multi-axpy kernels; the (restart+1) x restart Hessenberg least-squares
// problem is solved on the host with Givens rotations.
//
// SMH_SOLVE_FP64: restarted GMRES with FP64 coupling matvecs, x0 = rhs and
//   stopping rule ||rhs - M x|| <= tol ||rhs|| (smuthi's scipy call,
//   linear_system.py:285-286).
// SMH_SOLVE_MIXED: FP64 iterative refinement from x = 0: an inner GMRES with
//   FP32 coupling matvecs (FP64 Arnoldi) solves M d = r to
//   max(tol/2 ||rhs||, kInnerRel ||r||), x += d, r = rhs - M x with the FP64
//   matvec, until the FP64 residual meets tol. kInnerRel sits at the FP32
//   operator's resolution, so for tolerances down to ~1e-6 the first sweep
//   usually suffices and one FP64 matvec checks it. Inner solves run one
//   restart cycle per sweep, so the FP64 residual is taken after every cycle
//   (a reduced-precision operator never restarts on its own residuals).
// SMH_SOLVE_MIXED16: the same refinement with FP16-operand matvecs in the
//   inner GMRES (batched radial solves), whose target kInnerRel16 sits at
//   that operator's resolution (a few 1e-4): each sweep gains about three
//   digits, so tolerances near 1e-6 take two sweeps. Between sweeps the
//   residual comes from the FP32 matvec; the FP64 residual certifies the
//   result once that meets the tolerance.
#include <chrono>
#include <cmath>
#include <complex>

#include "smh_op.h"

namespace {

typedef std::complex<double> zd;

int blocks_for(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

constexpr int kDotThreads = 256;
constexpr int kDotBlocks = 512;
constexpr int kMaxBasis = 128;
constexpr double kInnerRel = 2e-7;    // inner (FP32-matvec) GMRES target relative to the current FP64 residual
constexpr double kInnerRel16 = 1e-3;  // the same for FP16-operand matvecs
// FP32 residuals drive FP16 sweeps while above this fraction of ||b|| (about
// 30 times the FP32 matvec's error)
constexpr double kFp32ResidualFloor = 1e-5;


// partial[b][j] = sum over this block's slice of conj(V_j) . w
__global__ void multi_dot_kernel(const double2* __restrict__ V, int64_t ld, int nv, const double2* __restrict__ w,
                                 int64_t n, double2* partial) {
  __shared__ double2 red[kDotThreads];
  for (int j = 0; j < nv; ++j) {
    const double2* v = V + (size_t)j * ld;
    double re = 0, im = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      const double2 a = v[i], b = w[i];
      re += a.x * b.x + a.y * b.y;
      im += a.x * b.y - a.y * b.x;
    }
    red[threadIdx.x] = make_double2(re, im);
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) {
        red[threadIdx.x].x += red[threadIdx.x + s].x;
        red[threadIdx.x].y += red[threadIdx.x + s].y;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * kMaxBasis + j] = red[0];
    __syncthreads();
  }
}

__global__ void reduce_partials_kernel(const double2* __restrict__ partial, int nblocks, int nv, double2* out) {
  const int j = threadIdx.x;
  if (j >= nv) return;
  double re = 0, im = 0;
  for (int b = 0; b < nblocks; ++b) {
    re += partial[(size_t)b * kMaxBasis + j].x;
    im += partial[(size_t)b * kMaxBasis + j].y;
  }
  out[j] = make_double2(re, im);
}

// w -= sum_j V_j c_j
__global__ void multi_axpy_kernel(const double2* __restrict__ V, int64_t ld, int nv, const double2* __restrict__ c,
                                  double2* w, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 acc = w[i];
  for (int j = 0; j < nv; ++j) {
    const double2 v = V[(size_t)j * ld + i], cj = c[j];
    acc.x -= v.x * cj.x - v.y * cj.y;
    acc.y -= v.x * cj.y + v.y * cj.x;
  }
  w[i] = acc;
}

// x += sum_j V_j c_j
__global__ void multi_update_kernel(const double2* __restrict__ V, int64_t ld, int nv, const double2* __restrict__ c,
                                    double2* x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 acc = x[i];
  for (int j = 0; j < nv; ++j) {
    const double2 v = V[(size_t)j * ld + i], cj = c[j];
    acc.x += v.x * cj.x - v.y * cj.y;
    acc.y += v.x * cj.y + v.y * cj.x;
  }
  x[i] = acc;
}

__global__ void scale_copy_kernel(const double2* __restrict__ src, double s, double2* dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = make_double2(src[i].x * s, src[i].y * s);
}

// out = v - t o y  (master matrix 1 - T W applied after the coupling matvec)
__global__ void master_fp64_kernel(const double2* __restrict__ v, const double2* __restrict__ t,
                                   const double2* __restrict__ y, double2* out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 a = t[i], b = y[i];
  out[i] = make_double2(v[i].x - (a.x * b.x - a.y * b.y), v[i].y - (a.x * b.y + a.y * b.x));
}

__global__ void master_fp32_kernel(const double2* __restrict__ v, const double2* __restrict__ t,
                                   const float2* __restrict__ y, double2* out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 a = t[i];
  const double bx = y[i].x, by = y[i].y;
  out[i] = make_double2(v[i].x - (a.x * bx - a.y * by), v[i].y - (a.x * by + a.y * bx));
}

__global__ void to_fp32_kernel(const double2* __restrict__ v, float2* out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float2((float)v[i].x, (float)v[i].y);
}

__global__ void sub_kernel(const double2* __restrict__ a, const double2* __restrict__ b, double2* out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double2(a[i].x - b[i].x, a[i].y - b[i].y);
}

__global__ void add_kernel(double2* a, const double2* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = make_double2(a[i].x + b[i].x, a[i].y + b[i].y);
}

struct Solver {
  smh_op* op;
  smh_ctx* ctx;
  hipStream_t s;
  int64_t n;  // padded length
  int restart;
  double2* V = nullptr;        // (restart + 1) basis vectors
  double2* w = nullptr;
  double2* t = nullptr;        // T diagonal (padded)
  double2* partial = nullptr;  // [kDotBlocks][kMaxBasis]
  double2* coef = nullptr;     // kMaxBasis
  double2* y64 = nullptr;      // coupling output FP64
  float2* x32 = nullptr;       // FP32 staging of the matvec input
  int64_t mv64 = 0, mv32 = 0, iters = 0;

  int init() {
    SMH_HIP(ctx, hipMalloc(&V, (size_t)(restart + 1) * n * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&w, n * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&t, n * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&partial, (size_t)kDotBlocks * kMaxBasis * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&coef, kMaxBasis * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&y64, n * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&x32, n * sizeof(float2)));
    return SMH_OK;
  }
  void release() {
    void* p[] = {V, w, t, partial, coef, y64, x32};
    for (void* q : p)
      if (q) hipFree(q);
  }

  // out = M v (padded), precision of the coupling matvec as requested
  int master(const double2* v, double2* out, int precision) {
    int rc;
    if (precision == SMH_FP64) {
      if ((rc = smh::op_apply(op, SMH_FP64, v, y64, false))) return rc;
      hipLaunchKernelGGL(master_fp64_kernel, dim3(blocks_for(n)), dim3(256), 0, s, v, t, y64, out, n);
      ++mv64;
    } else {
      hipLaunchKernelGGL(to_fp32_kernel, dim3(blocks_for(n)), dim3(256), 0, s, v, x32, n);
      if ((rc = smh::op_apply(op, SMH_FP32, x32, op->d_y32, false))) return rc;
      hipLaunchKernelGGL(master_fp32_kernel, dim3(blocks_for(n)), dim3(256), 0, s, v, t, op->d_y32, out, n);
      ++mv32;
    }
    SMH_HIP(ctx, hipGetLastError());
    return SMH_OK;
  }

  // h[0..nv) = V[0..nv)^H w (device -> host)
  int dots(const double2* Vb, int nv, const double2* wv, zd* h) {
    if (nv < 1 || nv > kMaxBasis) return smh_fail(ctx, SMH_ERR_ARG, "basis size out of range");
    hipLaunchKernelGGL(multi_dot_kernel, dim3(kDotBlocks), dim3(kDotThreads), 0, s, Vb, n, nv, wv, n, partial);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kMaxBasis), 0, s, partial, kDotBlocks, nv, coef);
    SMH_HIP(ctx, hipMemcpyAsync(h, coef, nv * sizeof(double2), hipMemcpyDeviceToHost, s));
    SMH_HIP(ctx, hipStreamSynchronize(s));
    return SMH_OK;
  }
  int norm(const double2* v, double& out) {
    zd h;
    int rc = dots(v, 1, v, &h);
    out = std::sqrt(std::max(0.0, h.real()));
    return rc;
  }

  // Restarted GMRES on M x = b from x (in/out), stops at ||b - M x|| <= atol.
  // x_zero: x is zero on entry, so the first residual is b itself.
  // trust_estimate (inner solves of the refinement): one restart cycle at
  // most, returning as soon as the Arnoldi residual estimate reaches atol or
  // the cycle ends, without recomputing the true residual (the caller takes
  // it in FP64 and starts the next sweep: a reduced-precision operator whose
  // error sits above atol would otherwise stall the restarts).
  // On return res_out is the true residual of x (or the estimate, see above)
  // and steps the number of Arnoldi steps taken.
  int gmres(double2* x, const double2* b, double atol, int maxiter, int precision, double& res_out, int& steps,
            bool x_zero = false, bool trust_estimate = false) {
    std::vector<zd> H((size_t)(restart + 1) * restart), cs(restart), sn(restart), g(restart + 1), hcol(restart + 1),
        hcol2(restart + 1), yv(restart);
    int rc;
    steps = 0;
    double beta = 0;
    for (int outer = 0;; ++outer) {
      // r = b - M x -> V_0
      if (outer == 0 && x_zero) {
        SMH_HIP(ctx, hipMemcpyAsync(w, b, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
      } else {
        if ((rc = master(x, w, precision))) return rc;
        hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(n)), dim3(256), 0, s, b, w, w, n);
      }
      if ((rc = norm(w, beta))) return rc;
      res_out = beta;
      if (beta <= atol || steps >= maxiter || beta == 0.0) return SMH_OK;
      hipLaunchKernelGGL(scale_copy_kernel, dim3(blocks_for(n)), dim3(256), 0, s, w, 1.0 / beta, V, n);
      std::fill(g.begin(), g.end(), zd(0));
      g[0] = beta;
      int k = 0;
      for (; k < restart && steps < maxiter; ++k, ++steps) {
        double2* vk = V + (size_t)k * n;
        if ((rc = master(vk, w, precision))) return rc;
        // CGS2 against V_0..V_k
        if ((rc = dots(V, k + 1, w, hcol.data()))) return rc;
        SMH_HIP(ctx, hipMemcpyAsync(coef, hcol.data(), (k + 1) * sizeof(double2), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(multi_axpy_kernel, dim3(blocks_for(n)), dim3(256), 0, s, V, n, k + 1, coef, w, n);
        if ((rc = dots(V, k + 1, w, hcol2.data()))) return rc;
        SMH_HIP(ctx, hipMemcpyAsync(coef, hcol2.data(), (k + 1) * sizeof(double2), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(multi_axpy_kernel, dim3(blocks_for(n)), dim3(256), 0, s, V, n, k + 1, coef, w, n);
        for (int j = 0; j <= k; ++j) hcol[j] += hcol2[j];
        double hn;
        if ((rc = norm(w, hn))) return rc;
        hcol[k + 1] = hn;
        if (hn > 0)
          hipLaunchKernelGGL(scale_copy_kernel, dim3(blocks_for(n)), dim3(256), 0, s, w, 1.0 / hn,
                             V + (size_t)(k + 1) * n, n);
        // Givens rotations on the new column
        for (int j = 0; j < k; ++j) {
          const zd a = hcol[j], bb = hcol[j + 1];
          hcol[j] = std::conj(cs[j]) * a + std::conj(sn[j]) * bb;
          hcol[j + 1] = -sn[j] * a + cs[j] * bb;
        }
        const zd a = hcol[k], bb = hcol[k + 1];
        const double den = std::sqrt(std::norm(a) + std::norm(bb));
        cs[k] = den > 0 ? a / den : zd(1);
        sn[k] = den > 0 ? bb / den : zd(0);
        hcol[k] = std::conj(cs[k]) * a + std::conj(sn[k]) * bb;
        hcol[k + 1] = 0;
        g[k + 1] = -sn[k] * g[k];
        g[k] = std::conj(cs[k]) * g[k];
        for (int j = 0; j <= k; ++j) H[(size_t)j * restart + k] = hcol[j];
        ++iters;
        if (std::abs(g[k + 1]) <= atol || hn == 0.0) {
          ++k;
          ++steps;
          break;
        }
      }
      // back substitution and update x += V y
      for (int i = k - 1; i >= 0; --i) {
        zd sum = g[i];
        for (int j = i + 1; j < k; ++j) sum -= H[(size_t)i * restart + j] * yv[j];
        yv[i] = sum / H[(size_t)i * restart + i];
      }
      if (k > 0) {
        SMH_HIP(ctx, hipMemcpyAsync(coef, yv.data(), k * sizeof(double2), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(multi_update_kernel, dim3(blocks_for(n)), dim3(256), 0, s, V, n, k, coef, x, n);
      }
      if (trust_estimate) {
        res_out = std::abs(g[k]);
        return SMH_OK;
      }
    }
  }
};


// ---- kBatch right-hand sides interleaved: element e = i * kBatch + r ----
constexpr int kB = smh::kBatchRhs;

// partial[(b * kMaxBasis + j) * kB + r] = block b's share of conj(V_j[., r]) . w[., r]
__global__ void bmulti_dot_kernel(const double2* __restrict__ V, int64_t ld, int nv, const double2* __restrict__ w,
                                  int64_t n8, double2* partial) {
  __shared__ double2 red[kDotThreads];
  for (int j = 0; j < nv; ++j) {
    const double2* v = V + (size_t)j * ld;
    double re = 0, im = 0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n8; e += (int64_t)gridDim.x * blockDim.x) {
      const double2 a = v[e], b = w[e];
      re += a.x * b.x + a.y * b.y;
      im += a.x * b.y - a.y * b.x;
    }
    red[threadIdx.x] = make_double2(re, im);
    __syncthreads();
    for (int s = blockDim.x / 2; s >= kB; s >>= 1) {  // strides keep e % kB
      if ((int)threadIdx.x < s) {
        red[threadIdx.x].x += red[threadIdx.x + s].x;
        red[threadIdx.x].y += red[threadIdx.x + s].y;
      }
      __syncthreads();
    }
    if ((int)threadIdx.x < kB) partial[((size_t)blockIdx.x * kMaxBasis + j) * kB + threadIdx.x] = red[threadIdx.x];
    __syncthreads();
  }
}

// one workgroup per (basis vector, right-hand side), a fixed-order tree over the partials
__global__ void __launch_bounds__(256) breduce_partials_kernel(const double2* __restrict__ partial, int nblocks,
                                                               double2* out) {
  const int jr = blockIdx.x;
  __shared__ double2 red[256];
  double re = 0, im = 0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    const double2 p = partial[(size_t)b * kMaxBasis * kB + jr];
    re += p.x;
    im += p.y;
  }
  red[threadIdx.x] = make_double2(re, im);
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x].x += red[threadIdx.x + w].x;
      red[threadIdx.x].y += red[threadIdx.x + w].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[jr] = red[0];
}

// w[e] += sgn * sum_j V_j[e] c[j][e % kB]
__global__ void bmulti_axpy_kernel(const double2* __restrict__ V, int64_t ld, int nv, const double2* __restrict__ c,
                                   double sgn, double2* w, int64_t n8) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n8) return;
  const int r = (int)(e % kB);
  double2 acc = w[e];
  for (int j = 0; j < nv; ++j) {
    const double2 v = V[(size_t)j * ld + e], cj = c[j * kB + r];
    acc.x += sgn * (v.x * cj.x - v.y * cj.y);
    acc.y += sgn * (v.x * cj.y + v.y * cj.x);
  }
  w[e] = acc;
}

__global__ void bscale_copy_kernel(const double2* __restrict__ src, const double* __restrict__ sc, double2* dst,
                                   int64_t n8) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n8) return;
  const double f = sc[e % kB];
  dst[e] = make_double2(src[e].x * f, src[e].y * f);
}

// out = v - t o y with t per unknown, shared by the right-hand sides
__global__ void bmaster_fp64_kernel(const double2* __restrict__ v, const double2* __restrict__ t,
                                    const double2* __restrict__ y, double2* out, int64_t n8) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n8) return;
  const double2 a = t[e / kB], b = y[e];
  out[e] = make_double2(v[e].x - (a.x * b.x - a.y * b.y), v[e].y - (a.x * b.y + a.y * b.x));
}

// FP64 padded [n][kB] -> FP32 input of the batched kernel (b32_index layout)
__global__ void to_fp32_b32_kernel(const double2* __restrict__ v, float2* out, int64_t n8, int B, int S,
                                   const int* __restrict__ slot) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n8) return;
  const int r = (int)(e % kB);
  const int64_t ip = e / kB, p = ip / B;
  out[smh::b32_index(p, (int)(ip - p * B), r, S, slot)] = make_float2((float)v[e].x, (float)v[e].y);
}

__global__ void bmaster_fp32_kernel(const double2* __restrict__ v, const double2* __restrict__ t,
                                    const float2* __restrict__ y, double2* out, int64_t n8) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n8) return;
  const double2 a = t[e / kB];
  const double bx = y[e].x, by = y[e].y;
  out[e] = make_double2(v[e].x - (a.x * bx - a.y * by), v[e].y - (a.x * by + a.y * bx));
}

// Restarted GMRES on kB systems at once, sharing every coupling matvec.
struct BatchSolver {
  smh_op* op;
  smh_ctx* ctx;
  hipStream_t s;
  int64_t n, n8;  // padded length, times kB
  int restart;
  double2* V = nullptr;        // (restart + 1) basis blocks [n][kB]
  double2* w = nullptr;
  double2* t = nullptr;        // T diagonal (padded, [n])
  double2* partial = nullptr;  // [kDotBlocks][kMaxBasis][kB]
  double2* coef = nullptr;     // [kMaxBasis][kB]
  double* scl = nullptr;       // [kB]
  double2* y64 = nullptr;      // coupling output FP64 [n][kB]
  float2* x32 = nullptr;
  void* x16 = nullptr;  // FP16 input of the FP16-operand kernel (allocated on first use)
  float2* y32 = nullptr;
  int64_t mv64 = 0, mv32 = 0, mv16 = 0, iters = 0;

  int init() {
    SMH_HIP(ctx, hipMalloc(&V, (size_t)(restart + 1) * n8 * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&w, n8 * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&t, n * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&partial, (size_t)kDotBlocks * kMaxBasis * kB * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&coef, (size_t)kMaxBasis * kB * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&scl, kB * sizeof(double)));
    SMH_HIP(ctx, hipMalloc(&y64, n8 * sizeof(double2)));
    // FP32 input in the batched kernel's layout (padding slots stay zero)
    const size_t n32 = (size_t)op->n_part * op->b32_S * kB;
    SMH_HIP(ctx, hipMalloc(&x32, n32 * sizeof(float2)));
    SMH_HIP(ctx, hipMemsetAsync(x32, 0, n32 * sizeof(float2), s));
    SMH_HIP(ctx, hipMalloc(&y32, n8 * sizeof(float2)));
    return SMH_OK;
  }
  void release() {
    void* p[] = {V, w, t, partial, coef, scl, y64, x32, x16, y32};
    for (void* q : p)
      if (q) hipFree(q);
  }

  int master(const double2* v, double2* out, int precision) {
    int rc;
    if (precision == SMH_FP64) {
      if ((rc = smh::op_apply_batch(op, SMH_FP64, v, y64, false))) return rc;
      hipLaunchKernelGGL(bmaster_fp64_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, v, t, y64, out, n8);
      ++mv64;
    } else if (precision == SMH_FP16) {
      if (!x16) SMH_HIP(ctx, hipMalloc(&x16, (size_t)op->n_part * op->b32_S * kB * 2 * sizeof(_Float16)));
      if ((rc = smh::op_to_half_batch(op, v, x16))) return rc;
      if ((rc = smh::op_apply_batch(op, SMH_FP16, x16, y32, false))) return rc;
      hipLaunchKernelGGL(bmaster_fp32_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, v, t, y32, out, n8);
      ++mv16;
    } else {
      hipLaunchKernelGGL(to_fp32_b32_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, v, x32, n8, op->B, op->b32_S,
                         (const int*)op->d_b32_tab);
      if ((rc = smh::op_apply_batch(op, SMH_FP32, x32, y32, false))) return rc;
      hipLaunchKernelGGL(bmaster_fp32_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, v, t, y32, out, n8);
      ++mv32;
    }
    SMH_HIP(ctx, hipGetLastError());
    return SMH_OK;
  }

  // h[j][r] = V_j[., r]^H w[., r] for j < nv (device -> host)
  int dots(const double2* Vb, int nv, const double2* wv, zd* h) {
    if (nv < 1 || nv > kMaxBasis) return smh_fail(ctx, SMH_ERR_ARG, "basis size out of range");
    hipLaunchKernelGGL(bmulti_dot_kernel, dim3(kDotBlocks), dim3(kDotThreads), 0, s, Vb, n8, nv, wv, n8, partial);
    hipLaunchKernelGGL(breduce_partials_kernel, dim3(nv * kB), dim3(256), 0, s, partial, kDotBlocks, coef);
    SMH_HIP(ctx, hipMemcpyAsync(h, coef, (size_t)nv * kB * sizeof(double2), hipMemcpyDeviceToHost, s));
    SMH_HIP(ctx, hipStreamSynchronize(s));
    return SMH_OK;
  }
  int norms(const double2* v, double* out) {
    zd h[kB];
    int rc = dots(v, 1, v, h);
    for (int r = 0; r < kB; ++r) out[r] = std::sqrt(std::max(0.0, h[r].real()));
    return rc;
  }
  int set_scale(const double* f, const double2* src, double2* dst) {
    SMH_HIP(ctx, hipMemcpyAsync(scl, f, kB * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(bscale_copy_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, src, scl, dst, n8);
    return SMH_OK;
  }

  // GMRES on M x = b for the systems r with active[r] (others are left
  // untouched), each to ||b - M x|| <= atol[r]; x_zero / trust_estimate as
  // for the single-RHS solver. res_out[r]: residual (or estimate) of x.
  int gmres(double2* x, const double2* b, const double* atol, const bool* active, int maxiter, int precision,
            double* res_out, int& steps, bool x_zero, bool trust_estimate) {
    std::vector<zd> H((size_t)kB * (restart + 1) * restart), cs((size_t)kB * restart), sn((size_t)kB * restart),
        g((size_t)kB * (restart + 1)), yv((size_t)kMaxBasis * kB), hc((size_t)(restart + 1) * kB),
        hc2((size_t)(restart + 1) * kB);
    int rc;
    steps = 0;
    for (int outer = 0;; ++outer) {
      if (outer == 0 && x_zero) {
        SMH_HIP(ctx, hipMemcpyAsync(w, b, n8 * sizeof(double2), hipMemcpyDeviceToDevice, s));
      } else {
        if ((rc = master(x, w, precision))) return rc;
        hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, b, w, w, n8);
      }
      double beta[kB];
      if ((rc = norms(w, beta))) return rc;
      bool cont[kB], live[kB], any = false;
      int kr[kB];
      double f[kB];
      for (int r = 0; r < kB; ++r) {
        res_out[r] = beta[r];
        cont[r] = active[r] && beta[r] > atol[r] && beta[r] > 0.0;
        live[r] = cont[r];
        kr[r] = 0;
        f[r] = cont[r] ? 1.0 / beta[r] : 0.0;
        any = any || cont[r];
      }
      if (!any || steps >= maxiter) return SMH_OK;
      set_scale(f, w, V);
      std::fill(g.begin(), g.end(), zd(0));
      for (int r = 0; r < kB; ++r) g[(size_t)r * (restart + 1)] = cont[r] ? beta[r] : 0.0;
      int k = 0;
      for (; k < restart && steps < maxiter; ++k, ++steps) {
        double2* vk = V + (size_t)k * n8;
        if ((rc = master(vk, w, precision))) return rc;
        // CGS2 against V_0..V_k, all systems at once
        if ((rc = dots(V, k + 1, w, hc.data()))) return rc;
        SMH_HIP(ctx, hipMemcpyAsync(coef, hc.data(), (size_t)(k + 1) * kB * sizeof(double2), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(bmulti_axpy_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, V, n8, k + 1, coef, -1.0, w, n8);
        if ((rc = dots(V, k + 1, w, hc2.data()))) return rc;
        SMH_HIP(ctx, hipMemcpyAsync(coef, hc2.data(), (size_t)(k + 1) * kB * sizeof(double2), hipMemcpyHostToDevice,
                                    s));
        hipLaunchKernelGGL(bmulti_axpy_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, V, n8, k + 1, coef, -1.0, w, n8);
        double hn[kB];
        if ((rc = norms(w, hn))) return rc;
        for (int r = 0; r < kB; ++r) f[r] = (live[r] && hn[r] > 0.0) ? 1.0 / hn[r] : 0.0;
        set_scale(f, w, V + (size_t)(k + 1) * n8);
        bool any_live = false;
        for (int r = 0; r < kB; ++r) {
          if (!live[r]) continue;
          zd* Hr = H.data() + (size_t)r * (restart + 1) * restart;
          zd* csr = cs.data() + (size_t)r * restart;
          zd* snr = sn.data() + (size_t)r * restart;
          zd* gr = g.data() + (size_t)r * (restart + 1);
          std::vector<zd> col(k + 2);
          for (int j = 0; j <= k; ++j) col[j] = hc[(size_t)j * kB + r] + hc2[(size_t)j * kB + r];
          col[k + 1] = hn[r];
          for (int j = 0; j < k; ++j) {
            const zd a = col[j], bb = col[j + 1];
            col[j] = std::conj(csr[j]) * a + std::conj(snr[j]) * bb;
            col[j + 1] = -snr[j] * a + csr[j] * bb;
          }
          const zd a = col[k], bb = col[k + 1];
          const double den = std::sqrt(std::norm(a) + std::norm(bb));
          csr[k] = den > 0 ? a / den : zd(1);
          snr[k] = den > 0 ? bb / den : zd(0);
          col[k] = std::conj(csr[k]) * a + std::conj(snr[k]) * bb;
          gr[k + 1] = -snr[k] * gr[k];
          gr[k] = std::conj(csr[k]) * gr[k];
          for (int j = 0; j <= k; ++j) Hr[(size_t)j * restart + k] = col[j];
          kr[r] = k + 1;
          if (std::abs(gr[k + 1]) <= atol[r] || hn[r] == 0.0)
            live[r] = false;
          else
            any_live = true;
        }
        ++iters;
        if (!any_live) {
          ++k;
          ++steps;
          break;
        }
      }
      // per system back substitution, then x += V y for all systems at once
      std::fill(yv.begin(), yv.end(), zd(0));
      for (int r = 0; r < kB; ++r) {
        if (!cont[r]) continue;
        const zd* Hr = H.data() + (size_t)r * (restart + 1) * restart;
        const zd* gr = g.data() + (size_t)r * (restart + 1);
        const int m = kr[r];
        for (int i = m - 1; i >= 0; --i) {
          zd sum = gr[i];
          for (int j = i + 1; j < m; ++j) sum -= Hr[(size_t)i * restart + j] * yv[(size_t)j * kB + r];
          yv[(size_t)i * kB + r] = sum / Hr[(size_t)i * restart + i];
        }
        res_out[r] = std::abs(gr[m]);
      }
      if (k > 0) {
        SMH_HIP(ctx, hipMemcpyAsync(coef, yv.data(), (size_t)k * kB * sizeof(double2), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(bmulti_axpy_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, V, n8, k, coef, 1.0, x, n8);
      }
      if (trust_estimate) return SMH_OK;
    }
  }
};

}  // namespace

// Blocks of kB right-hand sides solved together (every coupling matvec is one
// batched launch for the whole block); same modes and stopping rule as below.
static int solve_batched(smh_op* op, const double* t_diag, const double* rhs, double* x, int64_t nrhs, double tol,
                         int restart, int maxiter, int mode, double* stats,
                         std::chrono::steady_clock::time_point t0) {
  smh_ctx* ctx = op->ctx;
  hipStream_t s = ctx->stream;
  BatchSolver S;
  S.op = op;
  S.ctx = ctx;
  S.s = s;
  S.n = op->n_part * op->B;
  S.n8 = S.n * kB;
  S.restart = restart;
  int rc = S.init();
  double2 *b_pad = nullptr, *x_pad = nullptr, *d_pad = nullptr, *r_pad = nullptr;
  auto done = [&](int code) {
    S.release();
    void* p[] = {b_pad, x_pad, d_pad, r_pad};
    for (void* q : p)
      if (q) hipFree(q);
    return code;
  };
  if (rc) return done(rc);
  if ((rc = smh::op_reserve_packed(op, kB))) return done(rc);
  if (hipMalloc(&b_pad, S.n8 * sizeof(double2)) != hipSuccess || hipMalloc(&x_pad, S.n8 * sizeof(double2)) != hipSuccess ||
      hipMalloc(&d_pad, S.n8 * sizeof(double2)) != hipSuccess || hipMalloc(&r_pad, S.n8 * sizeof(double2)) != hipSuccess)
    return done(smh_fail(ctx, SMH_ERR_HIP, "solver allocation failed"));
  SMH_HIP(ctx, hipMemcpyAsync(op->d_pk_in, t_diag, op->n_unk * sizeof(double2), hipMemcpyHostToDevice, s));
  if ((rc = smh::op_unpack(op, SMH_FP64, op->d_pk_in, S.t))) return done(rc);
  double worst = 0, sweeps_total = 0;
  int64_t steps_total = 0;
  bool converged_all = true;
  for (int64_t r0 = 0; r0 < nrhs; r0 += kB) {
    const int nb = (int)std::min<int64_t>(kB, nrhs - r0);
    SMH_HIP(ctx, hipMemcpyAsync(op->d_pk_in, rhs + 2 * r0 * op->n_unk, (size_t)nb * op->n_unk * sizeof(double2),
                                hipMemcpyHostToDevice, s));
    if ((rc = smh::op_unpack_batch(op, SMH_FP64, op->d_pk_in, nb, b_pad))) return done(rc);
    double bnorm[kB], atol[kB], res[kB];
    bool active[kB];
    if ((rc = S.norms(b_pad, bnorm))) return done(rc);
    for (int r = 0; r < kB; ++r) {
      atol[r] = tol * bnorm[r];
      active[r] = r < nb && bnorm[r] > 0.0;
    }
    int steps = 0;
    if (mode == SMH_SOLVE_FP64) {
      SMH_HIP(ctx, hipMemcpyAsync(x_pad, b_pad, S.n8 * sizeof(double2), hipMemcpyDeviceToDevice, s));
      if ((rc = S.gmres(x_pad, b_pad, atol, active, maxiter, SMH_FP64, res, steps, false, false))) return done(rc);
      steps_total += steps;
      sweeps_total += 1;
    } else {
      SMH_HIP(ctx, hipMemsetAsync(x_pad, 0, S.n8 * sizeof(double2), s));
      SMH_HIP(ctx, hipMemcpyAsync(r_pad, b_pad, S.n8 * sizeof(double2), hipMemcpyDeviceToDevice, s));
      for (int r = 0; r < kB; ++r) res[r] = bnorm[r];
      bool fresh = true;
      bool fp64_residuals = mode != SMH_SOLVE_MIXED16;
      double prev_res[kB];  // residual each system's last sweep started from
      const int inner_prec = mode == SMH_SOLVE_MIXED16 ? SMH_FP16 : SMH_FP32;
      int budget = maxiter;
      for (int sweep = 0; sweep < 50; ++sweep) {
        if (sweep > 0) {
          // FP16 inner solves: the residual that drives the next sweep comes
          // from the FP32 matvec (accurate to a few 1e-7 of ||b||, far below
          // the FP16 sweep's own error) while it is well above that floor;
          // once it meets the tolerance or nears the floor, FP64 residuals
          // certify the result and drive any further sweeps
          bool certified = fp64_residuals;
          if (!certified) {
            if ((rc = S.master(x_pad, r_pad, SMH_FP32))) return done(rc);
            hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(S.n8)), dim3(256), 0, s, b_pad, r_pad, r_pad, S.n8);
            if ((rc = S.norms(r_pad, res))) return done(rc);
            bool need64 = false, last = true;
            for (int r = 0; r < kB; ++r) {
              if (!active[r]) continue;
              need64 = need64 || res[r] <= std::max(atol[r], kFp32ResidualFloor * bnorm[r]);
              // the next sweep gains about the factor this one did: if that
              // takes every system to the tolerance or to the FP32 floor, its
              // residual is taken in FP64 directly
              last = last && prev_res[r] > 0.0 &&
                     res[r] * (res[r] / prev_res[r]) <= std::max(0.5 * atol[r], kFp32ResidualFloor * bnorm[r]);
            }
            certified = need64;
            fp64_residuals = need64 || last;
          }
          if (certified) {
            if ((rc = S.master(x_pad, r_pad, SMH_FP64))) return done(rc);
            hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(S.n8)), dim3(256), 0, s, b_pad, r_pad, r_pad, S.n8);
            if ((rc = S.norms(r_pad, res))) return done(rc);
            fresh = true;
          }
        }
        bool act[kB], any = false;
        double inner_atol[kB];
        for (int r = 0; r < kB; ++r) {
          prev_res[r] = res[r];
          act[r] = active[r] && res[r] > atol[r];
          inner_atol[r] = std::max(atol[r] * 0.5, (mode == SMH_SOLVE_MIXED16 ? kInnerRel16 : kInnerRel) * res[r]);
          any = any || act[r];
        }
        if (!any || budget <= 0) break;
        SMH_HIP(ctx, hipMemsetAsync(d_pad, 0, S.n8 * sizeof(double2), s));
        double inner_res[kB];
        int inner_steps = 0;
        if ((rc = S.gmres(d_pad, r_pad, inner_atol, act, budget, inner_prec, inner_res, inner_steps, true, true)))
          return done(rc);
        budget -= inner_steps;
        steps_total += inner_steps;
        hipLaunchKernelGGL(add_kernel, dim3(blocks_for(S.n8)), dim3(256), 0, s, x_pad, d_pad, S.n8);
        sweeps_total += 1;
        fresh = false;
      }
      if (!fresh) {
        if ((rc = S.master(x_pad, r_pad, SMH_FP64))) return done(rc);
        hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(S.n8)), dim3(256), 0, s, b_pad, r_pad, r_pad, S.n8);
        if ((rc = S.norms(r_pad, res))) return done(rc);
      }
    }
    for (int r = 0; r < nb; ++r) {
      const double rel = bnorm[r] > 0 ? res[r] / bnorm[r] : 0.0;
      worst = std::max(worst, rel);
      if (rel > tol) converged_all = false;
    }
    if ((rc = smh::op_pack_batch(op, SMH_FP64, x_pad, nb, op->d_pk_out))) return done(rc);
    SMH_HIP(ctx, hipMemcpyAsync(x + 2 * r0 * op->n_unk, op->d_pk_out, (size_t)nb * op->n_unk * sizeof(double2),
                                hipMemcpyDeviceToHost, s));
  }
  SMH_HIP(ctx, hipStreamSynchronize(s));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    stats[0] = (double)steps_total;
    stats[1] = (double)S.mv64;
    stats[2] = (double)S.mv32;
    stats[3] = worst;
    stats[4] = secs;
    stats[5] = sweeps_total;
    stats[6] = (double)kB;  // right-hand sides per batched matvec
    stats[7] = (double)S.mv16;
  }
  if (!converged_all) {
    smh_fail(ctx, SMH_ERR_NOCONV, "GMRES stopped at relative residual %.3e (tol %.1e)", worst, tol);
    return done(SMH_ERR_NOCONV);
  }
  return done(SMH_OK);
}

extern "C" int smh_solve(smh_op* op, const double* t_diag, const double* rhs, double* x, int64_t nrhs, double tol,
                         int restart, int maxiter, int mode, double* stats) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, t_diag && rhs && x && nrhs >= 1, SMH_ERR_ARG, "invalid solve arguments");
  SMH_CHECK(ctx, tol > 0 && restart >= 1 && restart < kMaxBasis && maxiter >= 1, SMH_ERR_ARG,
            "invalid tolerance / restart / maxiter");
  SMH_CHECK(ctx, mode == SMH_SOLVE_FP64 || mode == SMH_SOLVE_MIXED || mode == SMH_SOLVE_MIXED16, SMH_ERR_ARG,
            "invalid mode %d", mode);
  SMH_CHECK(ctx, mode != SMH_SOLVE_MIXED16 || (smh::op_has_batch(op, SMH_FP16) && smh::op_has_batch(op, SMH_FP64)),
            SMH_ERR_UNSUPPORTED, "FP16-operand solves need a radial operator with the batched FP16 kernel");
  SMH_CHECK(ctx, op->has_table, SMH_ERR_STATE, "coupling table not set");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  Solver S;
  S.op = op;
  S.ctx = ctx;
  S.s = ctx->stream;
  S.n = op->n_part * op->B;
  S.restart = restart;
  int rc = S.init();
  if (rc) {
    S.release();
    return rc;
  }
  auto done = [&](int code) {
    S.release();
    return code;
  };
  rc = smh::op_reserve_packed(op, 1);
  if (rc) return done(rc);
  double2 *b_pad = nullptr, *x_pad = nullptr, *d_pad = nullptr, *r_pad = nullptr;
  if (hipMalloc(&b_pad, S.n * sizeof(double2)) != hipSuccess || hipMalloc(&x_pad, S.n * sizeof(double2)) != hipSuccess ||
      hipMalloc(&d_pad, S.n * sizeof(double2)) != hipSuccess || hipMalloc(&r_pad, S.n * sizeof(double2)) != hipSuccess)
    return done(smh_fail(ctx, SMH_ERR_HIP, "solver allocation failed"));
  auto done2 = [&](int code) {
    hipFree(b_pad);
    hipFree(x_pad);
    hipFree(d_pad);
    hipFree(r_pad);
    return done(code);
  };
  hipStream_t s = ctx->stream;
  if ((nrhs > 1 && smh::op_has_batch(op, SMH_FP32) && smh::op_has_batch(op, SMH_FP64)) || mode == SMH_SOLVE_MIXED16) {
    const int code = solve_batched(op, t_diag, rhs, x, nrhs, tol, restart, maxiter, mode, stats, t0);
    return done2(code);
  }
  // T diagonal in padded layout
  if (hipMemcpyAsync(op->d_pk_in, t_diag, op->n_unk * sizeof(double2), hipMemcpyHostToDevice, s) != hipSuccess)
    return done2(smh_fail(ctx, SMH_ERR_HIP, "upload failed"));
  if ((rc = smh::op_unpack(op, SMH_FP64, op->d_pk_in, S.t))) return done2(rc);
  double worst = 0, sweeps_total = 0;
  int64_t steps_total = 0;
  bool converged_all = true;
  for (int64_t r = 0; r < nrhs; ++r) {
    if (hipMemcpyAsync(op->d_pk_in, rhs + 2 * r * op->n_unk, op->n_unk * sizeof(double2), hipMemcpyHostToDevice, s) !=
        hipSuccess)
      return done2(smh_fail(ctx, SMH_ERR_HIP, "upload failed"));
    if ((rc = smh::op_unpack(op, SMH_FP64, op->d_pk_in, b_pad))) return done2(rc);
    double bnorm;
    if ((rc = S.norm(b_pad, bnorm))) return done2(rc);
    const double atol = tol * bnorm;
    double res = 0;
    int steps = 0;
    if (mode == SMH_SOLVE_FP64) {
      // x0 = rhs, as smuthi passes rhs as initial guess; GMRES returns the true residual of x
      if (hipMemcpyAsync(x_pad, b_pad, S.n * sizeof(double2), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return done2(smh_fail(ctx, SMH_ERR_HIP, "copy failed"));
      if ((rc = S.gmres(x_pad, b_pad, atol, maxiter, SMH_FP64, res, steps))) return done2(rc);
      steps_total += steps;
      sweeps_total += 1;
    } else {
      // iterative refinement from x0 = 0: FP64 residual, FP32-matvec inner GMRES
      SMH_HIP(ctx, hipMemsetAsync(x_pad, 0, S.n * sizeof(double2), s));
      SMH_HIP(ctx, hipMemcpyAsync(r_pad, b_pad, S.n * sizeof(double2), hipMemcpyDeviceToDevice, s));
      res = bnorm;
      bool fresh = true;  // res is the FP64 residual of the current x
      int budget = maxiter;
      for (int sweep = 0; sweep < 50; ++sweep) {
        if (sweep > 0) {
          if ((rc = S.master(x_pad, r_pad, SMH_FP64))) return done2(rc);
          hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(S.n)), dim3(256), 0, s, b_pad, r_pad, r_pad, S.n);
          if ((rc = S.norm(r_pad, res))) return done2(rc);
          fresh = true;
        }
        if (res <= atol || budget <= 0) break;
        SMH_HIP(ctx, hipMemsetAsync(d_pad, 0, S.n * sizeof(double2), s));
        double inner_res = 0;
        int inner_steps = 0;
        const double inner_atol = std::max(atol * 0.5, kInnerRel * res);
        if ((rc = S.gmres(d_pad, r_pad, inner_atol, budget, SMH_FP32, inner_res, inner_steps, true, true)))
          return done2(rc);
        budget -= inner_steps;
        steps_total += inner_steps;
        hipLaunchKernelGGL(add_kernel, dim3(blocks_for(S.n)), dim3(256), 0, s, x_pad, d_pad, S.n);
        sweeps_total += 1;
        fresh = false;
      }
      if (!fresh) {
        if ((rc = S.master(x_pad, r_pad, SMH_FP64))) return done2(rc);
        hipLaunchKernelGGL(sub_kernel, dim3(blocks_for(S.n)), dim3(256), 0, s, b_pad, r_pad, r_pad, S.n);
        if ((rc = S.norm(r_pad, res))) return done2(rc);
      }
    }
    const double rel = bnorm > 0 ? res / bnorm : 0.0;
    worst = std::max(worst, rel);
    if (rel > tol) converged_all = false;
    if ((rc = smh::op_pack(op, SMH_FP64, x_pad, op->d_pk_out))) return done2(rc);
    if (hipMemcpyAsync(x + 2 * r * op->n_unk, op->d_pk_out, op->n_unk * sizeof(double2), hipMemcpyDeviceToHost, s) !=
        hipSuccess)
      return done2(smh_fail(ctx, SMH_ERR_HIP, "download failed"));
  }
  SMH_HIP(ctx, hipStreamSynchronize(s));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) {
    stats[0] = (double)steps_total;
    stats[1] = (double)S.mv64;
    stats[2] = (double)S.mv32;
    stats[3] = worst;
    stats[4] = secs;
    stats[5] = sweeps_total;
    stats[6] = 0;
    stats[7] = 0;
  }
  if (!converged_all) {
    smh_fail(ctx, SMH_ERR_NOCONV, "GMRES stopped at relative residual %.3e (tol %.1e)", worst, tol);
    return done2(SMH_ERR_NOCONV);
  }
  return done2(SMH_OK);
}
