// Direct solve of (1 - T W) b = T a for the explicit operator: the
// solver_type='LU' branch of LinearSystem.solve (linearsystem/linear_system.py:
// 259-268), which factors the stored master matrix once with LAPACK
// (scipy lu_factor), keeps the factors and solves every right-hand side with
// them (lu_solve). Here the master matrix M = 1 - diag(t) W is formed in FP64
// from the device-resident explicit matrix, factored with partial pivoting by
// rocSOLVER's getrf and solved by getrs, all on the device; the factors stay
// on the operator for later right-hand sides with the same T. Every solution
// is certified by its FP64 residual ||b - M x|| / ||b||, computed with the
// explicit matvec.
//
// The explicit matrix is row-major on the padded index (i * B + n), which
// rocSOLVER (column-major) sees as M^T: getrf factors that matrix and getrs
// solves with its transpose, i.e. with M itself. Padding slots carry t = 0, so
// their rows and columns of M are those of the identity.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <chrono>
#include <cmath>

#include "smh_op.h"

namespace {

int blocks_of(int64_t n, int t = 256) { return (int)std::min<int64_t>((n + t - 1) / t, 1 << 20); }

// M = I - diag(t) W on the padded index, row-major [n][n]
__global__ void master_dense_kernel(const double2* __restrict__ W, const double2* __restrict__ t, int64_t n,
                                    double2* __restrict__ M) {
  const int64_t total = n * n;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = f / n, c = f - r * n;
    const double2 w = W[f], a = t[r];
    double re = -(a.x * w.x - a.y * w.y);
    const double im = -(a.x * w.y + a.y * w.x);
    if (r == c) re += 1.0;
    M[f] = make_double2(re, im);
  }
}

// r = b - (x - t o y) for nrhs padded vectors [nrhs][n] (y = W x); per block
// partial sums of |r|^2 and |b|^2 for each right-hand side
__global__ void residual_parts_kernel(const double2* __restrict__ b, const double2* __restrict__ x,
                                      const double2* __restrict__ y, const double2* __restrict__ t, int64_t n,
                                      double* __restrict__ parts) {
  const int rhs = blockIdx.y;
  __shared__ double red[2][256];
  double rr = 0, bb = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = (int64_t)rhs * n + i;
    const double2 a = t[i], yv = y[e], xv = x[e], bv = b[e];
    const double mr = xv.x - (a.x * yv.x - a.y * yv.y), mi = xv.y - (a.x * yv.y + a.y * yv.x);
    const double dr = bv.x - mr, di = bv.y - mi;
    rr += dr * dr + di * di;
    bb += bv.x * bv.x + bv.y * bv.y;
  }
  red[0][threadIdx.x] = rr;
  red[1][threadIdx.x] = bb;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    parts[((size_t)rhs * gridDim.x + blockIdx.x) * 2] = red[0][0];
    parts[((size_t)rhs * gridDim.x + blockIdx.x) * 2 + 1] = red[1][0];
  }
}

}  // namespace

namespace smh {

void lu_release(smh_op* op) {
  if (op->d_lu) (void)hipFree(op->d_lu);
  if (op->d_piv) (void)hipFree(op->d_piv);
  if (op->blas) (void)rocblas_destroy_handle((rocblas_handle)op->blas);
  op->d_lu = nullptr;
  op->d_piv = nullptr;
  op->blas = nullptr;
  op->lu_t.clear();
}

int lu_solve(smh_op* op, const double* t_diag, const double* rhs, double* x, int64_t nrhs, double* stats) {
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, op->kind == 2 && op->d_W64, SMH_ERR_UNSUPPORTED,
            "LU factorization needs the stored (explicit) coupling matrix");
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = ctx->stream;
  const int64_t n = op->n_part * op->B;
  SMH_CHECK(ctx, n < (int64_t)1 << 31 && n * n < (int64_t)1 << 62, SMH_ERR_ARG, "matrix too large");
  if (!op->blas) {
    rocblas_handle h = nullptr;
    SMH_CHECK(ctx, rocblas_create_handle(&h) == rocblas_status_success, SMH_ERR_HIP, "rocblas_create_handle failed");
    op->blas = h;
  }
  rocblas_handle h = (rocblas_handle)op->blas;
  SMH_CHECK(ctx, rocblas_set_stream(h, s) == rocblas_status_success, SMH_ERR_HIP, "rocblas_set_stream failed");
  int rc = op_reserve_packed(op, nrhs);
  if (rc) return rc;
  double2 *d_t = nullptr, *d_b = nullptr, *d_x = nullptr, *d_y = nullptr;
  double* d_parts = nullptr;
  int* d_info = nullptr;
  constexpr int kParts = 256;
  auto done = [&](int code) {
    void* p[] = {d_t, d_b, d_x, d_y, d_parts, d_info};
    for (void* q : p)
      if (q) (void)hipFree(q);
    return code;
  };
  if (hipMalloc(&d_t, n * sizeof(double2)) != hipSuccess || hipMalloc(&d_b, nrhs * n * sizeof(double2)) != hipSuccess ||
      hipMalloc(&d_x, nrhs * n * sizeof(double2)) != hipSuccess ||
      hipMalloc(&d_y, nrhs * n * sizeof(double2)) != hipSuccess ||
      hipMalloc(&d_parts, nrhs * kParts * 2 * sizeof(double)) != hipSuccess ||
      hipMalloc(&d_info, sizeof(int)) != hipSuccess)
    return done(smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (LU solve vectors)"));
#define LU_HIP(call)                                                                                           \
  do {                                                                                                         \
    hipError_t e__ = (call);                                                                                   \
    if (e__ != hipSuccess)                                                                                     \
      return done(smh_fail(ctx, SMH_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e__), __FILE__, \
                           __LINE__));                                                                         \
  } while (0)
#define LU_TRY(call)         \
  do {                       \
    int rc__ = (call);       \
    if (rc__) return done(rc__); \
  } while (0)
  LU_HIP(hipMemcpyAsync(op->d_pk_in, t_diag, op->n_unk * sizeof(double2), hipMemcpyDefault, s));
  LU_TRY(op_unpack(op, SMH_FP64, op->d_pk_in, d_t));
  // the factors of an earlier call serve while T is the same (the reference
  // keeps LU_piv on the master matrix)
  const std::vector<double> tkey(t_diag, t_diag + 2 * op->n_unk);
  if (!op->d_lu || op->lu_t != tkey) {
    if (!op->d_lu) {
      LU_HIP(hipMalloc(&op->d_lu, (size_t)n * n * sizeof(double2)));
      LU_HIP(hipMalloc(&op->d_piv, (size_t)n * sizeof(int)));
    }
    op->lu_t.clear();
    hipLaunchKernelGGL(master_dense_kernel, dim3(blocks_of(n * n)), dim3(256), 0, s, (const double2*)op->d_W64,
                       (const double2*)d_t, n, op->d_lu);
    LU_HIP(hipGetLastError());
    const rocblas_status st = rocsolver_zgetrf(h, (rocblas_int)n, (rocblas_int)n,
                                               reinterpret_cast<rocblas_double_complex*>(op->d_lu), (rocblas_int)n,
                                               op->d_piv, d_info);
    if (st != rocblas_status_success) return done(smh_fail(ctx, SMH_ERR_HIP, "rocsolver_zgetrf failed (%d)", (int)st));
    int info = 0;
    LU_HIP(hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, s));
    LU_HIP(hipStreamSynchronize(s));
    if (info != 0)
      return done(smh_fail(ctx, SMH_ERR_ARG, "the master matrix is singular (zero pivot %d of %lld)", info,
                           (long long)n));
    op->lu_t = tkey;
  }
  // right-hand sides in the padded layout, one column each
  LU_HIP(hipMemcpyAsync(op->d_pk_in, rhs, (size_t)nrhs * op->n_unk * sizeof(double2), hipMemcpyDefault, s));
  for (int64_t r = 0; r < nrhs; ++r) {
    LU_TRY(op_unpack(op, SMH_FP64, op->d_pk_in + r * op->n_unk, d_b + r * n));
  }
  LU_HIP(hipMemcpyAsync(d_x, d_b, (size_t)nrhs * n * sizeof(double2), hipMemcpyDeviceToDevice, s));
  const rocblas_status st = rocsolver_zgetrs(h, rocblas_operation_transpose, (rocblas_int)n, (rocblas_int)nrhs,
                                             reinterpret_cast<rocblas_double_complex*>(op->d_lu), (rocblas_int)n,
                                             op->d_piv, reinterpret_cast<rocblas_double_complex*>(d_x),
                                             (rocblas_int)n);
  if (st != rocblas_status_success) return done(smh_fail(ctx, SMH_ERR_HIP, "rocsolver_zgetrs failed (%d)", (int)st));
  // certificate: FP64 residual of every solution
  for (int64_t r = 0; r < nrhs; ++r) LU_TRY(explicit_apply(op, SMH_FP64, d_x + r * n, d_y + r * n, false));
  hipLaunchKernelGGL(residual_parts_kernel, dim3(kParts, (unsigned)nrhs), dim3(256), 0, s, (const double2*)d_b,
                     (const double2*)d_x, (const double2*)d_y, (const double2*)d_t, n, d_parts);
  LU_HIP(hipGetLastError());
  std::vector<double> parts((size_t)nrhs * kParts * 2);
  LU_HIP(hipMemcpyAsync(parts.data(), d_parts, parts.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  for (int64_t r = 0; r < nrhs; ++r) LU_TRY(op_pack(op, SMH_FP64, d_x + r * n, op->d_pk_out + r * op->n_unk));
  LU_HIP(hipMemcpyAsync(x, op->d_pk_out, (size_t)nrhs * op->n_unk * sizeof(double2), hipMemcpyDefault, s));
  LU_HIP(hipStreamSynchronize(s));
  double worst = 0;
  for (int64_t r = 0; r < nrhs; ++r) {
    double rr = 0, bb = 0;
    for (int k = 0; k < kParts; ++k) {
      rr += parts[((size_t)r * kParts + k) * 2];
      bb += parts[((size_t)r * kParts + k) * 2 + 1];
    }
    worst = std::max(worst, bb > 0 ? std::sqrt(rr / bb) : 0.0);
  }
  if (stats) {
    for (int k = 0; k < 8; ++k) stats[k] = 0;
    stats[1] = (double)nrhs;  // FP64 matvecs (the residual certificates)
    stats[3] = worst;
    stats[4] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
#undef LU_HIP
#undef LU_TRY
  return done(SMH_OK);
}

}  // namespace smh
