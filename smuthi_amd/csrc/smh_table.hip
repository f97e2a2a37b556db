// FP64 construction of the radial coupling lookup table on the GPU.
//
//   L[j, n1, n2] = c64( sum_ld direct_coef[n1,n2,ld] h_ld(k rho_j) )        (rho_j >= mask, else 0)
//                + sum_k J_{|m1-m2|}(kappa_k rho_j) lm_coef[k, n1, n2]
//
// The direct part is rounded to complex64 as smuthi stores it
// (prepare_lookup.py:301-340); the Sommerfeld integral is the GEMM-shaped
// contraction of the Bessel matrix J[j, k] with the per-kappa coefficients,
// which carry the trapezoid weights (prepare_lookup.py:392-459).
#include "smh_op.h"
#include "smh_special.h"

using smh::zc;

namespace {

__global__ void radial_direct_kernel(double rho_min, double h, int64_t len, int BB, int L2, double kre, double kim,
                                     double mask, const double2* __restrict__ coef, double2* table) {
  const int64_t j = blockIdx.x;
  if (j >= len) return;
  const double rho = rho_min + (double)j * h;
  __shared__ zc hl[smh::kMaxBesselOrder + 1];
  double2* row = table + j * BB;
  if (rho < mask) {
    for (int p = threadIdx.x; p < BB; p += blockDim.x) row[p] = make_double2(0.0, 0.0);
    return;
  }
  if (threadIdx.x == 0) smh::spherical_hankel_all(L2, zc{kre * rho, kim * rho}, hl);
  __syncthreads();
  for (int p = threadIdx.x; p < BB; p += blockDim.x) {
    zc s{0.0, 0.0};
    for (int ld = 0; ld <= L2; ++ld) {
      const double2 c = coef[(size_t)p * (L2 + 1) + ld];
      s = smh::zadd(s, smh::zmul(zc{c.x, c.y}, hl[ld]));
    }
    row[p] = make_double2((double)(float)s.re, (double)(float)s.im);
  }
}

constexpr int kRB = 16;  // table rows per block

template <int KC, int NJ>
__global__ void radial_layer_kernel(double rho_min, double h, int64_t len, int BB, int nmax,
                                    const int* __restrict__ dm_of_p, int64_t n_kappa, const double2* __restrict__ kappa,
                                    const double2* __restrict__ lmc, double2* table) {
  __shared__ zc Jt[KC][NJ][kRB];  // at most 57 KB of static LDS
  const int64_t j0 = (int64_t)blockIdx.x * kRB;
  for (int64_t k0 = 0; k0 < n_kappa; k0 += KC) {
    const int kc = (int)min<int64_t>(KC, n_kappa - k0);
    for (int e = threadIdx.x; e < KC * kRB; e += blockDim.x) {
      const int kk = e / kRB, rr = e - kk * kRB;
      zc buf[smh::kMaxBesselOrder + 1];
      if (kk < kc) {
        const double rho = rho_min + (double)(j0 + rr) * h;
        const double2 kp = kappa[k0 + kk];
        smh::bessel_j_all(nmax, zc{kp.x * rho, kp.y * rho}, buf);
        for (int n = 0; n <= nmax; ++n) Jt[kk][n][rr] = buf[n];
      }
    }
    __syncthreads();
    for (int p = threadIdx.x; p < BB; p += blockDim.x) {
      const int dm = dm_of_p[p];
      zc acc[kRB];
#pragma unroll
      for (int rr = 0; rr < kRB; ++rr) acc[rr] = zc{0.0, 0.0};
      for (int kk = 0; kk < kc; ++kk) {
        const double2 c = lmc[(size_t)(k0 + kk) * BB + p];
#pragma unroll
        for (int rr = 0; rr < kRB; ++rr) {
          const zc jv = Jt[kk][dm][rr];
          acc[rr].re += jv.re * c.x - jv.im * c.y;
          acc[rr].im += jv.re * c.y + jv.im * c.x;
        }
      }
#pragma unroll
      for (int rr = 0; rr < kRB; ++rr) {
        if (j0 + rr < len) {
          double2* t = table + (size_t)(j0 + rr) * BB + p;
          double2 v = *t;
          v.x += acc[rr].re;
          v.y += acc[rr].im;
          *t = v;
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace

extern "C" int smh_radial_build_table(smh_op* op, double k_re, double k_im, double mask_radius,
                                      const double* direct_coef, int64_t n_kappa, const double* kappa,
                                      const double* lm_coef) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, op->kind == 0, SMH_ERR_ARG, "not a radial operator");
  SMH_CHECK(ctx, direct_coef != nullptr, SMH_ERR_ARG, "direct_coef is NULL");
  SMH_CHECK(ctx, lm_coef == nullptr || (kappa != nullptr && n_kappa >= 2), SMH_ERR_ARG,
            "layer-mediated part needs at least two contour points");
  const int L2 = 2 * op->l_max, nmax = 2 * op->m_max;
  SMH_CHECK(ctx, L2 <= smh::kMaxBesselOrder && nmax <= 8, SMH_ERR_UNSUPPORTED, "l_max %d too large", op->l_max);
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int BB = op->B * op->B;
  const size_t tbytes = (size_t)op->len_rho * BB * sizeof(double2);
  if (!op->d_table) SMH_HIP(ctx, hipMalloc(&op->d_table, tbytes));
  double2* d_coef = nullptr;
  double2* d_kappa = nullptr;
  double2* d_lmc = nullptr;
  int* d_dm = nullptr;
  auto release = [&]() {
    if (d_coef) hipFree(d_coef);
    if (d_kappa) hipFree(d_kappa);
    if (d_lmc) hipFree(d_lmc);
    if (d_dm) hipFree(d_dm);
  };
  const size_t cbytes = (size_t)BB * (L2 + 1) * sizeof(double2);
  if (hipMalloc(&d_coef, cbytes) != hipSuccess ||
      hipMemcpyAsync(d_coef, direct_coef, cbytes, hipMemcpyHostToDevice, s) != hipSuccess) {
    release();
    return smh_fail(ctx, SMH_ERR_HIP, "direct coefficient upload failed");
  }
  hipLaunchKernelGGL(radial_direct_kernel, dim3((unsigned)op->len_rho), dim3(256), 0, s, op->rho_min, op->h,
                     op->len_rho, BB, L2, k_re, k_im, mask_radius, d_coef, op->d_table);
  if (hipGetLastError() != hipSuccess) {
    release();
    return smh_fail(ctx, SMH_ERR_HIP, "direct table kernel launch failed");
  }
  if (lm_coef) {
    std::vector<int> tau, ll, mm;
    smh_index_tables(op->l_max, op->m_max, tau, ll, mm);
    std::vector<int> dm(BB);
    for (int n1 = 0; n1 < op->B; ++n1)
      for (int n2 = 0; n2 < op->B; ++n2) dm[n1 * op->B + n2] = abs(mm[n1] - mm[n2]);
    const size_t lbytes = (size_t)n_kappa * BB * sizeof(double2);
    if (hipMalloc(&d_kappa, n_kappa * sizeof(double2)) != hipSuccess || hipMalloc(&d_lmc, lbytes) != hipSuccess ||
        hipMalloc(&d_dm, BB * sizeof(int)) != hipSuccess ||
        hipMemcpyAsync(d_kappa, kappa, n_kappa * sizeof(double2), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_lmc, lm_coef, lbytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_dm, dm.data(), BB * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess) {
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "layer-mediated coefficient upload failed");
    }
    const unsigned nblk = (unsigned)((op->len_rho + kRB - 1) / kRB);
    if (nmax <= 6)
      hipLaunchKernelGGL((radial_layer_kernel<32, 7>), dim3(nblk), dim3(256), 0, s, op->rho_min, op->h, op->len_rho, BB,
                         nmax, d_dm, n_kappa, d_kappa, d_lmc, op->d_table);
    else
      hipLaunchKernelGGL((radial_layer_kernel<24, 9>), dim3(nblk), dim3(256), 0, s, op->rho_min, op->h, op->len_rho, BB,
                         nmax, d_dm, n_kappa, d_kappa, d_lmc, op->d_table);
    if (hipGetLastError() != hipSuccess) {
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "layer-mediated table kernel launch failed");
    }
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    release();
    return smh_fail(ctx, SMH_ERR_HIP, "table build failed: %s", hipGetErrorString(hipGetLastError()));
  }
  release();
  return smh::op_make_taps(op);
}
