// Plane wave expansion -> regular spherical wave expansion at many points
// (pwe_to_swe_conversion, fields/transformations.py:77-130, with the
// continuous kappa / alpha quadrature): for every reference point r_p,
//
//   a_n(p) = sum_k w_k e^{i kz_k dz_p} sum_pol Bdag_{n,pol}(k)
//            sum_j w_j e^{-i m_n alpha_j} g_pol(k, j) e^{i kappa_k (cos alpha_j dx_p + sin alpha_j dy_p)},
//
// d = r_p - r_pwe, w_k / w_j the quadrature weights of the two trapezoid
// rules (the factor 4 kappa_k folded into w_k), Bdag the B-dagger
// transformation coefficients at (kz / k, kappa / k). FP64 throughout.
//
// Layout. One thread per point, 256 points per workgroup, the kappa range
// split over blockIdx.y so that small point sets still fill the chip; each
// split writes its partial a_n and a second pass sums the splits in a fixed
// order. Per kappa the alpha samples are staged through LDS in tiles of 64 as
// G[pol][m][j] = w_j g_pol e^{-i m alpha_j} (broadcast reads); per (point,
// kappa, alpha) one complex exponential and 2 (2 m_max + 1) complex
// multiply-adds; per (point, kappa) the B-dagger contraction into the B
// accumulators held in registers.
#include <algorithm>
#include <cmath>
#include <vector>

#include "smh_common.h"

namespace {

constexpr int kP2sPoints = 256;
constexpr int kP2sTile = 64;
constexpr int kP2sBlocks = 2048;  // target workgroups per launch (kappa splits)

struct P2sArgs {
  const double* pts;     // [n_pts][3], relative to the expansion's reference point
  const double2* kp;     // [n_k] kappa
  const double2* kz;     // [n_k] kz (sign of the expansion's kind)
  const double2* wk;     // [n_k] kappa weights (4 kappa trapezoid)
  const double2* bdag;   // [n_k][2][B]
  const double2* g;      // [2][n_k][n_a] expansion coefficients
  const double* wa;      // [n_a] alpha weights
  const double* alpha;   // [n_a]
  int64_t n_pts, n_k, n_a;
  int64_t k_per;         // kappa values per split
  double2* part;         // [n_split][n_pts][B]
};

template <int LMAX, int MMAX>
__global__ void __launch_bounds__(kP2sPoints) pwe_to_swe_kernel(P2sArgs a) {
  constexpr int NM = 2 * MMAX + 1;
  constexpr int B = 2 * (MMAX * (MMAX + 2) + (LMAX - MMAX) * (2 * MMAX + 1));
  using C = cplx<double>;
  __shared__ C G[2][NM][kP2sTile];
  __shared__ double CS[kP2sTile][2];

  const int64_t ip = (int64_t)blockIdx.x * kP2sPoints + threadIdx.x;
  const bool live = ip < a.n_pts;
  double dx = 0, dy = 0, dz = 0;
  if (live) {
    dx = a.pts[3 * ip];
    dy = a.pts[3 * ip + 1];
    dz = a.pts[3 * ip + 2];
  }
  // m of each multipole n in the (tau, l, m) order
  C acc[B];
#pragma unroll
  for (int n = 0; n < B; ++n) acc[n] = C{0.0, 0.0};

  const int64_t k0 = (int64_t)blockIdx.y * a.k_per;
  const int64_t k1 = min(a.n_k, k0 + a.k_per);
  for (int64_t k = k0; k < k1; ++k) {
    const double2 kp = a.kp[k];
    C S[2][NM];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int m = 0; m < NM; ++m) S[q][m] = C{0.0, 0.0};
    for (int64_t j0 = 0; j0 < a.n_a; j0 += kP2sTile) {
      const int nj = (int)min<int64_t>(kP2sTile, a.n_a - j0);
      __syncthreads();
      for (int t = threadIdx.x; t < 2 * NM * kP2sTile; t += kP2sPoints) {
        const int q = t / (NM * kP2sTile), r = t - q * NM * kP2sTile, mi = r / kP2sTile, j = r - mi * kP2sTile;
        C v{0.0, 0.0};
        if (j < nj) {
          const double al = a.alpha[j0 + j];
          double sn, cn;
          sincos((double)(mi - MMAX) * al, &sn, &cn);
          const double2 gv = a.g[((int64_t)q * a.n_k + k) * a.n_a + j0 + j];
          const double w = a.wa[j0 + j];
          // w g e^{-i m alpha}
          v = C{w * (gv.x * cn + gv.y * sn), w * (gv.y * cn - gv.x * sn)};
        }
        G[q][mi][j] = v;
      }
      for (int j = threadIdx.x; j < kP2sTile; j += kP2sPoints) {
        double sn = 0, cn = 0;
        if (j < nj) sincos(a.alpha[j0 + j], &sn, &cn);
        CS[j][0] = cn;
        CS[j][1] = sn;
      }
      __syncthreads();
      for (int j = 0; j < nj; ++j) {
        // e^{i kappa u}, u = cos(alpha) dx + sin(alpha) dy
        const double u = CS[j][0] * dx + CS[j][1] * dy;
        double sn, cn;
        sincos(kp.x * u, &sn, &cn);
        const double mag = kp.y == 0.0 ? 1.0 : exp(-kp.y * u);
        const C ph{mag * cn, mag * sn};
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
          for (int m = 0; m < NM; ++m) {
            const C gq = G[q][m][j];
            S[q][m].re = fma(gq.re, ph.re, fma(-gq.im, ph.im, S[q][m].re));
            S[q][m].im = fma(gq.re, ph.im, fma(gq.im, ph.re, S[q][m].im));
          }
      }
    }
    // w_k e^{i kz dz}, then the B-dagger contraction
    const double2 kzv = a.kz[k], wkv = a.wk[k];
    double sn, cn;
    sincos(kzv.x * dz, &sn, &cn);
    const double mag = exp(-kzv.y * dz);
    const C e = cmul(C{wkv.x, wkv.y}, C{mag * cn, mag * sn});
    const double2* bd = a.bdag + k * 2 * B;
    int n = 0;
#pragma unroll
    for (int tau = 0; tau < 2; ++tau)
#pragma unroll
      for (int l = 1; l <= LMAX; ++l) {
        const int ml = l < MMAX ? l : MMAX;
#pragma unroll
        for (int m = -MMAX; m <= MMAX; ++m) {
          if (m < -ml || m > ml) continue;
          const double2 b0 = bd[n], b1 = bd[B + n];
          const C t = cadd(cmul(C{b0.x, b0.y}, S[0][m + MMAX]), cmul(C{b1.x, b1.y}, S[1][m + MMAX]));
          acc[n] = cadd(acc[n], cmul(e, t));
          ++n;
        }
      }
  }
  if (live) {
    double2* o = a.part + ((int64_t)blockIdx.y * a.n_pts + ip) * B;
#pragma unroll
    for (int n = 0; n < B; ++n) o[n] = make_double2(acc[n].re, acc[n].im);
  }
}

// Any (l_max, m_max): the orders m in chunks of kP2sOrders, one launch each
// (m0 .. m0 + nm - 1); per (point, kappa) the chunk's alpha sums as above,
// then the B-dagger contraction of the chunk's multipoles (nlist: [nm][LM]
// indices n, -1 padded) added to the point's partial row, which the thread
// owns (in global memory: B values do not fit registers at large l_max; the
// rows start zeroed). The unrolled kernels cover l_max <= 4.
constexpr int kP2sOrders = 5;
constexpr int kP2sMaxL = 20;

__global__ void __launch_bounds__(kP2sPoints)
    pwe_to_swe_any_kernel(P2sArgs a, int m0, int nm, const int* __restrict__ nlist, int LM, int B) {
  using C = cplx<double>;
  __shared__ C G[2][kP2sOrders][kP2sTile];
  __shared__ double CS[kP2sTile][2];
  const int64_t ip = (int64_t)blockIdx.x * kP2sPoints + threadIdx.x;
  const bool live = ip < a.n_pts;
  double dx = 0, dy = 0, dz = 0;
  if (live) {
    dx = a.pts[3 * ip];
    dy = a.pts[3 * ip + 1];
    dz = a.pts[3 * ip + 2];
  }
  double2* o = a.part + ((int64_t)blockIdx.y * a.n_pts + (live ? ip : 0)) * B;
  const int64_t k0 = (int64_t)blockIdx.y * a.k_per;
  const int64_t k1 = min(a.n_k, k0 + a.k_per);
  for (int64_t k = k0; k < k1; ++k) {
    const double2 kp = a.kp[k];
    C S[2][kP2sOrders];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int m = 0; m < kP2sOrders; ++m) S[q][m] = C{0.0, 0.0};
    for (int64_t j0 = 0; j0 < a.n_a; j0 += kP2sTile) {
      const int nj = (int)min<int64_t>(kP2sTile, a.n_a - j0);
      __syncthreads();
      for (int t = threadIdx.x; t < 2 * kP2sOrders * kP2sTile; t += kP2sPoints) {
        const int q = t / (kP2sOrders * kP2sTile), r = t - q * kP2sOrders * kP2sTile, mi = r / kP2sTile,
                  j = r - mi * kP2sTile;
        C v{0.0, 0.0};
        if (j < nj && mi < nm) {
          const double al = a.alpha[j0 + j];
          double sn, cn;
          sincos((double)(m0 + mi) * al, &sn, &cn);
          const double2 gv = a.g[((int64_t)q * a.n_k + k) * a.n_a + j0 + j];
          const double w = a.wa[j0 + j];
          v = C{w * (gv.x * cn + gv.y * sn), w * (gv.y * cn - gv.x * sn)};
        }
        G[q][mi][j] = v;
      }
      for (int j = threadIdx.x; j < kP2sTile; j += kP2sPoints) {
        double sn = 0, cn = 0;
        if (j < nj) sincos(a.alpha[j0 + j], &sn, &cn);
        CS[j][0] = cn;
        CS[j][1] = sn;
      }
      __syncthreads();
      for (int j = 0; j < nj; ++j) {
        const double u = CS[j][0] * dx + CS[j][1] * dy;
        double sn, cn;
        sincos(kp.x * u, &sn, &cn);
        const double mag = kp.y == 0.0 ? 1.0 : exp(-kp.y * u);
        const C ph{mag * cn, mag * sn};
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
          for (int m = 0; m < kP2sOrders; ++m) {
            const C gq = G[q][m][j];
            S[q][m].re = fma(gq.re, ph.re, fma(-gq.im, ph.im, S[q][m].re));
            S[q][m].im = fma(gq.re, ph.im, fma(gq.im, ph.re, S[q][m].im));
          }
      }
    }
    if (!live) continue;
    const double2 kzv = a.kz[k], wkv = a.wk[k];
    double sn, cn;
    sincos(kzv.x * dz, &sn, &cn);
    const double mag = exp(-kzv.y * dz);
    const C e = cmul(C{wkv.x, wkv.y}, C{mag * cn, mag * sn});
    const double2* bd = a.bdag + k * 2 * B;
#pragma unroll
    for (int mi = 0; mi < kP2sOrders; ++mi) {
      if (mi >= nm) break;
      for (int jj = 0; jj < LM; ++jj) {
        const int n = nlist[mi * LM + jj];
        if (n < 0) break;
        const double2 b0 = bd[n], b1 = bd[B + n];
        const C t = cadd(cmul(C{b0.x, b0.y}, S[0][mi]), cmul(C{b1.x, b1.y}, S[1][mi]));
        const C et = cmul(e, t);
        const double2 prev = o[n];
        o[n] = make_double2(prev.x + et.re, prev.y + et.im);
      }
    }
  }
}

__global__ void p2s_reduce_kernel(const double2* __restrict__ part, int n_split, int64_t count,
                                  double2* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  double re = 0, im = 0;
  for (int s = 0; s < n_split; ++s) {
    const double2 v = part[(int64_t)s * count + t];
    re += v.x;
    im += v.y;
  }
  out[t] = make_double2(re, im);
}

template <int L, int M>
void* p2s_fn() {
  return reinterpret_cast<void*>(&pwe_to_swe_kernel<L, M>);
}

void* p2s_kernel(int l, int m) {
#define SMH_P(L, Mx) \
  if (l == L && m == Mx) return p2s_fn<L, Mx>();
  SMH_P(1, 1)
  SMH_P(2, 1)
  SMH_P(2, 2)
  SMH_P(3, 1)
  SMH_P(3, 2)
  SMH_P(3, 3)
  SMH_P(4, 1)
  SMH_P(4, 2)
  SMH_P(4, 3)
  SMH_P(4, 4)
#undef SMH_P
  return nullptr;
}

}  // namespace

extern "C" int smh_pwe_to_swe(smh_ctx* ctx, int64_t n_k, const double* kp, const double* kz, const double* w_k,
                              int64_t n_a, const double* alpha, const double* w_a, const double* coef,
                              const double* bdag, int l_max, int m_max, int64_t n_pts, const double* pts,
                              double* out, double* kernel_ms) {
  if (!ctx) return SMH_ERR_ARG;
  SMH_CHECK(ctx, n_k >= 1 && n_a >= 1 && n_pts >= 0 && n_pts < (1 << 30), SMH_ERR_ARG, "invalid sizes");
  SMH_CHECK(ctx, kp && kz && w_k && alpha && w_a && coef && bdag, SMH_ERR_ARG, "expansion arrays are NULL");
  SMH_CHECK(ctx, l_max >= 1 && m_max >= 0 && m_max <= l_max && l_max <= kP2sMaxL, SMH_ERR_ARG,
            "l_max %d / m_max %d out of range (l_max <= %d)", l_max, m_max, kP2sMaxL);
  // unrolled kernels up to l_max 4, the runtime-order kernel above
  void* fn = p2s_kernel(l_max, m_max);
  if (kernel_ms) *kernel_ms = 0.0;
  if (n_pts == 0) return SMH_OK;
  SMH_CHECK(ctx, pts && out, SMH_ERR_ARG, "points / output are NULL");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  const int B = smh_blocksize(l_max, m_max);
  const int64_t pblocks = (n_pts + kP2sPoints - 1) / kP2sPoints;
  const int64_t n_split = std::max<int64_t>(1, std::min<int64_t>(n_k, (kP2sBlocks + pblocks - 1) / pblocks));
  const int64_t k_per = (n_k + n_split - 1) / n_split;
  const int64_t splits = (n_k + k_per - 1) / k_per;
  const size_t sz[10] = {(size_t)n_pts * 3 * 8, (size_t)n_k * 16,         (size_t)n_k * 16,
                         (size_t)n_k * 16,      (size_t)n_k * 2 * B * 16, (size_t)2 * n_k * n_a * 16,
                         (size_t)n_a * 8,       (size_t)n_a * 8,          (size_t)splits * n_pts * B * 16,
                         (size_t)n_pts * B * 16};
  const void* src[8] = {pts, kp, kz, w_k, bdag, coef, w_a, alpha};
  void* buf[10] = {};
  auto release = [&]() {
    for (void* q : buf)
      if (q) (void)hipFree(q);
  };
  hipStream_t s = ctx->stream;
  for (int i = 0; i < 10; ++i) {
    if (hipMalloc(&buf[i], sz[i]) != hipSuccess) {
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "PWE-to-SWE device allocation failed");
    }
    if (i < 8 && hipMemcpyAsync(buf[i], src[i], sz[i], hipMemcpyHostToDevice, s) != hipSuccess) {
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "PWE-to-SWE upload failed");
    }
  }
  P2sArgs a;
  a.pts = (const double*)buf[0];
  a.kp = (const double2*)buf[1];
  a.kz = (const double2*)buf[2];
  a.wk = (const double2*)buf[3];
  a.bdag = (const double2*)buf[4];
  a.g = (const double2*)buf[5];
  a.wa = (const double*)buf[6];
  a.alpha = (const double*)buf[7];
  a.n_pts = n_pts;
  a.n_k = n_k;
  a.n_a = n_a;
  a.k_per = k_per;
  a.part = (double2*)buf[8];
  const int64_t count = n_pts * B;
  // runtime-order kernel: the multipoles of each order m, in chunks of kP2sOrders orders
  const int n_orders = 2 * m_max + 1, LM = 2 * l_max;
  const int n_chunks = (n_orders + kP2sOrders - 1) / kP2sOrders;
  std::vector<int> nlist;
  void* d_nlist = nullptr;
  if (!fn) {
    nlist.assign((size_t)n_chunks * kP2sOrders * LM, -1);
    std::vector<int> fill((size_t)n_chunks * kP2sOrders, 0);
    for (int tau = 0; tau < 2; ++tau)
      for (int l = 1; l <= l_max; ++l) {
        const int ml = std::min(l, m_max);
        for (int m = -ml; m <= ml; ++m) {
          const int r = m + m_max;  // row: chunk r / kP2sOrders, order r % kP2sOrders in it
          nlist[(size_t)r * LM + fill[r]++] = smh_multi_to_single(tau, l, m, l_max, m_max);
        }
      }
    if (hipMalloc(&d_nlist, nlist.size() * sizeof(int)) != hipSuccess ||
        hipMemcpyAsync(d_nlist, nlist.data(), nlist.size() * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemsetAsync(buf[8], 0, sz[8], s) != hipSuccess) {
      if (d_nlist) (void)hipFree(d_nlist);
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "PWE-to-SWE device allocation failed");
    }
  }
  hipError_t e = hipEventRecord(ctx->ev0, s);
  void* args[] = {&a};
  if (fn) {
    if (e == hipSuccess) e = hipLaunchKernel(fn, dim3((unsigned)pblocks, (unsigned)splits), dim3(kP2sPoints), args, 0, s);
  } else {
    for (int c = 0; c < n_chunks && e == hipSuccess; ++c) {
      const int m0 = c * kP2sOrders - m_max, nm = std::min(kP2sOrders, n_orders - c * kP2sOrders);
      hipLaunchKernelGGL(pwe_to_swe_any_kernel, dim3((unsigned)pblocks, (unsigned)splits), dim3(kP2sPoints), 0, s, a,
                         m0, nm, (const int*)d_nlist + (size_t)c * kP2sOrders * LM, LM, B);
      e = hipGetLastError();
    }
  }
  if (e == hipSuccess) e = hipEventRecord(ctx->ev1, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(p2s_reduce_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                       (const double2*)buf[8], (int)splits, count, (double2*)buf[9]);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, buf[9], sz[9], hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  if (d_nlist) (void)hipFree(d_nlist);
  release();
  SMH_CHECK(ctx, e == hipSuccess, SMH_ERR_HIP, "PWE-to-SWE kernel failed: %s", hipGetErrorString(e));
  if (kernel_ms) *kernel_ms = ms;
  return SMH_OK;
}
