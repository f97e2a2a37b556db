// Plane wave expansion near field on the GPU: E (or omega H) of
//   F(r) = int kappa dkappa int dalpha f(kappa, alpha) e^{i (kx x + ky y + kz z)}
// by the nested trapezoid rule (alpha inner, kappa outer) of
// PlaneWaveExpansion.electric_field / magnetic_field (fields/expansions.py:635-782).
//
// Layout. Points are ordered into compact leaves of at most kLeaf points by
// recursive median splits; one workgroup evaluates one leaf, one thread one
// point. The leaf origin r0 splits every phase as e^{i k.r0} e^{i k.(r - r0)}:
// the first factor is folded into the (kappa, alpha) coefficients once per
// leaf in FP64 (staged in LDS for a chunk of kappa), so the per-term phase
// k.(r - r0) is small and FP32 keeps it accurate. Per term the FP32 kernel
// spends one native sin and cos (argument in revolutions) and, on the complex
// part of the contour only, one exp2, plus three complex multiply-adds.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <numeric>
#include <vector>

#include "smh_common.h"

namespace {

constexpr int kLeaf = 256;
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kLog2e = 1.4426950408889634073599246810019;

struct PweArgs {
  const double2* F;    // [3][n_k][n_a] weighted integrand coefficients
  const double2* kp;   // [n_k]
  const double2* kz;   // [n_k]
  const double* ca;    // cos alpha [n_a]
  const double* sa;    // sin alpha [n_a]
  const double* pts;   // [n_pts][3], leaf order
  const int* leaf;     // [n_leaves + 1] offsets into pts
  const double* org;   // [n_leaves][3] leaf origins
  double2* out;        // [3][n_pts], leaf order
  int n_k, n_a, kc;    // kc: kappa per LDS chunk
};

template <typename Real>
struct PweMath;

template <>
struct PweMath<float> {
  // e^{i 2 pi t} for |t| of a few revolutions (native v_sin / v_cos)
  static __device__ inline void cis_rev(float t, float& c, float& s) {
    s = __builtin_amdgcn_sinf(t);
    c = __builtin_amdgcn_cosf(t);
  }
  static __device__ inline float exp2_(float x) { return __builtin_amdgcn_exp2f(x); }
};

template <>
struct PweMath<double> {
  static __device__ inline void cis_rev(double t, double& c, double& s) { sincos(kTwoPi * t, &s, &c); }
  static __device__ inline double exp2_(double x) { return exp2(x); }
};

template <typename Real>
__global__ void __launch_bounds__(kLeaf) pwe_field_kernel(PweArgs a) {
  using M = PweMath<Real>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* cs = reinterpret_cast<Real*>(smem_raw);  // [n_a][2] cos, sin
  Real* fl = cs + 2 * a.n_a;                      // [kc][n_a][6] leaf coefficients (re, im) x 3
  const int leaf = blockIdx.x, tid = threadIdx.x;
  const int p0 = a.leaf[leaf], p1 = a.leaf[leaf + 1];
  const int p = p0 + tid;
  const bool active = p < p1;
  const double x0 = a.org[3 * leaf], y0 = a.org[3 * leaf + 1], z0 = a.org[3 * leaf + 2];
  Real dx = 0, dy = 0, dz = 0;
  if (active) {
    dx = (Real)(a.pts[3 * p] - x0);
    dy = (Real)(a.pts[3 * p + 1] - y0);
    dz = (Real)(a.pts[3 * p + 2] - z0);
  }
  for (int j = tid; j < a.n_a; j += blockDim.x) {
    cs[2 * j] = (Real)a.ca[j];
    cs[2 * j + 1] = (Real)a.sa[j];
  }
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int k0 = 0; k0 < a.n_k; k0 += a.kc) {
    const int kn = min(a.kc, a.n_k - k0);
    __syncthreads();
    // leaf coefficients: F(c, k, alpha) e^{i (kp (x0 cos + y0 sin) + kz z0)} in FP64
    for (int j = tid; j < kn * a.n_a; j += blockDim.x) {
      const int kk = j / a.n_a, ia = j - kk * a.n_a, k = k0 + kk;
      const double2 kp = a.kp[k], kz = a.kz[k];
      const double sxy = x0 * a.ca[ia] + y0 * a.sa[ia];
      // i (kp sxy + kz z0)
      const double ph_re = -(kp.y * sxy + kz.y * z0), ph_im = kp.x * sxy + kz.x * z0;
      double sn, cn;
      sincos(ph_im, &sn, &cn);
      const double mag = exp(ph_re);
      const double er = mag * cn, ei = mag * sn;
      Real* dst = fl + (size_t)j * 6;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double2 f = a.F[((size_t)c * a.n_k + k) * a.n_a + ia];
        dst[2 * c] = (Real)(f.x * er - f.y * ei);
        dst[2 * c + 1] = (Real)(f.x * ei + f.y * er);
      }
    }
    __syncthreads();
    for (int kk = 0; kk < kn; ++kk) {
      const double2 kpd = a.kp[k0 + kk], kzd = a.kz[k0 + kk];
      const Real kr = (Real)(kpd.x / kTwoPi);     // revolutions per nm
      const Real ki = (Real)(-kpd.y * kLog2e);    // log2 of the decay per nm
      const bool lossy = kpd.y != 0.0;
      Real in[6] = {0, 0, 0, 0, 0, 0};
      const Real* f = fl + (size_t)kk * a.n_a * 6;
      for (int ia = 0; ia < a.n_a; ++ia) {
        const Real s = dx * cs[2 * ia] + dy * cs[2 * ia + 1];
        Real pc, ps;
        M::cis_rev(kr * s, pc, ps);
        if (lossy) {
          const Real m = M::exp2_(ki * s);
          pc *= m;
          ps *= m;
        }
        const Real* fa = f + ia * 6;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          in[2 * c] += fa[2 * c] * pc - fa[2 * c + 1] * ps;
          in[2 * c + 1] += fa[2 * c] * ps + fa[2 * c + 1] * pc;
        }
      }
      // e^{i kz dz}
      Real zc, zs;
      M::cis_rev((Real)(kzd.x / kTwoPi) * dz, zc, zs);
      const Real zm = M::exp2_((Real)(-kzd.y * kLog2e) * dz);
      zc *= zm;
      zs *= zm;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc[2 * c] += (double)(in[2 * c] * zc - in[2 * c + 1] * zs);
        acc[2 * c + 1] += (double)(in[2 * c] * zs + in[2 * c + 1] * zc);
      }
    }
  }
  if (active) {
    const int64_t n = a.leaf[gridDim.x];
#pragma unroll
    for (int c = 0; c < 3; ++c) a.out[(size_t)c * n + p] = make_double2(acc[2 * c], acc[2 * c + 1]);
  }
}

// Recursive median split of idx[b, e) along the widest extent into leaves of
// at most `leaf` points.
void split_leaves(const double* xyz, std::vector<int>& idx, int b, int e, std::vector<int>& bounds, int leaf = kLeaf) {
  if (e - b <= leaf) {
    bounds.push_back(e);
    return;
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int j = b; j < e; ++j)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], xyz[3 * idx[j] + d]);
      hi[d] = std::max(hi[d], xyz[3 * idx[j] + d]);
    }
  int dim = 0;
  for (int d = 1; d < 3; ++d)
    if (hi[d] - lo[d] > hi[dim] - lo[dim]) dim = d;
  // split at a multiple of the leaf size so leaves stay full
  const int half = ((e - b + 1) / 2 + leaf - 1) / leaf * leaf;
  const int mid = std::min(e - 1, b + half);
  std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e,
                   [&](int u, int v) { return xyz[3 * u + dim] < xyz[3 * v + dim]; });
  split_leaves(xyz, idx, b, mid, bounds, leaf);
  split_leaves(xyz, idx, mid, e, bounds, leaf);
}

}  // namespace

extern "C" int smh_pwe_field(smh_ctx* ctx, int kind, int64_t n_k, const double* kp, const double* kz, int64_t n_a,
                             const double* alpha, const double* g, double k_re, double k_im, int64_t n_pts,
                             const double* xyz, double* out, int precision, double* kernel_ms) {
  if (!ctx) return SMH_ERR_ARG;
  SMH_CHECK(ctx, kind == 0 || kind == 1, SMH_ERR_ARG, "kind must be 0 (E) or 1 (omega H)");
  SMH_CHECK(ctx, n_k >= 1 && n_a >= 1 && n_pts >= 0 && kp && kz && alpha && g, SMH_ERR_ARG, "invalid PWE arguments");
  SMH_CHECK(ctx, precision == SMH_FP32 || precision == SMH_FP64, SMH_ERR_ARG, "invalid precision %d", precision);
  SMH_CHECK(ctx, n_pts < (1 << 30) && n_k * n_a < (1 << 30), SMH_ERR_ARG, "PWE too large");
  if (kernel_ms) *kernel_ms = 0.0;
  if (n_pts == 0) return SMH_OK;
  SMH_CHECK(ctx, xyz && out, SMH_ERR_ARG, "points / output are NULL");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  typedef std::complex<double> zd;
  const zd kk(k_re, k_im);
  SMH_CHECK(ctx, std::abs(kk) > 0, SMH_ERR_ARG, "wavenumber is zero");
  // weights: trapezoid over alpha (inner) and over kappa with the kappa factor (outer);
  // a single kappa sample is a discrete plane wave: its first alpha sample, weight 1
  std::vector<zd> wk(n_k);
  std::vector<double> wa(n_a);
  if (n_k == 1) {
    wk[0] = 1.0;
    std::fill(wa.begin(), wa.end(), 0.0);
    wa[0] = 1.0;
  } else {
    auto KP = [&](int64_t j) { return zd(kp[2 * j], kp[2 * j + 1]); };
    for (int64_t j = 0; j < n_k; ++j) {
      const zd d = (j + 1 < n_k ? KP(j + 1) : KP(j)) - (j > 0 ? KP(j - 1) : KP(j));
      wk[j] = 0.5 * d * KP(j);
    }
    for (int64_t j = 0; j < n_a; ++j)
      wa[j] = n_a == 1 ? 0.0 : 0.5 * ((j + 1 < n_a ? alpha[j + 1] : alpha[j]) - (j > 0 ? alpha[j - 1] : alpha[j]));
  }
  // integrand coefficients per component (expansions.py:676-699 for E, :745-766 for H)
  std::vector<zd> F((size_t)3 * n_k * n_a);
  std::vector<double> ca(n_a), sa(n_a);
  for (int64_t j = 0; j < n_a; ++j) {
    ca[j] = std::cos(alpha[j]);
    sa[j] = std::sin(alpha[j]);
  }
  for (int64_t k = 0; k < n_k; ++k) {
    const zd kpk(kp[2 * k], kp[2 * k + 1]), kzk(kz[2 * k], kz[2 * k + 1]);
    for (int64_t j = 0; j < n_a; ++j) {
      const size_t ij = (size_t)k * n_a + j;
      const zd g0(g[2 * ij], g[2 * ij + 1]);
      const zd g1(g[2 * ((size_t)n_k * n_a + ij)], g[2 * ((size_t)n_k * n_a + ij) + 1]);
      zd fx, fy, fz;
      if (kind == 0) {
        fx = -sa[j] * g0 + ca[j] * kzk / kk * g1;
        fy = ca[j] * g0 + sa[j] * kzk / kk * g1;
        fz = -kpk / kk * g1;
      } else {
        fx = -kzk * ca[j] * g0 - sa[j] * kk * g1;
        fy = -kzk * sa[j] * g0 + ca[j] * kk * g1;
        fz = kpk * g0;
      }
      const zd w = wk[k] * wa[j];
      F[ij] = w * fx;
      F[(size_t)n_k * n_a + ij] = w * fy;
      F[(size_t)2 * n_k * n_a + ij] = w * fz;
    }
  }
  // leaves and their origins; z origin on the side toward which every term decays
  std::vector<int> idx(n_pts);
  std::iota(idx.begin(), idx.end(), 0);
  std::vector<int> bounds{0};
  split_leaves(xyz, idx, 0, (int)n_pts, bounds);
  const int n_leaves = (int)bounds.size() - 1;
  bool up = true, down = true;
  for (int64_t k = 0; k < n_k; ++k) {
    if (kz[2 * k + 1] < 0) up = false;
    if (kz[2 * k + 1] > 0) down = false;
  }
  std::vector<double> pts((size_t)3 * n_pts), org((size_t)3 * n_leaves);
  for (int64_t j = 0; j < n_pts; ++j)
    for (int d = 0; d < 3; ++d) pts[3 * j + d] = xyz[3 * (size_t)idx[j] + d];
  for (int l = 0; l < n_leaves; ++l) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int j = bounds[l]; j < bounds[l + 1]; ++j)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], pts[3 * j + d]);
        hi[d] = std::max(hi[d], pts[3 * j + d]);
      }
    org[3 * l] = 0.5 * (lo[0] + hi[0]);
    org[3 * l + 1] = 0.5 * (lo[1] + hi[1]);
    org[3 * l + 2] = up ? lo[2] : (down ? hi[2] : 0.5 * (lo[2] + hi[2]));
  }
  // device buffers
  const size_t rs = precision == SMH_FP32 ? sizeof(float) : sizeof(double);
  const int kc = (int)std::max<int64_t>(1, std::min<int64_t>(16, (48 * 1024) / (n_a * 6 * (int64_t)rs)));
  const size_t lds = (size_t)2 * n_a * rs + (size_t)kc * n_a * 6 * rs;
  SMH_CHECK(ctx, lds <= 160 * 1024, SMH_ERR_UNSUPPORTED, "too many azimuth samples (%lld)", (long long)n_a);
  void* buf[9] = {nullptr};
  const size_t sizes[9] = {F.size() * sizeof(zd), (size_t)n_k * 16, (size_t)n_k * 16, (size_t)n_a * 8,
                           (size_t)n_a * 8, pts.size() * 8, bounds.size() * 4, org.size() * 8,
                           (size_t)3 * n_pts * 16};
  auto release = [&]() {
    for (void* b : buf)
      if (b) hipFree(b);
  };
  for (int i = 0; i < 9; ++i)
    if (hipMalloc(&buf[i], sizes[i]) != hipSuccess) {
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "PWE device allocation failed");
    }
  hipStream_t s = ctx->stream;
  const void* src[8] = {F.data(), kp, kz, ca.data(), sa.data(), pts.data(), bounds.data(), org.data()};
  for (int i = 0; i < 8; ++i)
    if (hipMemcpyAsync(buf[i], src[i], sizes[i], hipMemcpyHostToDevice, s) != hipSuccess) {
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "PWE upload failed");
    }
  PweArgs a;
  a.F = (const double2*)buf[0];
  a.kp = (const double2*)buf[1];
  a.kz = (const double2*)buf[2];
  a.ca = (const double*)buf[3];
  a.sa = (const double*)buf[4];
  a.pts = (const double*)buf[5];
  a.leaf = (const int*)buf[6];
  a.org = (const double*)buf[7];
  a.out = (double2*)buf[8];
  a.n_k = (int)n_k;
  a.n_a = (int)n_a;
  a.kc = kc;
  void* kfn = precision == SMH_FP32 ? reinterpret_cast<void*>(&pwe_field_kernel<float>)
                                    : reinterpret_cast<void*>(&pwe_field_kernel<double>);
  hipError_t e = hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev0, s);
  void* args[] = {&a};
  if (e == hipSuccess) e = hipLaunchKernel(kfn, dim3(n_leaves), dim3(kLeaf), args, lds, s);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev1, s);
  std::vector<std::complex<double>> res((size_t)3 * n_pts);
  if (e == hipSuccess) e = hipMemcpyAsync(res.data(), buf[8], sizes[8], hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  release();
  SMH_CHECK(ctx, e == hipSuccess, SMH_ERR_HIP, "PWE field kernel failed: %s", hipGetErrorString(e));
  for (int c = 0; c < 3; ++c)
    for (int64_t j = 0; j < n_pts; ++j) {
      const std::complex<double> v = res[(size_t)c * n_pts + j];
      out[2 * ((size_t)c * n_pts + idx[j])] = v.real();
      out[2 * ((size_t)c * n_pts + idx[j]) + 1] = v.imag();
    }
  if (kernel_ms) *kernel_ms = ms;
  return SMH_OK;
}

// ---- plane wave expansion of the scattered field --------------------------
// g_dir(pol, kappa, alpha) = sum_p e^{-i kappa (x_p cos a + y_p sin a)} e^{i kz_dir (z_ref - z_p)}
//                            * sum_m e^{i m a} c_{p,m,pol,dir}(kappa),
// c_{p,m,pol,dir} = sum_{tau,l} b_{p,tau l m} B_{tau l m,pol}(+-kz/k, kappa/k) / (2 pi kz k)
// (swe_to_pwe_conversion, fields/transformations.py:133-226, summed over the
// particles of one layer as scattered_field_pwe, postprocessing/scattered_field.py:80-150).
// One workgroup per kappa, one thread per alpha; particles are taken leaf by
// leaf, their c staged in LDS; the leaf centre's phase is applied in FP64.
// A launch covers NM consecutive orders m (its accumulators live in
// registers); more than kAsmOrders orders run as several launches, each adding
// its orders' share into g.
namespace {

constexpr int kAsmLeaf = 128;
constexpr int kAsmOrders = 9;  // orders per launch (m_max <= 4 in one launch)

struct AsmArgs {
  const double* pts;    // [n_part][3], leaf order
  const double2* coef;  // [n_part][B], leaf order
  const int* leaf;      // [n_leaves + 1]
  const double* ctr;    // [n_leaves][2] leaf centres (x, y)
  const double2* kp;    // [n_k]
  const double2* kz;    // [n_k] upgoing kz
  const double2* btab;  // [n_k][2 dir][2 pol][B], 1 / (2 pi kz k) included
  const int* nlist;     // [NM][LM] multipole indices n of order m = j + m_first (-1: none)
  const double* alpha;  // [n_a]
  double2* g;           // [2 dir][2 pol][n_k][n_a]
  double z_ref;
  int B, LM, n_leaves, n_k, n_a;
  int m_first;     // order of the launch's first row of nlist
  int accumulate;  // add into g (later launches of a chunked order range)
};

template <typename Real, int NM>
__global__ void __launch_bounds__(256) pwe_assemble_kernel(AsmArgs a) {
  using M = PweMath<Real>;
  constexpr int Q = 4;  // (dir, pol)
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* cl = reinterpret_cast<Real*>(smem_raw);  // [kAsmLeaf][NM][Q][2]
  Real* dxy = cl + (size_t)kAsmLeaf * NM * Q * 2;  // [kAsmLeaf][2]
  const int k = blockIdx.x, tid = threadIdx.x;
  const int ia = blockIdx.y * blockDim.x + tid;
  const bool active = ia < a.n_a;
  const double al = active ? a.alpha[ia] : 0.0;
  const double cad = cos(al), sad = sin(al);
  const Real ca = (Real)cad, sa = (Real)sad;
  const double2 kp = a.kp[k], kzu = a.kz[k];
  const Real kr = (Real)(-kp.x / kTwoPi);   // e^{-i kappa s}: revolutions per nm
  const Real ki = (Real)(kp.y * kLog2e);    // |e^{-i kappa s}| = 2^(ki s)
  const bool lossy = kp.y != 0.0;
  Real U[NM][Q][2];
#pragma unroll
  for (int m = 0; m < NM; ++m)
#pragma unroll
    for (int q = 0; q < Q; ++q) U[m][q][0] = U[m][q][1] = Real(0);
  for (int l = 0; l < a.n_leaves; ++l) {
    const int p0 = a.leaf[l], np = a.leaf[l + 1] - p0;
    const double X = a.ctr[2 * l], Y = a.ctr[2 * l + 1];
    __syncthreads();
    for (int w = tid; w < np * NM; w += blockDim.x) {
      const int pl = w / NM, m = w - pl * NM, p = p0 + pl;
      const double zp = a.pts[3 * p + 2];
      // e^{i kz (z_ref - z_p)} (up) and e^{-i kz (z_ref - z_p)} (down)
      const double dz = a.z_ref - zp;
      double su, cu;
      sincos(kzu.x * dz, &su, &cu);
      const double mu = exp(-kzu.y * dz), md = exp(kzu.y * dz);
      const double2 ph[2] = {make_double2(mu * cu, mu * su), make_double2(md * cu, -md * su)};
      double cr[Q] = {0, 0, 0, 0}, ci[Q] = {0, 0, 0, 0};
      for (int j = 0; j < a.LM; ++j) {
        const int n = a.nlist[m * a.LM + j];
        if (n < 0) break;
        const double2 b = a.coef[(size_t)p * a.B + n];
        for (int q = 0; q < Q; ++q) {
          const double2 t = a.btab[((size_t)k * Q + q) * a.B + n];
          cr[q] += b.x * t.x - b.y * t.y;
          ci[q] += b.x * t.y + b.y * t.x;
        }
      }
      Real* dst = cl + ((size_t)pl * NM + m) * Q * 2;
      for (int q = 0; q < Q; ++q) {
        const double2 e = ph[q >> 1];
        dst[2 * q] = (Real)(cr[q] * e.x - ci[q] * e.y);
        dst[2 * q + 1] = (Real)(cr[q] * e.y + ci[q] * e.x);
      }
      if (m == 0) {
        dxy[2 * pl] = (Real)(a.pts[3 * p] - X);
        dxy[2 * pl + 1] = (Real)(a.pts[3 * p + 1] - Y);
      }
    }
    __syncthreads();
    Real S[NM][Q][2];
#pragma unroll
    for (int m = 0; m < NM; ++m)
#pragma unroll
      for (int q = 0; q < Q; ++q) S[m][q][0] = S[m][q][1] = Real(0);
    for (int pl = 0; pl < np; ++pl) {
      const Real s = dxy[2 * pl] * ca + dxy[2 * pl + 1] * sa;
      Real pc, ps;
      M::cis_rev(kr * s, pc, ps);
      if (lossy) {
        const Real mg = M::exp2_(ki * s);
        pc *= mg;
        ps *= mg;
      }
      const Real* c = cl + (size_t)pl * NM * Q * 2;
#pragma unroll
      for (int m = 0; m < NM; ++m)
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const Real xr = c[(m * Q + q) * 2], xi = c[(m * Q + q) * 2 + 1];
          S[m][q][0] += xr * pc - xi * ps;
          S[m][q][1] += xr * ps + xi * pc;
        }
    }
    // leaf centre phase e^{-i kappa (X cos a + Y sin a)} in FP64
    const double sc = X * cad + Y * sad;
    double sn, cn;
    sincos(-kp.x * sc, &sn, &cn);
    const double mg = exp(kp.y * sc);
    const Real lr = (Real)(mg * cn), li = (Real)(mg * sn);
#pragma unroll
    for (int m = 0; m < NM; ++m)
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        U[m][q][0] += S[m][q][0] * lr - S[m][q][1] * li;
        U[m][q][1] += S[m][q][0] * li + S[m][q][1] * lr;
      }
  }
  if (!active) return;
  for (int q = 0; q < Q; ++q) {
    double gr = 0, gi = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
      double em_s, em_c;
      sincos((double)(m + a.m_first) * al, &em_s, &em_c);
      gr += (double)U[m][q][0] * em_c - (double)U[m][q][1] * em_s;
      gi += (double)U[m][q][0] * em_s + (double)U[m][q][1] * em_c;
    }
    double2* dst = a.g + ((size_t)q * a.n_k + k) * a.n_a + ia;
    if (a.accumulate) {
      const double2 o = *dst;
      gr += o.x;
      gi += o.y;
    }
    *dst = make_double2(gr, gi);
  }
}

template <typename Real>
void* assemble_kernel_for(int nm) {
  switch (nm) {
    case 3: return reinterpret_cast<void*>(&pwe_assemble_kernel<Real, 3>);
    case 5: return reinterpret_cast<void*>(&pwe_assemble_kernel<Real, 5>);
    case 7: return reinterpret_cast<void*>(&pwe_assemble_kernel<Real, 7>);
    case 9: return reinterpret_cast<void*>(&pwe_assemble_kernel<Real, 9>);
    default: return nullptr;
  }
}

}  // namespace

extern "C" int smh_pwe_assemble(smh_ctx* ctx, int64_t n_part, const double* xyz, int l_max, int m_max,
                                const double* coef, double z_ref, int64_t n_k, const double* kp, const double* kz,
                                const double* btab, int64_t n_a, const double* alpha, double* g, int precision,
                                double* kernel_ms) {
  if (!ctx) return SMH_ERR_ARG;
  SMH_CHECK(ctx, n_part >= 1 && xyz && coef && kp && kz && btab && alpha && g, SMH_ERR_ARG,
            "invalid PWE assembly arguments");
  SMH_CHECK(ctx, l_max >= 1 && m_max >= 1 && m_max <= l_max && l_max <= 64, SMH_ERR_ARG,
            "invalid l_max %d / m_max %d", l_max, m_max);
  SMH_CHECK(ctx, n_k >= 1 && n_a >= 1 && n_k < (1 << 30) && n_a < (1 << 30) && n_part < (1 << 30), SMH_ERR_ARG,
            "invalid sizes");
  SMH_CHECK(ctx, precision == SMH_FP32 || precision == SMH_FP64, SMH_ERR_ARG, "invalid precision %d", precision);
  if (kernel_ms) *kernel_ms = 0.0;
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  const int B = smh_blocksize(l_max, m_max), n_orders = 2 * m_max + 1, LM = 2 * l_max;
  // orders per launch: all of them when they fit, else chunks of kAsmOrders
  // (the last one padded with empty orders, whose rows of nlist hold -1)
  const int NM = n_orders <= kAsmOrders ? n_orders : kAsmOrders;
  const int n_chunks = (n_orders + NM - 1) / NM;
  // multipole indices of each order m (row j: order j - m_max)
  std::vector<int> nlist((size_t)n_chunks * NM * LM, -1), fill(n_chunks * NM, 0);
  for (int tau = 0; tau < 2; ++tau)
    for (int l = 1; l <= l_max; ++l) {
      const int ml = std::min(l, m_max);
      for (int m = -ml; m <= ml; ++m) {
        const int j = m + m_max;
        nlist[(size_t)j * LM + fill[j]++] = smh_multi_to_single(tau, l, m, l_max, m_max);
      }
    }
  // particle leaves
  std::vector<int> idx(n_part);
  std::iota(idx.begin(), idx.end(), 0);
  std::vector<int> bounds{0};
  split_leaves(xyz, idx, 0, (int)n_part, bounds, kAsmLeaf);
  const int n_leaves = (int)bounds.size() - 1;
  std::vector<double> pts((size_t)3 * n_part), ctr((size_t)2 * n_leaves);
  std::vector<double> cf((size_t)2 * n_part * B);
  for (int64_t j = 0; j < n_part; ++j) {
    for (int d = 0; d < 3; ++d) pts[3 * j + d] = xyz[3 * (size_t)idx[j] + d];
    std::memcpy(&cf[(size_t)2 * j * B], &coef[(size_t)2 * idx[j] * B], sizeof(double) * 2 * B);
  }
  for (int l = 0; l < n_leaves; ++l) {
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    for (int j = bounds[l]; j < bounds[l + 1]; ++j)
      for (int d = 0; d < 2; ++d) {
        lo[d] = std::min(lo[d], pts[3 * j + d]);
        hi[d] = std::max(hi[d], pts[3 * j + d]);
      }
    ctr[2 * l] = 0.5 * (lo[0] + hi[0]);
    ctr[2 * l + 1] = 0.5 * (lo[1] + hi[1]);
  }
  const size_t rs = precision == SMH_FP32 ? sizeof(float) : sizeof(double);
  const size_t lds = (size_t)kAsmLeaf * NM * 4 * 2 * rs + (size_t)kAsmLeaf * 2 * rs;
  void* buf[10] = {nullptr};
  const size_t sizes[10] = {pts.size() * 8, cf.size() * 8, bounds.size() * 4, ctr.size() * 8,
                            (size_t)n_k * 16, (size_t)n_k * 16, (size_t)n_k * 4 * B * 16, nlist.size() * 4,
                            (size_t)n_a * 8, (size_t)4 * n_k * n_a * 16};
  auto release = [&]() {
    for (void* b : buf)
      if (b) hipFree(b);
  };
  for (int i = 0; i < 10; ++i)
    if (hipMalloc(&buf[i], sizes[i]) != hipSuccess) {
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "PWE assembly allocation failed");
    }
  hipStream_t s = ctx->stream;
  const void* src[9] = {pts.data(), cf.data(), bounds.data(), ctr.data(), kp, kz, btab, nlist.data(), alpha};
  for (int i = 0; i < 9; ++i)
    if (hipMemcpyAsync(buf[i], src[i], sizes[i], hipMemcpyHostToDevice, s) != hipSuccess) {
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "PWE assembly upload failed");
    }
  AsmArgs a;
  a.pts = (const double*)buf[0];
  a.coef = (const double2*)buf[1];
  a.leaf = (const int*)buf[2];
  a.ctr = (const double*)buf[3];
  a.kp = (const double2*)buf[4];
  a.kz = (const double2*)buf[5];
  a.btab = (const double2*)buf[6];
  a.nlist = (const int*)buf[7];
  a.alpha = (const double*)buf[8];
  a.g = (double2*)buf[9];
  a.z_ref = z_ref;
  a.B = B;
  a.LM = LM;
  a.n_leaves = n_leaves;
  a.n_k = (int)n_k;
  a.n_a = (int)n_a;
  void* kfn = precision == SMH_FP32 ? assemble_kernel_for<float>(NM) : assemble_kernel_for<double>(NM);
  const int threads = (int)std::min<int64_t>(256, (n_a + 63) / 64 * 64);
  const unsigned a_blocks = (unsigned)((n_a + threads - 1) / threads);
  hipError_t e = hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev0, s);
  void* args[] = {&a};
  const int* nlist_dev = a.nlist;
  for (int c = 0; c < n_chunks && e == hipSuccess; ++c) {
    a.nlist = nlist_dev + (size_t)c * NM * LM;
    a.m_first = c * NM - m_max;
    a.accumulate = c > 0;
    e = hipLaunchKernel(kfn, dim3((unsigned)n_k, a_blocks), dim3(threads), args, lds, s);
  }
  if (e == hipSuccess) e = hipEventRecord(ctx->ev1, s);
  if (e == hipSuccess) e = hipMemcpyAsync(g, buf[9], sizes[9], hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  release();
  SMH_CHECK(ctx, e == hipSuccess, SMH_ERR_HIP, "PWE assembly kernel failed: %s", hipGetErrorString(e));
  if (kernel_ms) *kernel_ms = ms;
  return SMH_OK;
}
