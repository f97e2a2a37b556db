// Explicit coupling operator (CouplingMatrixExplicit, linearsystem/linear_system.py:441-486):
// the dense coupling matrix of every particle pair, built on the GPU in FP64,
//
//   W[i B + n1][j B + n2] = e^{i (m2 - m1) phi_ij} (D_ij[n1][n2] + R_ij[n1][n2]),
//
//   D_ij = sum_ld c[n1][n2][ld] h_ld(k_i d_ij) P_ld^|m1-m2|(dz_ij / d_ij)
//          (direct_coupling.py:18-105 through svwf_translate: pairs i != j in
//          one layer, or every pair i != j when the layer system is one medium),
//   R_ij = sum_k J_|m1-m2|(kappa_k rho_ij) sum_{a,b} C^{(L_i, L_j)}_{ab}[k][n1][n2]
//            e^{i s_a kz^{L_i}_k z_i} e^{-i s_b kz^{L_j}_k z_j}
//          (layer_mediated_coupling.py:215-340; s_0 = +1, s_1 = -1, z relative
//          to the particle layer's reference height, L_i the layer of particle
//          i; the Jacobian, the layer response and the quadrature weights of
//          the contour are folded into C on the host).
//
// Build (FP64 throughout), four kernels so that the accumulating one carries
// no special-function code:
//   1. phases T[j][k] = (e^{+i kz_k z_j}, e^{-i kz_k z_j}) of every particle;
//   2. J_n(kappa_k rho_ij), n = 0..2 m_max, for a slab of receivers and every
//      emitter (the slab is sized to keep this table near 1 GB);
//   3. the direct part times e^{i (m2 - m1) phi_ij}, written to W;
//   4. per (receiver, group of kEG emitters of one layer) block, each thread
//      owns one entry (n1, n2) and kEG accumulators; per kappa the receiver
//      side d_b = sum_a C_ab T_i[a] is formed once and reused by every emitter
//      of the group; W += e^{i (m2 - m1) phi} sum_k J (d_0 T_j[1] + d_1 T_j[0]).
// A last pass writes the complex64 copy used by FP32 matvecs.
//
// Matvec: y = W x, a dense complex GEMV over the padded vectors, one
// wavefront per row (HBM-bound on W).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "smh_op.h"
#include "smh_special.h"

namespace {

using smh::zc;

constexpr int kEG = 16;        // emitters per block of the layer kernel
constexpr int kNJ = 17;        // Bessel orders 0..2 m_max, m_max <= 8
constexpr int kMaxL2 = 16;     // 2 l_max
constexpr int kThreads = 1024;

struct ExplArgs {
  const double* xyz;     // [n][3]
  const int* layer;      // [n] index of the particle's layer among the occupied layers
  const int* emit;       // emitters sorted by layer
  const int2* groups;    // (first index into emit, count <= kEG), one layer per group
  int n_groups;
  int e0, n_emit;        // emitters are particles e0 .. e0 + n_emit - 1 (receivers 0 .. n_recv - 1)
  int64_t ldw;           // row length of W: n_emit * B
  int B, L2, nj, n_layers;  // nj = 2 m_max + 1 Bessel orders
  const int* mdiff;      // [B*B] m2 - m1
  const double2* dcoef;  // [B*B][L2+1] direct coefficients (a5 / b5)
  const double2* klay;   // [n_layers] wavenumber of each occupied layer
  int direct_all;        // direct coupling between layers too (one medium)
  int64_t n_kappa;
  const double2* C;      // [n_layers][n_layers][n_kappa][4][B*B], ab = 2a + b
  int i0;                // first receiver of the slab (layer kernel)
  double2* W64;
};

__device__ inline zc cis_z(double2 q, double t) {  // e^{i q t}
  const double m = exp(-q.y * t);
  double sn, cn;
  sincos(q.x * t, &sn, &cn);
  return zc{m * cn, m * sn};
}

__device__ inline zc zfma(zc a, zc b, zc c) {  // a * b + c
  return zc{fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im))};
}

__device__ inline zc ld(const double2* p) {
  const double2 v = *p;
  return zc{v.x, v.y};
}

// 1. T[j][k][0] = e^{+i kz^{L_j}_k z_j}, T[j][k][1] = e^{-i kz^{L_j}_k z_j}
__global__ void explicit_phase_kernel(int64_t n_part, int64_t n_kappa, const int* __restrict__ layer,
                                      const double* __restrict__ zrel, const double2* __restrict__ kz,
                                      double2* __restrict__ T) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_part * n_kappa) return;
  const int64_t j = t / n_kappa, k = t - j * n_kappa;
  const double2 q = kz[(int64_t)layer[j] * n_kappa + k];
  const zc p = cis_z(q, zrel[j]), m = cis_z(q, -zrel[j]);
  T[2 * t] = make_double2(p.re, p.im);
  T[2 * t + 1] = make_double2(m.re, m.im);
}

// 2. Jt[r][j - e0][k][n] = J_n(kappa_k rho_{i0 + r, j}) for one slab of receivers
__global__ void explicit_bessel_kernel(int i0, int slab, int e0, int64_t n_emit, int64_t n_kappa, int nj,
                                       const double* __restrict__ xyz, const double2* __restrict__ kappa,
                                       double2* __restrict__ Jt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)slab * n_emit * n_kappa) return;
  const int64_t rj = t / n_kappa, k = t - rj * n_kappa;
  const int64_t r = rj / n_emit, j = e0 + (rj - r * n_emit);
  const int64_t i = i0 + r;
  const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1];
  const double rho = sqrt(dx * dx + dy * dy);
  const double2 kp = kappa[k];
  zc buf[kNJ];
  smh::bessel_j_all(nj - 1, zc{kp.x * rho, kp.y * rho}, buf);
  for (int n = 0; n < nj; ++n) Jt[t * nj + n] = make_double2(buf[n].re, buf[n].im);
}

// e^{i q phi_ij}, q = 0..kNJ-1, of the block's emitters (threads < kEG)
__device__ inline void stage_phases(const ExplArgs& a, int i, int2 grp, zc (*PH)[kNJ], int* jidx) {
  if (threadIdx.x < kEG) {
    const int e = threadIdx.x;
    const int j = e < grp.y ? a.emit[grp.x + e] : -1;
    jidx[e] = j;
    double phi = 0.0;
    if (j >= 0) phi = atan2(a.xyz[3 * i + 1] - a.xyz[3 * j + 1], a.xyz[3 * i] - a.xyz[3 * j]);
    for (int q = 0; q < kNJ; ++q) {
      double sn, cn;
      sincos(q * phi, &sn, &cn);
      PH[e][q] = zc{cn, sn};
    }
  }
}

__device__ inline zc phase_of(const zc* ph, int md) {
  zc v = ph[md < 0 ? -md : md];
  if (md < 0) v.im = -v.im;
  return v;
}

// 3. W[i B + n1][j B + n2] = e^{i (m2 - m1) phi_ij} D_ij[n1][n2] (zero where no direct part)
__global__ void __launch_bounds__(kThreads) explicit_direct_kernel(ExplArgs a) {
  __shared__ zc H[kEG][kMaxL2 + 1];  // h_ld(k d)
  __shared__ double P[kEG][(kMaxL2 + 1) * (kMaxL2 + 1)];
  __shared__ zc PH[kEG][kNJ];
  __shared__ int jidx[kEG];
  __shared__ int has[kEG];
  const int i = blockIdx.x / a.n_groups;
  const int2 grp = a.groups[blockIdx.x - i * a.n_groups];
  const int li = a.layer[i];
  stage_phases(a, i, grp, PH, jidx);
  __syncthreads();
  if (threadIdx.x < kEG) {
    const int e = threadIdx.x;
    const int j = jidx[e];
    const bool direct = j >= 0 && j != i && (a.direct_all || a.layer[j] == li);
    has[e] = direct;
    if (direct) {
      const double dx = a.xyz[3 * i] - a.xyz[3 * j], dy = a.xyz[3 * i + 1] - a.xyz[3 * j + 1];
      const double dz = a.xyz[3 * i + 2] - a.xyz[3 * j + 2];
      const double rho = sqrt(dx * dx + dy * dy), d = sqrt(rho * rho + dz * dz);
      const double2 k = a.klay[li];
      smh::spherical_hankel_all(a.L2, zc{k.x * d, k.y * d}, H[e]);
      smh::legendre_p(a.L2, dz / d, rho / d, P[e]);
    }
  }
  __syncthreads();
  const int BB = a.B * a.B, W = a.L2 + 1;
  const int64_t ldw = a.ldw;
  for (int p = threadIdx.x; p < BB; p += blockDim.x) {
    const int md = a.mdiff[p], dm = md < 0 ? -md : md;
    const int n1 = p / a.B, n2 = p - n1 * a.B;
    const double2* cp = a.dcoef + (size_t)p * W;
    for (int e = 0; e < kEG; ++e) {
      const int j = jidx[e];
      if (j < 0) break;
      zc v{0.0, 0.0};
      if (has[e]) {
        for (int l = dm; l <= a.L2; ++l) {
          const double2 c = cp[l];
          const double pl = P[e][l * W + dm];
          v = zfma(zc{c.x * pl, c.y * pl}, H[e][l], v);
        }
        v = smh::zmul(v, phase_of(PH[e], md));
      }
      a.W64[((int64_t)i * a.B + n1) * ldw + (int64_t)(j - a.e0) * a.B + n2] = make_double2(v.re, v.im);
    }
  }
}

// 4. W += e^{i (m2 - m1) phi} R for the receivers of one slab. The phases T
// of the block's receiver and emitters are wave-uniform (scalar loads); the
// Bessel values, indexed by each entry's |m2 - m1|, are staged in LDS per
// chunk of kKC kappas.
constexpr int kKC = 12;
__global__ void __launch_bounds__(kThreads)
    explicit_layer_kernel(ExplArgs a, const double2* __restrict__ T, const double2* __restrict__ Jt,
                          const double2* __restrict__ C, const int* __restrict__ emit, double2* __restrict__ W64) {
  __shared__ zc PH[kEG][kNJ];
  __shared__ int jidx[kEG];
  __shared__ zc J[kEG][kKC][kNJ];
  const int r = blockIdx.x / a.n_groups;
  const int i = a.i0 + r;
  const int2 grp = a.groups[blockIdx.x - r * a.n_groups];
  const int li = a.layer[i];
  const int le = a.layer[emit[grp.x]];
  stage_phases(a, i, grp, PH, jidx);
  const int BB = a.B * a.B, nj = a.nj;
  const int64_t nk = a.n_kappa;
  const double2* Cp = C + (size_t)(li * a.n_layers + le) * nk * 4 * BB;
  const double2* Ti = T + (size_t)i * nk * 2;
  // emitters beyond the group's count repeat its first one (results discarded)
  const double2* Tj[kEG];
#pragma unroll
  for (int e = 0; e < kEG; ++e) Tj[e] = T + (size_t)emit[grp.x + (e < grp.y ? e : 0)] * nk * 2;
  const double2* Jr = Jt + (size_t)r * a.n_emit * nk * nj;
  for (int p0 = 0; p0 < BB; p0 += blockDim.x) {
    const int p = p0 + threadIdx.x;
    const bool act = p < BB;
    const int md = act ? a.mdiff[p] : 0, dm = md < 0 ? -md : md;
    zc acc[kEG];
#pragma unroll
    for (int e = 0; e < kEG; ++e) acc[e] = zc{0.0, 0.0};
    const double2* c = Cp + (act ? p : 0);
    for (int64_t k0 = 0; k0 < nk; k0 += kKC) {
      const int kc = (int)min<int64_t>(kKC, nk - k0);
      __syncthreads();
      for (int t = threadIdx.x; t < kEG * kKC * nj; t += blockDim.x) {
        const int e = t / (kKC * nj), q = t - e * (kKC * nj), kk = q / nj, n = q - kk * nj;
        const int j = jidx[e];
        zc v{0.0, 0.0};
        if (j >= 0 && kk < kc) v = ld(Jr + ((size_t)(j - a.e0) * nk + k0 + kk) * nj + n);
        J[e][kk][n] = v;
      }
      __syncthreads();
      if (act) {
        // the next kappa's coefficients are loaded while this one is summed
        const double2* ck = c + k0 * 4 * BB;
        zc c00 = ld(ck), c01 = ld(ck + BB), c10 = ld(ck + 2 * BB), c11 = ld(ck + 3 * BB);
#pragma unroll 1
        for (int kk = 0; kk < kc; ++kk) {
          const int64_t k = k0 + kk;
          const zc r0 = ld(Ti + 2 * k), r1 = ld(Ti + 2 * k + 1);
          const zc d0 = zfma(c00, r0, smh::zmul(c10, r1));
          const zc d1 = zfma(c01, r0, smh::zmul(c11, r1));
          if (kk + 1 < kc) {
            ck += 4 * BB;
            c00 = ld(ck);
            c01 = ld(ck + BB);
            c10 = ld(ck + 2 * BB);
            c11 = ld(ck + 3 * BB);
          }
#pragma unroll
          for (int e = 0; e < kEG; ++e) {
            const zc t = zfma(d0, ld(Tj[e] + 2 * k + 1), smh::zmul(d1, ld(Tj[e] + 2 * k)));
            acc[e] = zfma(J[e][kk][dm], t, acc[e]);
          }
        }
      }
    }
    if (act) {
      const int n1 = p / a.B, n2 = p - n1 * a.B;
      const int64_t ldw = a.ldw;
      for (int e = 0; e < kEG; ++e) {
        const int j = jidx[e];
        if (j < 0) break;
        const zc v = smh::zmul(acc[e], phase_of(PH[e], md));
        double2* o = W64 + ((int64_t)i * a.B + n1) * ldw + (int64_t)(j - a.e0) * a.B + n2;
        const double2 w = *o;
        *o = make_double2(w.x + v.re, w.y + v.im);
      }
    }
  }
}

__global__ void to_c64_kernel(const double2* __restrict__ src, float2* __restrict__ dst, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = src[t];
    dst[t] = make_float2((float)v.x, (float)v.y);
  }
}

// y[r] = sum_c W[r][c] x[c] for W [rows][n]; one wavefront per row, 4 rows per block.
template <typename Real>
__global__ void __launch_bounds__(256) explicit_gemv_kernel(const void* __restrict__ Wv, const void* __restrict__ xv,
                                                           void* __restrict__ yv, int64_t rows, int64_t n) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  Real re = 0, im = 0;
  if constexpr (sizeof(Real) == 8) {
    const double2* W = (const double2*)Wv + row * n;
    const double2* x = (const double2*)xv;
    int64_t c = lane;
    for (; c + 192 < n; c += 256) {
      double2 w[4], v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        w[u] = W[c + 64 * u];
        v[u] = x[c + 64 * u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        re = fma(w[u].x, v[u].x, fma(-w[u].y, v[u].y, re));
        im = fma(w[u].x, v[u].y, fma(w[u].y, v[u].x, im));
      }
    }
    for (; c < n; c += 64) {
      const double2 w = W[c], v = x[c];
      re = fma(w.x, v.x, fma(-w.y, v.y, re));
      im = fma(w.x, v.y, fma(w.y, v.x, im));
    }
  } else {
    // two complex entries per lane and load (n is even: blocksizes are)
    const float4* W = (const float4*)((const float2*)Wv + row * n);
    const float4* x = (const float4*)xv;
    const int64_t n2 = n >> 1;
    int64_t c = lane;
    for (; c + 192 < n2; c += 256) {
      float4 w[4], v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        w[u] = W[c + 64 * u];
        v[u] = x[c + 64 * u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        re = fmaf(w[u].x, v[u].x, fmaf(-w[u].y, v[u].y, re));
        im = fmaf(w[u].x, v[u].y, fmaf(w[u].y, v[u].x, im));
        re = fmaf(w[u].z, v[u].z, fmaf(-w[u].w, v[u].w, re));
        im = fmaf(w[u].z, v[u].w, fmaf(w[u].w, v[u].z, im));
      }
    }
    for (; c < n2; c += 64) {
      const float4 w = W[c], v = x[c];
      re = fmaf(w.x, v.x, fmaf(-w.y, v.y, re));
      im = fmaf(w.x, v.y, fmaf(w.y, v.x, im));
      re = fmaf(w.z, v.z, fmaf(-w.w, v.w, re));
      im = fmaf(w.z, v.w, fmaf(w.w, v.z, im));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    re += __shfl_xor(re, off);
    im += __shfl_xor(im, off);
  }
  if (lane == 0) {
    if constexpr (sizeof(Real) == 8)
      ((double2*)yv)[row] = make_double2(re, im);
    else
      ((float2*)yv)[row] = make_float2(re, im);
  }
}

}  // namespace

namespace smh {

int explicit_apply(smh_op* op, int precision, const void* x_pad, void* y_pad, bool time_kernel) {
  smh_ctx* ctx = op->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n = op->n_part * op->B;
  const dim3 grid((unsigned)((n + 3) / 4));
  if (time_kernel) SMH_HIP(ctx, hipEventRecord(ctx->ev0, s));
  if (precision == SMH_FP32)
    hipLaunchKernelGGL(explicit_gemv_kernel<float>, grid, dim3(256), 0, s, (const void*)op->d_W32, x_pad, y_pad, n,
                       n);
  else
    hipLaunchKernelGGL(explicit_gemv_kernel<double>, grid, dim3(256), 0, s, (const void*)op->d_W64, x_pad, y_pad, n,
                       n);
  SMH_HIP(ctx, hipGetLastError());
  if (time_kernel) {
    SMH_HIP(ctx, hipEventRecord(ctx->ev1, s));
    SMH_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0;
    SMH_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    op->last_kernel_ms = ms;
  }
  return SMH_OK;
}

int explicit_download(smh_op* op, double* W) {
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, op->d_W64 != nullptr, SMH_ERR_STATE, "no matrix");
  const int64_t n = op->n_part * op->B;
  SMH_HIP(ctx, hipMemcpy(W, op->d_W64, (size_t)n * n * sizeof(double2), hipMemcpyDeviceToHost));
  return SMH_OK;
}

}  // namespace smh

namespace {

// The explicit blocks of receivers 0 .. n_recv - 1 and emitters e0 .. e0 +
// n_emit - 1 of one particle set (n_tot = max(n_recv, e0 + n_emit) entries of
// xyz / layer / z_rel) into the device matrix W64 [n_recv B][n_emit B].
int explicit_build(smh_ctx* ctx, int64_t n_recv, int64_t e0, int64_t n_emit, const double* xyz, const int* layer,
                   const double* z_rel, int l_max, int m_max, int n_layers, const double* k_layer, int direct_all,
                   const double* direct_coef, int64_t n_kappa, const double* kappa, const double* kz,
                   const double* layer_coef, double2* W64, float* build_ms) {
  hipStream_t s = ctx->stream;
  const int64_t n_tot = std::max<int64_t>(n_recv, e0 + n_emit);
  const int B = smh_blocksize(l_max, m_max), BB = B * B, L2 = 2 * l_max;
  std::vector<int> tau, ll, mm;
  smh_index_tables(l_max, m_max, tau, ll, mm);
  std::vector<int> mdiff(BB);
  for (int n1 = 0; n1 < B; ++n1)
    for (int n2 = 0; n2 < B; ++n2) mdiff[n1 * B + n2] = mm[n2] - mm[n1];
  // emitters sorted by layer, cut into groups of kEG of one layer
  std::vector<int> emit(n_emit);
  for (int64_t j = 0; j < n_emit; ++j) emit[j] = (int)(e0 + j);
  std::stable_sort(emit.begin(), emit.end(), [&](int x, int y) { return layer[x] < layer[y]; });
  std::vector<int2> groups;
  for (int64_t g0 = 0; g0 < n_emit;) {
    int cnt = 1;
    while (cnt < kEG && g0 + cnt < n_emit && layer[emit[g0 + cnt]] == layer[emit[g0]]) ++cnt;
    groups.push_back(make_int2((int)g0, cnt));
    g0 += cnt;
  }
  const int64_t blocks = n_recv * (int64_t)groups.size();
  SMH_CHECK(ctx, blocks < (int64_t)1 << 31, SMH_ERR_ARG, "too many particle pairs");
  std::vector<void*> tmp;
  auto dev = [&](const void* src, size_t bytes) -> void* {
    void* d = nullptr;
    if (hipMalloc(&d, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    tmp.push_back(d);
    if (src && hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return nullptr;
    return d;
  };
  auto release = [&]() {
    for (void* p : tmp) (void)hipFree(p);
    tmp.clear();
  };
  ExplArgs a;
  a.xyz = (const double*)dev(xyz, n_tot * 3 * sizeof(double));
  const double* d_zrel = (const double*)dev(z_rel, n_tot * sizeof(double));
  a.layer = (const int*)dev(layer, n_tot * sizeof(int));
  a.emit = (const int*)dev(emit.data(), n_emit * sizeof(int));
  a.groups = (const int2*)dev(groups.data(), groups.size() * sizeof(int2));
  a.mdiff = (const int*)dev(mdiff.data(), BB * sizeof(int));
  a.dcoef = (const double2*)dev(direct_coef, (size_t)BB * (L2 + 1) * sizeof(double2));
  a.klay = (const double2*)dev(k_layer, n_layers * sizeof(double2));
  const int nj = 2 * m_max + 1;
  // receivers per slab of the Bessel table: enough for 2048 blocks of the
  // layer kernel, at most 8 GB of table
  const int64_t per_recv = std::max<int64_t>(1, n_emit * n_kappa * nj * (int64_t)sizeof(double2));
  const int64_t want = (2048 + (int64_t)groups.size() - 1) / (int64_t)groups.size();
  const int slab = (int)std::max<int64_t>(1, std::min<int64_t>({n_recv, want, ((int64_t)8 << 30) / per_recv}));
  const double2* d_kappa = nullptr;
  const double2* d_kz = nullptr;
  double2* d_T = nullptr;
  double2* d_J = nullptr;
  bool ok = a.xyz && d_zrel && a.layer && a.emit && a.groups && a.mdiff && a.dcoef && a.klay;
  if (ok && n_kappa > 0) {
    d_kappa = (const double2*)dev(kappa, n_kappa * sizeof(double2));
    d_kz = (const double2*)dev(kz, (size_t)n_layers * n_kappa * sizeof(double2));
    a.C = (const double2*)dev(layer_coef, (size_t)n_layers * n_layers * n_kappa * 4 * BB * sizeof(double2));
    d_T = (double2*)dev(nullptr, (size_t)n_tot * n_kappa * 2 * sizeof(double2));
    d_J = (double2*)dev(nullptr, (size_t)slab * per_recv);
    ok = d_kappa && d_kz && a.C && d_T && d_J;
  }
  if (!ok) {
    release();
    return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (explicit coupling coefficients)");
  }
  a.n_groups = (int)groups.size();
  a.e0 = (int)e0;
  a.n_emit = (int)n_emit;
  a.ldw = n_emit * B;
  a.B = B;
  a.L2 = L2;
  a.nj = nj;
  a.n_layers = n_layers;
  a.direct_all = direct_all;
  a.n_kappa = n_kappa;
  a.W64 = W64;
  const int threads = std::min(kThreads, (BB + 63) / 64 * 64);
  hipEvent_t ev0, ev1;
  SMH_HIP(ctx, hipEventCreate(&ev0));
  SMH_HIP(ctx, hipEventCreate(&ev1));
  SMH_HIP(ctx, hipEventRecord(ev0, s));
  hipLaunchKernelGGL(explicit_direct_kernel, dim3((unsigned)blocks), dim3(threads), 0, s, a);
  if (n_kappa > 0) {
    const int64_t nt = n_tot * n_kappa;
    hipLaunchKernelGGL(explicit_phase_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, n_tot, n_kappa,
                       a.layer, d_zrel, d_kz, d_T);
    for (int i0 = 0; i0 < n_recv; i0 += slab) {
      const int cnt = (int)std::min<int64_t>(slab, n_recv - i0);
      const int64_t nb = (int64_t)cnt * n_emit * n_kappa;
      hipLaunchKernelGGL(explicit_bessel_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, i0, cnt,
                         (int)e0, n_emit, n_kappa, nj, a.xyz, d_kappa, d_J);
      a.i0 = i0;
      hipLaunchKernelGGL(explicit_layer_kernel, dim3((unsigned)(cnt * a.n_groups)), dim3(threads), 0, s, a,
                         (const double2*)d_T, (const double2*)d_J, a.C, a.emit, W64);
    }
  }
  SMH_HIP(ctx, hipEventRecord(ev1, s));
  const hipError_t le = hipGetLastError();
  const hipError_t se = hipStreamSynchronize(s);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ev0, ev1);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  release();
  if (le != hipSuccess || se != hipSuccess)
    return smh_fail(ctx, SMH_ERR_HIP, "explicit coupling kernels failed: %s",
                    hipGetErrorString(le != hipSuccess ? le : se));
  if (build_ms) *build_ms = ms;
  return SMH_OK;
}

int explicit_check(smh_ctx* ctx, int64_t n_tot, const int* layer, int l_max, int m_max, int n_layers,
                   const double* direct_coef, int64_t n_kappa, const double* kappa, const double* kz,
                   const double* layer_coef) {
  SMH_CHECK(ctx, l_max >= 1 && m_max >= 1 && m_max <= l_max && 2 * l_max <= kMaxL2 && 2 * m_max < kNJ, SMH_ERR_ARG,
            "invalid l_max %d / m_max %d (explicit coupling: l_max <= 8)", l_max, m_max);
  SMH_CHECK(ctx, n_layers >= 1, SMH_ERR_ARG, "need at least one layer");
  SMH_CHECK(ctx, direct_coef != nullptr, SMH_ERR_ARG, "direct coefficients are NULL");
  SMH_CHECK(ctx, n_kappa == 0 || (kappa && kz && layer_coef && n_kappa >= 2), SMH_ERR_ARG,
            "layer-mediated part needs kappa, kz and coefficients");
  for (int64_t i = 0; i < n_tot; ++i)
    SMH_CHECK(ctx, layer[i] >= 0 && layer[i] < n_layers, SMH_ERR_ARG, "particle %lld: layer index %d out of range",
              (long long)i, layer[i]);
  return SMH_OK;
}

}  // namespace

extern "C" int smh_explicit_create(smh_ctx* ctx, int64_t n_part, const double* xyz, const int* lmax,
                                   const int* mmax, int l_max, int m_max, const int* layer, int n_layers,
                                   const double* k_layer, const double* z_rel, int direct_all,
                                   const double* direct_coef, int64_t n_kappa, const double* kappa,
                                   const double* kz, const double* layer_coef, double* build_ms, smh_op** out) {
  if (!ctx || !out) return SMH_ERR_ARG;
  *out = nullptr;
  SMH_CHECK(ctx, n_part >= 1 && n_part < (1 << 24) && xyz && lmax && mmax && layer && z_rel && k_layer, SMH_ERR_ARG,
            "invalid particle arguments");
  if (int code = explicit_check(ctx, n_part, layer, l_max, m_max, n_layers, direct_coef, n_kappa, kappa, kz,
                                layer_coef))
    return code;
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  smh_op* op = new smh_op();
  op->ctx = ctx;
  op->kind = 2;
  op->n_part = n_part;
  op->l_max = l_max;
  op->m_max = m_max;
  op->B = smh_blocksize(l_max, m_max);
  op->Bh = (op->B + 15) / 16 * 16;
  op->MP = 2 * op->Bh;
  auto fail = [&](int code) {
    smh_op_free(op);
    delete op;
    return code;
  };
  if (int code = smh::op_index_map(op, lmax, mmax)) return fail(code);
  const int64_t n = n_part * op->B;
  if (hipMalloc(&op->d_W64, (size_t)n * n * sizeof(double2)) != hipSuccess ||
      hipMalloc(&op->d_W32, (size_t)n * n * sizeof(float2)) != hipSuccess ||
      hipMalloc(&op->d_x32, n * sizeof(float2)) != hipSuccess || hipMalloc(&op->d_y32, n * sizeof(float2)) != hipSuccess ||
      hipMalloc(&op->d_x64, n * sizeof(double2)) != hipSuccess || hipMalloc(&op->d_y64, n * sizeof(double2)) != hipSuccess) {
    smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (dense matrix of %lld x %lld needs %.1f GB)", (long long)n,
             (long long)n, (double)n * n * 24 / 1e9);
    return fail(SMH_ERR_HIP);
  }
  float ms = 0;
  if (int code = explicit_build(ctx, n_part, 0, n_part, xyz, layer, z_rel, l_max, m_max, n_layers, k_layer,
                                direct_all, direct_coef, n_kappa, kappa, kz, layer_coef, op->d_W64, &ms))
    return fail(code);
  hipLaunchKernelGGL(to_c64_kernel, dim3(4096), dim3(256), 0, ctx->stream, (const double2*)op->d_W64, op->d_W32, n * n);
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    smh_fail(ctx, SMH_ERR_HIP, "explicit matrix conversion failed: %s", hipGetErrorString(e));
    return fail(SMH_ERR_HIP);
  }
  if (build_ms) *build_ms = ms;
  op->has_table = true;
  ctx->live_ops++;
  *out = op;
  return SMH_OK;
}

extern "C" int smh_explicit_blocks(smh_ctx* ctx, int64_t n_recv, int64_t n_emit, const double* xyz, int l_max,
                                   int m_max, const int* layer, int n_layers, const double* k_layer,
                                   const double* z_rel, int direct_all, const double* direct_coef, int64_t n_kappa,
                                   const double* kappa, const double* kz, const double* layer_coef, const double* x,
                                   double* out, double* build_ms) {
  if (!ctx) return SMH_ERR_ARG;
  SMH_CHECK(ctx, n_recv >= 1 && n_emit >= 1 && n_recv + n_emit < (1 << 24) && xyz && layer && z_rel && k_layer && out,
            SMH_ERR_ARG, "invalid particle arguments");
  if (int code = explicit_check(ctx, n_recv + n_emit, layer, l_max, m_max, n_layers, direct_coef, n_kappa, kappa, kz,
                                layer_coef))
    return code;
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  const int B = smh_blocksize(l_max, m_max);
  const size_t bytes = (size_t)n_recv * B * n_emit * B * sizeof(double2);
  double2* d_W = nullptr;
  SMH_CHECK(ctx, hipMalloc(&d_W, bytes) == hipSuccess, SMH_ERR_HIP, "device allocation failed (%.1f GB of blocks)",
            bytes / 1e9);
  float ms = 0;
  int code = explicit_build(ctx, n_recv, n_recv, n_emit, xyz, layer, z_rel, l_max, m_max, n_layers, k_layer,
                            direct_all, direct_coef, n_kappa, kappa, kz, layer_coef, d_W, &ms);
  if (code == SMH_OK && !x && hipMemcpy(out, d_W, bytes, hipMemcpyDeviceToHost) != hipSuccess)
    code = smh_fail(ctx, SMH_ERR_HIP, "block download failed");
  if (code == SMH_OK && x) {
    // out = W x on the device
    const int64_t rows = n_recv * B, cols = n_emit * B;
    double2 *d_x = nullptr, *d_y = nullptr;
    hipError_t e = hipMalloc(&d_x, cols * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc(&d_y, rows * sizeof(double2));
    if (e == hipSuccess) e = hipMemcpyAsync(d_x, x, cols * sizeof(double2), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(explicit_gemv_kernel<double>, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, ctx->stream,
                         (const void*)d_W, (const void*)d_x, (void*)d_y, rows, cols);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_y, rows * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (d_x) (void)hipFree(d_x);
    if (d_y) (void)hipFree(d_y);
    if (e != hipSuccess) code = smh_fail(ctx, SMH_ERR_HIP, "block product failed: %s", hipGetErrorString(e));
  }
  (void)hipFree(d_W);
  if (code == SMH_OK && build_ms) *build_ms = ms;
  return code;
}
