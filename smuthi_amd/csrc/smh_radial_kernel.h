// Radial-lookup coupling matvec on CDNA4 MFMA.
//
// Work decomposition. Particle pairs (i1 receiving, i2 emitting) are binned
// by the table interval r = floor((rho12 - rho_min) / h) and grouped by
// receiver chunk (Rc consecutive receivers). A work item is one chunk and a
// range of bins. For a bin r every pair needs the same T table rows ("taps"),
// so the bin is a GEMM:
//
//   Y[(re|im, n1), (pair, re|im)] = A_r[(re|im, n1), (n2, t)] * Z[(n2, t), (pair, re|im)]
//   A_r = [Re; Im] of the T tap rows,   Z[(n2,t), p] = w_t(rho_p) e^{i m2 phi_p} x_{i2(p)}[n2]
//
// done with v_mfma 16x16x4 (FP32 or FP64). The tap rows of a period of S
// consecutive bins are staged in LDS once and shared by all pairs of the
// chunk in those bins. Per pair the result is rotated by e^{-i m1 phi} and
// accumulated into an LDS accumulator of the chunk's receivers, which is
// written once per work item (no global atomics; a second kernel sums the
// work items of a chunk in fixed order).
#pragma once

#include <hip/hip_runtime.h>

#include "smh_common.h"

namespace smh {

template <int LMAX, int MMAX>
struct SvwfIdx {
  static constexpr int B = 2 * (MMAX * (MMAX + 2) + (LMAX - MMAX) * (2 * MMAX + 1));
  static constexpr int Bh = (B + 15) / 16 * 16;  // rows of the Re (or Im) half of A, padded to 16
  static constexpr int MP = 2 * Bh;              // A rows
  static constexpr int MT = MP / 16;             // MFMA row tiles
  // order m of multipole index n (tau-major, then l, then m)
  __host__ __device__ static constexpr int m_of(int n) {
    int tbs = B / 2;
    int k = n % tbs;
    for (int l = 1; l <= LMAX; ++l) {
      int ml = l < MMAX ? l : MMAX;
      if (k < 2 * ml + 1) return k - ml;
      k -= 2 * ml + 1;
    }
    return 0;
  }
};

template <typename Real>
struct Mfma;
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 op(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  // accumulator register v of lane l holds row 4*(l/16)+v
  static constexpr bool kRowIs4q = true;
};
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 op(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  // accumulator register v of lane l holds row 4*v + l/16
  static constexpr bool kRowIs4q = false;
};

template <typename Real>
struct Vec2;
template <>
struct Vec2<float> {
  typedef float2 t;
};
template <>
struct Vec2<double> {
  typedef double2 t;
};

struct RadialKArgs {
  const double2* xy;     // particle (x, y)
  const uint2* pairs;    // (i1, i2) sorted by (chunk, bin)
  const int* off;        // CSR over (chunk, bin): n_chunks*len_rho + 1 entries
  const int4* items;     // (chunk, bin begin, bin end, -)
  const void* tt;        // tap rows [n_tt][MP][B] of Real
  const void* x;         // padded input [n_part][B] of Real2
  void* ypart;           // per work item accumulators [n_items][Rc][2][Bh] of Real
  double rho_min, h;
  int len_rho, n_tt, s_mode, o_mode, interp, S, Rc, rows_cap;
};

// Geometry of a pair, identical in the binning and the matvec kernels.
__host__ __device__ inline void pair_geometry(double dx, double dy, double rho_min, double h, int& r, double& w,
                                              double& rho) {
  rho = sqrt(dx * dx + dy * dy);
  double t = (rho - rho_min) / h;
  double fl = floor(t);
  r = (int)fl;
  w = t - fl;
}

// Tap weights: linear (2 taps), Catmull-Rom (4 taps over rows r-1..r+2),
// not-a-knot spline (4 taps over rows y_r, Mh_r, y_{r+1}, Mh_{r+1} where
// Mh = h^2/6 S'').
__host__ __device__ inline double tap_weight(int interp, int t, double w) {
  if (interp == SMH_INTERP_LINEAR) return t == 0 ? 1.0 - w : w;
  if (interp == SMH_INTERP_CATMULL_ROM) {
    double w2 = w * w, w3 = w2 * w;
    switch (t) {
      case 0: return 0.5 * (-w3 + 2.0 * w2 - w);
      case 1: return 0.5 * (3.0 * w3 - 5.0 * w2 + 2.0);
      case 2: return 0.5 * (-3.0 * w3 + 4.0 * w2 + w);
      default: return 0.5 * (w3 - w2);
    }
  }
  double v = 1.0 - w;
  switch (t) {
    case 0: return v;
    case 1: return v * v * v - v;
    case 2: return w;
    default: return w * w * w - w;
  }
}

template <typename Real, int LMAX, int MMAX, int T>
__global__ void __launch_bounds__(512) radial_matvec_kernel(RadialKArgs a) {
  using I = SvwfIdx<LMAX, MMAX>;
  using M = Mfma<Real>;
  using R2 = typename Vec2<Real>::t;
  constexpr int B = I::B, Bh = I::Bh, MP = I::MP, MT = I::MT, LDB = B + 1;
  constexpr int KS = T * B / 4;  // K steps of 4
  constexpr int NE = 2 * MMAX + 1;
  static_assert((T * B) % 4 == 0, "K must be a multiple of 4");

  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* win = reinterpret_cast<Real*>(smem_raw);
  Real* yacc = win + (size_t)a.rows_cap * MP * LDB;
  int* tpre = reinterpret_cast<int*>(yacc + (size_t)a.Rc * 2 * Bh);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  const int4 item = a.items[blockIdx.x];
  const int chunk = item.x, rb = item.y, re = item.z;
  const int i1_base = chunk * a.Rc;
  const int* off = a.off + (size_t)chunk * a.len_rho;
  for (int j = tid; j < a.Rc * 2 * Bh; j += blockDim.x) yacc[j] = Real(0);

  const int col = lane & 15, pl = col >> 1, part = col & 1, q = lane >> 4;
  const int tl = (T == 4) ? q : (q & 1);

  for (int r0 = rb; r0 < re; r0 += a.S) {
    const int r1 = min(r0 + a.S, re);
    const int pb = off[r0], pe = off[r1];
    if (pb == pe) continue;
    const int nb = r1 - r0;
    __syncthreads();  // all tiles of the previous period are done with win / tpre
    if (tid == 0) {
      // tile prefix over the period's bins and its first / last non-empty bin
      int acc = 0, rf = -1, rl = -1;
      tpre[0] = 0;
      for (int k = 0; k < nb; ++k) {
        const int cnt = off[r0 + k + 1] - off[r0 + k];
        if (cnt > 0) {
          if (rf < 0) rf = r0 + k;
          rl = r0 + k;
        }
        acc += (cnt + 7) >> 3;
        tpre[k + 1] = acc;
      }
      tpre[62] = rf;
      tpre[63] = rl;
    }
    __syncthreads();
    const int rf = tpre[62], rl = tpre[63];
    {
      // tap rows of the non-empty bins [rf, rl]; all of them were range-checked when binning
      const int row0 = a.s_mode * rf + a.o_mode;
      const int nrows = a.s_mode * (rl - rf) + T;
      const Real* g = reinterpret_cast<const Real*>(a.tt) + (size_t)row0 * MP * B;
      const int total = nrows * MP * B;
      for (int j = tid; j < total; j += blockDim.x) {
        int row = j / (MP * B);
        int rem = j - row * (MP * B);
        int m = rem / B;
        int n2 = rem - m * B;
        win[(row * MP + m) * LDB + n2] = g[j];
      }
    }
    __syncthreads();
    const int ntiles = tpre[nb];
    for (int t = wave; t < ntiles; t += nwaves) {
      int k = 0;
      while (tpre[k + 1] <= t) ++k;
      const int r = r0 + k;
      const int p0 = off[r] + (t - tpre[k]) * 8;
      const int pend = off[r + 1];
      const int pidx = p0 + pl;
      const bool valid = pidx < pend;
      const uint2 pr = a.pairs[valid ? pidx : p0];

      const double2 xa = a.xy[pr.x], xb = a.xy[pr.y];
      const double dx = xa.x - xb.x, dy = xa.y - xb.y;
      int rr;
      double w, rho;
      pair_geometry(dx, dy, a.rho_min, a.h, rr, w, rho);
      double cphi = 1.0, sphi = 0.0;
      if (rho > 0.0) {
        cphi = dx / rho;
        sphi = dy / rho;
      }
      const Real wq = valid ? Real(tap_weight(a.interp, tl, w)) : Real(0);

      cplx<Real> E[NE];
      E[MMAX] = {Real(1), Real(0)};
      E[MMAX + 1] = {Real(cphi), Real(sphi)};
#pragma unroll
      for (int m = 2; m <= MMAX; ++m) E[MMAX + m] = cmul(E[MMAX + m - 1], E[MMAX + 1]);
#pragma unroll
      for (int m = 1; m <= MMAX; ++m) E[MMAX - m] = {E[MMAX + m].re, -E[MMAX + m].im};

      // ur[n2] = Re or Im of e^{i m2 phi} x_{i2}[n2]
      Real ur[B];
      const R2* xp = reinterpret_cast<const R2*>(a.x) + (size_t)pr.y * B;
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) {
        const R2 xv = xp[n2];
        const cplx<Real> e = E[MMAX + I::m_of(n2)];
        ur[n2] = part ? (e.re * xv.y + e.im * xv.x) : (e.re * xv.x - e.im * xv.y);
      }

      typename M::v4 acc[MT];
#pragma unroll
      for (int i = 0; i < MT; ++i) acc[i] = typename M::v4{0, 0, 0, 0};
      const int trow = a.s_mode * (r - rf) + tl;
      const Real* ap = win + (trow * MP + col) * LDB + ((T == 2) ? (q >> 1) : 0);
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        Real b;
        if (T == 4)
          b = wq * ur[s];
        else
          b = wq * ((q >> 1) ? ur[2 * s + 1] : ur[2 * s]);
        const int n2off = (T == 4) ? s : 2 * s;
#pragma unroll
        for (int i = 0; i < MT; ++i) acc[i] = M::op(ap[16 * i * LDB + n2off], b, acc[i]);
      }

      const int i1l = (int)pr.x - i1_base;
#pragma unroll
      for (int i = 0; i < Bh / 16; ++i) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const Real arz = acc[i][v], aiz = acc[i + Bh / 16][v];
          const Real arzp = __shfl_xor(arz, 1), aizp = __shfl_xor(aiz, 1);
          const Real yr = part ? (arzp - aiz) : (arz - aizp);
          const Real yi = part ? (arz + aizp) : (arzp + aiz);
          // row n1 of this accumulator element; m(n1) among the four q candidates
          int n1;
          cplx<Real> e;
          if (M::kRowIs4q) {
            n1 = 16 * i + 4 * q + v;
            const int c0 = 16 * i + v, c1 = c0 + 4, c2 = c0 + 8, c3 = c0 + 12;
            const cplx<Real> e0 = E[MMAX + (c0 < B ? I::m_of(c0) : 0)];
            const cplx<Real> e1 = E[MMAX + (c1 < B ? I::m_of(c1) : 0)];
            const cplx<Real> e2 = E[MMAX + (c2 < B ? I::m_of(c2) : 0)];
            const cplx<Real> e3 = E[MMAX + (c3 < B ? I::m_of(c3) : 0)];
            e = q == 0 ? e0 : (q == 1 ? e1 : (q == 2 ? e2 : e3));
          } else {
            n1 = 16 * i + 4 * v + q;
            const int c0 = 16 * i + 4 * v, c1 = c0 + 1, c2 = c0 + 2, c3 = c0 + 3;
            const cplx<Real> e0 = E[MMAX + (c0 < B ? I::m_of(c0) : 0)];
            const cplx<Real> e1 = E[MMAX + (c1 < B ? I::m_of(c1) : 0)];
            const cplx<Real> e2 = E[MMAX + (c2 < B ? I::m_of(c2) : 0)];
            const cplx<Real> e3 = E[MMAX + (c3 < B ? I::m_of(c3) : 0)];
            e = q == 0 ? e0 : (q == 1 ? e1 : (q == 2 ? e2 : e3));
          }
          // (yr + i yi) * conj(e)
          const Real val = part ? (yi * e.re - yr * e.im) : (yr * e.re + yi * e.im);
          if (valid && n1 < B) atomicAdd(&yacc[(i1l * 2 + part) * Bh + n1], val);
        }
      }
    }
  }
  __syncthreads();
  Real* yp = reinterpret_cast<Real*>(a.ypart) + (size_t)blockIdx.x * a.Rc * 2 * Bh;
  for (int j = tid; j < a.Rc * 2 * Bh; j += blockDim.x) yp[j] = yacc[j];
}

}  // namespace smh
