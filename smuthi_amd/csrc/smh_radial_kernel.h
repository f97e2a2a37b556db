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

// One binned pair: emitting particle i2 and receiving particle i1 - chunk base
// packed as i2 | (i1l << 24), the interpolation weight inside its table
// interval and the direction (cos phi, sin phi) of r1 - r2.
template <typename Real>
struct PairRec;
template <>
struct PairRec<float> {
  uint32_t ij;
  float w, c, s;
};
template <>
struct PairRec<double> {
  uint32_t ij, pad;
  double w, c, s;
};

struct RadialKArgs {
  const void* recs;        // PairRec<Real> sorted by (chunk, bin)
  const int* off;          // CSR over (chunk, bin): n_chunks*len_rho + 1 entries
  const int4* items;       // (chunk, first period, end period, -)
  const int4* periods;     // (first bin rf, bins nb, last non-empty bin rl, tiles)
  const void* tt;          // tap rows [n_tt][MP][B] of Real
  const void* x;           // padded input [n_part][B] of Real2
  void* ypart;             // per work item accumulators [n_items][Rc][2][Bh] of Real
  int len_rho, s_mode, o_mode, interp, Rc, rows_cap;
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

// Period bookkeeping in LDS: [0..63] tile prefix, [64..127] bin offsets.
constexpr int kMetaInts = 128;
// accumulator row stride: Bh + 1 keeps the LDS atomics of one instruction on distinct banks
template <int Bh>
struct YStride {
  static constexpr int v = Bh + 1;
};

template <typename Real, int LMAX, int MMAX, int T>
__global__ void __launch_bounds__(256) radial_matvec_kernel(RadialKArgs a) {
  using I = SvwfIdx<LMAX, MMAX>;
  using M = Mfma<Real>;
  using R2 = typename Vec2<Real>::t;
  using Rec = PairRec<Real>;
  constexpr int B = I::B, Bh = I::Bh, MP = I::MP, MT = I::MT, LDB = B + 1;
  constexpr int YS = YStride<Bh>::v;
  constexpr int KS = T * B / 4;  // K steps of 4
  constexpr int NE = 2 * MMAX + 1;
  constexpr int VW = 16 / sizeof(Real);  // reals per 16-byte staging load
  constexpr int MAXU = 20;               // staging loads per thread per period
  constexpr bool kPrefetchX = sizeof(Real) == 4;
  static_assert((T * B) % 4 == 0, "K must be a multiple of 4");
  static_assert((MP * B) % VW == 0, "tap rows must be a whole number of 16-byte vectors");

  extern __shared__ __align__(16) unsigned char smem_raw[];
  Real* win = reinterpret_cast<Real*>(smem_raw);
  Real* yacc = win + (size_t)a.rows_cap * MP * LDB;
  int* meta_a = reinterpret_cast<int*>(yacc + (size_t)a.Rc * 2 * YS);
  int* meta_b = meta_a + kMetaInts;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  const int4 item = a.items[blockIdx.x];
  const int chunk = item.x, pbeg = item.y, pend = item.z;
  const int* off = a.off + (size_t)chunk * a.len_rho;
  const Rec* recs = reinterpret_cast<const Rec*>(a.recs);
  const R2* xg = reinterpret_cast<const R2*>(a.x);
  const Real* ttg = reinterpret_cast<const Real*>(a.tt);
  for (int j = tid; j < a.Rc * 2 * YS; j += blockDim.x) yacc[j] = Real(0);

  const int col = lane & 15, pl = col >> 1, part = col & 1, q = lane >> 4;
  const int tl = (T == 4) ? q : (q & 1);

  // window rows of a period: [s*rf + o, s*rl + o + T)
  auto stage_load = [&](const int4 ph, Real (&st)[MAXU][VW]) {
    const int row0 = a.s_mode * ph.x + a.o_mode;
    const int nvec = (a.s_mode * (ph.z - ph.x) + T) * (MP * B / VW);
    const Real* g = ttg + (size_t)row0 * MP * B;
#pragma unroll
    for (int u = 0; u < MAXU; ++u) {
      const int v = tid + u * blockDim.x;
      if (v < nvec) {
        const R2* src = reinterpret_cast<const R2*>(g + (size_t)v * VW);
#pragma unroll
        for (int e = 0; e < VW / 2; ++e) {
          const R2 t2 = src[e];
          st[u][2 * e] = t2.x;
          st[u][2 * e + 1] = t2.y;
        }
      }
    }
  };
  auto stage_store = [&](const int4 ph, const Real (&st)[MAXU][VW]) {
    const int nvec = (a.s_mode * (ph.z - ph.x) + T) * (MP * B / VW);
#pragma unroll
    for (int u = 0; u < MAXU; ++u) {
      const int v = tid + u * blockDim.x;
      if (v < nvec) {
#pragma unroll
        for (int e = 0; e < VW; ++e) {
          const int j = v * VW + e;
          const int row = j / (MP * B);
          const int rem = j - row * (MP * B);
          const int m2 = rem / B;
          const int n2 = rem - m2 * B;
          win[(row * MP + m2) * LDB + n2] = st[u][e];
        }
      }
    }
  };
  // bin offsets of a period live in wave 0, one per lane (nb <= 48 < 64)
  auto meta_load = [&](const int4 ph) { return (wave == 0 && lane <= ph.y) ? off[ph.x + lane] : 0; };
  auto meta_store = [&](const int4 ph, int poff, int* m) {
    if (wave != 0) return;
    const int nxt = __shfl_down(poff, 1);
    int tiles = (lane < ph.y) ? ((nxt - poff + 7) >> 3) : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(tiles, d);
      if (lane >= d) tiles += o;
    }
    if (lane <= ph.y) m[64 + lane] = poff;
    if (lane < ph.y) m[lane + 1] = tiles;
    if (lane == 0) m[0] = 0;
  };

  Real stage[MAXU][VW];
  int4 ph_cur = pbeg < pend ? a.periods[pbeg] : int4{0, 0, 0, 0};
  if (pbeg < pend) {
    const int po = meta_load(ph_cur);
    stage_load(ph_cur, stage);
    meta_store(ph_cur, po, meta_a);
    stage_store(ph_cur, stage);
  }
  __syncthreads();
  int* cur = meta_a;
  int* nxt = meta_b;
  for (int p = pbeg; p < pend; ++p) {
    // prefetch the next period's tap rows and bin offsets while this one computes
    const bool has_next = p + 1 < pend;
    int4 ph_next = has_next ? a.periods[p + 1] : int4{0, 0, 0, 0};
    int po_next = 0;
    if (has_next) {
      po_next = meta_load(ph_next);
      stage_load(ph_next, stage);
    }
    const int ntiles = ph_cur.w;

    int kb = 0;  // bin of the tile being looked up; tiles of a wave come in increasing order
    auto tile_desc = [&](int t, int& kbin, int& p0, int& nval) {
      while (cur[kb + 1] <= t) ++kb;
      kbin = kb;
      p0 = cur[64 + kb] + (t - cur[kb]) * 8;
      nval = min(8, cur[64 + kb + 1] - p0);
    };

    int t = wave;
    Rec rec0{}, rec1{};
    int k0 = 0, p00 = 0, nv0 = 0, k1 = 0, p01 = 0, nv1 = 0;
    R2 x0[B];
    if (t < ntiles) {
      tile_desc(t, k0, p00, nv0);
      rec0 = recs[p00 + min(pl, nv0 - 1)];
      const R2* xp = xg + (size_t)(rec0.ij & 0xFFFFFFu) * B;
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) x0[n2] = xp[n2];
    }
    if (t + nwaves < ntiles) {
      tile_desc(t + nwaves, k1, p01, nv1);
      rec1 = recs[p01 + min(pl, nv1 - 1)];
    }
    for (; t < ntiles; t += nwaves) {
      R2 x1[B];
      const bool has1 = t + nwaves < ntiles;
      if (kPrefetchX && has1) {
        const R2* xp = xg + (size_t)(rec1.ij & 0xFFFFFFu) * B;
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) x1[n2] = xp[n2];
      }
      Rec rec2{};
      int k2 = 0, p02 = 0, nv2 = 0;
      if (t + 2 * nwaves < ntiles) {
        tile_desc(t + 2 * nwaves, k2, p02, nv2);
        rec2 = recs[p02 + min(pl, nv2 - 1)];
      }

      const bool valid = pl < nv0;
      const Real w = rec0.w;
      Real wq;
      if (T == 2) {
        wq = tl == 0 ? Real(1) - w : w;
      } else if (a.interp == SMH_INTERP_CATMULL_ROM) {
        const Real w2 = w * w, w3 = w2 * w;
        wq = tl == 0 ? Real(0.5) * (-w3 + Real(2) * w2 - w)
             : tl == 1 ? Real(0.5) * (Real(3) * w3 - Real(5) * w2 + Real(2))
             : tl == 2 ? Real(0.5) * (Real(-3) * w3 + Real(4) * w2 + w)
                       : Real(0.5) * (w3 - w2);
      } else {
        const Real v = Real(1) - w;
        wq = tl == 0 ? v : tl == 1 ? v * v * v - v : tl == 2 ? w : w * w * w - w;
      }
      if (!valid) wq = Real(0);

      cplx<Real> E[NE];
      E[MMAX] = {Real(1), Real(0)};
      E[MMAX + 1] = {rec0.c, rec0.s};
#pragma unroll
      for (int m = 2; m <= MMAX; ++m) E[MMAX + m] = cmul(E[MMAX + m - 1], E[MMAX + 1]);
#pragma unroll
      for (int m = 1; m <= MMAX; ++m) E[MMAX - m] = {E[MMAX + m].re, -E[MMAX + m].im};

      if (!kPrefetchX) {
        const R2* xp = xg + (size_t)(rec0.ij & 0xFFFFFFu) * B;
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) x0[n2] = xp[n2];
      }
      Real ur[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) {
        const cplx<Real> e = E[MMAX + I::m_of(n2)];
        ur[n2] = part ? (e.re * x0[n2].y + e.im * x0[n2].x) : (e.re * x0[n2].x - e.im * x0[n2].y);
      }

      typename M::v4 acc[MT];
#pragma unroll
      for (int i = 0; i < MT; ++i) acc[i] = typename M::v4{0, 0, 0, 0};
      // bin ph_cur.x + k0: its first tap row sits s_mode * k0 rows into the period's window
      const Real* ap = win + ((a.s_mode * k0 + tl) * MP + col) * LDB + ((T == 2) ? (q >> 1) : 0);
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

      Real* yrow = yacc + ((int)(rec0.ij >> 24) * 2 + part) * YS;
#pragma unroll
      for (int i = 0; i < Bh / 16; ++i) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const Real arz = acc[i][v], aiz = acc[i + Bh / 16][v];
          const Real arzp = __shfl_xor(arz, 1), aizp = __shfl_xor(aiz, 1);
          const Real yr = part ? (arzp - aiz) : (arz - aizp);
          const Real yi = part ? (arz + aizp) : (arzp + aiz);
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
          if (valid && n1 < B) atomicAdd(&yrow[n1], val);
        }
      }

      rec0 = rec1;
      k0 = k1;
      p00 = p01;
      nv0 = nv1;
      rec1 = rec2;
      k1 = k2;
      p01 = p02;
      nv1 = nv2;
      if (kPrefetchX && has1) {
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) x0[n2] = x1[n2];
      }
    }
    __syncthreads();  // every tile of this period is done with win and cur
    if (has_next) {
      stage_store(ph_next, stage);
      meta_store(ph_next, po_next, nxt);
    }
    __syncthreads();
    int* tmp = cur;
    cur = nxt;
    nxt = tmp;
    ph_cur = ph_next;
  }
  __syncthreads();
  Real* yp = reinterpret_cast<Real*>(a.ypart) + (size_t)blockIdx.x * a.Rc * 2 * Bh;
  for (int j = tid; j < a.Rc * 2 * Bh; j += blockDim.x) {
    const int rowi = j / Bh, n1 = j - rowi * Bh;
    yp[j] = yacc[rowi * YS + n1];
  }
}

}  // namespace smh
