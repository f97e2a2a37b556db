// Radial-lookup coupling matvec on CDNA4 MFMA.
//
// Work decomposition. Particle pairs (i1 receiving, i2 emitting) are binned
// by the table interval r = floor((rho12 - rho_min) / h) and grouped by
// receiver chunk (Rc consecutive receivers). For a bin r every pair needs the
// same T table rows ("taps"), so the bin is a GEMM over its pairs:
//
//   [y_re; y_im](n1, pair) = [Lr -Li; Li Lr](n1, (n2, t)) * [Zr; Zi]((n2, t), pair)
//   Z((n2, t), pair) = w_t(rho) e^{i m2 phi} x_{i2}[n2]
//
// run on v_mfma 16x16x4 (FP32 or FP64) with 16 pairs per tile. A wave owns a
// bin entry at a time and holds the bin's tap fragments in registers in the
// MFMA A-operand layout (the tap table is stored in exactly that order, so
// the loads are contiguous 16-byte vectors); the four blocks Lr, Li of the
// stacked matrix reuse the same registers, with -Zi supplied as a second B
// operand. Per pair the result is rotated by e^{-i m1 phi} and accumulated
// into an LDS accumulator of the chunk's receivers, written once per work
// item; a second kernel sums the work items of a chunk in fixed order (no
// global atomics, deterministic).
//
// K ordering inside one MFMA step: lane quarter q = lane / 16 takes tap
// t = q / G and multipole n2 = (q % G) * KS + s, G = 4 / T, at K-step s.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "smh_common.h"

namespace smh {

// Value of lane l ^ 1 (DPP quad permute [1, 0, 3, 2]: a VALU move, no LDS).
// The permutation never reads an absent lane; bound_ctrl lets the compiler
// fold the move into a following multiply (v_mul_f32_dpp).
__device__ inline float lane_xor1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, true));
}
__device__ inline double lane_xor1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int LMAX, int MMAX>
struct SvwfIdx {
  static constexpr int B = 2 * (MMAX * (MMAX + 2) + (LMAX - MMAX) * (2 * MMAX + 1));
  static constexpr int Bh = (B + 15) / 16 * 16;  // rows of Lr (or Li), padded to 16
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

// Fragment geometry of the tap table for blocksize B and T taps.
struct TapGeom {
  int G, KS, NG, MT, FR;
  __host__ __device__ constexpr TapGeom(int B, int T)
      : G(4 / T),
        KS((B + 4 / T - 1) / (4 / T)),
        NG(((B + 4 / T - 1) / (4 / T) + 3) / 4),
        MT(2 * ((B + 15) / 16)),
        FR(2 * ((B + 15) / 16) * (4 / T) * (((B + 4 / T - 1) / (4 / T) + 3) / 4) * 64) {}
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
// packed as i2 | (i1l << 23), the interpolation weight inside its table
// interval and the direction (cos phi, sin phi) of r1 - r2.
constexpr int kPairShift = 23;
constexpr uint32_t kPairMask = (1u << kPairShift) - 1u;
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

// A pair of the volume lookup: two interpolation weights (distance and z
// coordinate of the node cell).
template <typename Real>
struct PairRec2;
template <>
struct PairRec2<float> {
  uint32_t ij;
  float w, w2, c, s;
};
template <>
struct PairRec2<double> {
  uint32_t ij, pad;
  double w, w2, c, s;
};

// Largest number of pairs in one bin entry (a bin of a chunk with more pairs
// is split into several entries so waves stay balanced).
constexpr int kEntryPairs = 256;

struct RadialKArgs {
  const void* recs;     // PairRec<Real> sorted by (chunk, bin)
  const int4* entries;  // (first tap row, first pair, pairs, -) per bin entry, grouped by item
  const int2* items;    // (chunk, first entry); entries of item k end where item k+1 begins
  const void* tt;       // tap table [n_tt][FR] of Real in fragment order
  const void* x;        // padded input [n_part][B] of Real2
  void* ypart;          // per work item accumulators [n_items][Rc][2][B] of Real
  const float* xmax;        // FP32: max_j |Re x_j| + |Im x_j| of this matvec's input
  const float* item_lsum;   // FP32: per item, largest tap-row sum of |Re L| + |Im L| over its bins
  const int* item_en;       // FP32: per item, ceil(log2) of the most pairs one receiver has in it
  int n_items, n_entries, interp, Rc;
  int len_z;                // 2-D lookups: table rows per distance node (tap t sits at (t/2) len_z + t%2)
  int row_s = 1, row_o = 0;  // entry tap row = row_s * bin + row_o (kernels on per-bin coefficient rows)
  int vmode = 1, vslice = 0;  // volume lookups: interpolation mode and z slice (vol_tap_offset)
  // block -> work item (null: identity). Items sorted by first table row and
  // dealt so that the blocks one XCD runs (blockIdx mod 8) walk neighbouring
  // bins together and share the tap rows in that XCD's L2.
  const int* order = nullptr;
  // single-vector kernel without a launch order: block b runs item item0 + b
  // and writes its partial at ypart[b] (launches over a range of the items)
  int item0 = 0;
};

// FP32 receiver accumulation in LDS is fixed point: a contribution v is split
// exactly into v = h / S_hi + r with integer h, and r is rounded to l / S_lo,
// S_hi and S_lo powers of two chosen per work item from a bound on |v| and
// on the number of contributions a receiver gets, so neither int32 sum can
// overflow. Both parts are added with integer LDS atomics (an order of
// magnitude cheaper than float LDS atomics on gfx950, and order independent).
// Bound: |v| <= V = 8 * lsum * xmax (sum_t |w_t| <= 2, a factor 2 for the
// rotation and 2 for rounding); with at most 2^eN contributions,
//   S_hi = 2^floor(30 - eN - log2 V),  S_lo = S_hi * 2^(31 - eN),
// and the accumulated error is below 2^(2 eN - 62) V.
__device__ inline void fixed_scales(float lsum, float xmax, int eN, float& sh, float& ish, float& sl, float& isl) {
  const double V = 8.0 * (double)lsum * (double)xmax;
  int kh = 0;
  if (V > 0.0 && V < 1e300) kh = (int)floor(30.0 - eN - log2(V));
  kh = max(-120, min(kh, 120 - (31 - eN)));
  const int kl = kh + 31 - eN;
  sh = ldexpf(1.0f, kh);
  ish = ldexpf(1.0f, -kh);
  sl = ldexpf(1.0f, kl);
  isl = ldexpf(1.0f, -kl);
}
__device__ inline void fixed_add(int* p, int lo_off, float v, float sh, float ish, float sl) {
  const float h = rintf(v * sh);
  const float r = fmaf(-h, ish, v);  // exact: h / S_hi is within half an ulp of S_hi of v
  const float l = rintf(r * sl);
  atomicAdd(p, (int)h);
  atomicAdd(p + lo_off, (int)l);
}

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
template <typename Real>
__host__ __device__ inline Real tap_weight(int interp, int t, Real w) {
  if (interp == SMH_INTERP_LINEAR) return t == 0 ? Real(1) - w : w;
  if (interp == SMH_INTERP_CATMULL_ROM) {
    const Real w2 = w * w, w3 = w2 * w;
    switch (t) {
      case 0: return Real(0.5) * (-w3 + Real(2) * w2 - w);
      case 1: return Real(0.5) * (Real(3) * w3 - Real(5) * w2 + Real(2));
      case 2: return Real(0.5) * (Real(-3) * w3 + Real(4) * w2 + w);
      default: return Real(0.5) * (w3 - w2);
    }
  }
  const Real v = Real(1) - w;
  switch (t) {
    case 0: return v;
    case 1: return v * v * v - v;
    case 2: return w;
    default: return w * w * w - w;
  }
}

// Launch shape of a kernel instance: split the n2 range into KSPLIT parts so
// the tap fragments of one part fit the register budget of the occupancy
// wanted (FP32: 2 waves per SIMD, FP64: 1).
#ifndef SMH_FP32_TAP_REGS
#define SMH_FP32_TAP_REGS 128
#endif
#ifndef SMH_FP32_WAVES
#define SMH_FP32_WAVES 8
#endif
#ifndef SMH_X_PREFETCH
#define SMH_X_PREFETCH 0
#endif
template <typename Real, int B, int T>
struct RadialShape {
  static constexpr TapGeom g{B, T};
  static constexpr int kRegBudget = sizeof(Real) == 4 ? SMH_FP32_TAP_REGS : 192;
  static constexpr int regs_for(int split) { return g.MT * ((g.NG + split - 1) / split) * 4 * (int)(sizeof(Real) / 4); }
  static constexpr int pick() {
    int s = 1;
    while (regs_for(s) > kRegBudget && s < g.NG) ++s;
    return s;
  }
  static constexpr int KSPLIT = pick();
  static constexpr int NGS = (g.NG + KSPLIT - 1) / KSPLIT;  // n2 groups per part
  static constexpr int kWaves = (sizeof(Real) == 4 && regs_for(KSPLIT) <= 128) ? SMH_FP32_WAVES : 4;
  static constexpr int kThreads = 64 * kWaves;
  // accumulator words per receiver (odd stride spreads LDS banks): FP32 keeps
  // hi and lo int32 words for 2B values, FP64 keeps 2B doubles
  static constexpr int SY = sizeof(Real) == 4 ? 4 * B + 1 : 2 * B + 1;
};

// Volume lookups over nodes (distance r, z index j). vmode 1, bilinear: one
// launch per table, nodes at row r * len_z + j, taps (r, j), (r, j+1),
// (r+1, j), (r+1, j+1). vmode 2 (tensor Catmull-Rom) and 3 (tensor not-a-knot
// spline) run as four launches per table, one per z tap s ("slice"), each a
// 4-tap interpolation along rho weighted by the slice's z weight:
//   Catmull-Rom: tap t at node (r + t - 1, j + s - 1), one row per node;
//   spline: four rows per node (value, rho moment, z moment, mixed moment,
//   moments Mh = h^2/6 S''), tap t = (node r + t/2, rho kind t%2), slice
//   s = (node j + s/2, z kind s%2), weights of tap_weight() in each direction.
// Offsets are relative to the entry's row (node (r, j)).
__host__ __device__ inline int vol_tap_offset(int vmode, int s, int t, int len_z) {
  if (vmode == 2) return (t - 1) * len_z + (s - 1);
  if (vmode == 3) return 4 * ((t >> 1) * len_z + (s >> 1)) + (t & 1) + 2 * (s & 1);
  return (t >> 1) * len_z + (t & 1);
}
template <typename Real>
__host__ __device__ inline Real vol_tap_weight(int vmode, int s, int t, Real w, Real w2) {
  if (vmode == 2)
    return tap_weight<Real>(SMH_INTERP_CATMULL_ROM, t, w) * tap_weight<Real>(SMH_INTERP_CATMULL_ROM, s, w2);
  if (vmode == 3) return tap_weight<Real>(SMH_INTERP_SPLINE, t, w) * tap_weight<Real>(SMH_INTERP_SPLINE, s, w2);
  return ((t >> 1) ? w : Real(1) - w) * ((t & 1) ? w2 : Real(1) - w2);
}

// Lookup matvec. VOL = false: radial lookup, taps are consecutive tap-table
// rows, weights from the 1-D interpolant. VOL = true: volume lookup, taps and
// weights of vol_tap_offset / vol_tap_weight for a.vmode and a.vslice.
// TapT: storage type of the tap table (float for FP32, and for FP64 over
// complex64 tables, which float holds exactly; double for FP64 tables).
template <typename Real, typename TapT, int LMAX, int MMAX, int T, bool VOL>
__global__ void __launch_bounds__((RadialShape<Real, SvwfIdx<LMAX, MMAX>::B, T>::kThreads))
    lookup_matvec_kernel(RadialKArgs a) {
  using I = SvwfIdx<LMAX, MMAX>;
  using M = Mfma<Real>;
  using R2 = typename Vec2<Real>::t;
  using Rec = typename std::conditional<VOL, PairRec2<Real>, PairRec<Real>>::type;
  using Sh = RadialShape<Real, I::B, T>;
  constexpr int B = I::B, Bh = I::Bh, MMX = MMAX;
  constexpr TapGeom g{B, T};
  constexpr int G = g.G, KS = g.KS, NG = g.NG, MT = g.MT, MH = MT / 2, FR = g.FR;
  constexpr int KSPLIT = Sh::KSPLIT, NGS = Sh::NGS, SY = Sh::SY;
  constexpr int NE = 2 * MMX + 1;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr bool kFixed = sizeof(Real) == 4;
  using AccT = typename std::conditional<kFixed, int, double>::type;
  AccT* yacc = reinterpret_cast<AccT*>(smem_raw);
  int* next_entry = reinterpret_cast<int*>(yacc + (size_t)a.Rc * SY);
  // work item of this block (launch order maps blocks to items, see RadialKArgs::order)
  const int item = a.order ? a.order[blockIdx.x] : a.item0 + (int)blockIdx.x;
  float sh = 0.f, ish = 0.f, sl = 0.f, isl = 0.f;
  // headroom for the additions: 2^eN pairs per receiver, KSPLIT parts each
  const int eN = a.item_en[item] + (KSPLIT > 1 ? 32 - __builtin_clz(KSPLIT - 1) : 0);
  if (kFixed) fixed_scales(a.item_lsum[item], *a.xmax, eN, sh, ish, sl, isl);

  const int tid = threadIdx.x, lane = tid & 63;
  const int2 it = a.items[item];
  const int e_end = item + 1 < a.n_items ? a.items[item + 1].y : a.n_entries;
  const Rec* recs = reinterpret_cast<const Rec*>(a.recs);
  const R2* xg = reinterpret_cast<const R2*>(a.x);
  const TapT* ttg = reinterpret_cast<const TapT*>(a.tt);
  for (int j = tid; j < a.Rc * SY; j += blockDim.x) yacc[j] = AccT(0);
  if (tid == 0) *next_entry = it.y;
  __syncthreads();

  const int col = lane & 15, q = lane >> 4;
  const int tap = q / G, hq = q % G;

  for (;;) {
    int e = 0;
    if (lane == 0) e = atomicAdd(next_entry, 1);
    e = __shfl(e, 0);
    if (e >= e_end) break;
    const int4 en = a.entries[e];  // (first tap row, first pair, pairs)
    const int toff = VOL ? vol_tap_offset(a.vmode, a.vslice, tap, a.len_z) : tap;
    const TapT* trow = ttg + (size_t)(en.x + toff) * FR;
    const int ntiles = (en.z + 15) >> 4;

#pragma unroll
    for (int sp = 0; sp < KSPLIT; ++sp) {
      // tap fragments of this n2 part: A(row 16 i + col, n2 = hq*KS + 4 gg + v)
      Real af[MT][NGS][4];
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int gl = 0; gl < NGS; ++gl) {
          const int gg = sp * NGS + gl;
          if (gg < NG) {
            const TapT* src = trow + ((size_t)((i * G + hq) * NG + gg) * 16 + col) * 4;
            if constexpr (sizeof(TapT) == 4) {
              const float4 v = *reinterpret_cast<const float4*>(src);
              af[i][gl][0] = v.x;
              af[i][gl][1] = v.y;
              af[i][gl][2] = v.z;
              af[i][gl][3] = v.w;
            } else {
              const double2 v0 = reinterpret_cast<const double2*>(src)[0];
              const double2 v1 = reinterpret_cast<const double2*>(src)[1];
              af[i][gl][0] = v0.x;
              af[i][gl][1] = v0.y;
              af[i][gl][2] = v1.x;
              af[i][gl][3] = v1.y;
            }
          }
        }

      // software pipeline: pair records two tiles ahead, inputs x one tile ahead
      constexpr int NX = NGS * 4;
      auto rec_at = [&](int tl) -> Rec {
        const int ps = tl * 16 + col;
        return recs[en.y + (ps < en.z ? ps : en.z - 1)];
      };
      auto load_x = [&](const Rec& r, R2(&xs)[NX]) {
        const R2* xp = xg + (size_t)(r.ij & kPairMask) * B;
#pragma unroll
        for (int u = 0; u < NX; ++u) {
          const int s = sp * NX + u;
          if (s < KS) {
            const int n2 = (G == 2 && hq) ? KS + s : s;
            xs[u] = n2 < B ? xp[n2] : R2{Real(0), Real(0)};
          }
        }
      };
      Rec rec0 = rec_at(0);
      Rec rec1 = rec_at(ntiles > 1 ? 1 : 0);
      R2 x0[NX];
      load_x(rec0, x0);
      for (int tl = 0; tl < ntiles; ++tl) {
        const Rec rec2 = rec_at(tl + 2 < ntiles ? tl + 2 : ntiles - 1);
        R2 x1[NX];
        if (SMH_X_PREFETCH) load_x(rec1, x1);
        const Rec rec = rec0;
        const bool valid = tl * 16 + col < en.z;
        Real wq;
        if constexpr (VOL)
          wq = vol_tap_weight<Real>(a.vmode, a.vslice, tap, rec.w, rec.w2);
        else
          wq = tap_weight<Real>(a.interp, tap, rec.w);
        if (!valid) wq = Real(0);
        // e^{i m phi}, m = -MMAX..MMAX
        cplx<Real> E[NE];
        E[MMX] = {Real(1), Real(0)};
        E[MMX + 1] = {rec.c, rec.s};
#pragma unroll
        for (int m = 2; m <= MMX; ++m) E[MMX + m] = cmul(E[MMX + m - 1], E[MMX + 1]);
#pragma unroll
        for (int m = 1; m <= MMX; ++m) E[MMX - m] = {E[MMX + m].re, -E[MMX + m].im};

        typename M::v4 acc[MT];
#pragma unroll
        for (int i = 0; i < MT; ++i) acc[i] = typename M::v4{0, 0, 0, 0};
#pragma unroll
        for (int gl = 0; gl < NGS; ++gl) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int s = (sp * NGS + gl) * 4 + v;  // K-step
            if (s < KS) {
              // n2 of this lane: hq * KS + s; its order m is one of G compile-time values
              const int n2a = s, n2b = (G == 2) ? KS + s : s;
              const cplx<Real> ea = E[MMX + I::m_of(n2a < B ? n2a : 0)];
              const cplx<Real> eb = E[MMX + I::m_of(n2b < B ? n2b : 0)];
              const cplx<Real> em = hq ? eb : ea;
              R2 xv = x0[gl * 4 + v];
              xv.x *= wq;
              xv.y *= wq;
              const Real zr = em.re * xv.x - em.im * xv.y;
              const Real zi = em.re * xv.y + em.im * xv.x;
#pragma unroll
              for (int i = 0; i < MT; ++i) acc[i] = M::op(af[i][gl][v], zr, acc[i]);
#pragma unroll
              for (int i = 0; i < MH; ++i) acc[i] = M::op(af[i + MH][gl][v], -zi, acc[i]);
#pragma unroll
              for (int i = 0; i < MH; ++i) acc[i + MH] = M::op(af[i][gl][v], zi, acc[i + MH]);
            }
          }
        }

        // rotate by conj(e^{i m1 phi}) and accumulate into the receiver's LDS row
        AccT* yrow = yacc + (size_t)(rec.ij >> kPairShift) * SY;
#pragma unroll
        for (int i = 0; i < MH; ++i) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            int n1;
            cplx<Real> e;
            if (M::kRowIs4q) {
              n1 = 16 * i + 4 * q + v;
              const int c0 = 16 * i + v;
              const cplx<Real> e0 = E[MMX + (c0 < B ? I::m_of(c0) : 0)];
              const cplx<Real> e1 = E[MMX + (c0 + 4 < B ? I::m_of(c0 + 4) : 0)];
              const cplx<Real> e2 = E[MMX + (c0 + 8 < B ? I::m_of(c0 + 8) : 0)];
              const cplx<Real> e3 = E[MMX + (c0 + 12 < B ? I::m_of(c0 + 12) : 0)];
              e = q == 0 ? e0 : (q == 1 ? e1 : (q == 2 ? e2 : e3));
            } else {
              n1 = 16 * i + 4 * v + q;
              const int c0 = 16 * i + 4 * v;
              const cplx<Real> e0 = E[MMX + (c0 < B ? I::m_of(c0) : 0)];
              const cplx<Real> e1 = E[MMX + (c0 + 1 < B ? I::m_of(c0 + 1) : 0)];
              const cplx<Real> e2 = E[MMX + (c0 + 2 < B ? I::m_of(c0 + 2) : 0)];
              const cplx<Real> e3 = E[MMX + (c0 + 3 < B ? I::m_of(c0 + 3) : 0)];
              e = q == 0 ? e0 : (q == 1 ? e1 : (q == 2 ? e2 : e3));
            }
            const Real yr = acc[i][v], yi = acc[i + MH][v];
            if (valid && n1 < B) {
              const Real v0 = yr * e.re + yi * e.im, v1 = yi * e.re - yr * e.im;
              if constexpr (kFixed) {
                fixed_add(reinterpret_cast<int*>(&yrow[n1]), 2 * B, v0, sh, ish, sl);
                fixed_add(reinterpret_cast<int*>(&yrow[B + n1]), 2 * B, v1, sh, ish, sl);
              } else {
                atomicAdd(&yrow[n1], v0);
                atomicAdd(&yrow[B + n1], v1);
              }
            }
          }
        }
        rec0 = rec1;
        rec1 = rec2;
        if (SMH_X_PREFETCH) {
#pragma unroll
          for (int u = 0; u < NX; ++u) x0[u] = x1[u];
        } else if (tl + 1 < ntiles) {
          load_x(rec0, x0);
        }
      }
    }
  }
  __syncthreads();
  Real* yp = reinterpret_cast<Real*>(a.ypart) + (size_t)(item - a.item0) * a.Rc * 2 * B;
  for (int j = tid; j < a.Rc * 2 * B; j += blockDim.x) {
    const int il = j / (2 * B), k = j - il * 2 * B;
    if constexpr (kFixed)
      yp[j] = (Real)((double)yacc[il * SY + k] * (double)ish + (double)yacc[il * SY + 2 * B + k] * (double)isl);
    else
      yp[j] = (Real)yacc[il * SY + k];
  }
}

}  // namespace smh
