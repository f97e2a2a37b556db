// Lookup matvec for a block of kBatch right-hand sides at once.
//
// Per pair the lookup block is interpolated first and then applied to all
// right-hand sides with MFMA 16x16x4:
//   Y[(Re|Im L, n1), (rhs, re|im)] = sum_n2 A_pair[(Re|Im L, n1), n2] Z[n2, (rhs, re|im)]
//   A_pair = sum_t w_t(pair) A_t,   Z[n2, (rhs, .)] = e^{i m2 phi} x_{i2}[n2][rhs]
// A wave owns a bin entry at a time and holds the T tap rows' values it needs
// in registers: lane (q, rr) keeps A_t(16 i + rr, 4 s + q) for row tiles
// i < 4 and K-steps s, so the interpolation of its A fragment is T FMAs next
// to each MFMA. One pair is one tile (16 columns = kBatch RHS x re/im). The
// epilogue combines the re/im columns of a lane pair, rotates by
// e^{-i m1 phi} and adds to the receiver's LDS accumulator [Rc][kBatch][2][B]
// (FP32: two-level int32 fixed point as the single-RHS kernel; FP64: doubles).
// Volume lookups (VOL) take the taps and weights of vol_tap_offset /
// vol_tap_weight for a.vmode and the z slice a.vslice, one launch per table
// and slice as the single-vector volume kernel.
#pragma once

#include "smh_op.h"
#include "smh_radial_kernel.h"

namespace smh {

constexpr int kBatch = kBatchRhs;

// Launch shape: FP32 runs two waves per SIMD (tap values within 128
// registers), FP64 one; the n2 range is split into SPL parts (each with its
// own epilogue) when the tap values of a whole pair exceed that budget.
template <typename Real, int B, int T>
struct BatchShape {
  static constexpr int SY = sizeof(Real) == 4 ? 4 * kBatch * B + 1 : 2 * kBatch * B + 1;
  static constexpr int kThreads = sizeof(Real) == 4 ? 512 : 256;
  static constexpr int MT = 2 * ((B + 15) / 16), S8 = (B + 3) / 4;
  static constexpr int kBudget = sizeof(Real) == 4 ? 128 : 256;
  static constexpr int regs_for(int spl) { return T * MT * ((S8 + spl - 1) / spl) * (int)(sizeof(Real) / 4); }
  static constexpr int pick() {
    int sp = 1;
    while (regs_for(sp) > kBudget && sp < S8) ++sp;
    return sp;
  }
  static constexpr int SPL = pick();
  static constexpr int SP = (S8 + SPL - 1) / SPL;  // K-steps per part
};

template <typename Real, typename TapT, int LMAX, int MMAX, int T, bool VOL>
__global__ void __launch_bounds__((BatchShape<Real, SvwfIdx<LMAX, MMAX>::B, T>::kThreads))
    batch_matvec_kernel(RadialKArgs a) {
  using I = SvwfIdx<LMAX, MMAX>;
  using M = Mfma<Real>;
  using R2 = typename Vec2<Real>::t;
  using Rec = typename std::conditional<VOL, PairRec2<Real>, PairRec<Real>>::type;
  constexpr int B = I::B, Bh = I::Bh, MMX = MMAX;
  static_assert(Bh == 16 || Bh == 32 || Bh == 48, "blocksize");
  constexpr TapGeom g{B, T};
  constexpr int G = g.G, KS = g.KS, NG = g.NG, FR = g.FR;
  constexpr int MT = 2 * Bh / 16, MH = MT / 2;  // row tiles: [0, MH) Re L, [MH, MT) Im L
  using Sh = BatchShape<Real, B, T>;
  constexpr int SY = Sh::SY, SPL = Sh::SPL, SP = Sh::SP;
  constexpr int NE = 2 * MMX + 1;
  constexpr bool kFixed = sizeof(Real) == 4;
  using AccT = typename std::conditional<kFixed, int, double>::type;

  extern __shared__ __align__(16) unsigned char smem_raw[];
  AccT* yacc = reinterpret_cast<AccT*>(smem_raw);
  int* next_entry = reinterpret_cast<int*>(yacc + (size_t)a.Rc * SY);
  // work item of this block (launch order maps blocks to items, see RadialKArgs::order)
  const int item = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
  float sh = 0.f, ish = 0.f, sl = 0.f, isl = 0.f;
  // headroom for the additions: 2^eN pairs per receiver, SPL parts each
  const int eN = a.item_en[item] + (SPL > 1 ? 32 - __builtin_clz(SPL - 1) : 0);
  if (kFixed) fixed_scales(a.item_lsum[item], *a.xmax, eN, sh, ish, sl, isl);

  const int tid = threadIdx.x, lane = tid & 63;
  const int2 it = a.items[item];
  const int e_end = item + 1 < a.n_items ? a.items[item + 1].y : a.n_entries;
  const Rec* recs = reinterpret_cast<const Rec*>(a.recs);
  const R2* xg = reinterpret_cast<const R2*>(a.x);  // [n_part][B][kBatch]
  const TapT* ttg = reinterpret_cast<const TapT*>(a.tt);
  for (int j = tid; j < a.Rc * SY; j += blockDim.x) yacc[j] = AccT(0);
  if (tid == 0) *next_entry = it.y;
  __syncthreads();

  const int rr = lane & 15, q = lane >> 4;
  const int rhs = rr >> 1, zp = rr & 1;  // as a B-operand / result lane: column (rhs, re|im)

  // tap-table offset of A(m, n2) inside one tap row
  auto foff = [&](int m, int n2) {
    const int h = n2 / KS, k = n2 - h * KS;
    return ((((m >> 4) * G + h) * NG + (k >> 2)) * 16 + (m & 15)) * 4 + (k & 3);
  };

  for (;;) {
    int e = 0;
    if (lane == 0) e = atomicAdd(next_entry, 1);
    e = __shfl(e, 0);
    if (e >= e_end) break;
    const int4 en = a.entries[e];
#pragma unroll
    for (int sp = 0; sp < SPL; ++sp) {
    // tap values of this lane for the K-steps of part sp: tv[t][i][s] = A_t(16 i + rr, 4 (sp SP + s) + q)
    Real tv[T][MT][SP];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      const int toff = VOL ? vol_tap_offset(a.vmode, a.vslice, t, a.len_z) : t;
      const TapT* row = ttg + (size_t)(en.x + toff) * FR;
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int s = 0; s < SP; ++s) {
          const int n2 = 4 * (sp * SP + s) + q;
          tv[t][i][s] = n2 < B ? (Real)row[foff(16 * i + rr, n2)] : Real(0);
        }
    }
    auto rec_at = [&](int p) { return recs[en.y + (p < en.z ? p : en.z - 1)]; };
    // inputs of a pair for this lane: x_{i2}[4 s' + q][rhs] of the part's K-steps
    auto load_x = [&](const Rec& r, R2(&xs)[SP]) {
      const R2* xp = xg + (size_t)(r.ij & kPairMask) * B * kBatch + rhs;
#pragma unroll
      for (int s = 0; s < SP; ++s) {
        const int n2 = 4 * (sp * SP + s) + q;
        xs[s] = n2 < B ? xp[(size_t)n2 * kBatch] : R2{Real(0), Real(0)};
      }
    };
    // software pipeline: records two pairs ahead, inputs one pair ahead
    Rec rec_n = rec_at(0), rec_nn = rec_at(1);
    R2 xn[SP];
    load_x(rec_n, xn);
    for (int p = 0; p < en.z; ++p) {
      const Rec rec = rec_n;
      R2 xc[SP];
#pragma unroll
      for (int s = 0; s < SP; ++s) xc[s] = xn[s];
      rec_n = rec_nn;
      rec_nn = rec_at(p + 2);
      load_x(rec_n, xn);
      // interpolation weights of this pair
      Real w[T];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        if constexpr (VOL)
          w[t] = vol_tap_weight<Real>(a.vmode, a.vslice, t, rec.w, rec.w2);
        else
          w[t] = tap_weight<Real>(a.interp, t, rec.w);
      }
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
      for (int s = 0; s < SP; ++s) {
        // B operand: (e^{i m2 phi} x[n2][rhs]).re or .im, n2 = 4 s' + q
        const int n2 = 4 * (sp * SP + s) + q;
        Real z = Real(0);
        if (n2 < B) {
          const int c0 = 4 * (sp * SP + s);
          const cplx<Real> e0 = E[MMX + I::m_of(c0 < B ? c0 : 0)];
          const cplx<Real> e1 = E[MMX + I::m_of(c0 + 1 < B ? c0 + 1 : 0)];
          const cplx<Real> e2 = E[MMX + I::m_of(c0 + 2 < B ? c0 + 2 : 0)];
          const cplx<Real> e3 = E[MMX + I::m_of(c0 + 3 < B ? c0 + 3 : 0)];
          const cplx<Real> em = q == 0 ? e0 : (q == 1 ? e1 : (q == 2 ? e2 : e3));
          const R2 xv = xc[s];
          z = zp ? (em.re * xv.y + em.im * xv.x) : (em.re * xv.x - em.im * xv.y);
        }
#pragma unroll
        for (int i = 0; i < MT; ++i) {
          Real av = w[0] * tv[0][i][s];
#pragma unroll
          for (int t = 1; t < T; ++t) av += w[t] * tv[t][i][s];
          acc[i] = M::op(av, z, acc[i]);
        }
      }

      // y = (Lr Zr - Li Zi) + i (Lr Zi + Li Zr); each lane of a (re, im) column pair ends with both parts
      AccT* yrow = yacc + (size_t)(rec.ij >> kPairShift) * SY;
#pragma unroll
      for (int i = 0; i < MH; ++i) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const Real ar = acc[i][v], ai = acc[i + MH][v];  // Lr z_zp, Li z_zp
          const Real pr = lane_xor1(ar), pi = lane_xor1(ai);
          const Real yr = zp ? (pr - ai) : (ar - pi);
          const Real yi = zp ? (ar + pi) : (pr + ai);
          int n1;
          cplx<Real> em;
          if (M::kRowIs4q) {
            n1 = 16 * i + 4 * q + v;
            const int c0 = 16 * i + v;
            const cplx<Real> e0 = E[MMX + (c0 < B ? I::m_of(c0) : 0)];
            const cplx<Real> e1 = E[MMX + (c0 + 4 < B ? I::m_of(c0 + 4) : 0)];
            const cplx<Real> e2 = E[MMX + (c0 + 8 < B ? I::m_of(c0 + 8) : 0)];
            const cplx<Real> e3 = E[MMX + (c0 + 12 < B ? I::m_of(c0 + 12) : 0)];
            em = q == 0 ? e0 : (q == 1 ? e1 : (q == 2 ? e2 : e3));
          } else {
            n1 = 16 * i + 4 * v + q;
            const int c0 = 16 * i + 4 * v;
            const cplx<Real> e0 = E[MMX + (c0 < B ? I::m_of(c0) : 0)];
            const cplx<Real> e1 = E[MMX + (c0 + 1 < B ? I::m_of(c0 + 1) : 0)];
            const cplx<Real> e2 = E[MMX + (c0 + 2 < B ? I::m_of(c0 + 2) : 0)];
            const cplx<Real> e3 = E[MMX + (c0 + 3 < B ? I::m_of(c0 + 3) : 0)];
            em = q == 0 ? e0 : (q == 1 ? e1 : (q == 2 ? e2 : e3));
          }
          // (yr + i yi) conj(em); this lane keeps the part zp
          const Real val = zp ? (yi * em.re - yr * em.im) : (yr * em.re + yi * em.im);
          if (n1 < B) {
            const int k = (rhs * 2 + zp) * B + n1;
            if constexpr (kFixed)
              fixed_add(reinterpret_cast<int*>(&yrow[k]), 2 * kBatch * B, val, sh, ish, sl);
            else
              atomicAdd(&yrow[k], val);
          }
        }
      }
    }
    }
  }
  __syncthreads();
  // partial of this item: [Rc][kBatch][2][B]
  Real* yp = reinterpret_cast<Real*>(a.ypart) + (size_t)item * a.Rc * 2 * kBatch * B;
  for (int j = tid; j < a.Rc * 2 * kBatch * B; j += blockDim.x) {
    const int il = j / (2 * kBatch * B), k = j - il * 2 * kBatch * B;
    if constexpr (kFixed)
      yp[j] = (Real)((double)yacc[il * SY + k] * (double)ish + (double)yacc[il * SY + 2 * kBatch * B + k] * (double)isl);
    else
      yp[j] = (Real)yacc[il * SY + k];
  }
}

}  // namespace smh
