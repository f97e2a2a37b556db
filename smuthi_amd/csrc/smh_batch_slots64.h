// FP64 lookup matvec for kBatch right-hand sides in slot order.
//
// The FP64 counterpart of batch_slots_kernel: per pair the bin's interpolant
// (FP64 polynomial coefficients, Horner in the bin fraction) is applied to
// all right-hand sides with the FP64 MFMA 16x16x4 (columns = (rhs, re|im)),
// the multipoles permuted into the slot order of SlotOrder so that every
// azimuthal phase is known at compile time:
//
//   * K-step s covers slots 4 s .. 4 s + 3, one slot group, so the input
//     rotation e^{i m2 phi} x is z = cos(|m| phi) a + e sin(|m| phi) b with a
//     compile-time |m|; a lane loads its operands (a, b) as (Re x, Im x) on
//     the re column and (Im x, Re x) on the im column (two 8-byte loads per
//     slot), so no per-lane selects;
//   * the FP64 MFMA leaves row 4 v + q of tile i in register v of lane
//     (q, rr): slot 16 i + 4 v + q, i.e. slot group 4 i + v, so the receiver
//     rotation e^{-i m1 phi} of each register is compile-time as well (only
//     the sign of m1 in mixed groups depends on q);
//   * the re/im columns of a right-hand side are combined with one DPP
//     exchange (lane ^ 1) and each value is added to the receiver's LDS
//     accumulator with one FP64 LDS atomic;
//   * the K-steps run in SlotShape::P64 parts (two at l_max = m_max = 3, six
//     at 4), each a pass over the entry's pairs with its own epilogue, so one
//     part's coefficients (at most 128 registers) leave room for two waves
//     per SIMD.
//
// Layouts (S = slots per particle, NS = S / 4 K-steps):
//   coefficients [bins][T][MT][NS/2][64 lanes][2] FP64: lane (q, rr) of tile i
//         gets c_t(slot 16 (i mod MH) + rr, slot 4 s + q), s = 2 s2 + e, rows
//         of tiles i < MH from Re L, the others from Im L;
//   x     [n_part][4 (q)][kBatch][NS] double2 (b32_index);
//   LDS   [Rc][16 columns][S + 1] double;
//   ypart [n_items][Rc][16 columns][S] double.
#pragma once

#include "smh_batch_slots.h"

namespace smh {

__device__ inline double xor_sign64(double v, unsigned m) {
  const long long b = __double_as_longlong(v) ^ ((long long)m << 32);
  return __longlong_as_double(b);
}
// value of lane l ^ 1
__device__ inline double lane_xor1_64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// two waves per SIMD; the K-steps in SlotShape::P64 parts per pair (two at
// l_max = m_max = 3)
constexpr int kSlots64Threads = 512;

// cos(k phi), sin(k phi) for k <= K (angle addition)
template <int K>
__device__ inline void phase_powers64(double c1, double s1, double (&cs)[K + 1], double (&sn)[K + 1]) {
  cs[0] = 1.0;
  sn[0] = 0.0;
  if constexpr (K >= 1) {
    cs[1] = c1;
    sn[1] = s1;
  }
  if constexpr (K >= 2) {
    cs[2] = c1 * c1 - s1 * s1;
    sn[2] = 2.0 * c1 * s1;
  }
#pragma unroll
  for (int k = 2; k < K; ++k) {
    cs[k + 1] = cs[k] * c1 - sn[k] * s1;
    sn[k + 1] = sn[k] * c1 + cs[k] * s1;
  }
}

// VOL: volume lookups, as batch_slots_kernel's VOL form: the table holds FP64
// tap rows in this kernel's fragment order ([rows][MT][NS/2][64 lanes][2]),
// an entry's T taps of z slice a.vslice at rows en.x + vol_tap_offset, and a
// pair weights them with vol_tap_weight(w, w2). TAP32 (VOL, tables of plain
// FP32 values: no spline moments): the taps come from the FP32 kernel's rows
// ([rows][MT][NS/4][64 lanes][4], the same values exactly) and are widened in
// registers, half the tap-row bytes of the FP64 rows.
template <int LMAX, int MMAX, int T, bool VOL = false, bool TAP32 = false>
__global__ void __launch_bounds__(kSlots64Threads) batch_slots64_kernel(RadialKArgs a) {
  using Sh = SlotShape<LMAX, MMAX, T>;
  constexpr int S = Sh::S, NS = Sh::NS, MT = Sh::MT, MH = Sh::MH, SC = Sh::SC, LO = Sh::LO;
  constexpr int KB = kBatchRhs;
  static_assert(KB * 2 == 16, "16 result columns = kBatch right-hand sides x (re, im)");
    constexpr int NP = Sh::P64;  // parts of the K-steps
  static_assert(NP > 0, "coefficients exceed the register budget");
  constexpr int NSH = NS / NP;  // K-steps per part (even: coefficient pairs per lane)
  static_assert(MMAX <= 4, "phase selection covers |m| <= 4");
  static_assert(VOL || !TAP32, "FP32 tap rows are a volume form");
  constexpr SlotOrder<LMAX, MMAX> so{};
  static_assert(so.ok, "slot order");
  using V4 = typename Mfma<double>::v4;

  extern __shared__ __align__(16) double slot64_acc[];
  double* yacc = slot64_acc;  // [Rc][LO]
  int* next_entry = reinterpret_cast<int*>(yacc + (size_t)a.Rc * LO);
  const int item = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;

  const int tid = threadIdx.x, lane = tid & 63;
  const int2 it = a.items[item];
  const int e_end = item + 1 < a.n_items ? a.items[item + 1].y : a.n_entries;
  using Rec = typename std::conditional<VOL, PairRec2<double>, PairRec<double>>::type;
  const Rec* recs = reinterpret_cast<const Rec*>(a.recs);
  const double* xg = reinterpret_cast<const double*>(a.x);
  const double2* tt2 = reinterpret_cast<const double2*>(a.tt);
  for (int j = tid; j < a.Rc * LO; j += blockDim.x) yacc[j] = 0.0;
  if (tid == 0) *next_entry = it.y;
  __syncthreads();

  const int rr = lane & 15, q = lane >> 4;
  const int rhs = rr >> 1, zp = rr & 1;
  // input rotation: e = (zp ? +1 : -1) sign(m2) sin(|m2| phi) multiplies the
  // second operand; sign(m2) is uniform per K-step or + / - for q < 2 / q >= 2
  const unsigned zmask = zp ? 0u : 0x80000000u;
  const unsigned qmask = q >= 2 ? 0x80000000u : 0u;
  const double sgnz = zp ? 1.0 : -1.0;
  const int colbase = (rhs * 2 + zp) * SC + q;
  // the lane's first operand is Re x (re column) or Im x (im column)
  const int64_t xlane = (int64_t)(q * KB + rhs) * NS * 2 + zp, xother = 1 - 2 * zp;

  for (;;) {
    int e0 = 0;
    if (lane == 0) e0 = atomicAdd(next_entry, 1);
    const int e = __builtin_amdgcn_readfirstlane(e0);
    if (e >= e_end) break;
    const int4 en = a.entries[e];
    const int bin = (en.x - a.row_o) / a.row_s;
    const int p_end = en.z;
    // the K-steps in NP parts of NSH, each with its own pass over the
    // entry's pairs and its own epilogue (coefficients of one part in
    // registers: two waves per SIMD)
#pragma unroll
    for (int h = 0; h < NP; ++h) {
      double tv[T][MT][NSH];
#pragma unroll
      for (int t = 0; t < T; ++t)
#pragma unroll
        for (int i = 0; i < MT; ++i) {
          const size_t trow =
              VOL ? (size_t)(en.x + vol_tap_offset(a.vmode, a.vslice, t, a.len_z)) * MT + i : ((size_t)bin * T + t) * MT + i;
          if constexpr (TAP32) {
            // K-step k of the lane: float k % 4 of float4 (k / 4) * 64 + lane of the row
            const float* rowf = reinterpret_cast<const float*>(a.tt) + trow * (NS / 4) * 64 * 4;
#pragma unroll
            for (int sl = 0; sl < NSH; ++sl) {
              const int k = h * NSH + sl;
              tv[t][i][sl] = (double)rowf[((k >> 2) * 64 + lane) * 4 + (k & 3)];
            }
          } else {
            const double2* row = tt2 + trow * (NS / 2) * 64;
#pragma unroll
            for (int s2 = 0; s2 < NSH / 2; ++s2) {
              const double2 v = row[(h * (NSH / 2) + s2) * 64 + lane];
              tv[t][i][2 * s2] = v.x;
              tv[t][i][2 * s2 + 1] = v.y;
            }
          }
        }
      auto load_x = [&](const Rec& r, double (&xs)[2 * NSH]) {
        const double* xp = xg + (int64_t)(r.ij & kPairMask) * S * KB * 2 + xlane + 2 * h * NSH;
#pragma unroll
        for (int s = 0; s < NSH; ++s) {
          xs[2 * s] = xp[2 * s];
          xs[2 * s + 1] = xp[2 * s + xother];
        }
      };
      Rec rec_n = recs[en.y];
      double xc[2 * NSH];
      load_x(rec_n, xc);
      for (int p = 0; p < p_end; ++p) {
        const Rec rec = rec_n;
        rec_n = recs[en.y + (p + 1 < p_end ? p + 1 : p)];

        const double w = rec.w;
        double wt[T];  // VOL: the pair's tap weights
        if constexpr (VOL) {
#pragma unroll
          for (int t = 0; t < T; ++t) wt[t] = vol_tap_weight<double>(a.vmode, a.vslice, t, rec.w, rec.w2);
        }
        double cs[MMAX + 1], sn[MMAX + 1];
        phase_powers64<MMAX>(rec.c, rec.s, cs, sn);

        // B operand per K-step: (e^{i m2 phi} x[slot 4 s + q][rhs]).re or .im
        double z[NSH];
#pragma unroll
        for (int sl = 0; sl < NSH; ++sl) {
          const int s = h * NSH + sl;
          const int am = so.absm[s];
          const double av = xc[2 * sl], bv = xc[2 * sl + 1];
          if (am == 0) {
            z[sl] = av;
          } else {
            const int sg = so.sgn[s];
            const double ez = xor_sign64(sn[am], zmask);
            const double ee = sg > 0 ? ez : (sg < 0 ? -ez : xor_sign64(ez, qmask));
            z[sl] = fma(cs[am], av, ee * bv);
          }
        }
        load_x(rec_n, xc);

        V4 acc[MT];
#pragma unroll
        for (int i = 0; i < MT; ++i) acc[i] = V4{0, 0, 0, 0};
#pragma unroll
        for (int sl = 0; sl < NSH; ++sl) {
#pragma unroll
          for (int i = 0; i < MT; ++i) {
            double v;
            if constexpr (VOL) {
              v = wt[0] * tv[0][i][sl];
#pragma unroll
              for (int t = 1; t < T; ++t) v = fma(wt[t], tv[t][i][sl], v);
            } else {
              v = tv[T - 1][i][sl];
#pragma unroll
              for (int t = T - 2; t >= 0; --t) v = fma(v, w, tv[t][i][sl]);
            }
            acc[i] = Mfma<double>::op(v, z[sl], acc[i]);
          }
        }

        // register v of tile i: slot 16 i + 4 v + q (group 4 i + v), rotated by e^{-i m1 phi}
        double* yrow = yacc + (size_t)(rec.ij >> kPairShift) * LO + colbase;
#pragma unroll
        for (int i = 0; i < MH; ++i) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int g = 4 * i + v;
            const int am = so.absm[g], sg = so.sgn[g];
            const double er = cs[am];
            const double eps = am == 0 ? 0.0 : (sg > 0 ? sn[am] : (sg < 0 ? -sn[am] : xor_sign64(sn[am], qmask)));
            const double al = acc[i][v], be = acc[i + MH][v];
            const double own = fma(al, er, be * eps);
            const double snd = fma(be, er, -(al * eps));
            const double val = fma(sgnz, lane_xor1_64(snd), own);
            if (so.n_of[16 * i + 4 * v] >= 0 || so.n_of[16 * i + 4 * v + 1] >= 0 ||
                so.n_of[16 * i + 4 * v + 2] >= 0 || so.n_of[16 * i + 4 * v + 3] >= 0)
              atomicAdd(yrow + 16 * i + 4 * v, val);
          }
        }
      }
    }
  }
  __syncthreads();
  // partial of this item: [Rc][16 columns][S]
  double* yp = reinterpret_cast<double*>(a.ypart) + (size_t)item * a.Rc * 2 * KB * S;
  for (int j = tid; j < a.Rc * 2 * KB * S; j += blockDim.x) {
    const int il = j / (2 * KB * S), k = j - il * 2 * KB * S;
    const int c = k / S, sl2 = k - c * S;
    yp[j] = yacc[il * LO + c * SC + sl2];
  }
}

}  // namespace smh
