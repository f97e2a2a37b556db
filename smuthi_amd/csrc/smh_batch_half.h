// FP16-input lookup matvec for kBatch right-hand sides in slot order.
//
// The product of batch_slots_kernel (per pair: interpolate the lookup block,
// apply it to all right-hand sides, rotate, accumulate in fixed point) with
// FP16 operands on the FP16 MFMA (16x16x32, FP32 accumulation):
//
//   * the bin's interpolant is stored as FP16 polynomial coefficients scaled
//     by a per-bin power of two (hscale[bin]; |value| <= 2^12), evaluated by
//     Horner in packed FP16 (two slots per instruction);
//   * the input is stored in FP16 as x / xmax (re and im planes per 4 slots)
//     and rotated by e^{i m2 phi} in packed FP16;
//   * a chunk covers 16 slots (4 slot groups): lane (q, rr) holds the 4 slots
//     of group 4 c + q of chunk c as its K values, so its azimuthal phase
//     cos / sin(|m| phi) is a per-lane polynomial in cos phi (Chebyshev
//     coefficients of the lane's |m|), shared with the receiver rotation of
//     result tile c;
//   * the complex product runs inside the MFMA: one 16x16x32 MFMA per row
//     tile and chunk takes K = [Re L | Im L] of the chunk's slots against
//     [Re z; -Im z] (re columns) or [Im z; Re z] (im columns), so each lane
//     ends with its column of y = L z, and the receiver rotation
//     e^{-i m1 phi} needs one DPP exchange (lane ^ 1) per value;
//   * result rows and the partial-sum layout are those of batch_slots_kernel,
//     with the scales hscale[bin] * xmax folded into the rotation; the
//     receiver sums are those of batch_slots_kernel (one int64 fixed-point
//     word per result, fixed_scale64).
//
// The operands carry FP16 rounding (relative 2^-11 of the bin / input scale):
// a matvec is accurate to a few 1e-4 relative, which the FP64 iterative
// refinement of the mixed solve corrects.
//
// Layouts (S slots per particle, NC = S / 16 chunks):
//   coefficients [bins][T][MT][64 lanes][NC][4] FP16: lane (q, rr) of tile i
//         gets c_t(slot 16 (i mod MH) + rr, slot 16 c + 4 q + j), tiles
//         i < MH from Re L, the others from Im L;
//   x     [n_part][4 (q)][kBatch][NC][2 (re, im)][4 (j)] FP16 (b16_index): a
//         lane reads its chunk's 4 values of one plane as its first operand
//         (re for the re column, im for the im column) and the other plane
//         as the second, two 8-byte loads per chunk.
#pragma once

#include "smh_batch_slots.h"

namespace smh {

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));

// Coefficients of c^k in the Chebyshev polynomials T_m(c) (cos(m phi) =
// T_m(cos phi)) and U_{m-1}(c) (sin(m phi) = sin phi U_{m-1}(cos phi); zero
// for m = 0), m, k <= kChebMax, from T_{n+1} = 2 c T_n - T_{n-1} and the same
// recurrence for U.
constexpr int kChebMax = 4;
struct ChebTable {
  float t[kChebMax + 1][kChebMax + 1];
  float u[kChebMax + 1][kChebMax + 1];
  constexpr ChebTable() : t(), u() {
    t[0][0] = 1.f;
    t[1][1] = 1.f;
    u[1][0] = 1.f;  // U_0 = 1
    for (int m = 1; m < kChebMax; ++m)
      for (int k = 0; k <= kChebMax; ++k) {
        t[m + 1][k] = (k > 0 ? 2.f * t[m][k - 1] : 0.f) - t[m - 1][k];
        u[m + 1][k] = (k > 0 ? 2.f * u[m][k - 1] : 0.f) - u[m - 1][k];
      }
  }
};
__device__ constexpr ChebTable kCheb{};

// 16 waves per workgroup: the receiver accumulator takes the CU's LDS, so the
// workgroup is the CU's only one, and 4 waves per SIMD hide the gather latency
// (128 registers per lane).
constexpr int kHalfThreads = 1024;
// Parts of the result tiles (a divisor of the S / 16 tiles) such that one
// part's FP16 coefficients, T terms x its Re and Im tiles x S / 16 chunks x 2
// packed registers, take at most 64 registers per lane (0: none).
__host__ __device__ constexpr int half_tile_parts(int T, int B) {
  const int MH = (B + 15) / 16;
  for (int p = 1; p <= MH; ++p)
    if (MH % p == 0 && T * 2 * (MH / p) * MH * 2 <= 64) return p;
  return 0;
}

// VOL: volume lookups. The table holds the tap rows in this kernel's
// fragment order ([rows][MT][64 lanes][NC][4] FP16, one row per tap row of the
// volume lookup), each row divided by its own power of two hscale[row]; an
// entry's T taps of z slice a.vslice sit at rows en.x + vol_tap_offset, and a
// pair weights them with vol_tap_weight(w, w2) hscale[row] / (the largest of
// the entry's four scales), which then scales the result. The weights go to
// the input side (one MFMA per tap), so the taps are summed in FP32.
template <int LMAX, int MMAX, int T, bool VOL = false>
__global__ void __launch_bounds__(kHalfThreads) batch_half_kernel(RadialKArgs a, const float* __restrict__ hscale) {
  using Sh = SlotShape<LMAX, MMAX, T>;
  constexpr int S = Sh::S, MT = Sh::MT, MH = Sh::MH, SC = Sh::SC, LO = Sh::LO;
  constexpr int NC = S / 16;
  constexpr int KB = kBatchRhs;
  static_assert(KB * 2 == 16, "16 result columns = kBatch right-hand sides x (re, im)");
  static_assert(S % 16 == 0, "slots per particle must be a multiple of 16");
  static_assert(MMAX <= kChebMax, "Chebyshev phase coefficients cover |m| <= kChebMax");
  static_assert(MH == NC, "result tile c and input chunk c share the lane's slot group");
  // result tiles in TP parts of MHP (Re and Im rows of each), each a pass over
  // an entry's pairs, so that one part's FP16 coefficients take at most 64
  // registers per lane (one part at l_max <= 3)
  constexpr int TP = half_tile_parts(T, SvwfIdx<LMAX, MMAX>::B), MHP = MH / TP;
  static_assert(TP > 0 && MH % TP == 0, "tile parts");
  constexpr SlotOrder<LMAX, MMAX> so{};
  static_assert(so.ok, "slot order");

  // [Rc][LO] int64 fixed-point words, as batch_slots_kernel
  extern __shared__ __align__(16) unsigned long long half_acc[];
  unsigned long long* yacc = half_acc;
  int* next_entry = reinterpret_cast<int*>(yacc + (size_t)a.Rc * LO);
  int* n_adds = next_entry + 1;  // pairs accumulated per receiver
  const int item = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
  const float xmax = *a.xmax;
  float sh = 0.f;
  double ish = 0.0;
  fixed_scale64(a.item_lsum[item], xmax, a.item_en[item], sh, ish);

  const int tid = threadIdx.x, lane = tid & 63;
  const int2 it = a.items[item];
  const int e_end = item + 1 < a.n_items ? a.items[item + 1].y : a.n_entries;
  using Rec = typename std::conditional<VOL, PairRec2<float>, PairRec<float>>::type;
  const Rec* recs = reinterpret_cast<const Rec*>(a.recs);
  const _Float16* xg = reinterpret_cast<const _Float16*>(a.x);
  const uint4* tth = reinterpret_cast<const uint4*>(a.tt);
  for (int j = tid; j < a.Rc * LO; j += blockDim.x) yacc[j] = 0ull;
  for (int j = tid; j < a.Rc; j += blockDim.x) n_adds[j] = 0;
  if (tid == 0) *next_entry = it.y;
  __syncthreads();
  // every pair adds once to each word of its receiver: count them first
  for (int e = it.y + (tid >> 6); e < e_end; e += blockDim.x >> 6) {
    const int4 en = a.entries[e];
    for (int p = lane; p < en.z; p += 64) atomicAdd(n_adds + (recs[en.y + p].ij >> kPairShift), 1);
  }

  const int rr = lane & 15, q = lane >> 4;
  const int rhs = rr >> 1, zp = rr & 1;
  const unsigned zmask = zp ? 0u : 0x80000000u;  // z: sign of the b term for the re column
  // The rows of result tile i are slot group 4 i + q (receiver rotation, as
  // batch_slots_kernel) and the K values of chunk c are the 4 slots of group
  // 4 c + q (input rotation): tile c and chunk c share the lane's |m|. Per
  // chunk the lane evaluates cos(|m| phi) = T_|m|(cos phi) and sin(|m| phi) =
  // sin phi U_{|m|-1}(cos phi) as polynomials in cos phi whose coefficients
  // (Chebyshev T and U) it holds for its |m|. The volume form, short of
  // registers for these coefficients, selects among cos / sin(k phi) instead
  // (one-hot weights hot[c][k] = [|m| == k]).
  constexpr bool kChebPhase = !VOL;
  float tcf[NC][MMAX + 1], ucf[NC][MMAX + 1], hot[NC][MMAX + 1];
  unsigned smlo[MH], smhi[MH], zlo[NC], zhi[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int g0 = 4 * c;
    const int ra = q == 0 ? so.absm[g0] : (q == 1 ? so.absm[g0 + 1] : (q == 2 ? so.absm[g0 + 2] : so.absm[g0 + 3]));
    const int sg = q == 0 ? so.sgn[g0] : (q == 1 ? so.sgn[g0 + 1] : (q == 2 ? so.sgn[g0 + 2] : so.sgn[g0 + 3]));
#pragma unroll
    for (int k = 0; k <= MMAX; ++k) {
      if constexpr (kChebPhase) {
        tcf[c][k] = kCheb.t[ra][k];
        ucf[c][k] = kCheb.u[ra][k];
      } else {
        hot[c][k] = ra == k ? 1.f : 0.f;
      }
    }
    // (the im column takes -sin: y'_im = cos y_im - sin y_re)
    smlo[c] = (sg < 0 ? 0x80000000u : 0u) ^ (zp ? 0x80000000u : 0u);
    smhi[c] = (sg > 0 ? 0u : 0x80000000u) ^ (zp ? 0x80000000u : 0u);
    zlo[c] = (sg < 0 ? 0x80000000u : 0u) ^ zmask;  // slots j < 2
    zhi[c] = (sg > 0 ? 0u : 0x80000000u) ^ zmask;  // slots j >= 2 (m < 0 in mixed groups)
  }
  // the Lim half of K takes -Im z (re column) or Re z (im column):
  // (zp ? 1 : -1) (cz b - e a) with zb's cz and e, the sign as FP16 sign bits
  const unsigned hsign = zp ? 0u : 0x80008000u;
  const int colbase = (rhs * 2 + zp) * SC + 4 * q;
  // first / second operand plane of the lane: (re, im) for the re column, (im, re) for the im column
  const int64_t xlane = (int64_t)(q * KB + rhs) * NC * 8 + 4 * zp, xother = 4 - 8 * zp;

  typedef half2_t Coef[T][2 * MHP][NC][2];
  // polynomial coefficients of a bin for the tiles of part h: tv[t][ti][c] = 4
  // FP16 values (2 packed pairs) of tile h MHP + ti (ti < MHP, Re L) or
  // MH + h MHP + ti - MHP (Im L) (VOL: the values of the entry's tap rows rows[t])
  auto load_coef = [&](int bin, const int* rows, Coef& tv, int h) {
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
      for (int ti = 0; ti < 2 * MHP; ++ti) {
        const int i = ti < MHP ? h * MHP + ti : MH + h * MHP + ti - MHP;
        const size_t trow = VOL ? (size_t)rows[t] : (size_t)bin * T + t;
        if constexpr (NC % 2 == 0) {
          const uint4* row = tth + (trow * MT + i) * 64 * (NC / 2);
#pragma unroll
          for (int c2 = 0; c2 < NC / 2; ++c2) {
            const uint4 v = row[lane * (NC / 2) + c2];
            tv[t][ti][2 * c2][0] = __builtin_bit_cast(half2_t, v.x);
            tv[t][ti][2 * c2][1] = __builtin_bit_cast(half2_t, v.y);
            tv[t][ti][2 * c2 + 1][0] = __builtin_bit_cast(half2_t, v.z);
            tv[t][ti][2 * c2 + 1][1] = __builtin_bit_cast(half2_t, v.w);
          }
        } else {
          const uint2* row2 = reinterpret_cast<const uint2*>(tth) + (trow * MT + i) * 64 * NC;
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const uint2 v = row2[lane * NC + c];
            tv[t][ti][c][0] = __builtin_bit_cast(half2_t, v.x);
            tv[t][ti][c][1] = __builtin_bit_cast(half2_t, v.y);
          }
        }
      }
  };
  auto load_x = [&](const Rec& r, uint2 (&xs)[2 * NC]) {
    const _Float16* xp = xg + (int64_t)(r.ij & kPairMask) * S * KB * 2 + xlane;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      xs[2 * c] = *reinterpret_cast<const uint2*>(xp + 8 * c);
      xs[2 * c + 1] = *reinterpret_cast<const uint2*>(xp + 8 * c + xother);
    }
  };
  auto grab = [&]() {
    int e0 = 0;
    if (lane == 0) e0 = atomicAdd(next_entry, 1);
    return __builtin_amdgcn_readfirstlane(e0);
  };
  auto bin_of = [&](const int4& en) { return (en.x - a.row_o) / a.row_s; };
  uint2 xc[2 * NC];  // input of the next pair to run
  // the pairs of one entry (one bin) with its coefficients tv; rec_n is its
  // first pair, whose input is in xc
  auto run_entry = [&](const int4& en, const Coef& tv, Rec rec_n, const int* rows, int h) {
    float bscale, rel[T];
    if constexpr (VOL) {
      float hmax = 0.f;
#pragma unroll
      for (int t = 0; t < T; ++t) hmax = fmaxf(hmax, hscale[rows[t]]);
#pragma unroll
      for (int t = 0; t < T; ++t) rel[t] = hscale[rows[t]] / hmax;  // powers of two: exact
      bscale = hmax * xmax;
    } else {
      bscale = hscale[bin_of(en)] * xmax;
    }
    const int p_end = en.z;
    for (int p = 0; p < p_end; ++p) {
      const Rec rec = rec_n;
      rec_n = recs[en.y + (p + 1 < p_end ? p + 1 : p)];

      // cos / sin of |m| phi for the lane's group of each chunk (Horner in cos phi up to the
      // largest |m| of the chunk's four groups)
      const float c1 = rec.c, s1 = rec.s;
      float cph[NC], sph[NC];
      if constexpr (kChebPhase) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const int dg = max(max(so.absm[4 * c], so.absm[4 * c + 1]), max(so.absm[4 * c + 2], so.absm[4 * c + 3]));
          float v = tcf[c][dg];
#pragma unroll
          for (int k = MMAX - 1; k >= 0; --k)
            if (k < dg) v = fmaf(v, c1, tcf[c][k]);
          cph[c] = v;
          float u = dg > 0 ? ucf[c][dg - 1] : 0.f;
#pragma unroll
          for (int k = MMAX - 2; k >= 0; --k)
            if (k < dg - 1) u = fmaf(u, c1, ucf[c][k]);
          sph[c] = u * s1;
        }
      } else {
        float cs[MMAX + 1], sn[MMAX + 1];
        phase_powers<MMAX>(c1, s1, cs, sn);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          float cz = 0.f, sz = 0.f;
#pragma unroll
          for (int am = 0; am <= MMAX; ++am) {
            // |m| values outside the chunk's four groups drop out at compile time
            if (so.absm[4 * c] != am && so.absm[4 * c + 1] != am && so.absm[4 * c + 2] != am &&
                so.absm[4 * c + 3] != am)
              continue;
            cz = fmaf(hot[c][am], cs[am], cz);
            if (am > 0) sz = fmaf(hot[c][am], sn[am], sz);
          }
          cph[c] = cz;
          sph[c] = sz;
        }
      }

      // B operands per chunk, FP16: zb = the lane's column of z = e^{i m2 phi} x
      // (cz a + e b), zb2 = the column of the other part with the sign of the
      // complex product (-Im z for the re column, Re z for the im column)
      half4_t zb[NC], zb2[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const float cz = cph[c], sz = sph[c];
        const _Float16 czh = (_Float16)cz, eloh = (_Float16)xor_sign(sz, zlo[c]), ehih = (_Float16)xor_sign(sz, zhi[c]);
        const half2_t cz2 = {czh, czh}, elo2 = {eloh, eloh}, ehi2 = {ehih, ehih};
        const half2_t a01 = __builtin_bit_cast(half2_t, xc[2 * c].x), a23 = __builtin_bit_cast(half2_t, xc[2 * c].y);
        const half2_t b01 = __builtin_bit_cast(half2_t, xc[2 * c + 1].x);
        const half2_t b23 = __builtin_bit_cast(half2_t, xc[2 * c + 1].y);
        const half2_t z01 = a01 * cz2 + b01 * elo2, z23 = a23 * cz2 + b23 * ehi2;
        zb[c] = half4_t{z01[0], z01[1], z23[0], z23[1]};
        const half2_t y01 = b01 * cz2 - a01 * elo2, y23 = b23 * cz2 - a23 * ehi2;
        zb2[c] = __builtin_bit_cast(half4_t, __builtin_bit_cast(u32x2_t, half4_t{y01[0], y01[1], y23[0], y23[1]}) ^ hsign);
      }
      load_x(rec_n, xc);

      const _Float16 wh = (_Float16)rec.w;
      const half2_t w2 = {wh, wh};
      half4_t wt[T];  // VOL: the pair's tap weights relative to the entry's scale
      if constexpr (VOL) {
#pragma unroll
        for (int t = 0; t < T; ++t) {
          const _Float16 h = (_Float16)(vol_tap_weight<float>(a.vmode, a.vslice, t, rec.w, rec.w2) * rel[t]);
          wt[t] = half4_t{h, h, h, h};
        }
      }
      typename Mfma<float>::v4 acc[MHP];
#pragma unroll
      for (int ti = 0; ti < MHP; ++ti) acc[ti] = typename Mfma<float>::v4{0, 0, 0, 0};
      // interpolated coefficients of part tile i, chunk c (Horner in packed FP16)
      auto coef = [&](int i, int c) {
        half2_t v0 = tv[T - 1][i][c][0], v1 = tv[T - 1][i][c][1];
#pragma unroll
        for (int t = T - 2; t >= 0; --t) {
          v0 = v0 * w2 + tv[t][i][c][0];
          v1 = v1 * w2 + tv[t][i][c][1];
        }
        return half4_t{v0[0], v0[1], v1[0], v1[1]};
      };
      // one 16x16x32 MFMA (the full-rate FP16 form) per row tile and chunk:
      // the lane's 8 K values are its 4 slots of chunk c from Re L (times zb)
      // and the same 4 from Im L (times zb2), so the column sums are the
      // complex product's re or im part directly. VOL: one MFMA per tap, the
      // tap's weight applied to the input side (zb, zb2) and the taps summed
      // by the MFMA in FP32, so the weighted sum of the taps (Catmull-Rom
      // weights of both signs) is not rounded to FP16
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if constexpr (VOL) {
#pragma unroll
          for (int t = 0; t < T; ++t) {
            const half8_t bv = __builtin_shufflevector(zb[c] * wt[t], zb2[c] * wt[t], 0, 1, 2, 3, 4, 5, 6, 7);
#pragma unroll
            for (int ti = 0; ti < MHP; ++ti) {
              const half4_t re = __builtin_shufflevector(tv[t][ti][c][0], tv[t][ti][c][1], 0, 1, 2, 3);
              const half4_t im = __builtin_shufflevector(tv[t][ti + MHP][c][0], tv[t][ti + MHP][c][1], 0, 1, 2, 3);
              const half8_t av = __builtin_shufflevector(re, im, 0, 1, 2, 3, 4, 5, 6, 7);
              acc[ti] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, acc[ti], 0, 0, 0);
            }
          }
        } else {
          const half8_t bv = __builtin_shufflevector(zb[c], zb2[c], 0, 1, 2, 3, 4, 5, 6, 7);
#pragma unroll
          for (int ti = 0; ti < MHP; ++ti) {
            const half8_t av = __builtin_shufflevector(coef(ti, c), coef(ti + MHP, c), 0, 1, 2, 3, 4, 5, 6, 7);
            acc[ti] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, acc[ti], 0, 0, 0);
          }
        }
      }

      // y (rhs, re|im) of rows 16 i + 4 q + v, rotated by e^{-i m1 phi}, scaled by S_hi and the operand scales
      const int il = rec.ij >> kPairShift;
      unsigned long long* yrow = yacc + (size_t)il * LO + colbase;
      const float scl = sh * bscale;
#pragma unroll
      for (int ti = 0; ti < MHP; ++ti) {
        const int i = h * MHP + ti;
        const float er = cph[i] * scl, ei = sph[i] * scl;
        const float elo = xor_sign(ei, smlo[i]), ehi = xor_sign(ei, smhi[i]);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float eps = v < 2 ? elo : ehi;
          // y' = e^{-i m1 phi} y: er y_own + eps y_other, y_other from lane ^ 1
          // (the exchange folds into the multiply: v_mul_f32_dpp)
          const float val = fmaf(er, acc[ti][v], eps * lane_xor1(acc[ti][v]));
          atomicAdd(yrow + 16 * i + v, fixed_bits64(val));
        }
      }
    }
  };

  Coef tv;
  for (;;) {
    const int e = grab();
    if (e >= e_end) break;
    const int4 en = a.entries[e];
    int rows[T];
#pragma unroll
    for (int t = 0; t < T; ++t) rows[t] = VOL ? en.x + vol_tap_offset(a.vmode, a.vslice, t, a.len_z) : 0;
#pragma unroll
    for (int h = 0; h < TP; ++h) {
      load_coef(bin_of(en), rows, tv, h);
      const Rec rec0 = recs[en.y];
      load_x(rec0, xc);
      run_entry(en, tv, rec0, rows, h);
    }
  }
  __syncthreads();
  float* yp = reinterpret_cast<float*>(a.ypart) + (size_t)item * a.Rc * 2 * KB * S;
  for (int j = tid; j < a.Rc * 2 * KB * S; j += blockDim.x) {
    const int il = j / (2 * KB * S), k = j - il * 2 * KB * S;
    const int c = k / S, sl2 = k - c * S;
    const int w0 = il * LO + c * SC + sl2;
    const long long sum = (long long)(yacc[w0] - (unsigned long long)n_adds[il] * kMagic64Bits);  // modulo 2^64
    yp[j] = (float)((double)sum * ish);
  }
}

}  // namespace smh
