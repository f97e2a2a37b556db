// FP32 lookup matvec for kBatch right-hand sides in "slot" order.
//
// Per pair: interpolate the lookup block, then one MFMA 16x16x4 tile per
// K-step applies it to all right-hand sides (columns = (rhs, re|im)), with
// the multipoles permuted inside the kernel so that the azimuthal phases
// cost no per-lane selection:
//
//   slot order: multipoles grouped by |m| (m = 0 first); inside a group the
//   m > 0 ones before the m < 0 ones; every group padded to a multiple of 4
//   slots (padding slots have zero taps and zero inputs).
//
// An MFMA K-step covers 4 slots, so its |m| is known at compile time and
// e^{i m2 phi} x reduces to z = cos(|m| phi) a + sin(|m| phi) b with lane
// operands (a, b) = (Re x, Im x) or (Im x, Re x) and a sign that is uniform
// per step or + for slots 0, 1 and - for slots 2, 3 of the step. The 4 result
// rows a lane holds (4q + v) are one slot group as well, so the receiver
// rotation e^{-i m1 phi} needs one phase per lane and tile. The re/im
// combination of a (rhs, re) / (rhs, im) lane pair is one DPP exchange.
//
// Layouts (S = slots per particle = Bh, NS = S / 4 K-steps):
//   coefficients [bins][T][MT][NS/4][64 lanes][4]: the interpolant on a bin
//         as a polynomial in the bin fraction w (evaluated by Horner, T - 1
//         FMAs per value); lane (q, rr) of tile i gets
//         c_k(slot 16 (i mod MH) + rr, slot 4 s + q), s = 4 s4 + j, rows of
//         tiles i < MH from Re L, the others from Im L (one dwordx4 per s4);
//   x     [n_part][4 (q)][kBatch][NS] float2: the NS inputs a lane needs are
//         contiguous (NS/2 dwordx4 loads per pair);
//   LDS   [Rc][16 columns][S + 1] int64 (fixed point, fixed_scale64);
//   ypart [n_items][Rc][16 columns][S] float.
#pragma once

#include "smh_op.h"
#include "smh_radial_kernel.h"

namespace smh {

template <int L, int M>
struct SlotOrder {
  static constexpr int B = SvwfIdx<L, M>::B, S = SvwfIdx<L, M>::Bh, NS = S / 4;
  int n_of[S];     // multipole of slot k (-1: padding)
  int slot_of[B];  // slot of multipole n
  int absm[NS];    // |m| of K-step / row group g (slots 4g..4g+3)
  int sgn[NS];     // sign of m in group g: +1, -1, or 0 (+ for slots 4g, 4g+1, - for 4g+2, 4g+3)
  bool ok;
  constexpr SlotOrder() : n_of(), slot_of(), absm(), sgn(), ok(true) {
    int sg[S] = {};
    int k = 0;
    for (int a = 0; a <= M; ++a) {
      const int k0 = k;
      for (int pass = 0; pass < (a ? 2 : 1); ++pass)
        for (int n = 0; n < B; ++n) {
          const int m = SvwfIdx<L, M>::m_of(n);
          if (m == (pass ? -a : a) && k < S) {
            n_of[k] = n;
            slot_of[n] = k;
            sg[k] = pass ? -1 : 1;
            ++k;
          }
        }
      while (k % 4 && k < S) {
        n_of[k] = -1;
        sg[k] = 0;
        ++k;
      }
      for (int g = k0 / 4; g < k / 4; ++g) absm[g] = a;
    }
    if (k > S) ok = false;
    while (k < S) {
      n_of[k] = -1;
      sg[k] = 0;
      ++k;
    }
    for (int g = 0; g < NS; ++g) {
      bool pos = false, neg = false;
      for (int j = 0; j < 4; ++j) {
        pos = pos || sg[4 * g + j] > 0;
        neg = neg || sg[4 * g + j] < 0;
      }
      if (pos && neg) {
        sgn[g] = 0;
        // mixed groups must be (+, +, -, -)
        if (sg[4 * g] < 0 || sg[4 * g + 1] < 0 || sg[4 * g + 2] > 0 || sg[4 * g + 3] > 0) ok = false;
      } else {
        sgn[g] = neg ? -1 : 1;
      }
    }
    int cnt = 0;
    for (int s = 0; s < S; ++s) cnt += n_of[s] >= 0;
    if (cnt != B) ok = false;
  }
};

// Host copy of the slot order: slot_of[B] then n_of[S].
template <int L, int M>
void slot_tables(int* out) {
  constexpr SlotOrder<L, M> so{};
  static_assert(so.ok, "slot order");
  for (int n = 0; n < so.B; ++n) out[n] = so.slot_of[n];
  for (int k = 0; k < so.S; ++k) out[so.B + k] = so.n_of[k];
}

// Shape of the slot kernels. fits: the FP32 kernel's tap fragments of all
// K-steps fit one set of at most 128 registers per lane. Larger blocks split
// the K-steps into parts, each a pass over an entry's pairs with its own
// epilogue: P32 parts of whole float4 groups (4 K-steps) within 128 FP32
// registers, P64 parts of an even number of K-steps within 64 FP64 values
// (0: no such split).
template <int L, int M, int T>
struct SlotShape {
  static constexpr int S = SvwfIdx<L, M>::Bh, NS = S / 4, MT = 2 * S / 16, MH = MT / 2;
  static constexpr int SC = S + 1;          // LDS column stride (words)
  static constexpr int LO = 2 * kBatchRhs * SC;  // int64 accumulator words per receiver
  static constexpr int SY = 2 * LO;         // LDS int32 words per receiver
  static constexpr int FRB = MT * NS * 64;  // floats per tap row
  static constexpr bool fits = T * MT * NS <= 128;
  static constexpr int parts(int budget, int step) {
    for (int p = 1; p <= NS; ++p)
      if (NS % p == 0 && (NS / p) % step == 0 && T * MT * (NS / p) <= budget) return p;
    return 0;
  }
  static constexpr int P32 = parts(128, 4), P64 = parts(64, 2);
};

// cos(k phi), sin(k phi) for k <= K from cos phi, sin phi (angle addition)
template <int K>
__device__ inline void phase_powers(float c1, float s1, float (&cs)[K + 1], float (&sn)[K + 1]) {
  cs[0] = 1.f;
  sn[0] = 0.f;
  if constexpr (K >= 1) {
    cs[1] = c1;
    sn[1] = s1;
  }
  if constexpr (K >= 2) {
    cs[2] = c1 * c1 - s1 * s1;
    sn[2] = 2.f * c1 * s1;
  }
#pragma unroll
  for (int k = 2; k < K; ++k) {
    cs[k + 1] = cs[k] * c1 - sn[k] * s1;
    sn[k + 1] = sn[k] * c1 + cs[k] * s1;
  }
}

// Receiver sums of the slot kernels: one int64 fixed-point word per result.
// A value v of a work item (|v| <= V = 8 lsum xmax) is added as
// rint(v sh) through kMagic64 (1.5 * 2^52: the bits of u + kMagic64 are
// kMagic64Bits + rint(u) for |u| <= 2^51), sh = 2^k with |v| sh <=
// 2^min(51, 62 - eN), so 2^eN additions stay within 2^62; the words
// accumulate those bits and the finish subtracts kMagic64Bits once per
// accumulated pair (modulo 2^64). Resolution V' 2^-min(51, 62 - eN), V' the
// power of two above V: far below the FP32 rounding of the values, and the
// sums are order independent.
constexpr double kMagic64 = 6755399441055744.0;
constexpr unsigned long long kMagic64Bits = 0x4338000000000000ull;
__device__ inline void fixed_scale64(float lsum, float xmax, int eN, float& sh, double& ish) {
  const double V = 8.0 * (double)lsum * (double)xmax;
  int k = 0;
  if (V > 0.0 && V < 1e300) k = min(51, 62 - eN) - ((int)floor(log2(V)) + 1);
  k = max(-126, min(k, 126));  // the float range of sh
  sh = ldexpf(1.f, k);
  ish = ldexp(1.0, -k);
}
__device__ inline unsigned long long fixed_bits64(float v) {
  return (unsigned long long)__double_as_longlong((double)v + kMagic64);
}

__device__ inline float xor_sign(float v, unsigned m) { return __uint_as_float(__float_as_uint(v) ^ m); }
// v + c * (value of lane l ^ 1)
__device__ inline float fma_xor1(float v, float c, float other) {
  return fmaf(c, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(other), 0xB1, 0xF, 0xF, false)), v);
}

// VOL: volume lookups. The table holds tap rows in the same slot fragment
// order ([rows][MT][NS/4][64 lanes][4], one row per tap row of the volume
// lookup: node * vq + moment kind); an entry's T taps for z slice a.vslice sit
// at rows en.x + vol_tap_offset(a.vmode, a.vslice, t, a.len_z), and a pair
// weights them with vol_tap_weight(w, w2) instead of Horner in w.
template <int LMAX, int MMAX, int T, bool VOL = false>
__global__ void __launch_bounds__(512) batch_slots_kernel(RadialKArgs a) {
  using Sh = SlotShape<LMAX, MMAX, T>;
  constexpr int S = Sh::S, NS = Sh::NS, MT = Sh::MT, MH = Sh::MH, SC = Sh::SC, LO = Sh::LO;
  constexpr int KB = kBatchRhs;
  static_assert(KB * 2 == 16, "16 result columns = kBatch right-hand sides x (re, im)");
  // K-steps in P parts (one register set of tap fragments per part)
  constexpr int P = Sh::P32, NSP = NS / (P > 0 ? P : 1);
  static_assert(P > 0, "tap fragments exceed the register budget");
  static_assert(MMAX <= 4, "phase selection covers |m| <= 4");
  constexpr SlotOrder<LMAX, MMAX> so{};
  static_assert(so.ok, "slot order");

  extern __shared__ __align__(16) unsigned long long slot_acc[];
  unsigned long long* yacc = slot_acc;  // [Rc][LO] int64 words
  int* next_entry = reinterpret_cast<int*>(yacc + (size_t)a.Rc * LO);
  int* n_adds = next_entry + 1;  // pairs accumulated per receiver
  // work item of this block (launch order maps blocks to items, see RadialKArgs::order)
  const int item = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
  float sh = 0.f;
  double ish = 0.0;
  // each pair adds once per part to each word of its receiver
  constexpr int PB = P <= 1 ? 0 : (P <= 2 ? 1 : (P <= 4 ? 2 : (P <= 8 ? 3 : 4)));  // ceil(log2 P)
  fixed_scale64(a.item_lsum[item], *a.xmax, a.item_en[item] + PB, sh, ish);

  const int tid = threadIdx.x, lane = tid & 63;
  const int2 it = a.items[item];
  const int e_end = item + 1 < a.n_items ? a.items[item + 1].y : a.n_entries;
  using Rec = typename std::conditional<VOL, PairRec2<float>, PairRec<float>>::type;
  const Rec* recs = reinterpret_cast<const Rec*>(a.recs);
  const float2* xg = reinterpret_cast<const float2*>(a.x);
  const float4* tt4 = reinterpret_cast<const float4*>(a.tt);
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
  const unsigned zmask = zp ? 0u : 0x80000000u;      // z: sign of the b term for the re column
  const unsigned qmask = q >= 2 ? 0x80000000u : 0u;  // mixed-sign steps: slots 2, 3 have m < 0
  const float sgnz = zp ? 1.f : -1.f;
  // result rows of tile i are slot group 4 i + q: its |m| and the sign masks of rows v < 2 / v >= 2
  // (as one-hot weights hot[i][a] = [|m| == a], so the per-pair phase of a lane is
  // a short FMA chain over the |m| values that occur in the tile, no branches)
  float hot[MH][MMAX + 1];
  unsigned smlo[MH], smhi[MH];
#pragma unroll
  for (int i = 0; i < MH; ++i) {
    const int g0 = 4 * i;
    const int ra = q == 0 ? so.absm[g0] : (q == 1 ? so.absm[g0 + 1] : (q == 2 ? so.absm[g0 + 2] : so.absm[g0 + 3]));
    const int sg = q == 0 ? so.sgn[g0] : (q == 1 ? so.sgn[g0 + 1] : (q == 2 ? so.sgn[g0 + 2] : so.sgn[g0 + 3]));
#pragma unroll
    for (int am = 0; am <= MMAX; ++am) hot[i][am] = ra == am ? 1.f : 0.f;
    smlo[i] = sg < 0 ? 0x80000000u : 0u;
    smhi[i] = sg > 0 ? 0u : 0x80000000u;
  }
  const int colbase = (rhs * 2 + zp) * SC + 4 * q;
  const int64_t xlane = (int64_t)(q * KB + rhs) * NS;

  for (;;) {
    int e0 = 0;
    if (lane == 0) e0 = atomicAdd(next_entry, 1);
    const int e = __builtin_amdgcn_readfirstlane(e0);
    if (e >= e_end) break;
    const int4 en = a.entries[e];
    const int bin = (en.x - a.row_o) / a.row_s;
    const int p_end = en.z;
#pragma unroll
    for (int h = 0; h < P; ++h) {
      // K-steps s0 .. s0 + NSP - 1 of this part
      const int s0 = h * NSP;
      // polynomial coefficients tv[t][i][s] = c_t(tile i row rr, slot 4 (s0 + s) + q) of the entry's bin
      // (VOL: the values of the entry's tap t)
      float tv[T][MT][NSP];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const size_t trow =
            VOL ? (size_t)(en.x + vol_tap_offset(a.vmode, a.vslice, t, a.len_z)) : (size_t)bin * T + t;
        const float4* row = tt4 + trow * (Sh::FRB / 4);
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
          for (int s4 = 0; s4 < NSP / 4; ++s4) {
            const float4 v = row[(i * (NS / 4) + s0 / 4 + s4) * 64 + lane];
            tv[t][i][4 * s4 + 0] = v.x;
            tv[t][i][4 * s4 + 1] = v.y;
            tv[t][i][4 * s4 + 2] = v.z;
            tv[t][i][4 * s4 + 3] = v.w;
          }
      }
      auto load_x = [&](const Rec& r, float4 (&xs)[NSP / 2]) {
        const float4* xp =
            reinterpret_cast<const float4*>(xg + (int64_t)(r.ij & kPairMask) * S * KB + xlane) + s0 / 2;
#pragma unroll
        for (int k = 0; k < NSP / 2; ++k) xs[k] = xp[k];
      };
      Rec rec_n = recs[en.y];
      float4 xc[NSP / 2];
      load_x(rec_n, xc);
      for (int p = 0; p < p_end; ++p) {
        const Rec rec = rec_n;
        rec_n = recs[en.y + (p + 1 < p_end ? p + 1 : p)];

        const float w = rec.w;
        float wt[T];  // VOL: the pair's tap weights
        if constexpr (VOL) {
#pragma unroll
          for (int t = 0; t < T; ++t) wt[t] = vol_tap_weight<float>(a.vmode, a.vslice, t, rec.w, rec.w2);
        }
        // cos / sin of k phi, k = 1..MMAX
        float cs[MMAX + 1], sn[MMAX + 1];
        phase_powers<MMAX>(rec.c, rec.s, cs, sn);
        float ez[MMAX + 1];  // sin(k phi) with the re-column sign
        ez[0] = 0.f;
#pragma unroll
        for (int k = 1; k <= MMAX; ++k) ez[k] = xor_sign(sn[k], zmask);

        // B operand per K-step: (e^{i m2 phi} x[slot 4 s + q][rhs]).re or .im
        float z[NSP];
#pragma unroll
        for (int s = 0; s < NSP; ++s) {
          const float4 x4 = xc[s >> 1];
          const float xr = (s & 1) ? x4.z : x4.x, xi = (s & 1) ? x4.w : x4.y;
          const float av0 = zp ? xi : xr, bv0 = zp ? xr : xi;
          const int am = so.absm[s0 + s];
          if (am == 0) {
            z[s] = av0;
          } else {
            const int sg = so.sgn[s0 + s];
            const float e = sg > 0 ? ez[am] : (sg < 0 ? -ez[am] : xor_sign(ez[am], qmask));
            z[s] = fmaf(cs[am], av0, e * bv0);
          }
        }
        // the input of the next pair goes into the same registers
        load_x(rec_n, xc);

        typename Mfma<float>::v4 acc[MT];
#pragma unroll
        for (int i = 0; i < MT; ++i) acc[i] = typename Mfma<float>::v4{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < NSP; ++s) {
#pragma unroll
          for (int i = 0; i < MT; ++i) {
            float v;
            if constexpr (VOL) {
              v = wt[0] * tv[0][i][s];
#pragma unroll
              for (int t = 1; t < T; ++t) v = fmaf(wt[t], tv[t][i][s], v);
            } else {
              v = tv[T - 1][i][s];
#pragma unroll
              for (int t = T - 2; t >= 0; --t) v = fmaf(v, w, tv[t][i][s]);
            }
            acc[i] = Mfma<float>::op(v, z[s], acc[i]);
          }
        }

        // y (rhs, re|im) of rows 16 i + 4 q + v, rotated by e^{-i m1 phi}, scaled by S_hi
        unsigned long long* yrow = yacc + (size_t)(rec.ij >> kPairShift) * LO + colbase;
#pragma unroll
        for (int i = 0; i < MH; ++i) {
          float er = 0.f, ei = 0.f;
#pragma unroll
          for (int am = 0; am <= MMAX; ++am) {
            // |m| values outside this tile's four row groups drop out at compile time
            if (so.absm[4 * i] != am && so.absm[4 * i + 1] != am && so.absm[4 * i + 2] != am &&
                so.absm[4 * i + 3] != am)
              continue;
            er = fmaf(hot[i][am], cs[am], er);
            if (am > 0) ei = fmaf(hot[i][am], sn[am], ei);
          }
          er *= sh;
          ei *= sh;
          const float elo = xor_sign(ei, smlo[i]), ehi = xor_sign(ei, smhi[i]);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float eps = v < 2 ? elo : ehi;
            const float al = acc[i][v], be = acc[i + MH][v];
            const float own = fmaf(al, er, be * eps);
            const float snd = fmaf(be, er, -(al * eps));
            const float val = fma_xor1(own, sgnz, snd);
            atomicAdd(yrow + 16 * i + v, fixed_bits64(val));
          }
        }
      }
    }
  }
  __syncthreads();
  // partial of this item: [Rc][16 columns][S]
  float* yp = reinterpret_cast<float*>(a.ypart) + (size_t)item * a.Rc * 2 * KB * S;
  for (int j = tid; j < a.Rc * 2 * KB * S; j += blockDim.x) {
    const int il = j / (2 * KB * S), k = j - il * 2 * KB * S;
    const int c = k / S, sl2 = k - c * S;
    const int w0 = il * LO + c * SC + sl2;
    const long long sum = (long long)(yacc[w0] - (unsigned long long)n_adds[il] * P * kMagic64Bits);  // modulo 2^64
    yp[j] = (float)((double)sum * ish);
  }
}

}  // namespace smh
