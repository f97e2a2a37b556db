// Radial-lookup coupling operator: pair binning, tap tables, matvec driver.
#include <algorithm>
#include <cmath>

#include "smh_op.h"
#include "smh_batch_half.h"
#include "smh_batch_slots.h"
#include "smh_batch_slots64.h"
#include "smh_lookup.h"

using namespace smh;

namespace {

constexpr int kThreads = 256;

// Grid: (receiver i1, block of 256 emitters of i1's ensemble); pairs couple
// particles of one ensemble only (seg_of null: one ensemble).
__global__ void pair_bin_kernel(const double2* __restrict__ xy, int n, double rho_min, double h, int len_rho,
                                int s_mode, int o_mode, int T, int64_t n_tt, int Rc, const int* __restrict__ seg_of,
                                const int64_t* __restrict__ seg_off, int* cnt, const int* off, int* cursor,
                                uint2* pairs, int* err) {
  const int i1 = blockIdx.x;
  const int sg = seg_of ? seg_of[i1] : 0;
  const int64_t i2l = (int64_t)seg_off[sg] + (int64_t)blockIdx.y * 256 + threadIdx.x;
  if (i1 >= n || i2l >= seg_off[sg + 1]) return;
  const int i2 = (int)i2l;
  const double dx = xy[i1].x - xy[i2].x, dy = xy[i1].y - xy[i2].y;
  int r;
  double w, rho;
  pair_geometry(dx, dy, rho_min, h, r, w, rho);
  const int64_t row0 = (int64_t)s_mode * r + o_mode;
  if (r < 0 || r >= len_rho - 1 || row0 < 0 || row0 + T > n_tt) {
    atomicOr(err, 1);
    return;
  }
  const int64_t key = (int64_t)(i1 / Rc) * len_rho + r;
  if (pairs == nullptr) {
    atomicAdd(&cnt[key], 1);
  } else {
    const int pos = off[key] + atomicAdd(&cursor[key], 1);
    pairs[pos] = make_uint2((unsigned)i1, (unsigned)i2);
  }
}

// Geometry of every binned pair in FP64, stored in the kernel's precision.
template <typename Real>
__global__ void pair_record_kernel(const uint2* __restrict__ pairs, int64_t n_pairs, const double2* __restrict__ xy,
                                   double rho_min, double h, int Rc, PairRec<Real>* rec) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  const uint2 pr = pairs[p];
  const double dx = xy[pr.x].x - xy[pr.y].x, dy = xy[pr.x].y - xy[pr.y].y;
  int r;
  double w, rho;
  pair_geometry(dx, dy, rho_min, h, r, w, rho);
  double c = 1.0, sn = 0.0;
  if (rho > 0.0) {
    c = dx / rho;
    sn = dy / rho;
  }
  PairRec<Real> out{};
  out.ij = pr.y | ((pr.x % (unsigned)Rc) << kPairShift);
  out.w = (Real)w;
  out.c = (Real)c;
  out.s = (Real)sn;
  rec[p] = out;
}

template <typename Real>
__global__ void scatter_kernel(const double2* __restrict__ pk, const int* __restrict__ pk2pad, int64_t n_unk,
                               typename Vec2<Real>::t* pad) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n_unk) {
    const double2 v = pk[j];
    pad[pk2pad[j]] = typename Vec2<Real>::t{Real(v.x), Real(v.y)};
  }
}

template <typename Real>
__global__ void pack_kernel(const typename Vec2<Real>::t* __restrict__ pad, const int* __restrict__ pk2pad,
                            int64_t n_unk, double2* pk) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n_unk) {
    const typename Vec2<Real>::t v = pad[pk2pad[j]];
    pk[j] = make_double2((double)v.x, (double)v.y);
  }
}

// Tap table in MFMA fragment order: row j holds reals
//   F[i][h][g][rr][e] = A_j(16 i + rr, n2 = h KS + 4 g + e),  i < MT, h < G, g < NG,
// where A_j rows [0, Bh) are Re L_j(n1, n2) and rows [Bh, 2 Bh) are Im L_j
// (zero outside n1 < B, n2 < B, 4 g + e < KS). Spline taps interleave values
// and moments (tap row 2k: y_k, row 2k+1: Mh_k).
__global__ void taps_kernel(const double2* __restrict__ table, const double2* __restrict__ mh, int64_t len, int B,
                            int Bh, int G, int KS, int NG, int FR, int interleave, double* tt64, float* tt32) {
  const int64_t n_tt = interleave ? 2 * len : len;
  const int64_t total = n_tt * FR;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = f / FR;
    int rem = (int)(f - row * FR);
    const int e = rem & 3;
    const int rr = (rem >> 2) & 15;
    rem >>= 6;
    const int g = rem % NG;
    rem /= NG;
    const int hh = rem % G;
    const int i = rem / G;
    const int m = 16 * i + rr;
    const int k = 4 * g + e;
    const int n2 = hh * KS + k;
    const double2* src = table;
    int64_t j = row;
    if (interleave) {
      j = row >> 1;
      if (row & 1) src = mh;
    }
    double v = 0.0;
    if (k < KS && n2 < B) {
      if (m < B)
        v = src[(j * B + m) * B + n2].x;
      else if (m >= Bh && m < Bh + B)
        v = src[(j * B + (m - Bh)) * B + n2].y;
    }
    tt64[f] = v;
    tt32[f] = (float)v;
  }
}

// Coefficient table of the FP32 slot-order kBatch-RHS kernel: the
// interpolant on bin r (rho in [rho_r, rho_r + h), w = (rho - rho_r) / h) as
// a polynomial sum_k c_k w^k, k < T, stored as rows r T + k of floats
//   F[i][s4][lane][e] = c_k(slot 16 (i mod MH) + rr, slot 4 (4 s4 + e) + q),  lane = 16 q + rr,
// rows from Re L for i < MH and from Im L for i >= MH, zero for padding slots
// (n_of[slot] < 0). The coefficients expand the tap weights of tap_weight():
//   linear       c0 = y_r, c1 = y_{r+1} - y_r;
//   Catmull-Rom  c0 = y_r, c1 = (y_{r+1} - y_{r-1}) / 2,
//                c2 = y_{r-1} - 5/2 y_r + 2 y_{r+1} - 1/2 y_{r+2},
//                c3 = (-y_{r-1} + 3 y_r - 3 y_{r+1} + y_{r+2}) / 2;
//   spline       c0 = y_r, c1 = y_{r+1} - y_r - 2 Mh_r - Mh_{r+1}, c2 = 3 Mh_r, c3 = Mh_{r+1} - Mh_r.
// Bins no pair uses (at the table ends) read clamped rows.
// Coefficient k of the bin-r interpolant of table entry nn (Re part if re, else Im).
__device__ inline double poly_coef(const double2* __restrict__ table, const double2* __restrict__ mh, int64_t len,
                                   int BB, int64_t nn, bool re, int interp, int64_t r, int k) {
  auto Y = [&](int64_t j) {
    j = j < 0 ? 0 : (j >= len ? len - 1 : j);
    const double2 c = table[j * BB + nn];
    return re ? c.x : c.y;
  };
  auto Mh = [&](int64_t j) {
    j = j < 0 ? 0 : (j >= len ? len - 1 : j);
    const double2 c = mh[j * BB + nn];
    return re ? c.x : c.y;
  };
  if (interp == SMH_INTERP_LINEAR) return k == 0 ? Y(r) : Y(r + 1) - Y(r);
  if (interp == SMH_INTERP_CATMULL_ROM) {
    const double p0 = Y(r - 1), p1 = Y(r), p2 = Y(r + 1), p3 = Y(r + 2);
    return k == 0 ? p1
                  : (k == 1 ? 0.5 * (p2 - p0)
                            : (k == 2 ? p0 - 2.5 * p1 + 2.0 * p2 - 0.5 * p3 : 0.5 * (-p0 + 3.0 * p1 - 3.0 * p2 + p3)));
  }
  const double y0 = Y(r), y1 = Y(r + 1), m0 = Mh(r), m1 = Mh(r + 1);
  return k == 0 ? y0 : (k == 1 ? y1 - y0 - 2.0 * m0 - m1 : (k == 2 ? 3.0 * m0 : m1 - m0));
}

__global__ void poly_slot_kernel(const double2* __restrict__ table, const double2* __restrict__ mh, int64_t len,
                                 int B, int S, const int* __restrict__ n_of, int interp, int T, float* tt) {
  const int NS = S / 4, MH = S / 16, FRB = 2 * MH * NS * 64;
  const int64_t total = len * T * FRB;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = f / FRB;
    int rem = (int)(f - row * FRB);
    const int e = rem & 3;
    const int ln = (rem >> 2) & 63;
    rem >>= 8;
    const int s4 = rem % (NS / 4);
    const int i = rem / (NS / 4);
    const int q = ln >> 4, rr = ln & 15;
    const int n1 = n_of[16 * (i % MH) + rr];
    const int n2 = n_of[4 * (4 * s4 + e) + q];
    const int64_t r = row / T;
    const int k = (int)(row - r * T);
    float v = 0.f;
    if (n1 >= 0 && n2 >= 0) v = (float)poly_coef(table, mh, len, B * B, (int64_t)n1 * B + n2, i < MH, interp, r, k);
    tt[f] = v;
  }
}

// The same coefficients in FP64 for the FP64 slot-order kernel, layout
// [rows][MT][NS/2][64 lanes][2]: lane (q, rr) of tile i holds
// c_k(slot 16 (i mod MH) + rr, slot 4 (2 s2 + e) + q).
__global__ void poly_slot64_kernel(const double2* __restrict__ table, const double2* __restrict__ mh, int64_t len,
                                   int B, int S, const int* __restrict__ n_of, int interp, int T, double* tt) {
  const int NS = S / 4, MH = S / 16, FRB = 2 * MH * NS * 64;
  const int64_t total = len * T * FRB;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = f / FRB;
    int rem = (int)(f - row * FRB);
    const int e = rem & 1;
    const int ln = (rem >> 1) & 63;
    rem >>= 7;
    const int s2 = rem % (NS / 2);
    const int i = rem / (NS / 2);
    const int q = ln >> 4, rr = ln & 15;
    const int n1 = n_of[16 * (i % MH) + rr];
    const int n2 = n_of[4 * (2 * s2 + e) + q];
    const int64_t r = row / T;
    const int k = (int)(row - r * T);
    tt[f] = (n1 >= 0 && n2 >= 0) ? poly_coef(table, mh, len, B * B, (int64_t)n1 * B + n2, i < MH, interp, r, k) : 0.0;
  }
}

// FP16 coefficients of the FP16-input batched kernel, one block per bin r:
// layout [bins][T][MT][64 lanes][NC][4], lane (q, rr) of tile i holding
// c_t(slot 16 (i mod MH) + rr, slot 16 c + 4 q + j); the bin's values are
// divided by hscale[r] = 2^e, the smallest power of two with
// T * max |c| / 2^e <= 2^12 (every Horner partial sum stays far inside the
// FP16 range).
__global__ void poly_half_kernel(const double2* __restrict__ table, const double2* __restrict__ mh, int64_t len,
                                 int B, int S, const int* __restrict__ n_of, int interp, int T, _Float16* th,
                                 float* hscale) {
  const int MH = S / 16, MT = 2 * MH, NC = S / 16;
  const int per_t = MT * 64 * NC * 4;
  const int per_bin = T * per_t;
  const int64_t r = blockIdx.x;
  __shared__ double red[256];
  auto value = [&](int f) {
    const int k = f / per_t;
    int rem = f - k * per_t;
    const int j = rem & 3;
    rem >>= 2;
    const int c = rem % NC;
    rem /= NC;
    const int ln = rem & 63;
    const int i = rem >> 6;
    const int q = ln >> 4, rr = ln & 15;
    const int n1 = n_of[16 * (i % MH) + rr];
    const int n2 = n_of[16 * c + 4 * q + j];
    if (n1 < 0 || n2 < 0) return 0.0;
    return poly_coef(table, mh, len, B * B, (int64_t)n1 * B + n2, i < MH, interp, r, k);
  };
  double m = 0.0;
  for (int f = threadIdx.x; f < per_bin; f += blockDim.x) m = fmax(m, fabs(value(f)));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  const double mx = red[0] * T;
  const int e = mx > 0.0 ? (int)ceil(log2(mx)) - 12 : 0;
  const double inv = ldexp(1.0, -e);
  if (threadIdx.x == 0) hscale[r] = ldexpf(1.0f, e);
  for (int f = threadIdx.x; f < per_bin; f += blockDim.x) th[(size_t)r * per_bin + f] = (_Float16)(value(f) * inv);
}

// Linear interpolation error of each bin relative to its scale: bin r
// (between nodes r and r + 1) gets max_n max(|d2_r|, |d2_{r+1}|) / 8 over
// max_n max(|y_r|, |y_{r+1}|), d2_j = y_{j-1} - 2 y_j + y_{j+1} (h^2 y'' at
// node j) of the Re and Im parts; one block per bin.
__global__ void lin_ratio_kernel(const double2* __restrict__ table, int64_t len, int BB, float* ratio) {
  const int64_t r = blockIdx.x;
  __shared__ double rd[256], ry[256];
  auto d2 = [&](int64_t j, int n) {
    if (j < 1 || j > len - 2) return 0.0;
    const double2 a = table[(j - 1) * BB + n], b = table[j * BB + n], c = table[(j + 1) * BB + n];
    return fmax(fabs(a.x - 2.0 * b.x + c.x), fabs(a.y - 2.0 * b.y + c.y));
  };
  double md = 0.0, my = 0.0;
  for (int n = threadIdx.x; n < BB; n += blockDim.x) {
    md = fmax(md, fmax(d2(r, n), d2(r + 1, n)));
    const double2 y0 = table[r * BB + n], y1 = table[(r + 1 < len ? r + 1 : r) * BB + n];
    my = fmax(my, fmax(fmax(fabs(y0.x), fabs(y0.y)), fmax(fabs(y1.x), fabs(y1.y))));
  }
  rd[threadIdx.x] = md;
  ry[threadIdx.x] = my;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      rd[threadIdx.x] = fmax(rd[threadIdx.x], rd[threadIdx.x + w]);
      ry[threadIdx.x] = fmax(ry[threadIdx.x], ry[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) ratio[r] = ry[0] > 0.0 ? (float)(rd[0] / (8.0 * ry[0])) : 0.f;
}

// Largest ratio[] over the bins the entries use (one block).
__global__ void used_max_kernel(const int4* __restrict__ entries, int n_entries, int row_o, int row_s,
                                const float* __restrict__ ratio, int64_t len, float* out) {
  __shared__ float red[1024];
  float m = 0.f;
  for (int e = threadIdx.x; e < n_entries; e += blockDim.x) {
    const int64_t bin = (entries[e].x - row_o) / row_s;
    if (bin >= 0 && bin < len) m = fmaxf(m, ratio[bin]);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Not-a-knot cubic spline moments, scaled: Mh_j = h^2/6 S''(rho_j), one
// thread per complex column (n1, n2). With uniform spacing the system is
//   Mh_{j-1} + 4 Mh_j + Mh_{j+1} = y_{j-1} - 2 y_j + y_{j+1},  j = 1..n-2,
// closed by Mh_0 - 2 Mh_1 + Mh_2 = 0 and Mh_{n-3} - 2 Mh_{n-2} + Mh_{n-1} = 0,
// which gives Mh_1 = rhs_1 / 6, Mh_{n-2} = rhs_{n-2} / 6 and a tridiagonal
// system for Mh_2..Mh_{n-3} (Thomas algorithm, c' precomputed on the host).
__global__ void spline_kernel(const double2* __restrict__ y, const double* __restrict__ cprime, int64_t n, int BB,
                              double2* mh) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= BB) return;
  auto Y = [&](int64_t j) { return y[j * BB + col]; };
  auto rhs = [&](int64_t j) {
    const double2 a = Y(j - 1), b = Y(j), c = Y(j + 1);
    return make_double2(a.x - 2.0 * b.x + c.x, a.y - 2.0 * b.y + c.y);
  };
  const double2 r1 = rhs(1), rn2 = rhs(n - 2);
  const double2 m1 = make_double2(r1.x / 6.0, r1.y / 6.0);
  const double2 mn2 = make_double2(rn2.x / 6.0, rn2.y / 6.0);
  // forward sweep over j = 2..n-3, d' stored in mh[j]
  double2 dprev = make_double2(0, 0);
  for (int64_t j = 2; j <= n - 3; ++j) {
    double2 d = rhs(j);
    if (j == 2) {
      d.x -= m1.x;
      d.y -= m1.y;
    }
    if (j == n - 3) {
      d.x -= mn2.x;
      d.y -= mn2.y;
    }
    const double den = (j == 2) ? 4.0 : 4.0 - cprime[j - 1];
    const double2 dp = (j == 2) ? make_double2(d.x / 4.0, d.y / 4.0)
                                : make_double2((d.x - dprev.x) / den, (d.y - dprev.y) / den);
    mh[j * BB + col] = dp;
    dprev = dp;
  }
  double2 next = mh[(n - 3) * BB + col];
  for (int64_t j = n - 4; j >= 2; --j) {
    const double2 dp = mh[j * BB + col];
    const double2 v = make_double2(dp.x - cprime[j] * next.x, dp.y - cprime[j] * next.y);
    mh[j * BB + col] = v;
    next = v;
  }
  const double2 m2 = mh[2 * BB + col], mn3 = mh[(n - 3) * BB + col];
  mh[1 * BB + col] = m1;
  mh[(n - 2) * BB + col] = mn2;
  mh[0 * BB + col] = make_double2(2.0 * m1.x - m2.x, 2.0 * m1.y - m2.y);
  mh[(n - 1) * BB + col] = make_double2(2.0 * mn2.x - mn3.x, 2.0 * mn2.y - mn3.y);
}

template <int L, int Mx, int T, typename Real>
KernelInfo kernel_info() {
  using Sh = RadialShape<Real, SvwfIdx<L, Mx>::B, T>;
  KernelInfo k;
  k.fn = reinterpret_cast<void*>(&lookup_matvec_kernel<Real, Real, L, Mx, T, false>);
  k.threads = Sh::kThreads;
  k.sy = Sh::SY;
  return k;
}

template <int T, typename Real>
KernelInfo pick(int l, int m) {
#define SMH_K(L, Mx) \
  if (l == L && m == Mx) return kernel_info<L, Mx, T, Real>();
  SMH_K(1, 1)
  SMH_K(2, 1)
  SMH_K(2, 2)
  SMH_K(3, 1)
  SMH_K(3, 2)
  SMH_K(3, 3)
  SMH_K(4, 1)
  SMH_K(4, 2)
  SMH_K(4, 3)
  SMH_K(4, 4)
#undef SMH_K
  return KernelInfo{};
}

// The slot-order kBatch-RHS kernels (FP32 and FP64), their coefficient
// fragments in whole parts of the K-steps (every cut-off up to l_max = m_max
// = 4; none otherwise).
template <int L, int Mx, int T, typename Real>
KernelInfo batch_kernel_info() {
  KernelInfo k;
  if constexpr (std::is_same<Real, float>::value && SlotShape<L, Mx, T>::P32 > 0 && Mx <= 4) {
    k.fn = reinterpret_cast<void*>(&batch_slots_kernel<L, Mx, T>);
    k.threads = 512;
    k.sy = SlotShape<L, Mx, T>::SY;  // int32 words (one int64 word per result)
    k.slots = SlotShape<L, Mx, T>::S;
    k.slot_tables = &slot_tables<L, Mx>;
  } else if constexpr (std::is_same<Real, double>::value && SlotShape<L, Mx, T>::P64 > 0 && Mx <= 4) {
    k.fn = reinterpret_cast<void*>(&batch_slots64_kernel<L, Mx, T>);
    k.threads = kSlots64Threads;
    k.sy = SlotShape<L, Mx, T>::LO;  // doubles
    k.slots = SlotShape<L, Mx, T>::S;
    k.slot_tables = &slot_tables<L, Mx>;
  }
  return k;
}

template <int T, typename Real>
KernelInfo pick_batch(int l, int m) {
#define SMH_K(L, Mx) \
  if (l == L && m == Mx) return batch_kernel_info<L, Mx, T, Real>();
  SMH_K(1, 1)
  SMH_K(2, 1)
  SMH_K(2, 2)
  SMH_K(3, 1)
  SMH_K(3, 2)
  SMH_K(3, 3)
  SMH_K(4, 1)
  SMH_K(4, 2)
  SMH_K(4, 3)
  SMH_K(4, 4)
#undef SMH_K
  return KernelInfo{};
}

// FP16-operand kBatch-RHS kernel: where the FP32 slot-order kernel is used
// (its binning and slot order)
template <int L, int Mx, int T>
void* half_fn() {
  if constexpr (SlotShape<L, Mx, 4>::P32 > 0 && Mx <= 4 && half_tile_parts(T, SvwfIdx<L, Mx>::B) > 0)
    return reinterpret_cast<void*>(&batch_half_kernel<L, Mx, T>);
  else
    return nullptr;
}

void* radial_half_kernel(int T, int l, int m) {
#define SMH_K(L, Mx) \
  if (l == L && m == Mx) return T == 4 ? half_fn<L, Mx, 4>() : half_fn<L, Mx, 2>();
  SMH_K(1, 1)
  SMH_K(2, 1)
  SMH_K(2, 2)
  SMH_K(3, 1)
  SMH_K(3, 2)
  SMH_K(3, 3)
  SMH_K(4, 1)
  SMH_K(4, 2)
  SMH_K(4, 3)
  SMH_K(4, 4)
#undef SMH_K
  return nullptr;
}

KernelInfo radial_batch_kernel(int precision, int T, int l, int m) {
  if (precision == SMH_FP32) return T == 4 ? pick_batch<4, float>(l, m) : pick_batch<2, float>(l, m);
  return T == 4 ? pick_batch<4, double>(l, m) : pick_batch<2, double>(l, m);
}

KernelInfo radial_kernel(int precision, int T, int l, int m) {
  if (precision == SMH_FP32) return T == 4 ? pick<4, float>(l, m) : pick<2, float>(l, m);
  return T == 4 ? pick<4, double>(l, m) : pick<2, double>(l, m);
}


}  // namespace

namespace smh {

int op_set_lds(smh_op* op, void* fn, size_t lds) {
  for (void* f : op->attr_set)
    if (f == fn) return SMH_OK;
  // the device's limit, so launches of this kernel by any operator fit
  const size_t lim = std::max(lds, op->ctx->lds_per_block);
  SMH_HIP(op->ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim));
  op->attr_set.push_back(fn);
  return SMH_OK;
}

int op_reserve_packed(smh_op* op, int64_t nrhs) {
  const int64_t need = op->n_unk * nrhs;
  if (need <= op->pk_cap) return SMH_OK;
  if (op->d_pk_in) hipFree(op->d_pk_in);
  if (op->d_pk_out) hipFree(op->d_pk_out);
  op->d_pk_in = op->d_pk_out = nullptr;
  SMH_HIP(op->ctx, hipMalloc(&op->d_pk_in, need * sizeof(double2)));
  SMH_HIP(op->ctx, hipMalloc(&op->d_pk_out, need * sizeof(double2)));
  op->pk_cap = need;
  return SMH_OK;
}

int op_unpack(smh_op* op, int precision, const double2* pk, void* pad) {
  hipStream_t s = op->ctx->stream;
  const int64_t n_pad = op->n_part * op->B;
  if (precision == SMH_FP32) {
    SMH_HIP(op->ctx, hipMemsetAsync(pad, 0, n_pad * sizeof(float2), s));
    hipLaunchKernelGGL(scatter_kernel<float>, dim3(blocks_for(op->n_unk)), dim3(256), 0, s, pk, op->d_pk2pad,
                       op->n_unk, (float2*)pad);
  } else {
    SMH_HIP(op->ctx, hipMemsetAsync(pad, 0, n_pad * sizeof(double2), s));
    hipLaunchKernelGGL(scatter_kernel<double>, dim3(blocks_for(op->n_unk)), dim3(256), 0, s, pk, op->d_pk2pad,
                       op->n_unk, (double2*)pad);
  }
  SMH_HIP(op->ctx, hipGetLastError());
  return SMH_OK;
}

int op_pack(smh_op* op, int precision, const void* pad, double2* pk) {
  hipStream_t s = op->ctx->stream;
  if (precision == SMH_FP32)
    hipLaunchKernelGGL(pack_kernel<float>, dim3(blocks_for(op->n_unk)), dim3(256), 0, s, (const float2*)pad,
                       op->d_pk2pad, op->n_unk, pk);
  else
    hipLaunchKernelGGL(pack_kernel<double>, dim3(blocks_for(op->n_unk)), dim3(256), 0, s, (const double2*)pad,
                       op->d_pk2pad, op->n_unk, pk);
  SMH_HIP(op->ctx, hipGetLastError());
  return SMH_OK;
}

int op_make_taps(smh_op* op) {
  smh_ctx* ctx = op->ctx;
  hipStream_t s = ctx->stream;
  const int64_t len = op->len_rho;
  const int BB = op->B * op->B;
  double2* d_mh = nullptr;
  double* d_cp = nullptr;
  if (op->interp == SMH_INTERP_SPLINE) {
    SMH_CHECK(ctx, len >= 7, SMH_ERR_ARG, "spline interpolation needs at least 7 table rows (got %lld)",
              (long long)len);
    std::vector<double> cp(len, 0.0);
    for (int64_t j = 2; j <= len - 3; ++j) cp[j] = (j == 2) ? 0.25 : 1.0 / (4.0 - cp[j - 1]);
    SMH_HIP(ctx, hipMalloc(&d_mh, len * BB * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&d_cp, len * sizeof(double)));
    SMH_HIP(ctx, hipMemcpyAsync(d_cp, cp.data(), len * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(spline_kernel, dim3(blocks_for(BB, 64)), dim3(64), 0, s, op->d_table, d_cp, len, BB, d_mh);
    SMH_HIP(ctx, hipGetLastError());
  }
  const TapGeom g(op->B, op->T);
  op->FR = g.FR;
  if (!op->d_tt64) {
    SMH_HIP(ctx, hipMalloc(&op->d_tt64, (size_t)op->n_tt * g.FR * sizeof(double)));
    SMH_HIP(ctx, hipMalloc(&op->d_tt32, (size_t)op->n_tt * g.FR * sizeof(float)));
  }
  hipLaunchKernelGGL(taps_kernel, dim3(4096), dim3(256), 0, s, op->d_table, d_mh, len, op->B, op->Bh, g.G, g.KS, g.NG,
                     g.FR, op->interp == SMH_INTERP_SPLINE ? 1 : 0, op->d_tt64, op->d_tt32);
  if (op->d_b32_tab) {
    const size_t frb = (size_t)2 * op->b32_S * op->b32_S;  // MT * NS * 64 floats
    if (!op->d_tt32b) SMH_HIP(ctx, hipMalloc(&op->d_tt32b, (size_t)len * op->T * frb * sizeof(float)));
    hipLaunchKernelGGL(poly_slot_kernel, dim3(4096), dim3(256), 0, s, op->d_table, d_mh, len, op->B, op->b32_S,
                       op->d_b32_tab + op->B, op->interp, op->T, op->d_tt32b);
    if (radial_batch_kernel(SMH_FP64, op->T, op->l_max, op->m_max).slots == op->b32_S) {
      if (!op->d_tt64b) SMH_HIP(ctx, hipMalloc(&op->d_tt64b, (size_t)len * op->T * frb * sizeof(double)));
      hipLaunchKernelGGL(poly_slot64_kernel, dim3(4096), dim3(256), 0, s, op->d_table, d_mh, len, op->B, op->b32_S,
                         op->d_b32_tab + op->B, op->interp, op->T, op->d_tt64b);
    }
    // The FP16 operator (inner solves only) interpolates linearly between
    // the table nodes when that error stays within two FP16 roundings (2^-10
    // of each bin's scale) on every bin a pair uses: half the Horner work and
    // coefficient traffic, at an operator error the inner solves' 1e-3
    // target does not see (the FP32 / FP64 residuals of the refinement use
    // the operator's own interpolation). Otherwise it keeps that one.
    op->T_half = op->T;
    const smh_op::Binning& hb = op->bbin[0];
    if (op->T > 2 && hb.n_entries > 0 && radial_half_kernel(2, op->l_max, op->m_max)) {
      float* d_ratio = nullptr;
      SMH_HIP(ctx, hipMalloc(&d_ratio, (size_t)(len + 1) * sizeof(float)));
      hipLaunchKernelGGL(lin_ratio_kernel, dim3((unsigned)len), dim3(256), 0, s, op->d_table, len, op->B * op->B,
                         d_ratio);
      hipLaunchKernelGGL(used_max_kernel, dim3(1), dim3(1024), 0, s, hb.d_entries, hb.n_entries, op->o_mode,
                         op->s_mode, d_ratio, len, d_ratio + len);
      float worst = 1.f;
      const hipError_t e1 = hipMemcpyAsync(&worst, d_ratio + len, sizeof(float), hipMemcpyDeviceToHost, s);
      const hipError_t e2 = hipStreamSynchronize(s);
      hipFree(d_ratio);
      SMH_CHECK(ctx, e1 == hipSuccess && e2 == hipSuccess, SMH_ERR_HIP, "interpolation check failed");
      op->lin_ratio = worst;
      if (worst <= 0x1p-10f) op->T_half = 2;
    }
    if (radial_half_kernel(op->T_half, op->l_max, op->m_max)) {
      // T * MT * 64 * NC * 4 halves, MT = 2 S / 16, NC = S / 16: 2 S^2 per tap
      // (allocated for the operator's T, the larger of the two)
      const size_t per_bin = (size_t)std::max(op->T, 2) * 2 * op->b32_S * op->b32_S;
      if (!op->d_tth) {
        SMH_HIP(ctx, hipMalloc(&op->d_tth, (size_t)len * per_bin * sizeof(_Float16)));
        SMH_HIP(ctx, hipMalloc(&op->d_hscale, (size_t)len * sizeof(float)));
      }
      hipLaunchKernelGGL(poly_half_kernel, dim3((unsigned)len), dim3(256), 0, s, op->d_table, d_mh, len, op->B,
                         op->b32_S, op->d_b32_tab + op->B, op->T_half == 2 ? (int)SMH_INTERP_LINEAR : op->interp,
                         op->T_half, op->d_tth, op->d_hscale);
    }
  }
  if (!op->d_rowsum) {
    SMH_HIP(ctx, hipMalloc(&op->d_rowsum, (size_t)op->n_tt * sizeof(float)));
    SMH_HIP(ctx, hipMalloc(&op->d_xmax, (1 + kXmaxBlocks) * sizeof(float)));
  }
  hipLaunchKernelGGL(tap_rowsum_kernel<double2>, dim3(blocks_for(op->n_tt)), dim3(256), 0, s, op->d_table, d_mh,
                     op->n_tt, op->B, op->interp == SMH_INTERP_SPLINE ? 1 : 0, op->d_rowsum);
  for (smh_op::Binning* bn : {&op->bin[0], &op->bbin[0]}) set_item_lsum(*bn, op->d_rowsum, op->T, 0, 0, s);
  SMH_HIP(ctx, hipGetLastError());
  SMH_HIP(ctx, hipStreamSynchronize(s));
  if (d_mh) hipFree(d_mh);
  if (d_cp) hipFree(d_cp);
  op->has_table = true;
  return SMH_OK;
}

int op_apply(smh_op* op, int precision, const void* x_pad, void* y_pad, bool time_kernel) {
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, op->has_table, SMH_ERR_STATE, "coupling table not set");
  if (op->kind == 1) return vol_apply(op, precision, x_pad, y_pad, time_kernel);
  if (op->kind == 2) return explicit_apply(op, precision, x_pad, y_pad, time_kernel);
  const int pi = precision == SMH_FP32 ? 0 : 1;
  const smh_op::Binning& bn = op->bin[pi];
  SMH_CHECK(ctx, bn.Rc > 0, SMH_ERR_UNSUPPORTED, "no %s kernel fits LDS for blocksize %d (l_max %d, m_max %d)",
            pi == 0 ? "FP32" : "FP64", op->B, op->l_max, op->m_max);
  const KernelInfo k = radial_kernel(precision, op->T, op->l_max, op->m_max);
  SMH_CHECK(ctx, k.fn != nullptr, SMH_ERR_UNSUPPORTED, "no matvec kernel for l_max %d m_max %d", op->l_max,
            op->m_max);
  RadialKArgs a;
  a.recs = bn.d_rec;
  a.entries = bn.d_entries;
  a.items = bn.d_items;
  a.order = bn.d_order;
  a.tt = pi == 0 ? (const void*)op->d_tt32 : (const void*)op->d_tt64;
  a.x = x_pad;
  a.ypart = op->d_ypart;
  a.xmax = op->d_xmax;
  a.item_lsum = bn.d_item_lsum;
  a.item_en = bn.d_item_en;
  a.n_items = bn.n_items;
  a.n_entries = bn.n_entries;
  a.interp = op->interp;
  a.Rc = bn.Rc;
  hipStream_t s = ctx->stream;
  if (int rc_ = op_set_lds(op, k.fn, bn.lds)) return rc_;
  if (pi == 0)
    launch_xmax((const float2*)x_pad, op->n_part * op->B, op->d_xmax, s);
  if (bn.n_items > 0) {
    if (time_kernel) SMH_HIP(ctx, hipEventRecord(ctx->ev0, s));
    void* args[] = {&a};
    SMH_HIP(ctx, hipLaunchKernel(k.fn, dim3(bn.n_items), dim3(k.threads), args, bn.lds, s));
    if (time_kernel) SMH_HIP(ctx, hipEventRecord(ctx->ev1, s));
  }
  const int64_t n = op->n_part * op->B;
  if (pi == 0)
    hipLaunchKernelGGL(reduce_kernel<float>, dim3(blocks_for(n)), dim3(256), 0, s, (const float*)op->d_ypart,
                       bn.d_chunk_items, bn.Rc, op->B, op->n_part, (float2*)y_pad, 0);
  else
    hipLaunchKernelGGL(reduce_kernel<double>, dim3(blocks_for(n)), dim3(256), 0, s, (const double*)op->d_ypart,
                       bn.d_chunk_items, bn.Rc, op->B, op->n_part, (double2*)y_pad, 0);
  SMH_HIP(ctx, hipGetLastError());
  if (time_kernel && bn.n_items > 0) {
    SMH_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0;
    SMH_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    op->last_kernel_ms = ms;
  }
  return SMH_OK;
}

}  // namespace smh

namespace {

// kBatch-vector layouts: packed [nrhs][n_unk] (smuthi) <-> padded [n_part][B][kBatch]
// (right-hand sides nrhs..kBatch-1 are zero); the FP32 input may use the slot
// layout of b32_index() instead (S slots per particle, slot table given).
template <typename Real>
__global__ void scatter_batch_kernel(const double2* __restrict__ pk, const int* __restrict__ pk2pad, int64_t n_unk,
                                     int nrhs, int B, int S, const int* __restrict__ slot,
                                     typename Vec2<Real>::t* pad) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_unk * nrhs) return;
  const int64_t r = j / n_unk, u = j - r * n_unk;
  const double2 v = pk[j];
  const int ip = pk2pad[u];
  const int64_t p = ip / B;
  pad[b32_index(p, (int)(ip - p * B), (int)r, S, slot)] = typename Vec2<Real>::t{Real(v.x), Real(v.y)};
}


// FP64 padded [n_part][B][kBatch] -> FP16 input of the FP16-operand kernel
// (b16_index layout, values divided by *xmax)
__global__ void to_half_b16_kernel(const double2* __restrict__ v, int64_t n8, int B, int S,
                                   const int* __restrict__ slot, const float* __restrict__ xmax, _Float16* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n8) return;
  const int r = (int)(e % kBatch);
  const int64_t ip = e / kBatch, p = ip / B;
  const float m = *xmax;
  const double inv = m > 0.f ? 1.0 / (double)m : 0.0;
  const int64_t k = b16_index(p, (int)(ip - p * B), r, S, slot);
  out[k] = (_Float16)(v[e].x * inv);
  out[k + 4] = (_Float16)(v[e].y * inv);
}

template <typename Real>
__global__ void pack_batch_kernel(const typename Vec2<Real>::t* __restrict__ pad, const int* __restrict__ pk2pad,
                                  int64_t n_unk, int nrhs, double2* pk) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_unk * nrhs) return;
  const int64_t r = j / n_unk, u = j - r * n_unk;
  const typename Vec2<Real>::t v = pad[(int64_t)pk2pad[u] * kBatch + r];
  pk[j] = make_double2((double)v.x, (double)v.y);
}

}  // namespace

namespace smh {

int op_unpack_batch(smh_op* op, int precision, const double2* pk, int nrhs, void* pad) {
  hipStream_t s = op->ctx->stream;
  const int64_t n = op->n_unk * nrhs;
  if (precision == SMH_FP32) {
    const int* slot = op->d_b32_tab;
    SMH_HIP(op->ctx, hipMemsetAsync(pad, 0, (size_t)op->n_part * op->b32_S * kBatch * sizeof(float2), s));
    hipLaunchKernelGGL(scatter_batch_kernel<float>, dim3(blocks_for(n)), dim3(256), 0, s, pk, op->d_pk2pad,
                       op->n_unk, nrhs, op->B, op->b32_S, slot, (float2*)pad);
  } else if (precision == SMH_FP16) {
    // through the FP64 padded layout (in the output buffer, free until the matvec runs)
    SMH_CHECK(op->ctx, op->d_yb != nullptr, SMH_ERR_STATE, "batched buffers not reserved");
    SMH_HIP(op->ctx, hipMemsetAsync(op->d_yb, 0, (size_t)op->n_part * op->B * kBatch * sizeof(double2), s));
    hipLaunchKernelGGL(scatter_batch_kernel<double>, dim3(blocks_for(n)), dim3(256), 0, s, pk, op->d_pk2pad,
                       op->n_unk, nrhs, op->B, op->B, (const int*)nullptr, (double2*)op->d_yb);
    SMH_HIP(op->ctx, hipGetLastError());
    return op_to_half_batch(op, (const double2*)op->d_yb, pad);
  } else {
    SMH_HIP(op->ctx, hipMemsetAsync(pad, 0, (size_t)op->n_part * op->B * kBatch * sizeof(double2), s));
    hipLaunchKernelGGL(scatter_batch_kernel<double>, dim3(blocks_for(n)), dim3(256), 0, s, pk, op->d_pk2pad,
                       op->n_unk, nrhs, op->B, op->B, (const int*)nullptr, (double2*)pad);
  }
  SMH_HIP(op->ctx, hipGetLastError());
  return SMH_OK;
}

int op_to_half_batch(smh_op* op, const double2* v, void* out) {
  hipStream_t s = op->ctx->stream;
  const int64_t n8 = op->n_part * op->B * kBatch;
  launch_xmax(v, n8, op->d_xmax, s);
  fill_zero(out, (size_t)op->n_part * op->b32_S * kBatch * 2 * sizeof(_Float16), s);
  hipLaunchKernelGGL(to_half_b16_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, v, n8, op->B, op->b32_S,
                     (const int*)op->d_b32_tab, (const float*)op->d_xmax, (_Float16*)out);
  SMH_HIP(op->ctx, hipGetLastError());
  return SMH_OK;
}

int op_pack_batch(smh_op* op, int precision, const void* pad, int nrhs, double2* pk) {
  hipStream_t s = op->ctx->stream;
  const int64_t n = op->n_unk * nrhs;
  if (precision == SMH_FP32 || precision == SMH_FP16)
    hipLaunchKernelGGL(pack_batch_kernel<float>, dim3(blocks_for(n)), dim3(256), 0, s, (const float2*)pad,
                       op->d_pk2pad, op->n_unk, nrhs, pk);
  else
    hipLaunchKernelGGL(pack_batch_kernel<double>, dim3(blocks_for(n)), dim3(256), 0, s, (const double2*)pad,
                       op->d_pk2pad, op->n_unk, nrhs, pk);
  SMH_HIP(op->ctx, hipGetLastError());
  return SMH_OK;
}

int op_reserve_batch(smh_op* op) {
  if (op->d_xb) return SMH_OK;
  // FP64 padded vectors or the FP32 / FP16-kernel input in its slot layout, whichever is larger
  const size_t bytes = (size_t)op->n_part * kBatch *
                       std::max<size_t>((size_t)op->B * sizeof(double2), (size_t)op->b32_S * sizeof(float2));
  SMH_HIP(op->ctx, hipMalloc(&op->d_xb, bytes));
  SMH_HIP(op->ctx, hipMalloc(&op->d_yb, bytes));
  return SMH_OK;
}

bool op_has_batch(const smh_op* op, int precision) {
  if (precision == SMH_FP16)
    return (op->kind == 0 && op->bbin[0].Rc > 0 && op->d_tth != nullptr) ||
           (op->kind == 1 && op->vbbin[0][0].Rc > 0 && op->d_vtth[0] != nullptr && op->d_vtth[1] != nullptr);
  if (op->kind == 1) return op->vbbin[0][precision == SMH_FP32 ? 0 : 1].Rc > 0;
  return op->kind == 0 && op->bbin[precision == SMH_FP32 ? 0 : 1].Rc > 0;
}

int op_set_segments(smh_op* op, int n_ens, const int64_t* ens_off) {
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, n_ens >= 1 && n_ens <= (1 << 16) && ens_off[0] == 0 && ens_off[n_ens] == op->n_part, SMH_ERR_ARG,
            "ensemble offsets must run from 0 to the particle count");
  for (int e = 0; e < n_ens; ++e)
    SMH_CHECK(ctx, ens_off[e + 1] > ens_off[e], SMH_ERR_ARG, "ensemble %d is empty", e);
  op->n_seg = n_ens;
  op->seg_off.assign(ens_off, ens_off + n_ens + 1);
  std::vector<int> seg_of(op->n_part);
  for (int e = 0; e < n_ens; ++e)
    for (int64_t i = ens_off[e]; i < ens_off[e + 1]; ++i) seg_of[i] = e;
  if (op->d_seg_off) hipFree(op->d_seg_off);
  op->d_seg_off = nullptr;
  SMH_HIP(ctx, hipMalloc(&op->d_seg_off, (n_ens + 1) * sizeof(int64_t)));
  SMH_HIP(ctx, hipMalloc(&op->d_seg_of, op->n_part * sizeof(int)));
  SMH_HIP(ctx, hipMemcpy(op->d_seg_off, op->seg_off.data(), (n_ens + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  SMH_HIP(ctx, hipMemcpy(op->d_seg_of, seg_of.data(), op->n_part * sizeof(int), hipMemcpyHostToDevice));
  return SMH_OK;
}

int op_apply_batch(smh_op* op, int precision, const void* x_pad, void* y_pad, bool time_kernel) {
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, op->has_table, SMH_ERR_STATE, "coupling table not set");
  if (op->kind == 1) {
    SMH_CHECK(ctx, precision != SMH_FP16 || op_has_batch(op, SMH_FP16), SMH_ERR_UNSUPPORTED,
              "no FP16 matvec for this volume operator");
    return vol_apply_batch(op, precision, x_pad, y_pad, time_kernel);
  }
  const bool half = precision == SMH_FP16;
  const int pi = (precision == SMH_FP32 || half) ? 0 : 1;
  SMH_CHECK(ctx, op_has_batch(op, precision), SMH_ERR_UNSUPPORTED, "no batched matvec for this operator");
  const smh_op::Binning& bn = op->bbin[pi];
  const KernelInfo k = radial_batch_kernel(pi == 0 ? SMH_FP32 : SMH_FP64, op->T, op->l_max, op->m_max);
  void* fn = half ? radial_half_kernel(op->T_half, op->l_max, op->m_max) : k.fn;
  RadialKArgs a;
  a.recs = bn.d_rec;
  a.entries = bn.d_entries;
  a.items = bn.d_items;
  a.order = bn.d_order;
  const bool slots = k.slots > 0;
  SMH_CHECK(ctx, !slots || (op->d_b32_tab && op->b32_S == k.slots && (pi == 0 ? (void*)op->d_tt32b : (void*)op->d_tt64b)),
            SMH_ERR_STATE, "slot-order tap table missing");
  SMH_CHECK(ctx, !half || (fn && slots), SMH_ERR_STATE, "FP16 tap table missing");
  a.tt = half ? (const void*)op->d_tth
              : (pi == 0 ? (slots ? (const void*)op->d_tt32b : (const void*)op->d_tt32)
                         : (slots ? (const void*)op->d_tt64b : (const void*)op->d_tt64));
  a.x = x_pad;
  hipStream_t s = ctx->stream;
  if (pi == 1 && slots) {
    // FP64 padded [n_part][B][kBatch] -> the slot layout (padding slots stay zero)
    const int64_t n8 = op->n_part * op->B * kBatch;
    if (!op->d_x64s) {
      const size_t bytes = (size_t)op->n_part * k.slots * kBatch * sizeof(double2);
      SMH_HIP(ctx, hipMalloc(&op->d_x64s, bytes));
      SMH_HIP(ctx, hipMemsetAsync(op->d_x64s, 0, bytes, s));
    }
    hipLaunchKernelGGL(to_slots64_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, (const double2*)x_pad, n8, op->B,
                       k.slots, (const int*)op->d_b32_tab, op->d_x64s);
    a.x = op->d_x64s;
  }
  a.ypart = op->d_ypart;
  a.xmax = op->d_xmax;
  a.item_lsum = bn.d_item_lsum;
  a.item_en = bn.d_item_en;
  a.n_items = bn.n_items;
  a.n_entries = bn.n_entries;
  a.interp = op->interp;
  a.Rc = bn.Rc;
  a.len_z = 0;
  a.row_s = op->s_mode;
  a.row_o = op->o_mode;
  if (int rc_ = op_set_lds(op, fn, bn.lds)) return rc_;
  const int64_t nx = op->n_part * op->B * kBatch;
  const int S = slots ? k.slots : op->B;
  const int* slot = slots ? op->d_b32_tab : nullptr;
  // (the FP16 input was scaled by the xmax its conversion wrote)
  if (pi == 0 && !half)
    launch_xmax((const float2*)x_pad, op->n_part * S * kBatch, op->d_xmax, s);
  if (bn.n_items > 0) {
    if (time_kernel) SMH_HIP(ctx, hipEventRecord(ctx->ev0, s));
    const float* hs = op->d_hscale;
    void* args1[] = {&a};
    void* args2[] = {&a, &hs};
    const int threads = half ? kHalfThreads : k.threads;
    SMH_HIP(ctx, hipLaunchKernel(fn, dim3(bn.n_items), dim3(threads), half ? args2 : args1,
                                 bn.lds, s));
    if (time_kernel) SMH_HIP(ctx, hipEventRecord(ctx->ev1, s));
  }
  if (pi == 0)
    hipLaunchKernelGGL(reduce_batch_kernel<float>, dim3(blocks_for(nx)), dim3(256), 0, s, (const float*)op->d_ypart,
                       bn.d_chunk_items, bn.Rc, op->B, S, slot, op->n_part, (float2*)y_pad, 0);
  else
    hipLaunchKernelGGL(reduce_batch_kernel<double>, dim3(blocks_for(nx)), dim3(256), 0, s,
                       (const double*)op->d_ypart, bn.d_chunk_items, bn.Rc, op->B, S, slot, op->n_part,
                       (double2*)y_pad, 0);
  SMH_HIP(ctx, hipGetLastError());
  if (time_kernel && bn.n_items > 0) {
    SMH_HIP(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0;
    SMH_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    op->last_kernel_ms = ms;
  }
  return SMH_OK;
}

// System-vector layout: packed index (smuthi's particle-major blocks of
// blocksize(l_max_i, m_max_i)) -> padded index i * B + n of the operator's
// (l_max, m_max) order; uploaded to the device.
int op_index_map(smh_op* op, const int* lmax, const int* mmax) {
  smh_ctx* ctx = op->ctx;
  op->pk2pad.clear();
  int64_t n_unk = 0;
  for (int64_t i = 0; i < op->n_part; ++i) {
    if (lmax[i] < 1 || mmax[i] < 0 || mmax[i] > lmax[i] || lmax[i] > op->l_max || mmax[i] > op->m_max)
      return smh_fail(ctx, SMH_ERR_ARG, "particle %lld has l_max %d m_max %d outside the table's %d/%d",
                      (long long)i, lmax[i], mmax[i], op->l_max, op->m_max);
    for (int t = 0; t < 2; ++t)
      for (int l = 1; l <= lmax[i]; ++l) {
        const int ml = l < mmax[i] ? l : mmax[i];
        for (int m = -ml; m <= ml; ++m)
          op->pk2pad.push_back((int)(i * op->B + smh_multi_to_single(t, l, m, op->l_max, op->m_max)));
      }
    n_unk += smh_blocksize(lmax[i], mmax[i]);
  }
  op->n_unk = n_unk;
  // one ensemble unless the operator was created for several
  op->n_seg = 1;
  op->seg_off = {0, op->n_part};
  SMH_HIP(ctx, hipMalloc(&op->d_seg_off, 2 * sizeof(int64_t)));
  SMH_HIP(ctx, hipMemcpyAsync(op->d_seg_off, op->seg_off.data(), 2 * sizeof(int64_t), hipMemcpyHostToDevice,
                              ctx->stream));
  SMH_HIP(ctx, hipMalloc(&op->d_pk2pad, n_unk * sizeof(int)));
  SMH_HIP(ctx, hipMemcpyAsync(op->d_pk2pad, op->pk2pad.data(), n_unk * sizeof(int), hipMemcpyHostToDevice,
                              ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SMH_OK;
}

}  // namespace smh

void smh_op_free(smh_op* op) {
  smh::lu_release(op);
  smh::solver_release(op);
  void* sp[] = {op->d_seg_off, op->d_seg_of, op->d_P, op->d_Pnbr, op->d_near, op->d_near_nbr};
  for (void* p : sp)
    if (p) hipFree(p);
  void* ptrs[] = {op->d_xy, op->d_xyz, op->d_pk2pad, op->d_table, op->d_table_minus, op->d_tt64, op->d_tt32,
                  op->d_tt32b, op->d_tt64b, op->d_x64s, op->d_b32_tab, op->d_tth, op->d_hscale, op->d_W64, op->d_W32, op->d_ypart, op->d_pk_in, op->d_pk_out, op->d_x32, op->d_y32, op->d_x64, op->d_y64,
                  op->d_rowsum, op->d_xmax, op->d_xb, op->d_yb};
  for (void* p : ptrs)
    if (p) hipFree(p);
  auto free_bin = [](smh_op::Binning& b) {
    void* bp[] = {b.d_rec, b.d_entries, b.d_items, b.d_chunk_items, b.d_item_lsum, b.d_item_en, b.d_order};
    for (void* p : bp)
      if (p) hipFree(p);
  };
  for (auto& b : op->bin) free_bin(b);
  for (auto& b : op->bbin) free_bin(b);
  for (auto& row : op->vbin)
    for (auto& b : row) free_bin(b);
  for (auto& row : op->vbbin)
    for (auto& b : row) free_bin(b);
  for (int t = 0; t < 2; ++t) {
    void* vp[] = {op->d_vtab[t],   op->d_vtt[t], op->d_vtts[t], op->d_vtts64[t],
                  op->d_vtth[t],   op->d_vhs[t], op->d_vtt64[t], op->d_vrowsum[t]};
    for (void* p : vp)
      if (p) hipFree(p);
  }
}

// Bin all ordered pairs by (receiver chunk, table interval) for one precision,
// write their records and the bin entries / work items the matvec kernel walks.
template <typename Real>
static int bin_pairs(smh_op* op, smh_op::Binning& bn, int Rc, bool batch) {
  smh_ctx* ctx = op->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n_part = op->n_part, len_rho = op->len_rho;
  bn.Rc = Rc;
  bn.n_chunks = (int)((n_part + Rc - 1) / Rc);
  const int64_t nkeys = (int64_t)bn.n_chunks * len_rho;
  int *d_cnt = nullptr, *d_cur = nullptr, *d_err = nullptr, *d_off = nullptr;
  uint2* d_pairs = nullptr;
  auto cleanup = [&]() {
    if (d_cnt) hipFree(d_cnt);
    if (d_cur) hipFree(d_cur);
    if (d_err) hipFree(d_err);
    if (d_off) hipFree(d_off);
    if (d_pairs) hipFree(d_pairs);
  };
  if (hipMalloc(&d_cnt, (nkeys + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc(&d_cur, (nkeys + 1) * sizeof(int)) != hipSuccess || hipMalloc(&d_err, sizeof(int)) != hipSuccess ||
      hipMalloc(&d_off, (nkeys + 1) * sizeof(int)) != hipSuccess) {
    cleanup();
    smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (binning)");
    return SMH_ERR_HIP;
  }
  hipMemsetAsync(d_cnt, 0, (nkeys + 1) * sizeof(int), s);
  hipMemsetAsync(d_cur, 0, (nkeys + 1) * sizeof(int), s);
  hipMemsetAsync(d_err, 0, sizeof(int), s);
  int64_t max_seg = 0;
  for (int sg = 0; sg < op->n_seg; ++sg) max_seg = std::max<int64_t>(max_seg, op->seg_off[sg + 1] - op->seg_off[sg]);
  const dim3 grid((unsigned)n_part, (unsigned)((max_seg + 255) / 256));
  if (grid.y > 65535) {
    cleanup();
    return smh_fail(ctx, SMH_ERR_ARG, "ensemble too large for the pair binning");
  }
  hipLaunchKernelGGL(pair_bin_kernel, grid, dim3(256), 0, s, op->d_xy, (int)n_part, op->rho_min, op->h, (int)len_rho,
                     op->s_mode, op->o_mode, op->T, op->n_tt, Rc, (const int*)op->d_seg_of,
                     (const int64_t*)op->d_seg_off, d_cnt, (const int*)nullptr, d_cur, (uint2*)nullptr, d_err);
  std::vector<int> cnt(nkeys);
  int herr = 0;
  if (hipMemcpyAsync(cnt.data(), d_cnt, nkeys * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    cleanup();
    smh_fail(ctx, SMH_ERR_HIP, "binning count failed: %s", hipGetErrorString(hipGetLastError()));
    return SMH_ERR_HIP;
  }
  if (herr) {
    cleanup();
    smh_fail(ctx, SMH_ERR_ARG, "a particle distance lies outside the lookup table (rho_min %g, h %g, %lld rows)",
             op->rho_min, op->h, (long long)len_rho);
    return SMH_ERR_ARG;
  }
  std::vector<int> off(nkeys + 1);
  int64_t acc = 0;
  for (int64_t k2 = 0; k2 < nkeys; ++k2) {
    off[k2] = (int)acc;
    acc += cnt[k2];
    if (acc >= (int64_t)INT32_MAX) {
      cleanup();
      smh_fail(ctx, SMH_ERR_UNSUPPORTED, "too many particle pairs for 32-bit offsets");
      return SMH_ERR_UNSUPPORTED;
    }
  }
  off[nkeys] = (int)acc;
  op->n_pairs = acc;
  if (hipMalloc(&d_pairs, std::max<int64_t>(acc, 1) * sizeof(uint2)) != hipSuccess ||
      hipMemcpyAsync(d_off, off.data(), (nkeys + 1) * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess) {
    cleanup();
    smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (pairs)");
    return SMH_ERR_HIP;
  }
  hipLaunchKernelGGL(pair_bin_kernel, grid, dim3(256), 0, s, op->d_xy, (int)n_part, op->rho_min, op->h, (int)len_rho,
                     op->s_mode, op->o_mode, op->T, op->n_tt, Rc, (const int*)op->d_seg_of,
                     (const int64_t*)op->d_seg_off, d_cnt, (const int*)d_off, d_cur, d_pairs, d_err);
  if (hipMalloc(&bn.d_rec, std::max<int64_t>(acc, 1) * sizeof(PairRec<Real>)) != hipSuccess) {
    cleanup();
    smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (pair records, %lld pairs)", (long long)acc);
    return SMH_ERR_HIP;
  }
  if (acc > 0)
    hipLaunchKernelGGL(pair_record_kernel<Real>, dim3((unsigned)((acc + 255) / 256)), dim3(256), 0, s, d_pairs, acc,
                       op->d_xy, op->rho_min, op->h, Rc, (PairRec<Real>*)bn.d_rec);
  if (hipStreamSynchronize(s) != hipSuccess) {
    cleanup();
    smh_fail(ctx, SMH_ERR_HIP, "pair records failed");
    return SMH_ERR_HIP;
  }
  cleanup();

  return build_work_items<PairRec<Real>>(
      op, bn, cnt, off, len_rho, [&](int64_t r) { return (int)(op->s_mode * r + op->o_mode); }, batch);
}

// Operator over n_ens ensembles (particles [ens_off[e], ens_off[e+1])), or one
// ensemble when ens_off is null.
static int radial_create(smh_ctx* ctx, int n_ens, const int64_t* ens_off, int64_t n_part, const double* xy,
                         const int* lmax, const int* mmax, int l_max, int m_max, double rho_min, double resolution,
                         int64_t len_rho, int interp, smh_op** out) {
  if (!ctx || !out) return SMH_ERR_ARG;
  *out = nullptr;
  SMH_CHECK(ctx, n_part >= 1 && xy && lmax && mmax, SMH_ERR_ARG, "need at least one particle");
  SMH_CHECK(ctx, n_part <= (int64_t)kPairMask, SMH_ERR_ARG, "too many particles (pair records index 23 bits)");
  SMH_CHECK(ctx, l_max >= 1 && m_max >= 1 && m_max <= l_max, SMH_ERR_ARG, "invalid l_max %d / m_max %d", l_max,
            m_max);
  SMH_CHECK(ctx, resolution > 0 && len_rho >= 4 && len_rho < (1 << 30), SMH_ERR_ARG, "invalid table grid");
  SMH_CHECK(ctx, interp >= 0 && interp <= 2, SMH_ERR_ARG, "invalid interpolation kind %d", interp);
  SMH_CHECK(ctx, radial_kernel(SMH_FP32, 4, l_max, m_max).fn != nullptr, SMH_ERR_UNSUPPORTED,
            "no kernel for l_max %d m_max %d", l_max, m_max);
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  smh_op* op = new smh_op();
  op->ctx = ctx;
  op->kind = 0;
  op->n_part = n_part;
  op->l_max = l_max;
  op->m_max = m_max;
  op->B = smh_blocksize(l_max, m_max);
  op->Bh = (op->B + 15) / 16 * 16;
  op->MP = 2 * op->Bh;
  op->interp = interp;
  op->rho_min = rho_min;
  op->h = resolution;
  op->len_rho = len_rho;
  if (interp == SMH_INTERP_LINEAR) {
    op->T = 2;
    op->s_mode = 1;
    op->o_mode = 0;
    op->n_tt = len_rho;
  } else if (interp == SMH_INTERP_CATMULL_ROM) {
    op->T = 4;
    op->s_mode = 1;
    op->o_mode = -1;
    op->n_tt = len_rho;
  } else {
    op->T = 4;
    op->s_mode = 2;
    op->o_mode = 0;
    op->n_tt = 2 * len_rho;
  }
  int rc = SMH_OK;
  auto fail = [&](int code) {
    smh_op_free(op);
    delete op;
    return code;
  };
  if (int code = op_index_map(op, lmax, mmax)) return fail(code);
  if (ens_off)
    if (int code = op_set_segments(op, n_ens, ens_off)) return fail(code);
  if (hipMalloc(&op->d_xy, n_part * sizeof(double2)) != hipSuccess) {
    smh_fail(ctx, SMH_ERR_HIP, "device allocation failed");
    return fail(SMH_ERR_HIP);
  }
  hipStream_t s = ctx->stream;
  std::vector<double2> hxy(n_part);
  for (int64_t i = 0; i < n_part; ++i) hxy[i] = make_double2(xy[2 * i], xy[2 * i + 1]);
  if (hipMemcpyAsync(op->d_xy, hxy.data(), n_part * sizeof(double2), hipMemcpyHostToDevice, s) != hipSuccess) {
    smh_fail(ctx, SMH_ERR_HIP, "upload failed");
    return fail(SMH_ERR_HIP);
  }

  // ---- pair binning per precision
  size_t ypart = 0;
  for (int pi = 0; pi < 2; ++pi) {
    const KernelInfo k = radial_kernel(pi == 0 ? SMH_FP32 : SMH_FP64, op->T, l_max, m_max);
    const size_t rs = pi == 0 ? sizeof(float) : sizeof(double);
    size_t lds = 0;
    const int Rc = chunk_size(op, rs, k.sy, lds);
    if (Rc == 0) continue;  // no kernel of this precision fits; matvecs in it report UNSUPPORTED
    op->bin[pi].lds = lds;
    const int code =
        pi == 0 ? bin_pairs<float>(op, op->bin[0], Rc, false) : bin_pairs<double>(op, op->bin[1], Rc, false);
    if (code) return fail(code);
    op->bin[pi].part_bytes = (size_t)Rc * 2 * op->B * rs;
    ypart = std::max(ypart, (size_t)op->bin[pi].n_items * op->bin[pi].part_bytes);
    // the same for the kBatch-RHS kernel, whose accumulator holds kBatch vectors per receiver
    const KernelInfo kb = radial_batch_kernel(pi == 0 ? SMH_FP32 : SMH_FP64, op->T, l_max, m_max);
    size_t ldsb = 0;
    const int Rb = kb.fn ? chunk_size(op, rs, kb.sy, ldsb) : 0;
    if (Rb == 0) continue;
    op->bbin[pi].lds = ldsb;
    const int codeb =
        pi == 0 ? bin_pairs<float>(op, op->bbin[0], Rb, true) : bin_pairs<double>(op, op->bbin[1], Rb, true);
    if (codeb) return fail(codeb);
    const int Sb = kb.slots > 0 ? kb.slots : op->B;
    if (pi == 0) {
      op->b32_S = Sb;
      if (kb.slots > 0) {
        std::vector<int> tab(op->B + Sb);
        kb.slot_tables(tab.data());
        if (hipMalloc(&op->d_b32_tab, tab.size() * sizeof(int)) != hipSuccess ||
            hipMemcpy(op->d_b32_tab, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
          smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (slot tables)");
          return fail(SMH_ERR_HIP);
        }
      }
    }
    op->bbin[pi].part_bytes = (size_t)Rb * 2 * kBatch * Sb * rs;
    ypart = std::max(ypart, (size_t)op->bbin[pi].n_items * op->bbin[pi].part_bytes);
  }
  op->ypart_bytes = std::max<size_t>(ypart, 16);
  const int64_t n_pad = n_part * op->B;
  if (hipMalloc(&op->d_ypart, op->ypart_bytes) != hipSuccess ||
      hipMalloc(&op->d_x32, n_pad * sizeof(float2)) != hipSuccess ||
      hipMalloc(&op->d_y32, n_pad * sizeof(float2)) != hipSuccess ||
      hipMalloc(&op->d_x64, n_pad * sizeof(double2)) != hipSuccess ||
      hipMalloc(&op->d_y64, n_pad * sizeof(double2)) != hipSuccess) {
    smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (vectors)");
    return fail(SMH_ERR_HIP);
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    smh_fail(ctx, SMH_ERR_HIP, "setup failed");
    return fail(SMH_ERR_HIP);
  }
  (void)rc;
  ctx->live_ops++;
  *out = op;
  return SMH_OK;
}

extern "C" int smh_radial_create(smh_ctx* ctx, int64_t n_part, const double* xy, const int* lmax, const int* mmax,
                                 int l_max, int m_max, double rho_min, double resolution, int64_t len_rho,
                                 int interp, smh_op** out) {
  return radial_create(ctx, 1, nullptr, n_part, xy, lmax, mmax, l_max, m_max, rho_min, resolution, len_rho, interp,
                       out);
}

extern "C" int smh_radial_create_ensembles(smh_ctx* ctx, int n_ens, const int64_t* ens_off, const double* xy,
                                           const int* lmax, const int* mmax, int l_max, int m_max, double rho_min,
                                           double resolution, int64_t len_rho, int interp, smh_op** out) {
  if (!ctx || !out) return SMH_ERR_ARG;
  SMH_CHECK(ctx, n_ens >= 1 && ens_off, SMH_ERR_ARG, "need at least one ensemble");
  return radial_create(ctx, n_ens, ens_off, ens_off[n_ens], xy, lmax, mmax, l_max, m_max, rho_min, resolution, len_rho,
                       interp, out);
}

extern "C" int smh_op_destroy(smh_op* op) {
  if (!op) return SMH_OK;
  smh_ctx* ctx = op->ctx;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  smh_op_free(op);
  delete op;
  ctx->live_ops--;
  smh_ctx_release_if_idle(ctx);
  return SMH_OK;
}

extern "C" int smh_op_upload_table(smh_op* op, const double* table, const double* table_minus) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, table != nullptr, SMH_ERR_ARG, "table is NULL");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  SMH_CHECK(ctx, op->kind != 2, SMH_ERR_ARG, "the explicit operator builds its matrix itself");
  if (op->kind == 1) {
    SMH_CHECK(ctx, table_minus != nullptr, SMH_ERR_ARG, "volume operator needs the minus table");
    return vol_upload(op, table, table_minus);
  }
  const size_t bytes = (size_t)op->len_rho * op->B * op->B * sizeof(double2);
  if (!op->d_table) SMH_HIP(ctx, hipMalloc(&op->d_table, bytes));
  SMH_HIP(ctx, hipMemcpyAsync(op->d_table, table, bytes, hipMemcpyHostToDevice, ctx->stream));
  return op_make_taps(op);
}

extern "C" int smh_op_download_table(smh_op* op, double* table, double* table_minus) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  if (op->kind == 1) return vol_download(op, table, table_minus);
  if (op->kind == 2) return explicit_download(op, table);
  SMH_CHECK(ctx, op->d_table != nullptr, SMH_ERR_STATE, "no table");
  const size_t bytes = (size_t)op->len_rho * op->B * op->B * sizeof(double2);
  SMH_HIP(ctx, hipMemcpyAsync(table, op->d_table, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SMH_OK;
}

extern "C" int smh_op_matvec(smh_op* op, const double* x, double* y, int64_t nrhs, int precision) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, x && y && nrhs >= 1, SMH_ERR_ARG, "invalid matvec arguments");
  SMH_CHECK(ctx, precision == SMH_FP32 || precision == SMH_FP64 || precision == SMH_FP16, SMH_ERR_ARG,
            "invalid precision %d", precision);
  SMH_CHECK(ctx, precision != SMH_FP16 || op_has_batch(op, SMH_FP16), SMH_ERR_UNSUPPORTED,
            "FP16 matvecs exist for radial operators with the slot-order kernel only");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  int rc = op_reserve_packed(op, nrhs);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  SMH_HIP(ctx, hipMemcpyAsync(op->d_pk_in, x, op->n_unk * nrhs * sizeof(double2), hipMemcpyHostToDevice, s));
  if ((nrhs > 1 || precision == SMH_FP16) && op_has_batch(op, precision)) {
    // blocks of kBatch right-hand sides through the batched kernel
    if ((rc = op_reserve_batch(op))) return rc;
    for (int64_t r0 = 0; r0 < nrhs; r0 += kBatch) {
      const int nb = (int)std::min<int64_t>(kBatch, nrhs - r0);
      if ((rc = op_unpack_batch(op, precision, op->d_pk_in + r0 * op->n_unk, nb, op->d_xb))) return rc;
      if ((rc = op_apply_batch(op, precision, op->d_xb, op->d_yb, false))) return rc;
      if ((rc = op_pack_batch(op, precision, op->d_yb, nb, op->d_pk_out + r0 * op->n_unk))) return rc;
    }
  } else {
    void* xp = precision == SMH_FP32 ? (void*)op->d_x32 : (void*)op->d_x64;
    void* yp = precision == SMH_FP32 ? (void*)op->d_y32 : (void*)op->d_y64;
    for (int64_t r = 0; r < nrhs; ++r) {
      if ((rc = op_unpack(op, precision, op->d_pk_in + r * op->n_unk, xp))) return rc;
      if ((rc = op_apply(op, precision, xp, yp, false))) return rc;
      if ((rc = op_pack(op, precision, yp, op->d_pk_out + r * op->n_unk))) return rc;
    }
  }
  SMH_HIP(ctx, hipMemcpyAsync(y, op->d_pk_out, op->n_unk * nrhs * sizeof(double2), hipMemcpyDeviceToHost, s));
  SMH_HIP(ctx, hipStreamSynchronize(s));
  return SMH_OK;
}

namespace {
// Sum the partials of items [k0, k1) of every receiver chunk onto y, in item
// order (partial of item w at ypart[w - k0]); with y zero before the first
// range, a sequence of ranges adds exactly what reduce_kernel adds at once.
template <typename Real>
__global__ void reduce_range_kernel(const Real* __restrict__ ypart, const int* __restrict__ chunk_items, int k0, int k1,
                                    int Rc, int B, int64_t n_part, typename Vec2<Real>::t* y) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_part * B) return;
  const int64_t i1 = j / B;
  const int n1 = (int)(j - i1 * B);
  const int c = (int)(i1 / Rc), il = (int)(i1 - (int64_t)c * Rc);
  const int w0 = max(chunk_items[c], k0), w1 = min(chunk_items[c + 1], k1);
  if (w0 >= w1) return;
  Real re = y[j].x, im = y[j].y;
  for (int w = w0; w < w1; ++w) {
    const Real* p = ypart + ((size_t)(w - k0) * Rc + il) * 2 * B;
    re += p[n1];
    im += p[B + n1];
  }
  y[j] = typename Vec2<Real>::t{re, im};
}

// Synthetic input: n entries into whichever of x32 / x64 is given.
__global__ void fill_pattern_kernel(float2* x32, double2* x64, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) {
    const double a = sin(0.37 * (double)j + 0.1), b = cos(0.11 * (double)j);
    if (x32) x32[j] = make_float2((float)a, (float)b);
    if (x64) x64[j] = make_double2(a, b);
  }
}
}  // namespace

extern "C" int smh_op_half_interp(smh_op* op, int* terms, double* lin_ratio) {
  if (!op) return SMH_ERR_ARG;
  if (terms) *terms = op->T_half;
  if (lin_ratio) *lin_ratio = (double)op->lin_ratio;
  return SMH_OK;
}

extern "C" int smh_op_bench_matvec_batch(smh_op* op, int precision, int iters, double* kernel_ms,
                                         double* total_ms) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, iters >= 1, SMH_ERR_ARG, "iters must be >= 1");
  SMH_CHECK(ctx, op_has_batch(op, precision), SMH_ERR_UNSUPPORTED, "no batched matvec for this operator");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  int rc = op_reserve_batch(op);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  // synthetic input in the precision's layout (the other buffer is scratch for the unused copy;
  // the FP16 input is converted from an FP64 pattern in the output buffer)
  const int64_t n = op->n_part * (precision == SMH_FP64 ? op->B : op->b32_S) * kBatch;
  if (precision == SMH_FP16) {
    const int64_t n64 = op->n_part * op->B * kBatch;
    hipLaunchKernelGGL(fill_pattern_kernel, dim3(blocks_for(n64)), dim3(256), 0, s, (float2*)nullptr,
                       (double2*)op->d_yb, n64);
    if ((rc = op_to_half_batch(op, (const double2*)op->d_yb, op->d_xb))) return rc;
  } else if (precision == SMH_FP32)
    hipLaunchKernelGGL(fill_pattern_kernel, dim3(blocks_for(n)), dim3(256), 0, s, (float2*)op->d_xb,
                       (double2*)nullptr, n);
  else
    hipLaunchKernelGGL(fill_pattern_kernel, dim3(blocks_for(n)), dim3(256), 0, s, (float2*)nullptr,
                       (double2*)op->d_xb, n);
  if ((rc = op_apply_batch(op, precision, op->d_xb, op->d_yb, false))) return rc;  // warm-up
  SMH_HIP(ctx, hipStreamSynchronize(s));
  double ksum = 0;
  hipEvent_t t0, t1;
  SMH_HIP(ctx, hipEventCreate(&t0));
  SMH_HIP(ctx, hipEventCreate(&t1));
  SMH_HIP(ctx, hipEventRecord(t0, s));
  for (int i = 0; i < iters; ++i) {
    if ((rc = op_apply_batch(op, precision, op->d_xb, op->d_yb, true))) return rc;
    ksum += op->last_kernel_ms;
  }
  SMH_HIP(ctx, hipEventRecord(t1, s));
  SMH_HIP(ctx, hipEventSynchronize(t1));
  float tot = 0;
  SMH_HIP(ctx, hipEventElapsedTime(&tot, t0, t1));
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  if (kernel_ms) *kernel_ms = ksum / iters;
  if (total_ms) *total_ms = tot / iters;
  return SMH_OK;
}

extern "C" int smh_op_bench_matvec(smh_op* op, int precision, int iters, double* kernel_ms, double* total_ms) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, iters >= 1, SMH_ERR_ARG, "iters must be >= 1");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t n = op->n_part * op->B;
  hipLaunchKernelGGL(fill_pattern_kernel, dim3(blocks_for(n)), dim3(256), 0, s, op->d_x32, op->d_x64, n);
  void* xp = precision == SMH_FP32 ? (void*)op->d_x32 : (void*)op->d_x64;
  void* yp = precision == SMH_FP32 ? (void*)op->d_y32 : (void*)op->d_y64;
  int rc = op_apply(op, precision, xp, yp, false);  // warm-up
  if (rc) return rc;
  SMH_HIP(ctx, hipStreamSynchronize(s));
  double ksum = 0;
  hipEvent_t t0, t1;
  SMH_HIP(ctx, hipEventCreate(&t0));
  SMH_HIP(ctx, hipEventCreate(&t1));
  SMH_HIP(ctx, hipEventRecord(t0, s));
  for (int i = 0; i < iters; ++i) {
    if ((rc = op_apply(op, precision, xp, yp, true))) return rc;
    ksum += op->last_kernel_ms;
  }
  SMH_HIP(ctx, hipEventRecord(t1, s));
  SMH_HIP(ctx, hipEventSynchronize(t1));
  float tot = 0;
  SMH_HIP(ctx, hipEventElapsedTime(&tot, t0, t1));
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  if (kernel_ms) *kernel_ms = ksum / iters;
  if (total_ms) *total_ms = tot / iters;
  return SMH_OK;
}

extern "C" int smh_op_matvec_noalloc(smh_op* op, const double* x, double* y, int precision) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, x && y, SMH_ERR_ARG, "invalid matvec arguments");
  SMH_CHECK(ctx, op->kind == 0, SMH_ERR_UNSUPPORTED, "the no-allocation matvec covers radial lookup operators");
  SMH_CHECK(ctx, precision == SMH_FP32 || precision == SMH_FP64, SMH_ERR_ARG, "invalid precision %d", precision);
  SMH_CHECK(ctx, op->has_table, SMH_ERR_STATE, "coupling table not set");
  SMH_CHECK(ctx, ctx->scratch != nullptr, SMH_ERR_STATE, "no scratch address on this context");
  const int pi = precision == SMH_FP32 ? 0 : 1;
  const smh_op::Binning& bn = op->bin[pi];
  SMH_CHECK(ctx, bn.Rc > 0, SMH_ERR_UNSUPPORTED, "no %s kernel fits LDS for blocksize %d", pi == 0 ? "FP32" : "FP64",
            op->B);
  const KernelInfo k = radial_kernel(precision, op->T, op->l_max, op->m_max);
  SMH_CHECK(ctx, k.fn != nullptr, SMH_ERR_UNSUPPORTED, "no matvec kernel for l_max %d m_max %d", op->l_max,
            op->m_max);
  // scratch layout: input bound (xmax), padded input, padded output, the
  // partials of as many work items as fit (launched in ranges of that size)
  const size_t rs = pi == 0 ? sizeof(float) : sizeof(double);
  const int64_t n_pad = op->n_part * op->B;
  const size_t head = 4096, vec = ((size_t)n_pad * 2 * rs + 255) / 256 * 256;
  const size_t per_item = (size_t)bn.Rc * 2 * op->B * rs;
  SMH_CHECK(ctx, head + 2 * vec + per_item <= kScratchBytes, SMH_ERR_ARG,
            "operator too large for the fixed scratch: %zu bytes needed (%lld padded unknowns, %zu per work item), "
            "%zu available",
            head + 2 * vec + per_item, (long long)n_pad, per_item, kScratchBytes);
  const int range = (int)std::min<size_t>((kScratchBytes - head - 2 * vec) / per_item, (size_t)1 << 20);
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  char* base = ctx->scratch;
  float* xmax = reinterpret_cast<float*>(base);
  void* xp = base + head;
  void* yp = base + head + vec;
  void* part = base + head + 2 * vec;
  fill_zero(xp, (size_t)n_pad * 2 * rs, s);
  fill_zero(yp, (size_t)n_pad * 2 * rs, s);
  if (pi == 0)
    hipLaunchKernelGGL(scatter_kernel<float>, dim3(blocks_for(op->n_unk)), dim3(256), 0, s, (const double2*)x,
                       op->d_pk2pad, op->n_unk, (float2*)xp);
  else
    hipLaunchKernelGGL(scatter_kernel<double>, dim3(blocks_for(op->n_unk)), dim3(256), 0, s, (const double2*)x,
                       op->d_pk2pad, op->n_unk, (double2*)xp);
  if (pi == 0) launch_xmax((const float2*)xp, n_pad, xmax, s);
  if (int rc_ = op_set_lds(op, k.fn, bn.lds)) return rc_;
  for (int k0 = 0; k0 < bn.n_items; k0 += range) {
    const int k1 = std::min(bn.n_items, k0 + range);
    RadialKArgs a;
    a.recs = bn.d_rec;
    a.entries = bn.d_entries;
    a.items = bn.d_items;
    a.order = nullptr;
    a.item0 = k0;
    a.tt = pi == 0 ? (const void*)op->d_tt32 : (const void*)op->d_tt64;
    a.x = xp;
    a.ypart = part;
    a.xmax = xmax;
    a.item_lsum = bn.d_item_lsum;
    a.item_en = bn.d_item_en;
    a.n_items = bn.n_items;
    a.n_entries = bn.n_entries;
    a.interp = op->interp;
    a.Rc = bn.Rc;
    void* args[] = {&a};
    SMH_HIP(ctx, hipLaunchKernel(k.fn, dim3(k1 - k0), dim3(k.threads), args, bn.lds, s));
    if (pi == 0)
      hipLaunchKernelGGL(reduce_range_kernel<float>, dim3(blocks_for(n_pad)), dim3(256), 0, s, (const float*)part,
                         bn.d_chunk_items, k0, k1, bn.Rc, op->B, op->n_part, (float2*)yp);
    else
      hipLaunchKernelGGL(reduce_range_kernel<double>, dim3(blocks_for(n_pad)), dim3(256), 0, s, (const double*)part,
                         bn.d_chunk_items, k0, k1, bn.Rc, op->B, op->n_part, (double2*)yp);
  }
  if (pi == 0)
    hipLaunchKernelGGL(pack_kernel<float>, dim3(blocks_for(op->n_unk)), dim3(256), 0, s, (const float2*)yp,
                       op->d_pk2pad, op->n_unk, (double2*)y);
  else
    hipLaunchKernelGGL(pack_kernel<double>, dim3(blocks_for(op->n_unk)), dim3(256), 0, s, (const double2*)yp,
                       op->d_pk2pad, op->n_unk, (double2*)y);
  SMH_HIP(ctx, hipGetLastError());
  SMH_HIP(ctx, hipStreamSynchronize(s));
  return SMH_OK;
}
