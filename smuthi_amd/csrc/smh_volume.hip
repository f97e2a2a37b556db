// Volume-lookup coupling operator (CouplingMatrixVolumeLookupCUDA,
// linearsystem/linear_system.py:688-783; tables of prepare_lookup.py:16-250):
// particles at arbitrary heights in one layer. The coupling block of a pair
// is Lpl(rho, z1 + z2) + Lmn(rho, z1 - z2), both complex64 tables on a
// uniform (rho, z) grid of one resolution, bilinearly interpolated.
//
// Table build (FP64, stored complex64 as the reference stores them):
//   Lpl[r, j] = float( sum_k J_|dm|(kappa_k rho_r) (c+pl_k e^{i kz (sz_j - 2 z_s)} + c-pl_k e^{-i kz (sz_j - 2 z_s)}) )
//   Lmn[r, j] = float(direct(rho_r, dz_j)) + float( same with c+-mn and e^{+-i kz dz_j} )
//   direct(rho, dz) = sum_ld coef[n1, n2, ld] h_ld(k d) P_ld^|dm|(dz / d),  d = sqrt(rho^2 + dz^2),
// zero for rho below the mask radius; the contour weights and layer response
// are folded into c+-pl/mn on the host.
//
// Matvec: the binned MFMA lookup kernel of the radial operator, once per
// table, with bilinear taps (r, j), (r, j+1), (r+1, j), (r+1, j+1): pairs are
// binned by (receiver chunk, r, j) so every bin entry shares its four tap
// rows. The FP64 kernel reads the same FP32 tap table (complex64 data, exact).
#include <cmath>

#include "smh_batch_slots.h"
#include "smh_batch_slots64.h"
#include "smh_batch_half.h"
#include "smh_lookup.h"
#include "smh_special.h"

namespace {

constexpr int kMaxL2 = 16;  // 2 l_max <= 16

// Direct part of the minus table, rounded to complex64; one block per node.
__global__ void vol_direct_kernel(double rho_min, double dz_min, double h, int64_t len_rho, int64_t len_z, int B,
                                  int L2, double kre, double kim, double mask, const double2* __restrict__ coef,
                                  const int* __restrict__ dm_of_p, float2* table) {
  const int64_t node = blockIdx.x;
  const int64_t r = node / len_z, j = node - r * len_z;
  const double rho = rho_min + (double)r * h, dz = dz_min + (double)j * h;
  const int BB = B * B;
  float2* row = table + node * BB;
  if (rho < mask) {
    for (int p = threadIdx.x; p < BB; p += blockDim.x) row[p] = make_float2(0.f, 0.f);
    return;
  }
  __shared__ zc hl[kMaxL2 + 1];
  __shared__ double P[(kMaxL2 + 1) * (kMaxL2 + 1)];
  if (threadIdx.x == 0) {
    double d = sqrt(rho * rho + dz * dz);
    if (d == 0.0) d = 1e-20;
    smh::spherical_hankel_all(L2, zc{kre * d, kim * d}, hl);
    smh::legendre_p(L2, dz / d, rho / d, P);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < BB; p += blockDim.x) {
    const int dm = dm_of_p[p];
    zc s{0.0, 0.0};
    for (int ld = dm; ld <= L2; ++ld) {
      const double2 c = coef[(size_t)p * (L2 + 1) + ld];
      const double pl = P[ld * (L2 + 1) + dm];
      s = smh::zadd(s, smh::zmul(zc{c.x * pl, c.y * pl}, hl[ld]));
    }
    row[p] = make_float2((float)s.re, (float)s.im);
  }
}

// Layer-mediated parts of both tables. Block: one rho node and kZT z nodes;
// per chunk of kappa the Bessel values and the z phases are staged in LDS,
// each thread owns one (n1, n2) entry. plus += float(layer), minus = float(minus) + float(layer).
constexpr int kZT = 8;
constexpr int kKC = 16;
constexpr int kNJ = 9;
__global__ void vol_layer_kernel(double rho_min, double sz_min, double dz_min, double h, int64_t len_z, int B,
                                 int nmax, double z_s, const int* __restrict__ dm_of_p, int64_t n_kappa,
                                 const double2* __restrict__ kappa, const double2* __restrict__ kz,
                                 const double2* __restrict__ c, float2* plus, float2* minus) {
  __shared__ zc J[kKC][kNJ];
  __shared__ zc E[kKC][4][kZT];  // e^{+i kz (sz - 2 z_s)}, e^{-i ...}, e^{+i kz dz}, e^{-i kz dz}
  const int64_t r = blockIdx.x;
  const int64_t j0 = (int64_t)blockIdx.y * kZT;
  const int nz = (int)min<int64_t>(kZT, len_z - j0);
  const double rho = rho_min + (double)r * h;
  const int BB = B * B;
  for (int p0 = 0; p0 < BB; p0 += blockDim.x) {
    const int p = p0 + threadIdx.x;
    const bool act = p < BB;
    const int dm = act ? dm_of_p[p] : 0;
    zc apl[kZT], amn[kZT];
    for (int z = 0; z < kZT; ++z) apl[z] = amn[z] = zc{0.0, 0.0};
    for (int64_t k0 = 0; k0 < n_kappa; k0 += kKC) {
      const int kc = (int)min<int64_t>(kKC, n_kappa - k0);
      __syncthreads();
      for (int e = threadIdx.x; e < kKC; e += blockDim.x) {
        if (e < kc) {
          zc buf[smh::kMaxBesselOrder + 1];
          const double2 kp = kappa[k0 + e];
          smh::bessel_j_all(nmax, zc{kp.x * rho, kp.y * rho}, buf);
          for (int n = 0; n <= nmax; ++n) J[e][n] = buf[n];
        }
      }
      for (int e = threadIdx.x; e < kKC * kZT; e += blockDim.x) {
        const int kk = e / kZT, z = e - kk * kZT;
        if (kk < kc && z < nz) {
          const double2 q = kz[k0 + kk];
          const double sz = sz_min + (double)(j0 + z) * h - 2.0 * z_s, dz = dz_min + (double)(j0 + z) * h;
          // e^{i q t} = e^{-Im q t} (cos(Re q t) + i sin(Re q t))
          auto cis = [&](double t, double sgn) {
            const double m = exp(-sgn * q.y * t);
            double sn, cn;
            sincos(sgn * q.x * t, &sn, &cn);
            return zc{m * cn, m * sn};
          };
          E[kk][0][z] = cis(sz, 1.0);
          E[kk][1][z] = cis(sz, -1.0);
          E[kk][2][z] = cis(dz, 1.0);
          E[kk][3][z] = cis(dz, -1.0);
        }
      }
      __syncthreads();
      if (act) {
        for (int kk = 0; kk < kc; ++kk) {
          const size_t base = ((size_t)(k0 + kk) * 4) * BB + p;
          const double2 cp = c[base], cm = c[base + BB], dp = c[base + 2 * (size_t)BB], dn = c[base + 3 * (size_t)BB];
          const zc jj = J[kk][dm];
          const zc jcp = smh::zmul(jj, zc{cp.x, cp.y}), jcm = smh::zmul(jj, zc{cm.x, cm.y});
          const zc jdp = smh::zmul(jj, zc{dp.x, dp.y}), jdn = smh::zmul(jj, zc{dn.x, dn.y});
          for (int z = 0; z < kZT; ++z) {
            apl[z] = smh::zadd(apl[z], smh::zadd(smh::zmul(jcp, E[kk][0][z]), smh::zmul(jcm, E[kk][1][z])));
            amn[z] = smh::zadd(amn[z], smh::zadd(smh::zmul(jdp, E[kk][2][z]), smh::zmul(jdn, E[kk][3][z])));
          }
        }
      }
    }
    if (act)
      for (int z = 0; z < nz; ++z) {
        const size_t o = ((size_t)r * len_z + j0 + z) * BB + p;
        plus[o] = make_float2((float)apl[z].re, (float)apl[z].im);
        const float2 d = minus[o];
        minus[o] = make_float2(d.x + (float)amn[z].re, d.y + (float)amn[z].im);
      }
  }
}

// FP32 tap table in fragment order (T = 4, one lane quarter per tap):
// row j holds F[i][g][rr][e] = A_j(16 i + rr, n2 = 4 g + e), A rows [0, Bh) = Re, [Bh, 2 Bh) = Im.
__global__ void vol_taps_kernel(const float2* __restrict__ table, int64_t rows, int B, int Bh, int NG, int FR,
                                float* tt) {
  const int64_t total = rows * FR;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = f / FR;
    int rem = (int)(f - row * FR);
    const int e = rem & 3, rr = (rem >> 2) & 15;
    rem >>= 6;
    const int g = rem % NG, i = rem / NG;
    const int m = 16 * i + rr, n2 = 4 * g + e;
    float v = 0.f;
    if (n2 < B) {
      if (m < B)
        v = table[(row * B + m) * B + n2].x;
      else if (m >= Bh && m < Bh + B)
        v = table[(row * B + (m - Bh)) * B + n2].y;
    }
    tt[f] = v;
  }
}

// Not-a-knot spline moments Mh = h^2/6 S'' along one axis of a complex array
// [outer][n][inner] (FP64), one thread per (outer, inner) column: the
// tridiagonal system of the uniform-grid spline (rows 1 and n-2 closed by the
// not-a-knot conditions, Thomas algorithm with c' from the host), as the
// radial lookup's spline_kernel.
template <typename Src>
__global__ void spline_axis_kernel(const Src* __restrict__ y, const double* __restrict__ cprime, int64_t n,
                                   int64_t inner, int64_t outer, double2* mh) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= inner * outer) return;
  const int64_t o = c / inner, ci = c - o * inner;
  const int64_t base = o * n * inner + ci;
  auto Y = [&](int64_t j) {
    const Src v = y[base + j * inner];
    return make_double2((double)v.x, (double)v.y);
  };
  auto M = [&](int64_t j) -> double2& { return mh[base + j * inner]; };
  auto rhs = [&](int64_t j) {
    const double2 a = Y(j - 1), b = Y(j), d = Y(j + 1);
    return make_double2(a.x - 2.0 * b.x + d.x, a.y - 2.0 * b.y + d.y);
  };
  const double2 r1 = rhs(1), rn2 = rhs(n - 2);
  const double2 m1 = make_double2(r1.x / 6.0, r1.y / 6.0);
  const double2 mn2 = make_double2(rn2.x / 6.0, rn2.y / 6.0);
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
    M(j) = dp;
    dprev = dp;
  }
  double2 next = M(n - 3);
  for (int64_t j = n - 4; j >= 2; --j) {
    const double2 dp = M(j);
    const double2 v = make_double2(dp.x - cprime[j] * next.x, dp.y - cprime[j] * next.y);
    M(j) = v;
    next = v;
  }
  const double2 m2 = M(2), mn3 = M(n - 3);
  M(1) = m1;
  M(n - 2) = mn2;
  M(0) = make_double2(2.0 * m1.x - m2.x, 2.0 * m1.y - m2.y);
  M(n - 1) = make_double2(2.0 * mn2.x - mn3.x, 2.0 * mn2.y - mn3.y);
}

// Tensor-spline tap tables: row 4 node + q holds quantity q (0 value,
// 1 rho moment, 2 z moment, 3 mixed moment) of the node in fragment order
// (as vol_taps_kernel), FP32 and FP64.
__global__ void vol_taps4_kernel(const float2* __restrict__ table, const double2* __restrict__ mr,
                                 const double2* __restrict__ mz, const double2* __restrict__ mrz, int64_t nodes, int B,
                                 int Bh, int NG, int FR, float* tt32, double* tt64) {
  const int64_t total = nodes * 4 * FR;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = f / FR;
    int rem = (int)(f - row * FR);
    const int e = rem & 3, rr = (rem >> 2) & 15;
    rem >>= 6;
    const int g = rem % NG, i = rem / NG;
    const int m = 16 * i + rr, n2 = 4 * g + e;
    const int64_t node = row >> 2;
    const int q = (int)(row & 3);
    double v = 0.0;
    if (n2 < B && (m < B || (m >= Bh && m < Bh + B))) {
      const int n1 = m < B ? m : m - Bh;
      const int64_t o = (node * B + n1) * B + n2;
      double2 c;
      if (q == 0) {
        const float2 t = table[o];
        c = make_double2((double)t.x, (double)t.y);
      } else {
        c = q == 1 ? mr[o] : (q == 2 ? mz[o] : mrz[o]);
      }
      v = m < B ? c.x : c.y;
    }
    tt32[f] = (float)v;
    tt64[f] = v;
  }
}

// The tap rows in the slot fragment order of batch_slots_kernel<..., VOL>:
// [rows][MT][NS/4][64 lanes][4] FP32, lane (q, rr) of tile i holding
// A(slot 16 (i mod MH) + rr, slot 4 (4 s4 + e) + q) of the row (tiles i < MH
// from Re, the others from Im); row = node * vq + kind, kind 0 the table
// value, 1..3 (spline) the moments along rho, along z and mixed.
__global__ void vol_slot_taps_kernel(const float2* __restrict__ table, const double2* __restrict__ mr,
                                     const double2* __restrict__ mz, const double2* __restrict__ mrz, int64_t rows,
                                     int vq, int B, int S, const int* __restrict__ n_of, float* tt) {
  const int NS = S / 4, MH = S / 16, FRB = 2 * MH * NS * 64;
  const int64_t total = rows * FRB;
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
    float v = 0.f;
    if (n1 >= 0 && n2 >= 0) {
      const int64_t node = row / vq;
      const int kind = (int)(row - node * vq);
      const int64_t o = (node * B + n1) * B + n2;
      double2 c;
      if (kind == 0) {
        const float2 t = table[o];
        c = make_double2((double)t.x, (double)t.y);
      } else {
        c = kind == 1 ? mr[o] : (kind == 2 ? mz[o] : mrz[o]);
      }
      v = (float)(i < MH ? c.x : c.y);
    }
    tt[f] = v;
  }
}

// The same rows in FP64 in the fragment order of batch_slots64_kernel<..., VOL>:
// [rows][MT][NS/2][64 lanes][2], lane (q, rr) of tile i holding
// A(slot 16 (i mod MH) + rr, slot 4 (2 s2 + e) + q).
__global__ void vol_slot_taps64_kernel(const float2* __restrict__ table, const double2* __restrict__ mr,
                                       const double2* __restrict__ mz, const double2* __restrict__ mrz, int64_t rows,
                                       int vq, int B, int S, const int* __restrict__ n_of, double* tt) {
  const int NS = S / 4, MH = S / 16, FRB = 2 * MH * NS * 64;
  const int64_t total = rows * FRB;
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
    double v = 0.0;
    if (n1 >= 0 && n2 >= 0) {
      const int64_t node = row / vq;
      const int kind = (int)(row - node * vq);
      const int64_t o = (node * B + n1) * B + n2;
      double2 c;
      if (kind == 0) {
        const float2 t = table[o];
        c = make_double2((double)t.x, (double)t.y);
      } else {
        c = kind == 1 ? mr[o] : (kind == 2 ? mz[o] : mrz[o]);
      }
      v = i < MH ? c.x : c.y;
    }
    tt[f] = v;
  }
}

// The same rows in FP16 in the fragment order of batch_half_kernel<..., VOL>:
// [rows][MT][64 lanes][NC][4], lane (q, rr) of tile i holding
// A(slot 16 (i mod MH) + rr, slot 16 c + 4 q + j) / hs[row], hs[row] the
// smallest power of two with max |A| / hs <= 2^12 over the row; one block
// per row.
__global__ void vol_half_taps_kernel(const float2* __restrict__ table, const double2* __restrict__ mr,
                                     const double2* __restrict__ mz, const double2* __restrict__ mrz, int vq, int B,
                                     int S, const int* __restrict__ n_of, _Float16* th, float* hs) {
  const int MH = S / 16, MT = 2 * MH, NC = S / 16;
  const int per_row = MT * 64 * NC * 4;
  const int64_t row = blockIdx.x;
  const int64_t node = row / vq;
  const int kind = (int)(row - node * vq);
  __shared__ double red[256];
  auto value = [&](int f) {
    int rem = f;
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
    const int64_t o = (node * B + n1) * B + n2;
    double2 v;
    if (kind == 0) {
      const float2 t = table[o];
      v = make_double2((double)t.x, (double)t.y);
    } else {
      v = kind == 1 ? mr[o] : (kind == 2 ? mz[o] : mrz[o]);
    }
    return i < MH ? v.x : v.y;
  };
  double m = 0.0;
  for (int f = threadIdx.x; f < per_row; f += blockDim.x) m = fmax(m, fabs(value(f)));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  const double mx = red[0];
  const int e = mx > 0.0 ? (int)ceil(log2(mx)) - 12 : 0;
  const double inv = ldexp(1.0, -e);
  if (threadIdx.x == 0) hs[row] = ldexpf(1.0f, e);
  for (int f = threadIdx.x; f < per_row; f += blockDim.x) th[row * per_row + f] = (_Float16)(value(f) * inv);
}

// Row sums of a fragment-order tap table: per row max over n1 of
// sum_n2 |Re A(n1, n2)| + |Im A(n1, n2)| (slightly enlarged), one block per row.
__global__ void frag_rowsum_kernel(const float* __restrict__ tt, int64_t rows, int B, int Bh, int NG, int FR,
                                   float* rowsum) {
  __shared__ float sums[64];
  __shared__ float red[256];
  for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
    for (int j = threadIdx.x; j < 64; j += blockDim.x) sums[j] = 0.f;
    __syncthreads();
    for (int rem0 = threadIdx.x; rem0 < FR; rem0 += blockDim.x) {
      int rem = rem0;
      const int rr = (rem >> 2) & 15;
      rem >>= 6;
      const int i = rem / NG;
      const int m = 16 * i + rr;
      const int n1 = m < Bh ? m : m - Bh;
      if (n1 < B) atomicAdd(&sums[n1], fabsf(tt[row * FR + rem0]));
    }
    __syncthreads();
    red[threadIdx.x] = (int)threadIdx.x < B ? sums[threadIdx.x] : 0.f;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
      __syncthreads();
    }
    if (threadIdx.x == 0) rowsum[row] = red[0] * 1.001f;
    __syncthreads();
  }
}

// Bilinear cell of a pair in one table: distance node r and z node j of
// zval = z1 + zsgn * z2, with weights inside the cell.
__host__ __device__ inline void vol_geometry(const double* a, const double* b, double rho_min, double h, double zmin,
                                             double zsgn, int& r, int& j, double& wr, double& wz, double& rho,
                                             double& dx, double& dy) {
  dx = a[0] - b[0];
  dy = a[1] - b[1];
  double w;
  pair_geometry(dx, dy, rho_min, h, r, w, rho);
  wr = w;
  const double t = (a[2] + zsgn * b[2] - zmin) / h;
  const double fl = floor(t);
  j = (int)fl;
  wz = t - fl;
}

// A pair is valid when its interpolation stencil lies inside the table: nodes
// r - lo .. r + hi - 1 in each direction (lo = 0, hi = 2 bilinear and spline;
// lo = 1, hi = 3 Catmull-Rom).
// Grid: (receiver i1, block of 256 emitters of i1's ensemble); pairs couple
// particles of one ensemble only (seg_of null: one ensemble).
__global__ void vol_pair_bin_kernel(const double* __restrict__ xyz, int n, double rho_min, double h, int len_rho,
                                    double zmin, double zsgn, int len_z, int lo, int hi, int Rc,
                                    const int* __restrict__ seg_of, const int64_t* __restrict__ seg_off, int* cnt,
                                    const int* off, int* cursor, uint2* pairs, int* err) {
  const int i1 = blockIdx.x;
  const int sg = seg_of ? seg_of[i1] : 0;
  const int64_t i2l = (int64_t)seg_off[sg] + (int64_t)blockIdx.y * 256 + threadIdx.x;
  if (i1 >= n || i2l >= seg_off[sg + 1]) return;
  const int i2 = (int)i2l;
  int r, j;
  double wr, wz, rho, dx, dy;
  vol_geometry(xyz + 3 * (size_t)i1, xyz + 3 * (size_t)i2, rho_min, h, zmin, zsgn, r, j, wr, wz, rho, dx, dy);
  if (r < lo || r > len_rho - hi || j < lo || j > len_z - hi) {
    atomicOr(err, 1);
    return;
  }
  const int64_t key = ((int64_t)(i1 / Rc) * len_rho + r) * len_z + j;
  if (pairs == nullptr) {
    atomicAdd(&cnt[key], 1);
  } else {
    const int pos = off[key] + atomicAdd(&cursor[key], 1);
    pairs[pos] = make_uint2((unsigned)i1, (unsigned)i2);
  }
}

template <typename Real>
__global__ void vol_pair_record_kernel(const uint2* __restrict__ pairs, int64_t n_pairs, const double* __restrict__ xyz,
                                       double rho_min, double h, double zmin, double zsgn, int Rc, PairRec2<Real>* rec) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  const uint2 pr = pairs[p];
  int r, j;
  double wr, wz, rho, dx, dy;
  vol_geometry(xyz + 3 * (size_t)pr.x, xyz + 3 * (size_t)pr.y, rho_min, h, zmin, zsgn, r, j, wr, wz, rho, dx, dy);
  double c = 1.0, sn = 0.0;
  if (rho > 0.0) {
    c = dx / rho;
    sn = dy / rho;
  }
  PairRec2<Real> out{};
  out.ij = pr.y | ((pr.x % (unsigned)Rc) << kPairShift);
  out.w = (Real)wr;
  out.w2 = (Real)wz;
  out.c = (Real)c;
  out.s = (Real)sn;
  rec[p] = out;
}

template <int L, int Mx, typename Real, typename TapT>
KernelInfo vol_kernel_info() {
  using Sh = RadialShape<Real, SvwfIdx<L, Mx>::B, 4>;
  KernelInfo k;
  k.fn = reinterpret_cast<void*>(&lookup_matvec_kernel<Real, TapT, L, Mx, 4, true>);
  k.threads = Sh::kThreads;
  k.sy = Sh::SY;
  return k;
}

// tap64: FP64 kernel over the FP64 tap table (spline moments); otherwise the
// FP32 tap table, which holds the complex64 table values exactly.
KernelInfo volume_kernel(int precision, int l, int m, bool tap64 = false) {
#define SMH_V(L, Mx)                                                                                    \
  if (l == L && m == Mx)                                                                                \
    return precision == SMH_FP32 ? vol_kernel_info<L, Mx, float, float>()                               \
                                 : (tap64 ? vol_kernel_info<L, Mx, double, double>()                     \
                                          : vol_kernel_info<L, Mx, double, float>());
  SMH_V(1, 1)
  SMH_V(2, 1)
  SMH_V(2, 2)
  SMH_V(3, 1)
  SMH_V(3, 2)
  SMH_V(3, 3)
  SMH_V(4, 1)
  SMH_V(4, 2)
  SMH_V(4, 3)
  SMH_V(4, 4)
#undef SMH_V
  return KernelInfo{};
}

// FP16 operands: where the FP32 slot-order kernel is used (same binning and
// input slot order)
template <int L, int Mx>
void* vol_half_fn() {
  if constexpr (SlotShape<L, Mx, 4>::P32 > 0 && Mx <= 4 && half_tile_parts(4, SvwfIdx<L, Mx>::B) > 0)
    return reinterpret_cast<void*>(&batch_half_kernel<L, Mx, 4, true>);
  else
    return nullptr;
}
void* vol_half_kernel(int l, int m) {
#define SMH_V(L, Mx) \
  if (l == L && m == Mx) return vol_half_fn<L, Mx>();
  SMH_V(1, 1)
  SMH_V(2, 1)
  SMH_V(2, 2)
  SMH_V(3, 1)
  SMH_V(3, 2)
  SMH_V(3, 3)
  SMH_V(4, 1)
  SMH_V(4, 2)
  SMH_V(4, 3)
  SMH_V(4, 4)
#undef SMH_V
  return nullptr;
}

// FP32 kBatch-RHS kernel: slot order, its tap fragments in whole parts of the
// K-steps (every cut-off up to l_max = m_max = 4)
template <int L, int Mx>
KernelInfo vol_batch32_kernel_info() {
  if constexpr (SlotShape<L, Mx, 4>::P32 > 0 && Mx <= 4) {
    KernelInfo k;
    k.fn = reinterpret_cast<void*>(&batch_slots_kernel<L, Mx, 4, true>);
    k.threads = 512;
    k.sy = SlotShape<L, Mx, 4>::SY;
    k.slots = SlotShape<L, Mx, 4>::S;
    k.slot_tables = &slot_tables<L, Mx>;
    return k;
  } else {
    return KernelInfo{};
  }
}
// FP64 kBatch-RHS kernel: slot order, its coefficients in parts of the
// K-steps (the tap rows in FP64 whatever the single-vector kernel's taps)
template <int L, int Mx>
KernelInfo vol_batch64_kernel_info() {
  if constexpr (SlotShape<L, Mx, 4>::P64 > 0 && Mx <= 4) {
    KernelInfo k;
    k.fn = reinterpret_cast<void*>(&batch_slots64_kernel<L, Mx, 4, true>);
    k.threads = kSlots64Threads;
    k.sy = SlotShape<L, Mx, 4>::LO;  // doubles
    k.slots = SlotShape<L, Mx, 4>::S;
    k.slot_tables = &slot_tables<L, Mx>;
    return k;
  } else {
    return KernelInfo{};
  }
}
// FP64 kBatch-RHS kernel over the FP32 slot-order tap rows (tables of plain
// values, vq = 1: the rows hold the FP32 table values exactly)
template <int L, int Mx>
void* vol_batch64_tap32_fn() {
  if constexpr (SlotShape<L, Mx, 4>::P64 > 0 && SlotShape<L, Mx, 4>::P32 > 0 && Mx <= 4)
    return reinterpret_cast<void*>(&batch_slots64_kernel<L, Mx, 4, true, true>);
  else
    return nullptr;
}
void* vol_batch64_tap32_kernel(int l, int m) {
#define SMH_V(L, Mx) \
  if (l == L && m == Mx) return vol_batch64_tap32_fn<L, Mx>();
  SMH_V(1, 1)
  SMH_V(2, 1)
  SMH_V(2, 2)
  SMH_V(3, 1)
  SMH_V(3, 2)
  SMH_V(3, 3)
  SMH_V(4, 1)
  SMH_V(4, 2)
  SMH_V(4, 3)
  SMH_V(4, 4)
#undef SMH_V
  return nullptr;
}

KernelInfo volume_batch_kernel(int precision, int l, int m) {
#define SMH_V(L, Mx)                                                                                    \
  if (l == L && m == Mx)                                                                                \
    return precision == SMH_FP32 ? vol_batch32_kernel_info<L, Mx>() : vol_batch64_kernel_info<L, Mx>();
  SMH_V(1, 1)
  SMH_V(2, 1)
  SMH_V(2, 2)
  SMH_V(3, 1)
  SMH_V(3, 2)
  SMH_V(3, 3)
  SMH_V(4, 1)
  SMH_V(4, 2)
  SMH_V(4, 3)
  SMH_V(4, 4)
#undef SMH_V
  return KernelInfo{};
}

// Bin the pairs of table tab (0: z1 + z2, 1: z1 - z2) for precision pi, for
// the single-vector kernel or (batch) the kBatch-RHS kernel.
template <typename Real>
int vol_bin_pairs(smh_op* op, int tab, int pi, int Rc, bool batch) {
  smh_ctx* ctx = op->ctx;
  smh_op::Binning& bn = batch ? op->vbbin[tab][pi] : op->vbin[tab][pi];
  hipStream_t s = ctx->stream;
  const int64_t n_part = op->n_part, len_rho = op->len_rho, len_z = op->len_z;
  const double zmin = tab == 0 ? op->sz_min : op->dz_min, zsgn = tab == 0 ? 1.0 : -1.0;
  const int lo = op->vmode == 2 ? 1 : 0, hi = op->vmode == 2 ? 3 : 2;
  bn.Rc = Rc;
  bn.n_chunks = (int)((n_part + Rc - 1) / Rc);
  const int64_t nkeys = (int64_t)bn.n_chunks * len_rho * len_z;
  SMH_CHECK(ctx, nkeys < ((int64_t)1 << 31), SMH_ERR_UNSUPPORTED,
            "volume table too large for the pair binning (%lld bins)", (long long)nkeys);
  int *d_cnt = nullptr, *d_cur = nullptr, *d_err = nullptr, *d_off = nullptr;
  uint2* d_pairs = nullptr;
  auto cleanup = [&]() {
    void* p[] = {d_cnt, d_cur, d_err, d_off, d_pairs};
    for (void* q : p)
      if (q) hipFree(q);
  };
  if (hipMalloc(&d_cnt, (nkeys + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc(&d_cur, (nkeys + 1) * sizeof(int)) != hipSuccess || hipMalloc(&d_err, sizeof(int)) != hipSuccess ||
      hipMalloc(&d_off, (nkeys + 1) * sizeof(int)) != hipSuccess) {
    cleanup();
    return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (volume binning)");
  }
  hipMemsetAsync(d_cnt, 0, (nkeys + 1) * sizeof(int), s);
  hipMemsetAsync(d_cur, 0, (nkeys + 1) * sizeof(int), s);
  hipMemsetAsync(d_err, 0, sizeof(int), s);
  int64_t max_seg = 0;
  for (int sg = 0; sg < op->n_seg; ++sg) max_seg = std::max<int64_t>(max_seg, op->seg_off[sg + 1] - op->seg_off[sg]);
  const dim3 grid((unsigned)n_part, (unsigned)((max_seg + 255) / 256));
  SMH_CHECK(ctx, grid.y <= 65535, SMH_ERR_ARG, "ensemble too large for the pair binning");
  hipLaunchKernelGGL(vol_pair_bin_kernel, grid, dim3(256), 0, s, op->d_xyz, (int)n_part, op->rho_min, op->h,
                     (int)len_rho, zmin, zsgn, (int)len_z, lo, hi, Rc, (const int*)op->d_seg_of,
                     (const int64_t*)op->d_seg_off, d_cnt, (const int*)nullptr, d_cur, (uint2*)nullptr, d_err);
  std::vector<int> cnt(nkeys);
  int herr = 0;
  if (hipMemcpyAsync(cnt.data(), d_cnt, nkeys * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    cleanup();
    return smh_fail(ctx, SMH_ERR_HIP, "volume binning count failed");
  }
  if (herr) {
    cleanup();
    return smh_fail(ctx, SMH_ERR_ARG, "a particle pair lies outside the volume lookup table");
  }
  std::vector<int> off(nkeys + 1);
  int64_t acc = 0;
  for (int64_t k2 = 0; k2 < nkeys; ++k2) {
    off[k2] = (int)acc;
    acc += cnt[k2];
  }
  off[nkeys] = (int)acc;
  if (acc >= (int64_t)INT32_MAX) {
    cleanup();
    return smh_fail(ctx, SMH_ERR_UNSUPPORTED, "too many particle pairs for 32-bit offsets");
  }
  if (hipMalloc(&d_pairs, std::max<int64_t>(acc, 1) * sizeof(uint2)) != hipSuccess ||
      hipMemcpyAsync(d_off, off.data(), (nkeys + 1) * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMalloc(&bn.d_rec, std::max<int64_t>(acc, 1) * sizeof(PairRec2<Real>)) != hipSuccess) {
    cleanup();
    return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (volume pairs)");
  }
  hipLaunchKernelGGL(vol_pair_bin_kernel, grid, dim3(256), 0, s, op->d_xyz, (int)n_part, op->rho_min, op->h,
                     (int)len_rho, zmin, zsgn, (int)len_z, lo, hi, Rc, (const int*)op->d_seg_of,
                     (const int64_t*)op->d_seg_off, d_cnt, (const int*)d_off, d_cur, d_pairs, d_err);
  if (acc > 0)
    hipLaunchKernelGGL(vol_pair_record_kernel<Real>, dim3((unsigned)((acc + 255) / 256)), dim3(256), 0, s, d_pairs,
                       acc, op->d_xyz, op->rho_min, op->h, zmin, zsgn, Rc, (PairRec2<Real>*)bn.d_rec);
  if (hipStreamSynchronize(s) != hipSuccess) {
    cleanup();
    return smh_fail(ctx, SMH_ERR_HIP, "volume pair records failed");
  }
  cleanup();
  op->n_pairs = acc;
  // a (chunk, r, j) key is a bin; its entry starts at the node's first tap row (r * len_z + j) * vq
  const int vq = op->vq;
  return build_work_items<PairRec2<Real>>(op, bn, cnt, off, len_rho * len_z,
                                          [vq](int64_t rj) { return (int)(rj * vq); }, batch);
}

// Tap tables, row sums and per-item statistics from the complex64 tables:
// bilinear / Catmull-Rom one FP32 row per node; spline four rows per node
// (value and FP64 moments along rho, along z and mixed), FP32 and FP64.
int vol_make_taps(smh_op* op) {
  smh_ctx* ctx = op->ctx;
  hipStream_t s = ctx->stream;
  const TapGeom g(op->B, 4);
  op->FR = g.FR;
  const int64_t nodes = op->len_rho * op->len_z, rows = nodes * op->vq;
  const int64_t BB = (int64_t)op->B * op->B;
  double2 *mr = nullptr, *mz = nullptr, *mrz = nullptr;
  double *cpr = nullptr, *cpz = nullptr;
  auto cleanup = [&]() {
    void* p[] = {mr, mz, mrz, cpr, cpz};
    for (void* q : p)
      if (q) hipFree(q);
  };
  if (op->vq == 4) {
    auto cprime = [](int64_t n) {
      std::vector<double> cp(n, 0.0);
      for (int64_t j = 2; j <= n - 3; ++j) cp[j] = (j == 2) ? 0.25 : 1.0 / (4.0 - cp[j - 1]);
      return cp;
    };
    const std::vector<double> hr = cprime(op->len_rho), hz = cprime(op->len_z);
    if (hipMalloc(&mr, nodes * BB * sizeof(double2)) != hipSuccess ||
        hipMalloc(&mz, nodes * BB * sizeof(double2)) != hipSuccess ||
        hipMalloc(&mrz, nodes * BB * sizeof(double2)) != hipSuccess ||
        hipMalloc(&cpr, hr.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&cpz, hz.size() * sizeof(double)) != hipSuccess ||
        hipMemcpyAsync(cpr, hr.data(), hr.size() * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(cpz, hz.data(), hz.size() * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) {
      cleanup();
      return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (spline moments)");
    }
  }
  for (int t = 0; t < 2; ++t) {
    if (!op->d_vtt[t]) {
      if (hipMalloc(&op->d_vtt[t], (size_t)rows * g.FR * sizeof(float)) != hipSuccess ||
          hipMalloc(&op->d_vrowsum[t], (size_t)rows * sizeof(float)) != hipSuccess ||
          (op->vq == 4 && hipMalloc(&op->d_vtt64[t], (size_t)rows * g.FR * sizeof(double)) != hipSuccess)) {
        cleanup();
        return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (volume tap tables)");
      }
    }
    if (op->vq == 1) {
      hipLaunchKernelGGL(vol_taps_kernel, dim3(4096), dim3(256), 0, s, op->d_vtab[t], rows, op->B, op->Bh, g.NG,
                         g.FR, op->d_vtt[t]);
    } else {
      const int64_t colr = op->len_z * BB, colz = op->len_rho * BB;
      hipLaunchKernelGGL(spline_axis_kernel<float2>, dim3(blocks_for(colr)), dim3(256), 0, s, op->d_vtab[t], cpr,
                         op->len_rho, colr, (int64_t)1, mr);
      hipLaunchKernelGGL(spline_axis_kernel<float2>, dim3(blocks_for(colz)), dim3(256), 0, s, op->d_vtab[t], cpz,
                         op->len_z, BB, op->len_rho, mz);
      hipLaunchKernelGGL(spline_axis_kernel<double2>, dim3(blocks_for(colz)), dim3(256), 0, s, (const double2*)mr,
                         cpz, op->len_z, BB, op->len_rho, mrz);
      hipLaunchKernelGGL(vol_taps4_kernel, dim3(4096), dim3(256), 0, s, op->d_vtab[t], mr, mz, mrz, nodes, op->B,
                         op->Bh, g.NG, g.FR, op->d_vtt[t], op->d_vtt64[t]);
    }
    if (op->d_b32_tab) {
      const int S = op->b32_S;
      if (!op->d_vtts[t] && hipMalloc(&op->d_vtts[t], (size_t)rows * 2 * S * S * sizeof(float)) != hipSuccess) {
        cleanup();
        return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (slot-order tap rows)");
      }
      hipLaunchKernelGGL(vol_slot_taps_kernel, dim3(4096), dim3(256), 0, s, op->d_vtab[t], mr, mz, mrz, rows, op->vq,
                         op->B, S, (const int*)op->d_b32_tab + op->B, op->d_vtts[t]);
      if (vol_half_kernel(op->l_max, op->m_max)) {
        if (!op->d_vtth[t] && (hipMalloc(&op->d_vtth[t], (size_t)rows * 2 * S * S * sizeof(_Float16)) != hipSuccess ||
                               hipMalloc(&op->d_vhs[t], (size_t)rows * sizeof(float)) != hipSuccess)) {
          cleanup();
          return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (FP16 tap rows)");
        }
        hipLaunchKernelGGL(vol_half_taps_kernel, dim3((unsigned)rows), dim3(256), 0, s, op->d_vtab[t], mr, mz, mrz,
                           op->vq, op->B, S, (const int*)op->d_b32_tab + op->B, op->d_vtth[t], op->d_vhs[t]);
      }
      // FP64 slot rows only for tables with spline moments (vq = 4); plain
      // values (vq = 1) reach the FP64 kernel through the FP32 rows above
      if (op->vq == 4 && volume_batch_kernel(SMH_FP64, op->l_max, op->m_max).slots == S) {
        if (!op->d_vtts64[t] &&
            hipMalloc(&op->d_vtts64[t], (size_t)rows * 2 * S * S * sizeof(double)) != hipSuccess) {
          cleanup();
          return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (FP64 slot-order tap rows)");
        }
        hipLaunchKernelGGL(vol_slot_taps64_kernel, dim3(4096), dim3(256), 0, s, op->d_vtab[t], mr, mz, mrz, rows,
                           op->vq, op->B, S, (const int*)op->d_b32_tab + op->B, op->d_vtts64[t]);
      }
    }
    hipLaunchKernelGGL(frag_rowsum_kernel, dim3((unsigned)std::min<int64_t>(rows, 65536)), dim3(256), 0, s,
                       op->d_vtt[t], rows, op->B, op->Bh, g.NG, g.FR, op->d_vrowsum[t]);
    for (smh_op::Binning* bn : {&op->vbin[t][0], &op->vbbin[t][0]})
      set_item_lsum(*bn, op->d_vrowsum[t], 4, op->vmode, (int)op->len_z, s);
    // the moments of this table must be consumed before the next one overwrites them
    if (hipStreamSynchronize(s) != hipSuccess) {
      cleanup();
      return smh_fail(ctx, SMH_ERR_HIP, "volume tap tables failed: %s", hipGetErrorString(hipGetLastError()));
    }
  }
  cleanup();
  op->has_table = true;
  return SMH_OK;
}

}  // namespace

namespace smh {

int vol_apply(smh_op* op, int precision, const void* x_pad, void* y_pad, bool time_kernel) {
  smh_ctx* ctx = op->ctx;
  const int pi = precision == SMH_FP32 ? 0 : 1;
  const bool tap64 = pi == 1 && op->vq == 4;
  const KernelInfo k = volume_kernel(precision, op->l_max, op->m_max, tap64);
  SMH_CHECK(ctx, k.fn != nullptr && op->vbin[0][pi].Rc > 0, SMH_ERR_UNSUPPORTED,
            "no volume matvec kernel for l_max %d m_max %d", op->l_max, op->m_max);
  hipStream_t s = ctx->stream;
  if (pi == 0)
    launch_xmax((const float2*)x_pad, op->n_part * op->B, op->d_xmax, s);
  double ms_total = 0;
  const int n_slices = op->vmode >= 2 ? 4 : 1;
  for (int t = 0; t < 2; ++t) {
    const smh_op::Binning& bn = op->vbin[t][pi];
    for (int sl = 0; sl < n_slices; ++sl) {
      RadialKArgs a;
      a.recs = bn.d_rec;
      a.entries = bn.d_entries;
      a.items = bn.d_items;
      a.order = bn.d_order;
      a.tt = tap64 ? (const void*)op->d_vtt64[t] : (const void*)op->d_vtt[t];
      a.x = x_pad;
      a.ypart = op->d_ypart;
      a.xmax = op->d_xmax;
      a.item_lsum = bn.d_item_lsum;
      a.item_en = bn.d_item_en;
      a.n_items = bn.n_items;
      a.n_entries = bn.n_entries;
      a.interp = op->interp;
      a.Rc = bn.Rc;
      a.len_z = (int)op->len_z;
      a.vmode = op->vmode;
      a.vslice = sl;
      if (int rc_ = op_set_lds(op, k.fn, bn.lds)) return rc_;
      if (bn.n_items > 0) {
        if (time_kernel) SMH_HIP(ctx, hipEventRecord(ctx->ev0, s));
        void* args[] = {&a};
        SMH_HIP(ctx, hipLaunchKernel(k.fn, dim3(bn.n_items), dim3(k.threads), args, bn.lds, s));
        if (time_kernel) {
          SMH_HIP(ctx, hipEventRecord(ctx->ev1, s));
          SMH_HIP(ctx, hipEventSynchronize(ctx->ev1));
          float ms = 0;
          SMH_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
          ms_total += ms;
        }
      }
      // the first launch writes y, the others add to it
      const int accumulate = (t > 0 || sl > 0) ? 1 : 0;
      const int64_t n = op->n_part * op->B;
      if (pi == 0)
        hipLaunchKernelGGL(reduce_kernel<float>, dim3(blocks_for(n)), dim3(256), 0, s, (const float*)op->d_ypart,
                           bn.d_chunk_items, bn.Rc, op->B, op->n_part, (float2*)y_pad, accumulate);
      else
        hipLaunchKernelGGL(reduce_kernel<double>, dim3(blocks_for(n)), dim3(256), 0, s, (const double*)op->d_ypart,
                           bn.d_chunk_items, bn.Rc, op->B, op->n_part, (double2*)y_pad, accumulate);
      SMH_HIP(ctx, hipGetLastError());
    }
  }
  if (time_kernel) op->last_kernel_ms = (float)ms_total;
  return SMH_OK;
}

// kBatch right-hand sides on the plain layout [n_part][B][kBatch] (FP32 input
// float2, FP64 double2): one batched launch per table and z slice, their
// partials summed into y in launch order.
int vol_apply_batch(smh_op* op, int precision, const void* x_pad, void* y_pad, bool time_kernel) {
  smh_ctx* ctx = op->ctx;
  // the FP16-operand kernel shares the FP32 kernel's binning, slot order and partial-sum layout
  const bool half = precision == SMH_FP16;
  const int pi = (precision == SMH_FP32 || half) ? 0 : 1;
  const bool tap64 = pi == 1 && op->vq == 4;
  const KernelInfo k = volume_batch_kernel(half ? SMH_FP32 : precision, op->l_max, op->m_max);
  // FP64 over tables of plain values: the FP32 slot rows, widened in the kernel
  const bool rows32 = pi == 1 && op->vq == 1 && k.slots > 0;
  void* fn = half ? vol_half_kernel(op->l_max, op->m_max)
                  : (rows32 ? vol_batch64_tap32_kernel(op->l_max, op->m_max) : k.fn);
  SMH_CHECK(ctx, fn != nullptr && op->vbbin[0][pi].Rc > 0, SMH_ERR_UNSUPPORTED,
            "no batched volume matvec kernel for l_max %d m_max %d", op->l_max, op->m_max);
  hipStream_t s = ctx->stream;
  const bool slots = k.slots > 0;
  SMH_CHECK(ctx,
            !slots || (op->d_b32_tab && op->b32_S == k.slots &&
                       (half ? op->d_vtth[0] && op->d_vtth[1]
                             : (pi == 0 || rows32 ? op->d_vtts[0] && op->d_vtts[1]
                                                  : op->d_vtts64[0] && op->d_vtts64[1]))),
            SMH_ERR_STATE, "slot-order tap rows missing");
  SMH_CHECK(ctx, !half || slots, SMH_ERR_STATE, "the FP16 kernel needs the slot order");
  const int S = slots ? k.slots : op->B;
  const int* slot = slots ? (const int*)op->d_b32_tab : nullptr;
  // (the FP16 input was scaled by the xmax its conversion wrote)
  if (pi == 0 && !half) launch_xmax((const float2*)x_pad, op->n_part * S * kBatchRhs, op->d_xmax, s);
  const void* xin = x_pad;
  if (pi == 1 && slots) {
    // FP64 padded [n_part][B][kBatch] -> the slot layout (padding slots stay zero)
    const int64_t n8 = op->n_part * op->B * kBatchRhs;
    if (!op->d_x64s) {
      const size_t bytes = (size_t)op->n_part * S * kBatchRhs * sizeof(double2);
      SMH_HIP(ctx, hipMalloc(&op->d_x64s, bytes));
      fill_zero(op->d_x64s, bytes, s);
    }
    hipLaunchKernelGGL(to_slots64_kernel, dim3(blocks_for(n8)), dim3(256), 0, s, (const double2*)x_pad, n8, op->B, S,
                       slot, op->d_x64s);
    xin = op->d_x64s;
  }
  double ms_total = 0;
  const int n_slices = op->vmode >= 2 ? 4 : 1;
  const int64_t nx = op->n_part * op->B * kBatchRhs;
  for (int t = 0; t < 2; ++t) {
    const smh_op::Binning& bn = op->vbbin[t][pi];
    for (int sl = 0; sl < n_slices; ++sl) {
      RadialKArgs a;
      a.recs = bn.d_rec;
      a.entries = bn.d_entries;
      a.items = bn.d_items;
      a.order = bn.d_order;
      a.tt = half ? (const void*)op->d_vtth[t]
                  : (slots ? (pi == 0 || rows32 ? (const void*)op->d_vtts[t] : (const void*)op->d_vtts64[t])
                           : (tap64 ? (const void*)op->d_vtt64[t] : (const void*)op->d_vtt[t]));
      a.x = xin;
      a.ypart = op->d_ypart;
      a.xmax = op->d_xmax;
      a.item_lsum = bn.d_item_lsum;
      a.item_en = bn.d_item_en;
      a.n_items = bn.n_items;
      a.n_entries = bn.n_entries;
      a.interp = op->interp;
      a.Rc = bn.Rc;
      a.len_z = (int)op->len_z;
      a.vmode = op->vmode;
      a.vslice = sl;
      a.row_s = 1;
      a.row_o = 0;
      if (int rc_ = op_set_lds(op, fn, bn.lds)) return rc_;
      if (bn.n_items > 0) {
        if (time_kernel) SMH_HIP(ctx, hipEventRecord(ctx->ev0, s));
        const float* hs = op->d_vhs[t];
        void* args1[] = {&a};
        void* args2[] = {&a, &hs};
        const int threads = half ? kHalfThreads : k.threads;
        SMH_HIP(ctx, hipLaunchKernel(fn, dim3(bn.n_items), dim3(threads), half ? args2 : args1,
                                     bn.lds, s));
        if (time_kernel) {
          SMH_HIP(ctx, hipEventRecord(ctx->ev1, s));
          SMH_HIP(ctx, hipEventSynchronize(ctx->ev1));
          float ms = 0;
          SMH_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
          ms_total += ms;
        }
      }
      const int accumulate = (t > 0 || sl > 0) ? 1 : 0;
      if (pi == 0)
        hipLaunchKernelGGL(reduce_batch_kernel<float>, dim3(blocks_for(nx)), dim3(256), 0, s,
                           (const float*)op->d_ypart, bn.d_chunk_items, bn.Rc, op->B, S, slot, op->n_part,
                           (float2*)y_pad, accumulate);
      else
        hipLaunchKernelGGL(reduce_batch_kernel<double>, dim3(blocks_for(nx)), dim3(256), 0, s,
                           (const double*)op->d_ypart, bn.d_chunk_items, bn.Rc, op->B, S, slot, op->n_part,
                           (double2*)y_pad, accumulate);
      SMH_HIP(ctx, hipGetLastError());
    }
  }
  if (time_kernel) op->last_kernel_ms = (float)ms_total;
  return SMH_OK;
}

int vol_upload(smh_op* op, const double* plus, const double* minus) {
  smh_ctx* ctx = op->ctx;
  const size_t n = (size_t)op->len_rho * op->len_z * op->B * op->B;
  std::vector<float2> h(n);
  const double* src[2] = {plus, minus};
  for (int t = 0; t < 2; ++t) {
    for (size_t i = 0; i < n; ++i) h[i] = make_float2((float)src[t][2 * i], (float)src[t][2 * i + 1]);
    if (!op->d_vtab[t]) SMH_HIP(ctx, hipMalloc(&op->d_vtab[t], n * sizeof(float2)));
    SMH_HIP(ctx, hipMemcpy(op->d_vtab[t], h.data(), n * sizeof(float2), hipMemcpyHostToDevice));
  }
  return vol_make_taps(op);
}

int vol_download(smh_op* op, double* plus, double* minus) {
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, op->d_vtab[0] != nullptr, SMH_ERR_STATE, "no table");
  const size_t n = (size_t)op->len_rho * op->len_z * op->B * op->B;
  std::vector<float2> h(n);
  double* dst[2] = {plus, minus};
  for (int t = 0; t < 2; ++t) {
    if (!dst[t]) continue;
    SMH_HIP(ctx, hipMemcpy(h.data(), op->d_vtab[t], n * sizeof(float2), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
      dst[t][2 * i] = h[i].x;
      dst[t][2 * i + 1] = h[i].y;
    }
  }
  return SMH_OK;
}

}  // namespace smh

// Operator over n_ens ensembles (particles [ens_off[e], ens_off[e+1])), or one
// ensemble when ens_off is null.
static int volume_create(smh_ctx* ctx, int n_ens, const int64_t* ens_off, int64_t n_part, const double* xyz,
                         const int* lmax, const int* mmax, int l_max, int m_max, double rho_min, double sz_min,
                         double dz_min, double resolution, int64_t len_rho, int64_t len_z, int interp, smh_op** out) {
  if (!ctx || !out) return SMH_ERR_ARG;
  *out = nullptr;
  SMH_CHECK(ctx, n_part >= 1 && xyz && lmax && mmax, SMH_ERR_ARG, "need at least one particle");
  SMH_CHECK(ctx, n_part <= (int64_t)kPairMask, SMH_ERR_ARG, "too many particles (pair records index 23 bits)");
  SMH_CHECK(ctx, l_max >= 1 && m_max >= 1 && m_max <= l_max && 2 * l_max <= kMaxL2, SMH_ERR_ARG,
            "invalid l_max %d / m_max %d", l_max, m_max);
  SMH_CHECK(ctx, resolution > 0 && len_rho >= 2 && len_z >= 2 && len_rho * len_z < (1 << 30), SMH_ERR_ARG,
            "invalid table grid");
  SMH_CHECK(ctx, interp == SMH_INTERP_LINEAR || interp == SMH_INTERP_CATMULL_ROM || interp == SMH_INTERP_SPLINE,
            SMH_ERR_ARG, "invalid interpolation %d", interp);
  SMH_CHECK(ctx, interp != SMH_INTERP_SPLINE || (len_rho >= 7 && len_z >= 7 && len_rho * len_z < (1 << 28)),
            SMH_ERR_ARG, "spline interpolation needs at least 7 nodes per axis (and < 2^28 nodes)");
  SMH_CHECK(ctx, interp != SMH_INTERP_CATMULL_ROM || (len_rho >= 4 && len_z >= 4), SMH_ERR_ARG,
            "Catmull-Rom interpolation needs at least 4 nodes per axis");
  SMH_CHECK(ctx, volume_kernel(SMH_FP32, l_max, m_max).fn != nullptr, SMH_ERR_UNSUPPORTED,
            "no kernel for l_max %d m_max %d", l_max, m_max);
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  smh_op* op = new smh_op();
  op->ctx = ctx;
  op->kind = 1;
  op->n_part = n_part;
  op->l_max = l_max;
  op->m_max = m_max;
  op->B = smh_blocksize(l_max, m_max);
  op->Bh = (op->B + 15) / 16 * 16;
  op->MP = 2 * op->Bh;
  op->interp = interp;
  op->T = 4;
  op->vmode = interp == SMH_INTERP_LINEAR ? 1 : (interp == SMH_INTERP_CATMULL_ROM ? 2 : 3);
  op->vq = interp == SMH_INTERP_SPLINE ? 4 : 1;
  op->rho_min = rho_min;
  op->sz_min = sz_min;
  op->dz_min = dz_min;
  op->h = resolution;
  op->len_rho = len_rho;
  op->len_z = len_z;
  auto fail = [&](int code) {
    smh_op_free(op);
    delete op;
    return code;
  };
  if (int code = op_index_map(op, lmax, mmax)) return fail(code);
  if (ens_off)
    if (int code = op_set_segments(op, n_ens, ens_off)) return fail(code);
  // the FP32 batched kernel's input: slot layout (b32_index) for the slot-order
  // kernel, else the plain layout [n_part][B][kBatch]
  op->b32_S = op->B;
  {
    const KernelInfo kb32 = volume_batch_kernel(SMH_FP32, l_max, m_max);
    if (kb32.fn && kb32.slots > 0) {
      std::vector<int> tab(op->B + kb32.slots);
      kb32.slot_tables(tab.data());
      if (hipMalloc(&op->d_b32_tab, tab.size() * sizeof(int)) != hipSuccess ||
          hipMemcpy(op->d_b32_tab, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (slot tables)");
        return fail(SMH_ERR_HIP);
      }
      op->b32_S = kb32.slots;
    }
  }
  hipStream_t s = ctx->stream;
  if (hipMalloc(&op->d_xyz, n_part * 3 * sizeof(double)) != hipSuccess ||
      hipMemcpyAsync(op->d_xyz, xyz, n_part * 3 * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMalloc(&op->d_xmax, (1 + kXmaxBlocks) * sizeof(float)) != hipSuccess) {
    smh_fail(ctx, SMH_ERR_HIP, "device allocation failed");
    return fail(SMH_ERR_HIP);
  }
  size_t ypart = 0;
  for (int pi = 0; pi < 2; ++pi) {
    const KernelInfo k = volume_kernel(pi == 0 ? SMH_FP32 : SMH_FP64, l_max, m_max);
    const size_t rs = pi == 0 ? sizeof(float) : sizeof(double);
    size_t lds = 0;
    const int Rc = chunk_size(op, rs, k.sy, lds);
    if (Rc == 0) continue;
    for (int t = 0; t < 2; ++t) {
      op->vbin[t][pi].lds = lds;
      const int code = pi == 0 ? vol_bin_pairs<float>(op, t, 0, Rc, false) : vol_bin_pairs<double>(op, t, 1, Rc, false);
      if (code) return fail(code);
      op->vbin[t][pi].part_bytes = (size_t)Rc * 2 * op->B * rs;
      ypart = std::max(ypart, (size_t)op->vbin[t][pi].n_items * op->vbin[t][pi].part_bytes);
    }
    // the same pairs for the kBatch-RHS kernel (its accumulator holds kBatch vectors per receiver)
    const KernelInfo kb = volume_batch_kernel(pi == 0 ? SMH_FP32 : SMH_FP64, l_max, m_max);
    size_t ldsb = 0;
    const int Rb = kb.fn ? chunk_size(op, rs, kb.sy, ldsb) : 0;
    if (Rb == 0) continue;
    for (int t = 0; t < 2; ++t) {
      op->vbbin[t][pi].lds = ldsb;
      const int code = pi == 0 ? vol_bin_pairs<float>(op, t, 0, Rb, true) : vol_bin_pairs<double>(op, t, 1, Rb, true);
      if (code) return fail(code);
      op->vbbin[t][pi].part_bytes = (size_t)Rb * 2 * kBatchRhs * (kb.slots > 0 ? kb.slots : op->B) * rs;
      ypart = std::max(ypart, (size_t)op->vbbin[t][pi].n_items * op->vbbin[t][pi].part_bytes);
    }
  }
  op->ypart_bytes = std::max<size_t>(ypart, 16);
  const int64_t n_pad = n_part * op->B;
  if (hipMalloc(&op->d_ypart, op->ypart_bytes) != hipSuccess ||
      hipMalloc(&op->d_x32, n_pad * sizeof(float2)) != hipSuccess ||
      hipMalloc(&op->d_y32, n_pad * sizeof(float2)) != hipSuccess ||
      hipMalloc(&op->d_x64, n_pad * sizeof(double2)) != hipSuccess ||
      hipMalloc(&op->d_y64, n_pad * sizeof(double2)) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (vectors)");
    return fail(SMH_ERR_HIP);
  }
  ctx->live_ops++;
  *out = op;
  return SMH_OK;
}

extern "C" int smh_volume_create(smh_ctx* ctx, int64_t n_part, const double* xyz, const int* lmax, const int* mmax,
                                 int l_max, int m_max, double rho_min, double sz_min, double dz_min, double resolution,
                                 int64_t len_rho, int64_t len_z, int interp, smh_op** out) {
  return volume_create(ctx, 1, nullptr, n_part, xyz, lmax, mmax, l_max, m_max, rho_min, sz_min, dz_min, resolution,
                       len_rho, len_z, interp, out);
}

extern "C" int smh_volume_create_ensembles(smh_ctx* ctx, int n_ens, const int64_t* ens_off, const double* xyz,
                                           const int* lmax, const int* mmax, int l_max, int m_max, double rho_min,
                                           double sz_min, double dz_min, double resolution, int64_t len_rho,
                                           int64_t len_z, int interp, smh_op** out) {
  if (!ctx || !out) return SMH_ERR_ARG;
  SMH_CHECK(ctx, n_ens >= 1 && ens_off, SMH_ERR_ARG, "need at least one ensemble");
  return volume_create(ctx, n_ens, ens_off, ens_off[n_ens], xyz, lmax, mmax, l_max, m_max, rho_min, sz_min, dz_min,
                       resolution, len_rho, len_z, interp, out);
}

extern "C" int smh_volume_build_table(smh_op* op, double k_re, double k_im, double mask_radius, double z_s,
                                      const double* direct_coef, int64_t n_kappa, const double* kappa,
                                      const double* kz, const double* layer_coef) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, op->kind == 1, SMH_ERR_ARG, "not a volume operator");
  SMH_CHECK(ctx, direct_coef != nullptr, SMH_ERR_ARG, "direct coefficients are NULL");
  SMH_CHECK(ctx, n_kappa == 0 || (kappa && kz && layer_coef && n_kappa >= 2), SMH_ERR_ARG,
            "layer-mediated part needs kappa, kz and coefficients");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int B = op->B, BB = B * B, L2 = 2 * op->l_max;
  const int64_t nodes = op->len_rho * op->len_z;
  for (int t = 0; t < 2; ++t)
    if (!op->d_vtab[t]) SMH_HIP(ctx, hipMalloc(&op->d_vtab[t], (size_t)nodes * BB * sizeof(float2)));
  std::vector<int> tau, ll, mm;
  smh_index_tables(op->l_max, op->m_max, tau, ll, mm);
  std::vector<int> dm(BB);
  for (int n1 = 0; n1 < B; ++n1)
    for (int n2 = 0; n2 < B; ++n2) dm[n1 * B + n2] = abs(mm[n1] - mm[n2]);
  double2 *d_coef = nullptr, *d_kappa = nullptr, *d_kz = nullptr, *d_lc = nullptr;
  int* d_dm = nullptr;
  auto release = [&]() {
    void* p[] = {d_coef, d_kappa, d_kz, d_lc, d_dm};
    for (void* q : p)
      if (q) hipFree(q);
  };
  const size_t cbytes = (size_t)BB * (L2 + 1) * sizeof(double2);
  if (hipMalloc(&d_coef, cbytes) != hipSuccess || hipMalloc(&d_dm, BB * sizeof(int)) != hipSuccess ||
      hipMemcpyAsync(d_coef, direct_coef, cbytes, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d_dm, dm.data(), BB * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess) {
    release();
    return smh_fail(ctx, SMH_ERR_HIP, "direct coefficient upload failed");
  }
  hipLaunchKernelGGL(vol_direct_kernel, dim3((unsigned)nodes), dim3(256), 0, s, op->rho_min, op->dz_min, op->h,
                     op->len_rho, op->len_z, B, L2, k_re, k_im, mask_radius, d_coef, d_dm, op->d_vtab[1]);
  if (n_kappa > 0) {
    const size_t lbytes = (size_t)n_kappa * 4 * BB * sizeof(double2);
    if (hipMalloc(&d_kappa, n_kappa * sizeof(double2)) != hipSuccess ||
        hipMalloc(&d_kz, n_kappa * sizeof(double2)) != hipSuccess || hipMalloc(&d_lc, lbytes) != hipSuccess ||
        hipMemcpyAsync(d_kappa, kappa, n_kappa * sizeof(double2), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_kz, kz, n_kappa * sizeof(double2), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_lc, layer_coef, lbytes, hipMemcpyHostToDevice, s) != hipSuccess) {
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "layer-mediated coefficient upload failed");
    }
    const int nmax = 2 * op->m_max;
    SMH_CHECK(ctx, nmax < kNJ, SMH_ERR_UNSUPPORTED, "m_max %d too large for the table kernel", op->m_max);
    dim3 grid((unsigned)op->len_rho, (unsigned)((op->len_z + kZT - 1) / kZT));
    hipLaunchKernelGGL(vol_layer_kernel, grid, dim3(256), 0, s, op->rho_min, op->sz_min, op->dz_min, op->h,
                       op->len_z, B, nmax, z_s, d_dm, n_kappa, d_kappa, d_kz, d_lc, op->d_vtab[0], op->d_vtab[1]);
  } else {
    SMH_HIP(ctx, hipMemsetAsync(op->d_vtab[0], 0, (size_t)nodes * BB * sizeof(float2), s));
  }
  const hipError_t e = hipStreamSynchronize(s);
  release();
  SMH_CHECK(ctx, e == hipSuccess, SMH_ERR_HIP, "volume table kernels failed: %s", hipGetErrorString(e));
  return vol_make_taps(op);
}
