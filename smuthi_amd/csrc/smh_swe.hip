// Direct field of outgoing spherical wave expansions on the GPU: at every
// point the sum over particles p of
//   E_p(r) = sum_{tau,l,m} b_{p,tau l m} N^(3)_{tau l m}(k (r - r_p))
// (SphericalWaveExpansion.electric_field, fields/expansions.py:348-381, with
// the SVWFs of fields/vector_wave_functions.py:64-155), or the magnetic field
// H_p = -i k / omega sum b_{p,(1-tau) l m} N_{tau l m} (expansions.py:383-424).
// A particle's expansion counts at points of its layer's z range outside its
// sphere, as the validity conditions the piecewise expansion of the scattered
// field gives it (postprocessing/scattered_field.py:72-75, particles.py:122-124).
//
// Layout. One thread per point, 256 points per workgroup; the particles are
// staged through LDS in chunks (centre, squared radius, z range, the B
// coefficients), every thread reading the same particle at the same time (LDS
// broadcast). Per (point, particle) the distance runs in FP64 and, in the FP32
// kernel, the phase Re(k) r is reduced to [-pi, pi] in FP64 (large k r stays
// accurate); the Hankel recurrence, the normalized Legendre functions and the
// multipole sum run in Real (FP32 or FP64). The sum is
// formed in spherical components per m, rotated by e^{i m phi} once per m and
// added in Cartesian components.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "smh_common.h"

namespace {

constexpr int kSwePoints = 256;
constexpr int kSweChunk = 32;

struct SweArgs {
  const double* part;   // [n_part][6]: x, y, z, radius^2, z lower, z upper
  const double2* coef;  // [n_part][B] in (tau, l, m) order; H: b_(1-tau) staged at tau
  const double* pts;    // [n_pts][3]
  double2* out;         // [3][n_pts]
  int64_t n_part, n_pts;
  double k_re, k_im;
};

// (tau, l, m) -> index, tau-major, then l, then m (fields/__init__.py:29-80)
__host__ __device__ constexpr int swe_index(int tau, int l, int m, int l_max, int m_max) {
  const int tb = m_max * (m_max + 2) + (l_max - m_max) * (2 * m_max + 1);
  int n = tau * tb;
  n += (l - 1) <= m_max ? (l - 1) * (l + 1) : m_max * (m_max + 2) + (l - 1 - m_max) * (2 * m_max + 1);
  return n + m + (l < m_max ? l : m_max);
}

// Prefactor of the normalized P_m^m = pref st^m (legendre_normalized, utility/math.py:38-104):
// res_m = prod_{t = 2, 4, .., 2m} (1 + 1/t), pref = sqrt(res_m / 2)
__host__ __device__ inline double legendre_mm_prefactor(int m) {
  double res = 1.0;
  for (int t = 2; t <= 2 * m; t += 2) res += res / t;
  return sqrt(res / 2.0);
}

template <typename Real, int LMAX, int MMAX>
__global__ void __launch_bounds__(kSwePoints) swe_field_kernel(SweArgs a) {
  constexpr int B = 2 * (MMAX * (MMAX + 2) + (LMAX - MMAX) * (2 * MMAX + 1));
  using C = cplx<Real>;
  __shared__ double sp[kSweChunk][6];
  __shared__ C sc[kSweChunk][B];

  const int64_t ip = (int64_t)blockIdx.x * kSwePoints + threadIdx.x;
  const bool live = ip < a.n_pts;
  double px = 0, py = 0, pz = 0;
  if (live) {
    px = a.pts[3 * ip];
    py = a.pts[3 * ip + 1];
    pz = a.pts[3 * ip + 2];
  }
  C ex{0, 0}, ey{0, 0}, ez{0, 0};

  for (int64_t p0 = 0; p0 < a.n_part; p0 += kSweChunk) {
    const int np = (int)(a.n_part - p0 < kSweChunk ? a.n_part - p0 : kSweChunk);
    __syncthreads();
    for (int j = threadIdx.x; j < np * 6; j += kSwePoints) sp[j / 6][j % 6] = a.part[(p0 + j / 6) * 6 + j % 6];
    for (int j = threadIdx.x; j < np * B; j += kSwePoints) {
      const double2 v = a.coef[(p0 + j / B) * B + j % B];
      sc[j / B][j % B] = C{(Real)v.x, (Real)v.y};
    }
    __syncthreads();
    if (!live) continue;
    for (int q = 0; q < np; ++q) {
      const double dx = px - sp[q][0], dy = py - sp[q][1], dz = pz - sp[q][2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (!(r2 > sp[q][3]) || pz < sp[q][4] || pz >= sp[q][5]) continue;
      const double r = sqrt(r2), rho = sqrt(dx * dx + dy * dy);
      const double ir = 1.0 / r;
      const Real ct = (Real)(dz * ir), st = (Real)(rho * ir);
      Real cp = 1, sn = 0;
      if (rho > 0) {
        const double irho = 1.0 / rho;
        cp = (Real)(dx * irho);
        sn = (Real)(dy * irho);
      }
      // spherical Hankel h_l(x), x = k r, and d/dx(x h_l) / x. FP64 kernel: all in FP64. FP32
      // kernel: the phase Re(k) r is reduced to [-pi, pi] in FP64, the rest runs in FP32.
      using Z = cplx<Real>;
      const Z x{(Real)(a.k_re * r), (Real)(a.k_im * r)};
      const Z ix = cdiv(Z{Real(1), Real(0)}, x);
      Z eix;
      if constexpr (sizeof(Real) == 8) {
        const double dmp = exp(-x.im);
        eix = Z{dmp * cos(x.re), dmp * sin(x.re)};
      } else {
        const double ph = a.k_re * r;
        const float red = (float)(ph - 6.283185307179586477 * rint(ph * 0.15915494309189533577));
        const float dmp = expf((float)(-a.k_im * r));
        float sr, cr;
        sincosf(red, &sr, &cr);
        eix = Z{dmp * cr, dmp * sr};
      }
      Z h[LMAX + 1];
      h[0] = cmul(Z{eix.im, -eix.re}, ix);  // -i e^{ix} / x
      h[1] = cscale(cmul(cmul(eix, Z{x.re, x.im + Real(1)}), cmul(ix, ix)), Real(-1));
#pragma unroll
      for (int l = 1; l < LMAX; ++l) h[l + 1] = csub(cscale(cmul(ix, h[l]), Real(2 * l + 1)), h[l - 1]);
      // radial factors per l: A = h / sqrt(2l(l+1)), Bv = l(l+1) h / x / sqrt(..), Cv = (x h_(l-1) - l h_l) / x / sqrt(..)
      C A[LMAX + 1], Bv[LMAX + 1], Cv[LMAX + 1];
#pragma unroll
      for (int l = 1; l <= LMAX; ++l) {
        const Real pf = (Real)(1.0 / sqrt(2.0 * l * (l + 1)));
        const Z hx = cmul(h[l], ix);
        A[l] = cscale(h[l], pf);
        Bv[l] = cscale(hx, pf * Real(l * (l + 1)));
        Cv[l] = cscale(csub(h[l - 1], cscale(hx, Real(l))), pf);
      }
      // normalized associated Legendre P, pi, tau of cos(theta) (utility/math.py:38-104, real argument)
      Real P[LMAX + 1][MMAX + 1], PI[LMAX + 1][MMAX + 1], TA[LMAX + 1][MMAX + 1];
      {
        Real pp[LMAX + 1];
        P[0][0] = (Real)(0.70710678118654752440);
        P[1][0] = (Real)sqrt(1.5) * ct;
        pp[0] = 0;
        pp[1] = (Real)sqrt(3.0) * P[0][0];
        TA[1][0] = -st * pp[1];
        PI[1][0] = 0;
#pragma unroll
        for (int l = 1; l < LMAX; ++l) {
          P[l + 1][0] = (Real)(sqrt((2.0 * l + 1) * (2.0 * l + 3)) / (l + 1)) * ct * P[l][0] -
                        (Real)(l / (l + 1.0) * sqrt((2.0 * l + 3) / (2.0 * l - 1))) * P[l - 1][0];
          pp[l + 1] = (Real)((l + 1) * sqrt((2.0 * l + 3) / (2.0 * l + 1))) * P[l][0] +
                      (Real)sqrt((2.0 * l + 3) / (2.0 * l + 1)) * ct * pp[l];
          TA[l + 1][0] = -st * pp[l + 1];
          PI[l + 1][0] = 0;
        }
#pragma unroll
        for (int m = 1; m <= MMAX; ++m) {
          const Real pref = (Real)legendre_mm_prefactor(m);
          Real stm1 = 1;
#pragma unroll
          for (int j = 1; j < m; ++j) stm1 *= st;
          P[m][m] = pref * stm1 * st;
          PI[m][m] = pref * stm1;
          TA[m][m] = Real(m) * ct * PI[m][m];
#pragma unroll
          for (int l = m; l < LMAX; ++l) {
            const Real c1 = (Real)sqrt((2.0 * l + 1) * (2.0 * l + 3) / ((l + 1.0 - m) * (l + 1.0 + m)));
            const Real c2 = (Real)sqrt((2.0 * l + 3) * (l - m) * (l + m) / ((2.0 * l - 1) * (l + 1.0 - m) * (l + 1.0 + m)));
            const Real pl1 = l > m ? P[l - 1][m] : Real(0), pil1 = l > m ? PI[l - 1][m] : Real(0);
            P[l + 1][m] = c1 * ct * P[l][m] - c2 * pl1;
            PI[l + 1][m] = c1 * ct * PI[l][m] - c2 * pil1;
            const Real c3 = (Real)((l + 1.0 + m) * sqrt((2.0 * l + 3) * (l + 1.0 - m) / ((2.0 * l + 1) * (l + 1.0 + m))));
            TA[l + 1][m] = Real(l + 1) * ct * PI[l + 1][m] - c3 * PI[l][m];
          }
        }
      }
      // unit vectors
      const Real erx = st * cp, ery = st * sn, erz = ct;
      const Real etx = ct * cp, ety = ct * sn, etz = -st;
      const Real epx = -sn, epy = cp;
      const C* b = sc[q];
      // e^{i m phi} for m >= 0
      C eim[MMAX + 1];
      eim[0] = C{1, 0};
#pragma unroll
      for (int m = 1; m <= MMAX; ++m) eim[m] = cmul(eim[m - 1], C{cp, sn});
#pragma unroll
      for (int m = -MMAX; m <= MMAX; ++m) {
        const int am = m < 0 ? -m : m;
        // spherical components of sum_l at this m (before e^{i m phi})
        C cr{0, 0}, cth{0, 0}, cph{0, 0};
#pragma unroll
        for (int l = (am > 1 ? am : 1); l <= LMAX; ++l) {
          const C b0 = b[swe_index(0, l, m, LMAX, MMAX)], b1 = b[swe_index(1, l, m, LMAX, MMAX)];
          const C u = cmul(b0, A[l]);   // tau = 0
          const C wb = cmul(b1, Bv[l]);  // tau = 1, radial part
          const C wc = cmul(b1, Cv[l]);  // tau = 1, transverse part
          const Real mpi = Real(m) * PI[l][am], ta = TA[l][am];
          // tau 0: u (i m pi e_theta - tau e_phi); tau 1: wb P e_r + wc (tau e_theta + i m pi e_phi)
          cr.re += wb.re * P[l][am];
          cr.im += wb.im * P[l][am];
          cth.re += -mpi * u.im + ta * wc.re;
          cth.im += mpi * u.re + ta * wc.im;
          cph.re += -ta * u.re - mpi * wc.im;
          cph.im += -ta * u.im + mpi * wc.re;
        }
        const C e = m >= 0 ? eim[am] : C{eim[am].re, -eim[am].im};
        const C r_ = cmul(cr, e), t_ = cmul(cth, e), f_ = cmul(cph, e);
        ex.re += r_.re * erx + t_.re * etx + f_.re * epx;
        ex.im += r_.im * erx + t_.im * etx + f_.im * epx;
        ey.re += r_.re * ery + t_.re * ety + f_.re * epy;
        ey.im += r_.im * ery + t_.im * ety + f_.im * epy;
        ez.re += r_.re * erz + t_.re * etz;
        ez.im += r_.im * erz + t_.im * etz;
      }
    }
  }
  if (live) {
    a.out[ip] = make_double2((double)ex.re, (double)ex.im);
    a.out[a.n_pts + ip] = make_double2((double)ey.re, (double)ey.im);
    a.out[2 * a.n_pts + ip] = make_double2((double)ez.re, (double)ez.im);
  }
}

// Any (l_max, m_max): the same sum with runtime orders. Per (point,
// particle) the orders |m| run one after the other; for each the Hankel and
// Legendre recurrences run up in l from their start (no per-l arrays), and
// m = +|m| and -|m| share them. The particle chunk is sized from B so that
// its coefficients fit the workgroup's LDS. Costs (2 m_max + 1) l_max
// recurrence steps per pair more than the unrolled kernels, which take
// l_max <= 4.
constexpr int kSweMaxL = 20;

template <typename Real>
__global__ void __launch_bounds__(kSwePoints) swe_field_any_kernel(SweArgs a, int lmax, int mmax, int B, int chunk) {
  using C = cplx<Real>;
  extern __shared__ __align__(16) unsigned char swe_smem[];
  double* sp = reinterpret_cast<double*>(swe_smem);  // [chunk][6]
  C* sc = reinterpret_cast<C*>(sp + (size_t)chunk * 6);  // [chunk][B]

  const int64_t ip = (int64_t)blockIdx.x * kSwePoints + threadIdx.x;
  const bool live = ip < a.n_pts;
  double px = 0, py = 0, pz = 0;
  if (live) {
    px = a.pts[3 * ip];
    py = a.pts[3 * ip + 1];
    pz = a.pts[3 * ip + 2];
  }
  C ex{0, 0}, ey{0, 0}, ez{0, 0};
  for (int64_t p0 = 0; p0 < a.n_part; p0 += chunk) {
    const int np = (int)(a.n_part - p0 < chunk ? a.n_part - p0 : chunk);
    __syncthreads();
    for (int j = threadIdx.x; j < np * 6; j += kSwePoints) sp[j] = a.part[p0 * 6 + j];
    for (int j = threadIdx.x; j < np * B; j += kSwePoints) {
      const double2 v = a.coef[p0 * B + j];
      sc[j] = C{(Real)v.x, (Real)v.y};
    }
    __syncthreads();
    if (!live) continue;
    for (int q = 0; q < np; ++q) {
      const double* pq = sp + 6 * q;
      const double dx = px - pq[0], dy = py - pq[1], dz = pz - pq[2];
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (!(r2 > pq[3]) || pz < pq[4] || pz >= pq[5]) continue;
      const double r = sqrt(r2), rho = sqrt(dx * dx + dy * dy);
      const double ir = 1.0 / r;
      const Real ct = (Real)(dz * ir), st = (Real)(rho * ir);
      Real cp = 1, sn = 0;
      if (rho > 0) {
        const double irho = 1.0 / rho;
        cp = (Real)(dx * irho);
        sn = (Real)(dy * irho);
      }
      const C x{(Real)(a.k_re * r), (Real)(a.k_im * r)};
      const C ix = cdiv(C{Real(1), Real(0)}, x);
      C eix;
      if constexpr (sizeof(Real) == 8) {
        const double dmp = exp(-x.im);
        eix = C{dmp * cos(x.re), dmp * sin(x.re)};
      } else {
        const double ph = a.k_re * r;
        const float red = (float)(ph - 6.283185307179586477 * rint(ph * 0.15915494309189533577));
        const float dmp = expf((float)(-a.k_im * r));
        float sr, cr;
        sincosf(red, &sr, &cr);
        eix = C{dmp * cr, dmp * sr};
      }
      const C h0 = cmul(C{eix.im, -eix.re}, ix);  // -i e^{ix} / x
      const C h1 = cscale(cmul(cmul(eix, C{x.re, x.im + Real(1)}), cmul(ix, ix)), Real(-1));
      const Real erx = st * cp, ery = st * sn, erz = ct;
      const Real etx = ct * cp, ety = ct * sn, etz = -st;
      const Real epx = -sn, epy = cp;
      const C* b = sc + (size_t)q * B;
      C eim{1, 0};  // e^{i am phi}
      for (int am = 0; am <= mmax; ++am) {
        if (am > 0) eim = cmul(eim, C{cp, sn});
        // spherical components at m = +am (s = 0) and m = -am (s = 1), before e^{i m phi}
        C cr[2] = {{0, 0}, {0, 0}}, cth[2] = {{0, 0}, {0, 0}}, cph[2] = {{0, 0}, {0, 0}};
        // Legendre state at l (P, pi, tau) and l - 1; Hankel h_l and h_(l-1)
        Real Pl, Pm1 = 0, PIl, PIm1 = 0, TAl, ppl = 0;
        int l0;
        if (am == 0) {
          // l = 1 (l = 0 carries no multipole)
          const Real P0 = (Real)0.70710678118654752440;
          Pm1 = P0;
          Pl = (Real)sqrt(1.5) * ct;
          ppl = (Real)sqrt(3.0) * P0;
          PIl = 0;
          TAl = -st * ppl;
          l0 = 1;
        } else {
          const Real pref = (Real)legendre_mm_prefactor(am);
          Real stm1 = 1;
          for (int j = 1; j < am; ++j) stm1 *= st;
          Pl = pref * stm1 * st;
          PIl = pref * stm1;
          TAl = Real(am) * ct * PIl;
          l0 = am;
        }
        C hm1 = h0, hl = h1;  // h_(l-1), h_l at l = 1
        for (int l = 1; l < l0; ++l) {
          const C hn = csub(cscale(cmul(ix, hl), Real(2 * l + 1)), hm1);
          hm1 = hl;
          hl = hn;
        }
        for (int l = l0; l <= lmax; ++l) {
          const Real pf = (Real)(1.0 / sqrt(2.0 * l * (l + 1)));
          const C hx = cmul(hl, ix);
          const C A = cscale(hl, pf);
          const C Bv = cscale(hx, pf * Real(l * (l + 1)));
          const C Cv = cscale(csub(hm1, cscale(hx, Real(l))), pf);
          for (int s = 0; s < (am > 0 ? 2 : 1); ++s) {
            const int m = s ? -am : am;
            const C b0 = b[swe_index(0, l, m, lmax, mmax)], b1 = b[swe_index(1, l, m, lmax, mmax)];
            const C u = cmul(b0, A), wb = cmul(b1, Bv), wc = cmul(b1, Cv);
            const Real mpi = Real(m) * PIl, ta = TAl;
            cr[s].re += wb.re * Pl;
            cr[s].im += wb.im * Pl;
            cth[s].re += -mpi * u.im + ta * wc.re;
            cth[s].im += mpi * u.re + ta * wc.im;
            cph[s].re += -ta * u.re - mpi * wc.im;
            cph[s].im += -ta * u.im + mpi * wc.re;
          }
          if (l == lmax) break;
          // step l -> l + 1
          const C hn = csub(cscale(cmul(ix, hl), Real(2 * l + 1)), hm1);
          hm1 = hl;
          hl = hn;
          if (am == 0) {
            const Real Pn = (Real)(sqrt((2.0 * l + 1) * (2.0 * l + 3)) / (l + 1)) * ct * Pl -
                            (Real)(l / (l + 1.0) * sqrt((2.0 * l + 3) / (2.0 * l - 1))) * Pm1;
            ppl = (Real)((l + 1) * sqrt((2.0 * l + 3) / (2.0 * l + 1))) * Pl +
                  (Real)sqrt((2.0 * l + 3) / (2.0 * l + 1)) * ct * ppl;
            Pm1 = Pl;
            Pl = Pn;
            TAl = -st * ppl;
          } else {
            const Real c1 = (Real)sqrt((2.0 * l + 1) * (2.0 * l + 3) / ((l + 1.0 - am) * (l + 1.0 + am)));
            const Real c2 =
                (Real)sqrt((2.0 * l + 3) * (l - am) * (l + am) / ((2.0 * l - 1) * (l + 1.0 - am) * (l + 1.0 + am)));
            const Real c3 = (Real)((l + 1.0 + am) * sqrt((2.0 * l + 3) * (l + 1.0 - am) / ((2.0 * l + 1) * (l + 1.0 + am))));
            const Real Pn = c1 * ct * Pl - c2 * Pm1;
            const Real PIn = c1 * ct * PIl - c2 * PIm1;
            TAl = Real(l + 1) * ct * PIn - c3 * PIl;
            Pm1 = Pl;
            Pl = Pn;
            PIm1 = PIl;
            PIl = PIn;
          }
        }
        for (int s = 0; s < (am > 0 ? 2 : 1); ++s) {
          const C e = s ? C{eim.re, -eim.im} : eim;
          const C r_ = cmul(cr[s], e), t_ = cmul(cth[s], e), f_ = cmul(cph[s], e);
          ex.re += r_.re * erx + t_.re * etx + f_.re * epx;
          ex.im += r_.im * erx + t_.im * etx + f_.im * epx;
          ey.re += r_.re * ery + t_.re * ety + f_.re * epy;
          ey.im += r_.im * ery + t_.im * ety + f_.im * epy;
          ez.re += r_.re * erz + t_.re * etz;
          ez.im += r_.im * erz + t_.im * etz;
        }
      }
    }
  }
  if (live) {
    a.out[ip] = make_double2((double)ex.re, (double)ex.im);
    a.out[a.n_pts + ip] = make_double2((double)ey.re, (double)ey.im);
    a.out[2 * a.n_pts + ip] = make_double2((double)ez.re, (double)ez.im);
  }
}

template <int L, int M>
void* swe_fn(int precision) {
  return precision == SMH_FP32 ? reinterpret_cast<void*>(&swe_field_kernel<float, L, M>)
                               : reinterpret_cast<void*>(&swe_field_kernel<double, L, M>);
}

void* swe_kernel(int precision, int l, int m) {
#define SMH_S(L, Mx) \
  if (l == L && m == Mx) return swe_fn<L, Mx>(precision);
  SMH_S(1, 1)
  SMH_S(2, 1)
  SMH_S(2, 2)
  SMH_S(3, 1)
  SMH_S(3, 2)
  SMH_S(3, 3)
  SMH_S(4, 1)
  SMH_S(4, 2)
  SMH_S(4, 3)
  SMH_S(4, 4)
#undef SMH_S
  return nullptr;
}

}  // namespace

extern "C" int smh_swe_field(smh_ctx* ctx, int64_t n_part, const double* pos, const double* radius,
                             const double* zlim, int l_max, int m_max, const double* coef, double k_re, double k_im,
                             int64_t n_pts, const double* pts, int field, double omega, int precision, double* out,
                             double* kernel_ms) {
  if (!ctx) return SMH_ERR_ARG;
  SMH_CHECK(ctx, field == 0 || field == 1, SMH_ERR_ARG, "field must be 0 (E) or 1 (H)");
  SMH_CHECK(ctx, precision == SMH_FP32 || precision == SMH_FP64, SMH_ERR_ARG, "invalid precision %d", precision);
  SMH_CHECK(ctx, n_part >= 0 && n_pts >= 0 && n_pts < (1 << 30), SMH_ERR_ARG, "invalid sizes");
  SMH_CHECK(ctx, field == 0 || omega != 0.0, SMH_ERR_ARG, "omega is zero");
  SMH_CHECK(ctx, l_max >= 1 && m_max >= 0 && m_max <= l_max && l_max <= kSweMaxL, SMH_ERR_ARG,
            "l_max %d / m_max %d out of range (l_max <= %d)", l_max, m_max, kSweMaxL);
  // unrolled kernels for l_max <= 4, the runtime-order kernel otherwise
  void* fn = swe_kernel(precision, l_max, m_max);
  if (kernel_ms) *kernel_ms = 0.0;
  if (n_pts == 0) return SMH_OK;
  SMH_CHECK(ctx, pts && out, SMH_ERR_ARG, "points / output are NULL");
  if (n_part == 0) {
    for (int64_t j = 0; j < 6 * n_pts; ++j) out[j] = 0.0;
    return SMH_OK;
  }
  SMH_CHECK(ctx, pos && radius && zlim && coef, SMH_ERR_ARG, "particle arrays are NULL");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  const int B = smh_blocksize(l_max, m_max);
  // particle records and coefficients (H: b_(1-tau, l, m) at tau)
  std::vector<double> part((size_t)6 * n_part);
  std::vector<double> cf((size_t)2 * n_part * B);
  const int half = B / 2;
  for (int64_t p = 0; p < n_part; ++p) {
    part[6 * p] = pos[3 * p];
    part[6 * p + 1] = pos[3 * p + 1];
    part[6 * p + 2] = pos[3 * p + 2];
    part[6 * p + 3] = radius[p] * radius[p];
    part[6 * p + 4] = zlim[2 * p];
    part[6 * p + 5] = zlim[2 * p + 1];
    for (int n = 0; n < B; ++n) {
      const int src = field == 0 ? n : (n < half ? n + half : n - half);
      cf[2 * ((size_t)p * B + n)] = coef[2 * ((size_t)p * B + src)];
      cf[2 * ((size_t)p * B + n) + 1] = coef[2 * ((size_t)p * B + src) + 1];
    }
  }
  void* buf[4] = {nullptr, nullptr, nullptr, nullptr};
  const size_t sizes[4] = {part.size() * 8, cf.size() * 8, (size_t)3 * n_pts * 8, (size_t)3 * n_pts * 16};
  auto release = [&]() {
    for (void* q : buf)
      if (q) hipFree(q);
  };
  for (int i = 0; i < 4; ++i)
    if (hipMalloc(&buf[i], sizes[i]) != hipSuccess) {
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "SWE field device allocation failed");
    }
  hipStream_t s = ctx->stream;
  const void* src[3] = {part.data(), cf.data(), pts};
  for (int i = 0; i < 3; ++i)
    if (hipMemcpyAsync(buf[i], src[i], sizes[i], hipMemcpyHostToDevice, s) != hipSuccess) {
      release();
      return smh_fail(ctx, SMH_ERR_HIP, "SWE field upload failed");
    }
  SweArgs a;
  a.part = (const double*)buf[0];
  a.coef = (const double2*)buf[1];
  a.pts = (const double*)buf[2];
  a.out = (double2*)buf[3];
  a.n_part = n_part;
  a.n_pts = n_pts;
  a.k_re = k_re;
  a.k_im = k_im;
  const unsigned nb = (unsigned)((n_pts + kSwePoints - 1) / kSwePoints);
  const size_t csize = precision == SMH_FP32 ? sizeof(float2) : sizeof(double2);
  int lmax_ = l_max, mmax_ = m_max, B_ = B;
  int chunk = (int)std::min<size_t>(32, 65536 / (6 * sizeof(double) + (size_t)B * csize));
  const size_t lds = fn ? 0 : (size_t)chunk * (6 * sizeof(double) + (size_t)B * csize);
  if (!fn)
    fn = precision == SMH_FP32 ? reinterpret_cast<void*>(&swe_field_any_kernel<float>)
                               : reinterpret_cast<void*>(&swe_field_any_kernel<double>);
  hipError_t e = hipEventRecord(ctx->ev0, s);
  void* args_fixed[] = {&a};
  void* args_any[] = {&a, &lmax_, &mmax_, &B_, &chunk};
  if (e == hipSuccess)
    e = hipLaunchKernel(fn, dim3(nb), dim3(kSwePoints), lds ? args_any : args_fixed, lds, s);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev1, s);
  if (e == hipSuccess) e = hipMemcpyAsync(out, buf[3], sizes[3], hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  release();
  SMH_CHECK(ctx, e == hipSuccess, SMH_ERR_HIP, "SWE field kernel failed: %s", hipGetErrorString(e));
  if (field == 1) {
    // H = -i k / omega * sum
    const std::complex<double> f = std::complex<double>(0, -1) * std::complex<double>(k_re, k_im) / omega;
    for (int64_t j = 0; j < 3 * n_pts; ++j) {
      const std::complex<double> v = f * std::complex<double>(out[2 * j], out[2 * j + 1]);
      out[2 * j] = v.real();
      out[2 * j + 1] = v.imag();
    }
  }
  if (kernel_ms) *kernel_ms = ms;
  return SMH_OK;
}
