// Special functions of complex argument in FP64, shared by the table kernels
// and the host check entry points:
//   * Bessel J_n(z), n = 0..nmax: power series (|z| <= 4), Miller's backward
//     recurrence normalised by J_0 + 2 sum J_2k = 1 (4 < |z| <= 25), Hankel's
//     asymptotic expansion for J_0, J_1 plus forward recurrence (|z| > 25).
//     Negative real part uses J_n(-z) = (-1)^n J_n(z).
//   * spherical Hankel h_n^(1)(z) by upward recurrence from h_0 = -i e^{iz}/z
//     and h_1 = -e^{iz} (z + i) / z^2.
//   * normalized associated Legendre P_l^m(cos t) of real argument.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace smh {

struct zc {
  double re, im;
};

__host__ __device__ inline zc zmul(zc a, zc b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline zc zadd(zc a, zc b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ inline zc zsub(zc a, zc b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ inline zc zscale(zc a, double s) { return {a.re * s, a.im * s}; }
__host__ __device__ inline double zabs(zc a) { return hypot(a.re, a.im); }
__host__ __device__ inline zc zdiv(zc a, zc b) {
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  double r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__host__ __device__ inline zc zexp(zc a) {
  double e = exp(a.re);
  double s, c;
  sincos(a.im, &s, &c);
  return {e * c, e * s};
}
__host__ __device__ inline zc zcos(zc a) {  // cos(x+iy) = cos x cosh y - i sin x sinh y
  double s, c;
  sincos(a.re, &s, &c);
  return {c * cosh(a.im), -s * sinh(a.im)};
}
__host__ __device__ inline zc zsin(zc a) {  // sin(x+iy) = sin x cosh y + i cos x sinh y
  double s, c;
  sincos(a.re, &s, &c);
  return {s * cosh(a.im), c * sinh(a.im)};
}
__host__ __device__ inline zc zsqrt(zc a) {
  double r = zabs(a);
  if (r == 0.0) return {0.0, 0.0};
  double t = sqrt(0.5 * (r + fabs(a.re)));
  if (a.re >= 0) return {t, a.im / (2 * t)};
  return {fabs(a.im) / (2 * t), a.im >= 0 ? t : -t};
}

constexpr int kMaxBesselOrder = 16;

// J_n(z) for n = 0..nmax (nmax <= kMaxBesselOrder).
__host__ __device__ inline void bessel_j_all(int nmax, zc z, zc* out) {
  double az = zabs(z);
  if (az == 0.0) {
    out[0] = {1.0, 0.0};
    for (int n = 1; n <= nmax; ++n) out[n] = {0.0, 0.0};
    return;
  }
  bool flip = z.re < 0.0;
  if (flip) z = {-z.re, -z.im};
  if (az <= 4.0) {
    zc mz2 = zscale(zmul(z, z), -0.25);
    zc half = zscale(z, 0.5);
    zc lead = {1.0, 0.0};  // (z/2)^n / n!
    for (int n = 0; n <= nmax; ++n) {
      if (n > 0) lead = zscale(zmul(lead, half), 1.0 / n);
      zc term = lead, sum = lead;
      for (int k = 1; k < 60; ++k) {
        term = zscale(zmul(term, mz2), 1.0 / (double(k) * double(n + k)));
        sum = zadd(sum, term);
        if (zabs(term) < 1e-18 * zabs(sum)) break;
      }
      out[n] = sum;
    }
  } else if (az <= 25.0) {
    int N = 2 * (int)((1.3 * az + 30.0) * 0.5);
    if (N < nmax + 20) N = nmax + 20 + (nmax & 1);
    // the unnormalised values are kept in out[] itself (no local array, so
    // that callers may pass LDS and keep their registers)
    zc fp1 = {0.0, 0.0}, f = {1e-30, 0.0};
    zc norm = {0.0, 0.0};
    for (int n = 0; n <= nmax; ++n) out[n] = {0.0, 0.0};
    zc inv_z = zdiv(zc{1.0, 0.0}, z);
    for (int n = N; n >= 1; --n) {
      // f_{n-1} = (2n/z) f_n - f_{n+1}
      zc fm1 = zsub(zscale(zmul(inv_z, f), 2.0 * n), fp1);
      fp1 = f;
      f = fm1;
      int idx = n - 1;
      if (idx <= nmax) out[idx] = f;
      if (idx > 0 && (idx % 2) == 0) norm = zadd(norm, zscale(f, 2.0));
      double a = zabs(f);
      if (a > 1e200) {
        f = zscale(f, 1e-200);
        fp1 = zscale(fp1, 1e-200);
        norm = zscale(norm, 1e-200);
        for (int q = 0; q <= nmax; ++q) out[q] = zscale(out[q], 1e-200);
      }
    }
    norm = zadd(norm, f);  // + f_0
    for (int n = 0; n <= nmax; ++n) out[n] = zdiv(out[n], norm);
  } else {
    // Hankel asymptotic expansion for J_0 and J_1.
    zc inv8z = zdiv(zc{1.0, 0.0}, zscale(z, 8.0));
    zc jv[2];
    for (int nu = 0; nu < 2; ++nu) {
      double mu = 4.0 * nu * nu;
      zc P = {1.0, 0.0}, Q = {0.0, 0.0};
      zc term = {1.0, 0.0};
      double prev = 1e300;
      for (int k = 1; k < 80; ++k) {
        double c = (mu - double(2 * k - 1) * double(2 * k - 1)) / double(k);
        term = zscale(zmul(term, inv8z), c);
        double at = zabs(term);
        if (at > prev) break;  // asymptotic series starts to diverge
        prev = at;
        // k odd -> Q with sign (-1)^((k-1)/2); k even -> P with sign (-1)^(k/2)
        if (k & 1)
          Q = ((k >> 1) & 1) ? zsub(Q, term) : zadd(Q, term);
        else
          P = ((k >> 1) & 1) ? zsub(P, term) : zadd(P, term);
        if (at < 1e-18) break;
      }
      zc chi = {z.re - (0.5 * nu + 0.25) * M_PI, z.im};
      zc val = zsub(zmul(P, zcos(chi)), zmul(Q, zsin(chi)));
      zc pre = zsqrt(zdiv(zc{2.0 / M_PI, 0.0}, z));
      jv[nu] = zmul(pre, val);
    }
    out[0] = jv[0];
    if (nmax >= 1) out[1] = jv[1];
    zc inv_z = zdiv(zc{1.0, 0.0}, z);
    for (int n = 1; n < nmax; ++n) out[n + 1] = zsub(zscale(zmul(inv_z, out[n]), 2.0 * n), out[n - 1]);
  }
  if (flip)
    for (int n = 1; n <= nmax; n += 2) out[n] = {-out[n].re, -out[n].im};
}

// Spherical Hankel h_n^(1)(z), n = 0..nmax, z != 0.
__host__ __device__ inline void spherical_hankel_all(int nmax, zc z, zc* out) {
  zc eiz = zexp(zc{-z.im, z.re});
  zc inv_z = zdiv(zc{1.0, 0.0}, z);
  zc h0 = zmul(zmul(zc{0.0, -1.0}, eiz), inv_z);
  out[0] = h0;
  if (nmax == 0) return;
  zc h1 = zmul(zscale(eiz, -1.0), zmul(zadd(z, zc{0.0, 1.0}), zmul(inv_z, inv_z)));
  out[1] = h1;
  for (int n = 1; n < nmax; ++n) out[n + 1] = zsub(zscale(zmul(inv_z, out[n]), 2.0 * n + 1.0), out[n - 1]);
}

// Normalized associated Legendre P_l^m(cos t) for real (ct, st), l <= lmax,
// m <= l, P[l * (lmax + 1) + m] (the normalization of smuthi's
// legendre_normalized, utility/math.py:38-104).
__host__ __device__ inline void legendre_p(int lmax, double ct, double st, double* P) {
  const int W = lmax + 1;
  for (int j = 0; j < W * W; ++j) P[j] = 0.0;
  P[0] = sqrt(0.5);
  if (lmax >= 1) P[W] = sqrt(1.5) * ct;
  for (int l = 1; l < lmax; ++l) {
    const double a = sqrt((2.0 * l + 1) * (2.0 * l + 3)) / (l + 1);
    const double b = (double)l / (l + 1) * sqrt((2.0 * l + 3) / (2.0 * l - 1));
    P[(l + 1) * W] = a * ct * P[l * W] - b * P[(l - 1) * W];
  }
  double prod = 1.0;
  for (int m = 1; m <= lmax; ++m) {
    prod *= (2.0 * m + 1) / (2.0 * m);
    double stm = 1.0;
    for (int i = 0; i < m; ++i) stm *= st;
    P[m * W + m] = sqrt(prod / 2.0) * stm;
    for (int l = m; l < lmax; ++l) {
      const double den = (double)(l + 1 - m) * (l + 1 + m);
      const double a = sqrt((2.0 * l + 1) * (2.0 * l + 3) / den);
      const double b = sqrt((2.0 * l + 3) * (l - m) * (l + m) / ((2.0 * l - 1) * den));
      P[(l + 1) * W + m] = a * ct * P[l * W + m] - (l - 1 >= m ? b * P[(l - 1) * W + m] : 0.0);
    }
  }
}

}  // namespace smh
