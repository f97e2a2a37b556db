// Shared definitions of the smh library: context, error reporting, complex
// helpers and the SVWF index tables used by the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/smh.h"

struct smh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string last_error;
  int num_cu = 0;
  size_t lds_per_block = 0;
  char arch[64] = {0};
  int live_ops = 0;               // operators created on this context and not yet destroyed
  bool destroy_requested = false; // smh_destroy called while operators were alive
  // device address of this device's fixed scratch array (smh_runtime.hip),
  // resolved once in smh_init; used by the entry points that allocate nothing
  char* scratch = nullptr;
};

// Size of the fixed per-device scratch array of the no-allocation entry points.
constexpr size_t kScratchBytes = (size_t)64 << 20;

// Release a context whose destruction was requested once its last operator is gone.
void smh_ctx_release_if_idle(smh_ctx* ctx);

// Record an error on the context and return its code.
int smh_fail(smh_ctx* ctx, int code, const char* fmt, ...);

#define SMH_HIP(ctx, call)                                                               \
  do {                                                                                   \
    hipError_t e__ = (call);                                                             \
    if (e__ != hipSuccess)                                                               \
      return smh_fail((ctx), SMH_ERR_HIP, "%s failed: %s (%s:%d)", #call,                \
                      hipGetErrorString(e__), __FILE__, __LINE__);                       \
  } while (0)

#define SMH_CHECK(ctx, cond, code, ...)                        \
  do {                                                         \
    if (!(cond)) return smh_fail((ctx), (code), __VA_ARGS__);  \
  } while (0)

// Multipole index n -> (tau, l, m) for given (l_max, m_max); tau-major, then l,
// then m (same enumeration as smuthi's multi_to_single_index).
inline int smh_blocksize(int l_max, int m_max) {
  return 2 * (m_max * (m_max + 2) + (l_max - m_max) * (2 * m_max + 1));
}

inline void smh_index_tables(int l_max, int m_max, std::vector<int>& tau, std::vector<int>& l,
                             std::vector<int>& m) {
  int bs = smh_blocksize(l_max, m_max);
  tau.assign(bs, 0);
  l.assign(bs, 0);
  m.assign(bs, 0);
  int n = 0;
  for (int t = 0; t < 2; ++t)
    for (int ll = 1; ll <= l_max; ++ll) {
      int ml = ll < m_max ? ll : m_max;
      for (int mm = -ml; mm <= ml; ++mm) {
        tau[n] = t;
        l[n] = ll;
        m[n] = mm;
        ++n;
      }
    }
}

// Index of (tau, l, m) in the (l_max, m_max) enumeration.
inline int smh_multi_to_single(int tau, int l, int m, int l_max, int m_max) {
  int tbs = m_max * (m_max + 2) + (l_max - m_max) * (2 * m_max + 1);
  int n = tau * tbs;
  if (l - 1 <= m_max)
    n += (l - 1) * (l + 1);
  else
    n += m_max * (m_max + 2) + (l - 1 - m_max) * (2 * m_max + 1);
  n += m + (l < m_max ? l : m_max);
  return n;
}

template <typename T>
struct cplx {
  T re, im;
};

template <typename T>
__host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename T>
__host__ __device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename T>
__host__ __device__ inline cplx<T> cscale(cplx<T> a, T s) {
  return {a.re * s, a.im * s};
}
template <typename T>
__host__ __device__ inline cplx<T> cdiv(cplx<T> a, cplx<T> b) {
  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    T r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    T r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}
