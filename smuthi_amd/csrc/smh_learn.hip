// Learned block-sparse right preconditioner for the coupling solve.
//
// No counterpart in smuthi, whose GMRES is unpreconditioned
// (linearsystem/linear_system.py:271-286). The preconditioner of particle i is
// the near-field correction of the T-matrix-scaled system M = 1 - T W:
//
//   (P v)_i = v_i + sum_{j in N(i)} S_ij A_ij v_j,   A_ij = diag(t_i) W_ij,
//
// over i itself and its K nearest neighbours N(i) (slot 0 = i), with W_ij
// the block the operator's lookup applies to that pair (the same taps and
// weights as its FP64 matvec) and S_ij = diag(s_ij[ch(n1)]) one complex scale
// per multipole channel ch = (tau, l) of the receiving multipole. s = 1 is the
// first-order near-field inverse of M (a Neumann step over the near pairs);
// a small network predicts s from each particle's neighbourhood.
//
// Network: 3 layers, x -> relu(x W1 + b1) -> relu(. W2 + b2) -> . W3 + b3,
// weights kept in FP32 (the optimiser's copy) and rounded to BF16 (round to
// nearest even) where they enter the products, as are the layer inputs and
// the back-propagated gradients; every product accumulates in FP32 on the
// 16x16x32 BF16 MFMA. Output y[row][slot][ch][re, im], s = 1 + y.
//
// Training (on the device): rows are the particles of training ensembles;
// for right-hand sides R and the solutions X = M^{-1} R the solver returned,
//   E_i,r = (P R_r)_i - X_i,r = (R - X)_i,r + sum_slot s_slot o U_i,slot,r,
//   U_i,slot,r = A_i,nbr(slot) R_nbr(slot),r,    L = sum |E|^2 / sum |X|^2,
// so P is fitted to the solution operator in the probes' directions (random
// right-hand sides make L an estimate of ||P - M^{-1}||_F over the pattern).
// dL/dy comes from one kernel per row, then the network's backward pass and
// Adam.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <random>
#include <vector>

#include "smh_op.h"
#include "smh_radial_kernel.h"

using namespace smh;

namespace {

int nblocks(int64_t n, int t = 256) { return (int)std::min<int64_t>((n + t - 1) / t, 1 << 20); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits, round to nearest even (finite inputs)
__host__ __device__ inline unsigned short bf16_bits(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// ---- BF16 MFMA GEMM ---------------------------------------------------------
// C[M][N] (+)= op(A) op(B): op(A)[m][k] = TA ? A[k * lda + m] : A[m * lda + k],
// op(B)[k][n] = TB ? B[n * ldb + k] : B[k * ldb + n]; FP32 in memory, rounded
// to BF16 in LDS, FP32 accumulation. Epilogue: + bias[n], relu, times
// (mask[m][n] > 0) (the relu derivative of a forward activation); with
// atomic set the K range is split over blockIdx.z and the tiles are added to
// C with atomics (C zeroed by the caller).
constexpr int kTM = 64, kTN = 64, kTK = 32, kLdsK = kTK + 8;

struct GemmArgs {
  const float* A;
  const float* B;
  float* C;
  const float* bias;
  const float* mask;
  int M, N, K, lda, ldb, ldc, ldm;
  int relu, atomic, kper;
};

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
  __shared__ unsigned short As[kTM][kLdsK];  // [m][k]
  __shared__ unsigned short Bs[kTN][kLdsK];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
  const int k_begin = blockIdx.z * g.kper, k_end = min(g.K, k_begin + g.kper);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // the wave's 32 x 32 sub-tile
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0, 0, 0, 0};
  const int rr = lane & 15, q = lane >> 4;
  for (int k0 = k_begin; k0 < k_end; k0 += kTK) {
    // stage: 64 x 32 of each operand, 8 values per thread
    for (int e = tid; e < kTM * kTK; e += 256) {
      const int mi = e / kTK, ki = e % kTK;
      const int m = m0 + mi, k = k0 + ki;
      float v = 0.f;
      if (m < g.M && k < k_end) v = TA ? g.A[(size_t)k * g.lda + m] : g.A[(size_t)m * g.lda + k];
      As[mi][ki] = bf16_bits(v);
    }
    for (int e = tid; e < kTN * kTK; e += 256) {
      const int ni = e / kTK, ki = e % kTK;
      const int n = n0 + ni, k = k0 + ki;
      float v = 0.f;
      if (n < g.N && k < k_end) v = TB ? g.B[(size_t)n * g.ldb + k] : g.B[(size_t)k * g.ldb + n];
      Bs[ni][ki] = bf16_bits(v);
    }
    __syncthreads();
    // lane (q, rr): A row rr of a 16-row tile with K values 8q .. 8q + 7, B column rr with the same K values
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const unsigned short* src = &As[wm + 16 * a + rr][8 * q];
#pragma unroll
      for (int e = 0; e < 8; ++e) af[a][e] = __builtin_bit_cast(__bf16, src[e]);
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const unsigned short* src = &Bs[wn + 16 * b + rr][8 * q];
#pragma unroll
      for (int e = 0; e < 8; ++e) bfr[b][e] = __builtin_bit_cast(__bf16, src[e]);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
  // result: register v of lane (q, rr) holds row 4 q + v, column rr of its 16 x 16 tile
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm + 16 * a + 4 * q + v, n = n0 + wn + 16 * b + rr;
        if (m >= g.M || n >= g.N) continue;
        float val = acc[a][b][v];
        if (g.atomic) {
          atomicAdd(&g.C[(size_t)m * g.ldc + n], val);
          continue;
        }
        if (g.bias) val += g.bias[n];
        if (g.relu) val = fmaxf(val, 0.f);
        if (g.mask && !(g.mask[(size_t)m * g.ldm + n] > 0.f)) val = 0.f;
        g.C[(size_t)m * g.ldc + n] = val;
      }
}

// column sums: out[n] = sum_m X[m][n] (bias gradients), one thread per column
__global__ void colsum_kernel(const float* __restrict__ X, int M, int N, float* out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += X[(size_t)m * N + n];
  out[n] = s;
}

__global__ void adam_kernel(float* p, const float* __restrict__ g, float* m, float* v, int64_t n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void zero_f_kernel(float* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.f;
}

// ---- near-field blocks A_ij = diag(t_i) W_ij --------------------------------
struct NearArgs {
  int kind;  // 0 radial, 1 volume
  int B, Bh, T, interp, s_mode, o_mode, vmode, vq, len_z;
  int G, KS, NG, FR;
  double rho_min, h, sz_min, dz_min;
  const double2* xy;      // radial
  const double* xyz;      // volume
  const double* tt64;     // radial FP64 tap table / volume FP64 spline tap tables
  const float* vtt[2];    // volume FP32 tap tables
  const double* vtt64[2];
  const int* m_of;        // azimuthal order of each padded multipole
  const double2* t;       // padded T diagonal
  const int* nbr;         // [n_part][k1], -1: empty slot
  int k1;
  double2* out;           // [n_part][k1][B][B]
};

__device__ inline int frag_off(const NearArgs& a, int m, int n2) {
  const int hh = n2 / a.KS, k = n2 - hh * a.KS;
  return ((((m >> 4) * a.G + hh) * a.NG + (k >> 2)) * 16 + (m & 15)) * 4 + (k & 3);
}

// one workgroup per (particle, slot), threads over the B x B entries
__global__ void __launch_bounds__(256) near_block_kernel(NearArgs a) {
  const int64_t i = blockIdx.x;
  const int slot = blockIdx.y;
  const int j = a.nbr[i * a.k1 + slot];
  const int BB = a.B * a.B;
  double2* out = a.out + ((size_t)i * a.k1 + slot) * BB;
  if (j < 0) {
    for (int e = threadIdx.x; e < BB; e += blockDim.x) out[e] = make_double2(0.0, 0.0);
    return;
  }
  double c = 1.0, s = 0.0;
  if (a.kind == 0) {
    const double dx = a.xy[i].x - a.xy[j].x, dy = a.xy[i].y - a.xy[j].y;
    int r;
    double w, rho;
    pair_geometry(dx, dy, a.rho_min, a.h, r, w, rho);
    if (rho > 0.0) {
      c = dx / rho;
      s = dy / rho;
    }
    const int64_t row0 = (int64_t)a.s_mode * r + a.o_mode;
    for (int e = threadIdx.x; e < BB; e += blockDim.x) {
      const int n1 = e / a.B, n2 = e - n1 * a.B;
      double re = 0.0, im = 0.0;
      for (int tp = 0; tp < a.T; ++tp) {
        const double wt = tap_weight<double>(a.interp, tp, w);
        const double* row = a.tt64 + (size_t)(row0 + tp) * a.FR;
        re += wt * row[frag_off(a, n1, n2)];
        im += wt * row[frag_off(a, a.Bh + n1, n2)];
      }
      // W = L e^{i (m2 - m1) phi}, then the row's T-matrix entry
      const int dm = a.m_of[n2] - a.m_of[n1];
      const double ph = atan2(s, c) * dm;
      const double pc = cos(ph), ps = sin(ph);
      const double wr = re * pc - im * ps, wi = re * ps + im * pc;
      const double2 tt = a.t[i * a.B + n1];
      out[e] = make_double2(tt.x * wr - tt.y * wi, tt.x * wi + tt.y * wr);
    }
  } else {
    const double* p1 = a.xyz + 3 * (size_t)i;
    const double* p2 = a.xyz + 3 * (size_t)j;
    const double dx = p1[0] - p2[0], dy = p1[1] - p2[1];
    const double rho = sqrt(dx * dx + dy * dy);
    if (rho > 0.0) {
      c = dx / rho;
      s = dy / rho;
    }
    const int ns = a.vmode >= 2 ? 4 : 1;
    for (int e = threadIdx.x; e < BB; e += blockDim.x) {
      const int n1 = e / a.B, n2 = e - n1 * a.B;
      double re = 0.0, im = 0.0;
      for (int tb = 0; tb < 2; ++tb) {
        const double zval = tb == 0 ? p1[2] + p2[2] : p1[2] - p2[2];
        const double zmin = tb == 0 ? a.sz_min : a.dz_min;
        int r;
        double wr, rr_;
        pair_geometry(dx, dy, a.rho_min, a.h, r, wr, rr_);
        const double tz = (zval - zmin) / a.h;
        const int jz = (int)floor(tz);
        const double wz = tz - jz;
        const int64_t node = ((int64_t)r * a.len_z + jz) * a.vq;
        for (int sl = 0; sl < ns; ++sl)
          for (int tp = 0; tp < 4; ++tp) {
            const double wt = vol_tap_weight<double>(a.vmode, sl, tp, wr, wz);
            const int64_t row = node + vol_tap_offset(a.vmode, sl, tp, a.len_z);
            if (a.vq == 4) {
              const double* rp = a.vtt64[tb] + (size_t)row * a.FR;
              re += wt * rp[frag_off(a, n1, n2)];
              im += wt * rp[frag_off(a, a.Bh + n1, n2)];
            } else {
              const float* rp = a.vtt[tb] + (size_t)row * a.FR;
              re += wt * (double)rp[frag_off(a, n1, n2)];
              im += wt * (double)rp[frag_off(a, a.Bh + n1, n2)];
            }
          }
      }
      const int dm = a.m_of[n2] - a.m_of[n1];
      const double ph = atan2(s, c) * dm;
      const double pc = cos(ph), ps = sin(ph);
      const double wr = re * pc - im * ps, wi = re * ps + im * pc;
      const double2 tt = a.t[i * a.B + n1];
      out[e] = make_double2(tt.x * wr - tt.y * wi, tt.x * wi + tt.y * wr);
    }
  }
}

// P[i][slot] = [slot == 0] I + diag(s[i][slot][ch(n1)]) A[i][slot]
__global__ void form_precond_kernel(const double2* __restrict__ A, const float* __restrict__ s, const int* __restrict__ ch,
                                    int64_t n_part, int k1, int B, int n_ch, double2* P) {
  const int64_t total = n_part * k1 * B * B;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (int64_t)gridDim.x * blockDim.x) {
    const int64_t blk = f / (B * B);
    const int e = (int)(f - blk * B * B), n1 = e / B, n2 = e - n1 * B;
    const int slot = (int)(blk % k1);
    const float* sp = s + ((size_t)blk * n_ch + ch[n1]) * 2;
    const double sr = (double)sp[0], si = (double)sp[1];
    const double2 a = A[f];
    double re = sr * a.x - si * a.y, im = sr * a.y + si * a.x;
    if (slot == 0 && n1 == n2) re += 1.0;
    P[f] = make_double2(re, im);
  }
}

// ---- training data and loss --------------------------------------------------
// U[row][slot][r][n] = sum_n2 A[i][slot][n][n2] R[r][nbr][n2] (complex64), E0[row][r][n] = R - X
// (R, X padded [nrhs][n_part][B] FP64); one workgroup per (particle, slot)
__global__ void __launch_bounds__(256) near_products_kernel(const double2* __restrict__ A, const int* __restrict__ nbr,
                                                            const double2* __restrict__ R, int64_t n_part, int k1,
                                                            int B, int nrhs, float2* U) {
  const int64_t i = blockIdx.x;
  const int slot = blockIdx.y;
  const int j = nbr[i * k1 + slot];
  const double2* Ab = A + ((size_t)i * k1 + slot) * B * B;
  float2* Ub = U + ((size_t)i * k1 + slot) * nrhs * B;
  for (int e = threadIdx.x; e < nrhs * B; e += blockDim.x) {
    const int r = e / B, n = e - r * B;
    double re = 0.0, im = 0.0;
    if (j >= 0) {
      const double2* x = R + ((size_t)r * n_part + j) * B;
      for (int n2 = 0; n2 < B; ++n2) {
        const double2 a = Ab[(size_t)n * B + n2], v = x[n2];
        re += a.x * v.x - a.y * v.y;
        im += a.x * v.y + a.y * v.x;
      }
    }
    Ub[e] = make_float2((float)re, (float)im);
  }
}

__global__ void residual0_kernel(const double2* __restrict__ R, const double2* __restrict__ X, int64_t n_part, int B,
                                 int nrhs, float2* E0) {
  const int64_t total = n_part * nrhs * B;
  for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < total; f += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = f / (nrhs * B);
    const int e = (int)(f - i * nrhs * B), r = e / B, n = e - r * B;
    const size_t src = ((size_t)r * n_part + i) * B + n;
    E0[f] = make_float2((float)(R[src].x - X[src].x), (float)(R[src].y - X[src].y));
  }
}

// Per row: E = E0 + sum_slot s_slot o U_slot with s = 1 + y, loss_row = sum |E|^2,
// dy[slot][ch] = 2 (Re, Im) sum_{r, n in ch} conj(U) E, both scaled by inv_norm.
__global__ void __launch_bounds__(256) loss_grad_kernel(const float* __restrict__ y, const float2* __restrict__ U,
                                                        const float2* __restrict__ E0, const int* __restrict__ ch,
                                                        int64_t rows, int k1, int B, int nrhs, int n_ch,
                                                        float inv_norm, float prior, float* dy, float* loss_row) {
  extern __shared__ float2 Es[];  // [nrhs][B]
  const int64_t i = blockIdx.x;
  const int ldy = k1 * n_ch * 2;
  const float* yi = y + (size_t)i * ldy;
  const float2* Ui = U + (size_t)i * k1 * nrhs * B;
  const float2* Ei = E0 + (size_t)i * nrhs * B;
  __shared__ float red[256];
  float l = 0.f;
  for (int e = threadIdx.x; e < nrhs * B; e += blockDim.x) {
    const int n = e % B;
    const int c = ch[n];
    float re = Ei[e].x, im = Ei[e].y;
    for (int slot = 0; slot < k1; ++slot) {
      const float sr = 1.f + yi[(slot * n_ch + c) * 2], si = yi[(slot * n_ch + c) * 2 + 1];
      const float2 u = Ui[(size_t)slot * nrhs * B + e];
      re += sr * u.x - si * u.y;
      im += sr * u.y + si * u.x;
    }
    Es[e] = make_float2(re, im);
    l += re * re + im * im;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  // the pull towards the first-order structure (y = 0): prior * |y_i|^2
  float lp = 0.f;
  for (int j = threadIdx.x; j < ldy; j += blockDim.x) lp += yi[j] * yi[j];
  __syncthreads();
  const float ld = red[0];
  __syncthreads();
  red[threadIdx.x] = lp;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_row[i] = ld * inv_norm + prior * red[0];
  for (int sc = threadIdx.x; sc < k1 * n_ch; sc += blockDim.x) {
    const int slot = sc / n_ch, c = sc - slot * n_ch;
    float gr = 0.f, gi = 0.f;
    for (int e = 0; e < nrhs * B; ++e) {
      if (ch[e % B] != c) continue;
      const float2 u = Ui[(size_t)slot * nrhs * B + e], ee = Es[e];
      gr += u.x * ee.x + u.y * ee.y;  // Re(conj(u) e)
      gi += u.x * ee.y - u.y * ee.x;  // Im(conj(u) e)
    }
    dy[(size_t)i * ldy + sc * 2] = 2.f * gr * inv_norm + 2.f * prior * yi[sc * 2];
    dy[(size_t)i * ldy + sc * 2 + 1] = 2.f * gi * inv_norm + 2.f * prior * yi[sc * 2 + 1];
  }
}

}  // namespace

struct smh_mlp {
  smh_ctx* ctx = nullptr;
  int n_in = 0, n_h = 0, n_out = 0;
  int64_t n_params = 0;
  float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
  int64_t adam_t = 0;
  // activations of `cap` rows
  int64_t cap = 0;
  float *d_x = nullptr, *d_h1 = nullptr, *d_h2 = nullptr, *d_y = nullptr, *d_dy = nullptr, *d_dh2 = nullptr,
        *d_dh1 = nullptr;
  // training set
  int64_t rows = 0, data_cap = 0;
  int k1 = 0, B = 0, nrhs = 0, n_ch = 0;
  float* d_feat = nullptr;
  float2* d_U = nullptr;
  float2* d_E0 = nullptr;
  int* d_ch = nullptr;
  float* d_loss = nullptr;
  double xnorm2 = 0.0;
  double prior_weight = 0.0;  // w of the pull towards y = 0, w / rows per row

  // parameter offsets: W1 [n_in][n_h], b1, W2 [n_h][n_h], b2, W3 [n_h][n_out], b3
  int64_t oW1() const { return 0; }
  int64_t ob1() const { return (int64_t)n_in * n_h; }
  int64_t oW2() const { return ob1() + n_h; }
  int64_t ob2() const { return oW2() + (int64_t)n_h * n_h; }
  int64_t oW3() const { return ob2() + n_h; }
  int64_t ob3() const { return oW3() + (int64_t)n_h * n_out; }
};

namespace {

void mlp_free_acts(smh_mlp* m) {
  float* p[] = {m->d_x, m->d_h1, m->d_h2, m->d_y, m->d_dy, m->d_dh2, m->d_dh1};
  for (float* q : p)
    if (q) (void)hipFree(q);
  m->d_x = m->d_h1 = m->d_h2 = m->d_y = m->d_dy = m->d_dh2 = m->d_dh1 = nullptr;
  m->cap = 0;
}

int mlp_reserve(smh_mlp* m, int64_t rows) {
  if (rows <= m->cap) return SMH_OK;
  mlp_free_acts(m);
  smh_ctx* ctx = m->ctx;
  auto al = [&](float** p, int64_t n) { return hipMalloc(p, (size_t)std::max<int64_t>(n, 1) * sizeof(float)) == hipSuccess; };
  if (!al(&m->d_x, rows * m->n_in) || !al(&m->d_h1, rows * m->n_h) || !al(&m->d_h2, rows * m->n_h) ||
      !al(&m->d_y, rows * m->n_out) || !al(&m->d_dy, rows * m->n_out) || !al(&m->d_dh2, rows * m->n_h) ||
      !al(&m->d_dh1, rows * m->n_h)) {
    mlp_free_acts(m);
    return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (network activations, %lld rows)", (long long)rows);
  }
  m->cap = rows;
  return SMH_OK;
}

int gemm(smh_ctx* ctx, bool ta, bool tb, const float* A, int lda, const float* B, int ldb, float* C, int ldc, int M,
         int N, int K, const float* bias, bool relu, const float* mask, int ldm, int splits = 1) {
  GemmArgs g{A, B, C, bias, mask, M, N, K, lda, ldb, ldc, ldm, relu ? 1 : 0, splits > 1 ? 1 : 0, 0};
  g.kper = (K + splits - 1) / splits;
  g.kper = (g.kper + kTK - 1) / kTK * kTK;
  const dim3 grid((unsigned)((N + kTN - 1) / kTN), (unsigned)((M + kTM - 1) / kTM), (unsigned)splits);
  hipStream_t s = ctx->stream;
  if (!ta && !tb) hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, dim3(256), 0, s, g);
  if (!ta && tb) hipLaunchKernelGGL((gemm_bf16_kernel<false, true>), grid, dim3(256), 0, s, g);
  if (ta && !tb) hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, dim3(256), 0, s, g);
  if (ta && tb) hipLaunchKernelGGL((gemm_bf16_kernel<true, true>), grid, dim3(256), 0, s, g);
  SMH_HIP(ctx, hipGetLastError());
  return SMH_OK;
}

// forward over rows of d_x (already on the device): h1, h2, y
int mlp_forward_dev(smh_mlp* m, int64_t rows, const float* x) {
  smh_ctx* ctx = m->ctx;
  const float* P = m->d_p;
  int rc;
  if ((rc = gemm(ctx, false, false, x, m->n_in, P + m->oW1(), m->n_h, m->d_h1, m->n_h, (int)rows, m->n_h, m->n_in,
                 P + m->ob1(), true, nullptr, 0)))
    return rc;
  if ((rc = gemm(ctx, false, false, m->d_h1, m->n_h, P + m->oW2(), m->n_h, m->d_h2, m->n_h, (int)rows, m->n_h, m->n_h,
                 P + m->ob2(), true, nullptr, 0)))
    return rc;
  return gemm(ctx, false, false, m->d_h2, m->n_h, P + m->oW3(), m->n_out, m->d_y, m->n_out, (int)rows, m->n_out,
              m->n_h, P + m->ob3(), false, nullptr, 0);
}

// gradient of sum(dy . y) over the parameters into d_g (forward activations present)
int mlp_backward_dev(smh_mlp* m, int64_t rows, const float* x, const float* dy) {
  smh_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  const float* P = m->d_p;
  float* G = m->d_g;
  int rc;
  hipLaunchKernelGGL(zero_f_kernel, dim3(1024), dim3(256), 0, s, G, m->n_params);
  const int splits = (int)std::max<int64_t>(1, std::min<int64_t>(64, rows / 2048));
  // layer 3: dW3 = h2^T dy, db3 = colsum(dy), dh2 = (dy W3^T) o [h2 > 0]
  if ((rc = gemm(ctx, true, false, m->d_h2, m->n_h, dy, m->n_out, G + m->oW3(), m->n_out, m->n_h, m->n_out, (int)rows,
                 nullptr, false, nullptr, 0, splits)))
    return rc;
  hipLaunchKernelGGL(colsum_kernel, dim3(nblocks(m->n_out)), dim3(256), 0, s, dy, (int)rows, m->n_out, G + m->ob3());
  if ((rc = gemm(ctx, false, true, dy, m->n_out, P + m->oW3(), m->n_out, m->d_dh2, m->n_h, (int)rows, m->n_h, m->n_out,
                 nullptr, false, m->d_h2, m->n_h)))
    return rc;
  // layer 2
  if ((rc = gemm(ctx, true, false, m->d_h1, m->n_h, m->d_dh2, m->n_h, G + m->oW2(), m->n_h, m->n_h, m->n_h, (int)rows,
                 nullptr, false, nullptr, 0, splits)))
    return rc;
  hipLaunchKernelGGL(colsum_kernel, dim3(nblocks(m->n_h)), dim3(256), 0, s, m->d_dh2, (int)rows, m->n_h, G + m->ob2());
  if ((rc = gemm(ctx, false, true, m->d_dh2, m->n_h, P + m->oW2(), m->n_h, m->d_dh1, m->n_h, (int)rows, m->n_h, m->n_h,
                 nullptr, false, m->d_h1, m->n_h)))
    return rc;
  // layer 1
  if ((rc = gemm(ctx, true, false, x, m->n_in, m->d_dh1, m->n_h, G + m->oW1(), m->n_h, m->n_in, m->n_h, (int)rows,
                 nullptr, false, nullptr, 0, splits)))
    return rc;
  hipLaunchKernelGGL(colsum_kernel, dim3(nblocks(m->n_h)), dim3(256), 0, s, m->d_dh1, (int)rows, m->n_h, G + m->ob1());
  SMH_HIP(ctx, hipGetLastError());
  return SMH_OK;
}

// channel (tau, l) of every padded multipole of (l_max, m_max)
std::vector<int> channel_table(int l_max, int m_max) {
  std::vector<int> tau, ll, mm;
  smh_index_tables(l_max, m_max, tau, ll, mm);
  std::vector<int> ch(tau.size());
  for (size_t n = 0; n < tau.size(); ++n) ch[n] = tau[n] * l_max + (ll[n] - 1);
  return ch;
}

}  // namespace

// ---- C ABI -------------------------------------------------------------------

extern "C" int smh_mlp_create(smh_ctx* ctx, int n_in, int n_hidden, int n_out, uint64_t seed, smh_mlp** out) {
  if (!ctx || !out) return SMH_ERR_ARG;
  *out = nullptr;
  SMH_CHECK(ctx, n_in >= 1 && n_hidden >= 1 && n_out >= 1 && n_in <= 4096 && n_hidden <= 4096 && n_out <= 4096,
            SMH_ERR_ARG, "invalid layer sizes %d / %d / %d", n_in, n_hidden, n_out);
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  smh_mlp* m = new smh_mlp();
  m->ctx = ctx;
  m->n_in = n_in;
  m->n_h = n_hidden;
  m->n_out = n_out;
  m->n_params = m->ob3() + n_out;
  // He-normal weights for the relu layers, the output layer and the biases zero:
  // the untrained network outputs y = 0, i.e. the first-order near-field inverse
  std::vector<float> p(m->n_params, 0.f);
  std::mt19937_64 gen(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  const float s1 = std::sqrt(2.f / n_in), s2 = std::sqrt(2.f / n_hidden);
  for (int64_t k = 0; k < (int64_t)n_in * n_hidden; ++k) p[m->oW1() + k] = s1 * nd(gen);
  for (int64_t k = 0; k < (int64_t)n_hidden * n_hidden; ++k) p[m->oW2() + k] = s2 * nd(gen);
  auto fail = [&](const char* what) {
    void* q[] = {m->d_p, m->d_g, m->d_m, m->d_v};
    for (void* r : q)
      if (r) (void)hipFree(r);
    delete m;
    return smh_fail(ctx, SMH_ERR_HIP, "%s", what);
  };
  const size_t bytes = (size_t)m->n_params * sizeof(float);
  if (hipMalloc(&m->d_p, bytes) != hipSuccess || hipMalloc(&m->d_g, bytes) != hipSuccess ||
      hipMalloc(&m->d_m, bytes) != hipSuccess || hipMalloc(&m->d_v, bytes) != hipSuccess)
    return fail("device allocation failed (network parameters)");
  if (hipMemcpy(m->d_p, p.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return fail("parameter upload failed");
  hipLaunchKernelGGL(zero_f_kernel, dim3(256), dim3(256), 0, ctx->stream, m->d_m, m->n_params);
  hipLaunchKernelGGL(zero_f_kernel, dim3(256), dim3(256), 0, ctx->stream, m->d_v, m->n_params);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail("network initialisation failed");
  ctx->live_ops++;
  *out = m;
  return SMH_OK;
}

extern "C" int smh_mlp_destroy(smh_mlp* m) {
  if (!m) return SMH_OK;
  smh_ctx* ctx = m->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  mlp_free_acts(m);
  void* q[] = {m->d_p, m->d_g, m->d_m, m->d_v, m->d_feat, m->d_U, m->d_E0, m->d_ch, m->d_loss};
  for (void* r : q)
    if (r) (void)hipFree(r);
  delete m;
  ctx->live_ops--;
  smh_ctx_release_if_idle(ctx);
  return SMH_OK;
}

extern "C" int64_t smh_mlp_num_params(const smh_mlp* m) { return m ? m->n_params : 0; }

extern "C" int smh_mlp_get_params(smh_mlp* m, float* out) {
  if (!m || !out) return SMH_ERR_ARG;
  smh_ctx* ctx = m->ctx;
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  SMH_HIP(ctx, hipMemcpyAsync(out, m->d_p, m->n_params * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SMH_OK;
}

extern "C" int smh_mlp_set_params(smh_mlp* m, const float* p) {
  if (!m || !p) return SMH_ERR_ARG;
  smh_ctx* ctx = m->ctx;
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  SMH_HIP(ctx, hipMemcpyAsync(m->d_p, p, m->n_params * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(zero_f_kernel, dim3(256), dim3(256), 0, ctx->stream, m->d_m, m->n_params);
  hipLaunchKernelGGL(zero_f_kernel, dim3(256), dim3(256), 0, ctx->stream, m->d_v, m->n_params);
  m->adam_t = 0;
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SMH_OK;
}

extern "C" int smh_mlp_forward(smh_mlp* m, int64_t rows, const float* x, float* y) {
  if (!m || !x || !y) return SMH_ERR_ARG;
  smh_ctx* ctx = m->ctx;
  SMH_CHECK(ctx, rows >= 1 && rows < (1 << 30), SMH_ERR_ARG, "invalid row count");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  int rc = mlp_reserve(m, rows);
  if (rc) return rc;
  SMH_HIP(ctx, hipMemcpyAsync(m->d_x, x, rows * m->n_in * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  if ((rc = mlp_forward_dev(m, rows, m->d_x))) return rc;
  SMH_HIP(ctx, hipMemcpyAsync(y, m->d_y, rows * m->n_out * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SMH_OK;
}

extern "C" int smh_mlp_backward(smh_mlp* m, int64_t rows, const float* x, const float* dy, float* grad) {
  if (!m || !x || !dy || !grad) return SMH_ERR_ARG;
  smh_ctx* ctx = m->ctx;
  SMH_CHECK(ctx, rows >= 1 && rows < (1 << 30), SMH_ERR_ARG, "invalid row count");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  int rc = mlp_reserve(m, rows);
  if (rc) return rc;
  SMH_HIP(ctx, hipMemcpyAsync(m->d_x, x, rows * m->n_in * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  if ((rc = mlp_forward_dev(m, rows, m->d_x))) return rc;
  SMH_HIP(ctx, hipMemcpyAsync(m->d_dy, dy, rows * m->n_out * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  if ((rc = mlp_backward_dev(m, rows, m->d_x, m->d_dy))) return rc;
  SMH_HIP(ctx, hipMemcpyAsync(grad, m->d_g, m->n_params * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SMH_OK;
}

extern "C" int smh_op_near_blocks(smh_op* op, int k1, const int* nbr, const double* t_diag) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, k1 >= 1 && k1 <= 64 && nbr && t_diag, SMH_ERR_ARG, "invalid near-block arguments");
  SMH_CHECK(ctx, op->kind == 0 || op->kind == 1, SMH_ERR_UNSUPPORTED, "near blocks need a lookup operator");
  SMH_CHECK(ctx, op->has_table, SMH_ERR_STATE, "coupling table not set");
  for (int64_t i = 0; i < op->n_part; ++i) {
    SMH_CHECK(ctx, nbr[i * k1] == i, SMH_ERR_ARG, "slot 0 of particle %lld must be the particle itself", (long long)i);
    for (int k = 0; k < k1; ++k) {
      const int j = nbr[i * k1 + k];
      SMH_CHECK(ctx, j >= -1 && j < op->n_part, SMH_ERR_ARG, "neighbour index %d out of range", j);
    }
  }
  // neighbours must share the particle's ensemble
  if (op->n_seg > 1) {
    std::vector<int> seg(op->n_part);
    for (int e = 0; e < op->n_seg; ++e)
      for (int64_t i = op->seg_off[e]; i < op->seg_off[e + 1]; ++i) seg[i] = e;
    for (int64_t i = 0; i < op->n_part; ++i)
      for (int k = 0; k < k1; ++k) {
        const int j = nbr[i * k1 + k];
        SMH_CHECK(ctx, j < 0 || seg[j] == seg[i], SMH_ERR_ARG, "neighbour %d of particle %lld is in another ensemble",
                  j, (long long)i);
      }
  }
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int B = op->B;
  const size_t nblk = (size_t)op->n_part * k1;
  if (op->near_k1 != k1) {
    if (op->d_near) (void)hipFree(op->d_near);
    if (op->d_near_nbr) (void)hipFree(op->d_near_nbr);
    op->d_near = nullptr;
    op->d_near_nbr = nullptr;
    SMH_HIP(ctx, hipMalloc(&op->d_near, nblk * B * B * sizeof(double2)));
    SMH_HIP(ctx, hipMalloc(&op->d_near_nbr, nblk * sizeof(int)));
    op->near_k1 = k1;
  }
  SMH_HIP(ctx, hipMemcpyAsync(op->d_near_nbr, nbr, nblk * sizeof(int), hipMemcpyHostToDevice, s));
  // T diagonal (padded) and the multipole orders
  int rc = smh::op_reserve_packed(op, 1);
  if (rc) return rc;
  double2* d_t = nullptr;
  int* d_m = nullptr;
  std::vector<int> tau, ll, mm;
  smh_index_tables(op->l_max, op->m_max, tau, ll, mm);
  auto done = [&](int code) {
    if (d_t) (void)hipFree(d_t);
    if (d_m) (void)hipFree(d_m);
    return code;
  };
  if (hipMalloc(&d_t, (size_t)op->n_part * B * sizeof(double2)) != hipSuccess ||
      hipMalloc(&d_m, B * sizeof(int)) != hipSuccess ||
      hipMemcpyAsync(d_m, mm.data(), B * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(op->d_pk_in, t_diag, op->n_unk * sizeof(double2), hipMemcpyHostToDevice, s) != hipSuccess)
    return done(smh_fail(ctx, SMH_ERR_HIP, "near-block staging failed"));
  if ((rc = smh::op_unpack(op, SMH_FP64, op->d_pk_in, d_t))) return done(rc);
  NearArgs a{};
  a.kind = op->kind;
  a.B = B;
  a.Bh = op->Bh;
  a.T = op->T;
  a.interp = op->interp;
  a.s_mode = op->s_mode;
  a.o_mode = op->o_mode;
  a.vmode = op->vmode;
  a.vq = op->vq;
  a.len_z = (int)op->len_z;
  const TapGeom g(B, op->T);
  a.G = g.G;
  a.KS = g.KS;
  a.NG = g.NG;
  a.FR = g.FR;
  a.rho_min = op->rho_min;
  a.h = op->h;
  a.sz_min = op->sz_min;
  a.dz_min = op->dz_min;
  a.xy = op->d_xy;
  a.xyz = op->d_xyz;
  a.tt64 = op->d_tt64;
  a.vtt[0] = op->d_vtt[0];
  a.vtt[1] = op->d_vtt[1];
  a.vtt64[0] = op->d_vtt64[0];
  a.vtt64[1] = op->d_vtt64[1];
  a.m_of = d_m;
  a.t = d_t;
  a.nbr = op->d_near_nbr;
  a.k1 = k1;
  a.out = op->d_near;
  SMH_CHECK(ctx, op->kind == 1 || a.tt64 != nullptr, SMH_ERR_STATE, "FP64 tap table missing");
  hipLaunchKernelGGL(near_block_kernel, dim3((unsigned)op->n_part, (unsigned)k1), dim3(256), 0, s, a);
  const hipError_t e1 = hipGetLastError(), e2 = hipStreamSynchronize(s);
  if (e1 != hipSuccess || e2 != hipSuccess)
    return done(smh_fail(ctx, SMH_ERR_HIP, "near blocks failed: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2)));
  return done(SMH_OK);
}

extern "C" int smh_op_download_near_blocks(smh_op* op, double* out) {
  if (!op || !out) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, op->d_near != nullptr, SMH_ERR_STATE, "near blocks not computed");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  SMH_HIP(ctx, hipMemcpy(out, op->d_near, (size_t)op->n_part * op->near_k1 * op->B * op->B * sizeof(double2),
                         hipMemcpyDeviceToHost));
  return SMH_OK;
}

extern "C" int smh_op_set_near_scaling(smh_op* op, const float* s) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  if (!s) return smh_op_set_preconditioner(op, 0, nullptr, nullptr);
  SMH_CHECK(ctx, op->d_near != nullptr, SMH_ERR_STATE, "near blocks not computed");
  SMH_CHECK(ctx, smh::op_has_batch(op, SMH_FP32) && smh::op_has_batch(op, SMH_FP64), SMH_ERR_UNSUPPORTED,
            "preconditioned solves need an operator with the batched kernels");
  const int k1 = op->near_k1, B = op->B, n_ch = 2 * op->l_max;
  const std::vector<int> ch = channel_table(op->l_max, op->m_max);
  const size_t ns = (size_t)op->n_part * k1 * n_ch * 2;
  float* d_s = nullptr;
  int* d_ch = nullptr;
  auto done = [&](int code) {
    if (d_s) (void)hipFree(d_s);
    if (d_ch) (void)hipFree(d_ch);
    return code;
  };
  hipStream_t st = ctx->stream;
  if (hipMalloc(&d_s, ns * sizeof(float)) != hipSuccess || hipMalloc(&d_ch, B * sizeof(int)) != hipSuccess ||
      hipMemcpyAsync(d_s, s, ns * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(d_ch, ch.data(), B * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess)
    return done(smh_fail(ctx, SMH_ERR_HIP, "scaling upload failed"));
  if (int rc = smh::op_reserve_precond(op, k1)) return done(rc);
  SMH_HIP(ctx, hipMemcpyAsync(op->d_Pnbr, op->d_near_nbr, (size_t)op->n_part * k1 * sizeof(int),
                              hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(form_precond_kernel, dim3(4096), dim3(256), 0, st, (const double2*)op->d_near, (const float*)d_s,
                     (const int*)d_ch, op->n_part, k1, B, n_ch, op->d_P);
  const hipError_t e1 = hipGetLastError(), e2 = hipStreamSynchronize(st);
  if (e1 != hipSuccess || e2 != hipSuccess) return done(smh_fail(ctx, SMH_ERR_HIP, "preconditioner forming failed"));
  return done(SMH_OK);
}

extern "C" int smh_mlp_add_data(smh_mlp* m, smh_op* op, const float* feat, int64_t nrhs, const double* rhs,
                                const double* sol) {
  if (!m || !op || !feat || !rhs || !sol) return SMH_ERR_ARG;
  smh_ctx* ctx = m->ctx;
  SMH_CHECK(ctx, op->ctx == ctx, SMH_ERR_ARG, "operator of another context");
  SMH_CHECK(ctx, op->d_near != nullptr, SMH_ERR_STATE, "near blocks not computed");
  const int k1 = op->near_k1, B = op->B, n_ch = 2 * op->l_max;
  SMH_CHECK(ctx, nrhs >= 1 && nrhs <= 64, SMH_ERR_ARG, "1 to 64 right-hand sides per ensemble");
  SMH_CHECK(ctx, (int64_t)k1 * n_ch * 2 == m->n_out, SMH_ERR_ARG,
            "network output %d does not match %d slots x %d channels x 2", m->n_out, k1, n_ch);
  SMH_CHECK(ctx, m->rows == 0 || (m->k1 == k1 && m->B == B && m->nrhs == (int)nrhs && m->n_ch == n_ch), SMH_ERR_ARG,
            "every ensemble of the training set needs the same slots, blocksize and right-hand side count");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t n = op->n_part, new_rows = m->rows + n;
  // grow the training buffers
  if (new_rows > m->data_cap) {
    const int64_t cap = std::max<int64_t>(new_rows, 2 * m->data_cap);
    float* f = nullptr;
    float2 *u = nullptr, *e0 = nullptr;
    if (hipMalloc(&f, (size_t)cap * m->n_in * sizeof(float)) != hipSuccess ||
        hipMalloc(&u, (size_t)cap * k1 * nrhs * B * sizeof(float2)) != hipSuccess ||
        hipMalloc(&e0, (size_t)cap * nrhs * B * sizeof(float2)) != hipSuccess) {
      if (f) (void)hipFree(f);
      if (u) (void)hipFree(u);
      if (e0) (void)hipFree(e0);
      return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (training set of %lld rows)", (long long)cap);
    }
    if (m->rows > 0) {
      SMH_HIP(ctx, hipMemcpyAsync(f, m->d_feat, (size_t)m->rows * m->n_in * sizeof(float), hipMemcpyDeviceToDevice, s));
      SMH_HIP(ctx, hipMemcpyAsync(u, m->d_U, (size_t)m->rows * k1 * nrhs * B * sizeof(float2), hipMemcpyDeviceToDevice,
                                  s));
      SMH_HIP(ctx, hipMemcpyAsync(e0, m->d_E0, (size_t)m->rows * nrhs * B * sizeof(float2), hipMemcpyDeviceToDevice,
                                  s));
    }
    SMH_HIP(ctx, hipStreamSynchronize(s));
    if (m->d_feat) (void)hipFree(m->d_feat);
    if (m->d_U) (void)hipFree(m->d_U);
    if (m->d_E0) (void)hipFree(m->d_E0);
    m->d_feat = f;
    m->d_U = u;
    m->d_E0 = e0;
    m->data_cap = cap;
  }
  if (!m->d_ch) {
    const std::vector<int> ch = channel_table(op->l_max, op->m_max);
    SMH_HIP(ctx, hipMalloc(&m->d_ch, B * sizeof(int)));
    SMH_HIP(ctx, hipMemcpy(m->d_ch, ch.data(), B * sizeof(int), hipMemcpyHostToDevice));
  }
  m->k1 = k1;
  m->B = B;
  m->nrhs = (int)nrhs;
  m->n_ch = n_ch;
  // padded right-hand sides and solutions [nrhs][n][B]
  int rc = smh::op_reserve_packed(op, 1);
  if (rc) return rc;
  double2 *R = nullptr, *X = nullptr;
  auto done = [&](int code) {
    if (R) (void)hipFree(R);
    if (X) (void)hipFree(X);
    return code;
  };
  const size_t vb = (size_t)n * B * sizeof(double2);
  if (hipMalloc(&R, nrhs * vb) != hipSuccess || hipMalloc(&X, nrhs * vb) != hipSuccess)
    return done(smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (training vectors)"));
  double xn = 0.0;
  for (int64_t r = 0; r < nrhs; ++r) {
    const double* sr = sol + 2 * r * op->n_unk;
    for (int64_t k = 0; k < 2 * op->n_unk; ++k) xn += sr[k] * sr[k];
    SMH_HIP(ctx, hipMemcpyAsync(op->d_pk_in, rhs + 2 * r * op->n_unk, op->n_unk * sizeof(double2),
                                hipMemcpyHostToDevice, s));
    if ((rc = smh::op_unpack(op, SMH_FP64, op->d_pk_in, R + r * n * B))) return done(rc);
    SMH_HIP(ctx, hipMemcpyAsync(op->d_pk_in, sol + 2 * r * op->n_unk, op->n_unk * sizeof(double2),
                                hipMemcpyHostToDevice, s));
    if ((rc = smh::op_unpack(op, SMH_FP64, op->d_pk_in, X + r * n * B))) return done(rc);
  }
  SMH_HIP(ctx, hipMemcpyAsync(m->d_feat + (size_t)m->rows * m->n_in, feat, (size_t)n * m->n_in * sizeof(float),
                              hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(near_products_kernel, dim3((unsigned)n, (unsigned)k1), dim3(256), 0, s,
                     (const double2*)op->d_near, (const int*)op->d_near_nbr, (const double2*)R, n, k1, B, (int)nrhs,
                     m->d_U + (size_t)m->rows * k1 * nrhs * B);
  hipLaunchKernelGGL(residual0_kernel, dim3(4096), dim3(256), 0, s, (const double2*)R, (const double2*)X, n, B,
                     (int)nrhs, m->d_E0 + (size_t)m->rows * nrhs * B);
  const hipError_t e1 = hipGetLastError(), e2 = hipStreamSynchronize(s);
  if (e1 != hipSuccess || e2 != hipSuccess) return done(smh_fail(ctx, SMH_ERR_HIP, "training data kernels failed"));
  m->rows = new_rows;
  m->xnorm2 += xn;
  if (m->d_loss) (void)hipFree(m->d_loss);
  m->d_loss = nullptr;
  SMH_HIP(ctx, hipMalloc(&m->d_loss, (size_t)m->rows * sizeof(float)));
  return done(SMH_OK);
}

namespace {
// loss and dL/dy at the current outputs d_y over the whole training set
int mlp_loss_dev(smh_mlp* m, double* loss) {
  smh_ctx* ctx = m->ctx;
  hipStream_t s = ctx->stream;
  const size_t lds = (size_t)m->nrhs * m->B * sizeof(float2);
  const float inv = (float)(1.0 / std::max(m->xnorm2, 1e-300));
  hipLaunchKernelGGL(loss_grad_kernel, dim3((unsigned)m->rows), dim3(256), lds, s, (const float*)m->d_y,
                     (const float2*)m->d_U, (const float2*)m->d_E0, (const int*)m->d_ch, m->rows, m->k1, m->B, m->nrhs,
                     m->n_ch, inv, (float)(m->prior_weight / (double)m->rows), m->d_dy, m->d_loss);
  SMH_HIP(ctx, hipGetLastError());
  if (loss) {
    std::vector<float> l(m->rows);
    SMH_HIP(ctx, hipMemcpyAsync(l.data(), m->d_loss, m->rows * sizeof(float), hipMemcpyDeviceToHost, s));
    SMH_HIP(ctx, hipStreamSynchronize(s));
    double t = 0;
    for (float v : l) t += v;
    *loss = t;
  }
  return SMH_OK;
}
}  // namespace

extern "C" int smh_mlp_set_prior_weight(smh_mlp* m, double w) {
  if (!m) return SMH_ERR_ARG;
  SMH_CHECK(m->ctx, w >= 0 && w < 1e30, SMH_ERR_ARG, "invalid prior weight %g", w);
  m->prior_weight = w;
  return SMH_OK;
}

extern "C" int smh_mlp_data_loss(smh_mlp* m, const float* y, float* dy, double* loss) {
  if (!m) return SMH_ERR_ARG;
  smh_ctx* ctx = m->ctx;
  SMH_CHECK(ctx, m->rows > 0, SMH_ERR_STATE, "no training data");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  int rc = mlp_reserve(m, m->rows);
  if (rc) return rc;
  if (y)
    SMH_HIP(ctx, hipMemcpyAsync(m->d_y, y, (size_t)m->rows * m->n_out * sizeof(float), hipMemcpyHostToDevice,
                                ctx->stream));
  else if ((rc = mlp_forward_dev(m, m->rows, m->d_feat)))
    return rc;
  if ((rc = mlp_loss_dev(m, loss))) return rc;
  if (dy)
    SMH_HIP(ctx, hipMemcpyAsync(dy, m->d_dy, (size_t)m->rows * m->n_out * sizeof(float), hipMemcpyDeviceToHost,
                                ctx->stream));
  SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return SMH_OK;
}

extern "C" int smh_mlp_train(smh_mlp* m, int steps, double lr, double* losses) {
  if (!m) return SMH_ERR_ARG;
  smh_ctx* ctx = m->ctx;
  SMH_CHECK(ctx, m->rows > 0, SMH_ERR_STATE, "no training data");
  SMH_CHECK(ctx, steps >= 1 && lr > 0, SMH_ERR_ARG, "invalid steps / learning rate");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  int rc = mlp_reserve(m, m->rows);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  for (int it = 0; it < steps; ++it) {
    if ((rc = mlp_forward_dev(m, m->rows, m->d_feat))) return rc;
    if ((rc = mlp_loss_dev(m, losses ? losses + it : nullptr))) return rc;
    if ((rc = mlp_backward_dev(m, m->rows, m->d_feat, m->d_dy))) return rc;
    ++m->adam_t;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
    hipLaunchKernelGGL(adam_kernel, dim3(nblocks(m->n_params)), dim3(256), 0, s, m->d_p, (const float*)m->d_g, m->d_m,
                       m->d_v, m->n_params, (float)lr, b1, b2, 1e-8f, c1, c2);
    SMH_HIP(ctx, hipGetLastError());
  }
  SMH_HIP(ctx, hipStreamSynchronize(s));
  return SMH_OK;
}
