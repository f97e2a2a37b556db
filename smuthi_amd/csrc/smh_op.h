// Coupling operator state shared by the operator, table and solver sources.
#pragma once

#include "smh_common.h"

struct smh_op {
  smh_ctx* ctx = nullptr;
  int kind = 0;  // 0 radial, 1 volume, 2 explicit (dense FP64 / FP32 matrix)
  int64_t n_part = 0, n_unk = 0;
  int l_max = 0, m_max = 0, B = 0, Bh = 0, MP = 0;
  int interp = 0, T = 0, s_mode = 1, o_mode = 0;
  double rho_min = 0, h = 0;
  int64_t len_rho = 0, n_tt = 0;
  // volume operators
  double sz_min = 0, dz_min = 0;
  int64_t len_z = 0;

  // particles and index maps
  double2* d_xy = nullptr;    // radial: (x, y)
  double* d_xyz = nullptr;    // volume: [n_part][3]
  int* d_pk2pad = nullptr;    // packed system index -> padded index i*B + n_global
  std::vector<int> pk2pad;

  // tables
  bool has_table = false;
  double2* d_table = nullptr;        // FP64 table [len_rho][(len_z)][B][B]
  double2* d_table_minus = nullptr;  // volume: minus table
  double* d_tt64 = nullptr;          // tap table [n_tt][FR] in MFMA fragment order
  float* d_tt32 = nullptr;
  int FR = 0;                        // reals per tap row
  // FP32 kBatch-RHS kernel in slot order (multipoles grouped by |m|): its tap
  // table, the slot tables (slot_of[B], n_of[b32_S]) and the slots per
  // particle of its input layout (b32_S = B and no tables: plain layout)
  float* d_tt32b = nullptr;
  int* d_b32_tab = nullptr;
  int b32_S = 0;
  // FP64 kBatch-RHS kernel in the same slot order: FP64 polynomial
  // coefficients and its input in the b32_index layout
  double* d_tt64b = nullptr;
  double2* d_x64s = nullptr;
  // FP16-operand kBatch-RHS kernel (slot order, b16_index input layout):
  // FP16 polynomial coefficients and one power-of-two scale per bin
  _Float16* d_tth = nullptr;
  float* d_hscale = nullptr;
  int T_half = 0;         // its Horner terms: 2 (linear between nodes) or T
  float lin_ratio = -1;   // largest linear-interpolation error / bin scale over the used bins
  float* d_rowsum = nullptr;         // per tap row: max_n1 sum_n2 |Re| + |Im|
  float* d_xmax = nullptr;           // [0] largest |Re| + |Im| of the current input; [1..256] partial maxima

  // pair binning, one per precision (index 0: FP32, 1: FP64); the receiver
  // chunk size follows the LDS accumulator of that precision
  struct Binning {
    int Rc = 0, n_chunks = 0, n_items = 0, n_entries = 0;
    void* d_rec = nullptr;           // PairRec<Real> sorted by (chunk, bin)
    int4* d_entries = nullptr;       // (first tap row, first pair, pairs, -)
    int2* d_items = nullptr;         // (chunk, first entry)
    int* d_chunk_items = nullptr;    // items of chunk c: [chunk_items[c], chunk_items[c+1])
    float* d_item_lsum = nullptr;    // per item: largest tap-row sum over its bins
    int* d_item_en = nullptr;        // per item: ceil(log2) of the most pairs of one receiver
    int* d_order = nullptr;          // block -> item launch order (XCD-aware)
    size_t lds = 0;
    // grouping of the entries into work items (regrouped by the autotuner):
    // host copies of the entries and their chunks, the item cost model
    // (batch: per pair, else per 16-pair tile), the target items per CU and
    // the launch order (0 chunk order, 1 XCD-dealt by table row, n >= 2 row
    // bands)
    std::vector<int4> h_entries;
    std::vector<int> h_ent_chunk;
    bool batch = false;
    int items_per_cu = 0;
    int order_mode = 1;
    void (*count_items)(Binning&, hipStream_t) = nullptr;  // item_count_kernel for the record type
    // item_lsum_kernel arguments once the table is set (rows == nullptr: not yet)
    const float* lsum_rows = nullptr;
    int lsum_T = 0, lsum_vmode = 0, lsum_len_z = 0;
    size_t part_bytes = 0;  // partial-result bytes one work item writes (d_ypart holds n_items of them)
  } bin[2];
  Binning bbin[2];  // the same pairs binned for the kBatch-RHS kernel (smaller receiver chunks)
  int64_t n_pairs = 0;

  // volume lookup: complex64 tables [len_rho][len_z][B][B] over (rho, z1+z2)
  // (index 0, "plus") and (rho, z1-z2) (index 1, "minus"), their FP32 tap
  // tables in fragment order, row sums, and a binning per table and precision
  float2* d_vtab[2] = {nullptr, nullptr};
  float* d_vtt[2] = {nullptr, nullptr};
  float* d_vtts[2] = {nullptr, nullptr};    // the FP32 tap rows in slot fragment order (8-RHS slot kernel)
  double* d_vtts64[2] = {nullptr, nullptr};  // and in FP64 for the FP64 8-RHS slot kernel
  _Float16* d_vtth[2] = {nullptr, nullptr};  // and in FP16 for the FP16-operand 8-RHS kernel,
  float* d_vhs[2] = {nullptr, nullptr};      // each row divided by its power of two d_vhs[row]
  double* d_vtt64[2] = {nullptr, nullptr};  // FP64 tap tables (spline: FP64 moments)
  int vmode = 1, vq = 1;  // interpolation mode (vol_tap_offset) and tap rows per (rho, z) node
  float* d_vrowsum[2] = {nullptr, nullptr};
  Binning vbin[2][2];
  Binning vbbin[2][2];  // the same pairs binned for the kBatch-RHS kernel

  // explicit operator: dense W [n_part * B][n_part * B] on padded vectors
  double2* d_W64 = nullptr;
  float2* d_W32 = nullptr;
  // LU factors of M = 1 - diag(t) W (explicit operator, solver_type 'LU'),
  // the T they were formed with and the rocBLAS handle of the factorization
  double2* d_lu = nullptr;
  int* d_piv = nullptr;
  void* blas = nullptr;
  std::vector<double> lu_t;

  // particle ensembles ("segments") of one operator: particles
  // [seg_off[s], seg_off[s+1]) form ensemble s; pairs couple particles of one
  // ensemble only and the solver treats every (ensemble, right-hand side) as
  // a system of its own. One segment unless created for several ensembles.
  int n_seg = 1;
  std::vector<int64_t> seg_off;  // host, n_seg + 1 particle offsets
  int64_t* d_seg_off = nullptr;  // device copy
  int* d_seg_of = nullptr;       // device: ensemble of every particle (null: one ensemble)

  // right block-sparse preconditioner: (P v)_i = sum_k P[i][k] v_{nbr[i][k]},
  // complex128 blocks [n_part][P_k1][B][B] on the padded index and emitters
  // d_Pnbr [n_part][P_k1] (-1: empty); null: none. GMRES then solves
  // M P y = b and returns x = P y.
  double2* d_P = nullptr;
  int* d_Pnbr = nullptr;
  int P_k1 = 0;
  // near-field blocks A_ij = diag(t_i) W_ij of each particle and its
  // near_k1 - 1 nearest neighbours (smh_learn.hip)
  double2* d_near = nullptr;
  int* d_near_nbr = nullptr;
  int near_k1 = 0;

  // solver workspace (smh_solver.hip), grown on demand, freed with the operator
  void* solver_ws = nullptr;
  // kernels whose dynamic LDS limit has been raised for this operator
  std::vector<void*> attr_set;

  // work buffers
  void* d_ypart = nullptr;
  size_t ypart_bytes = 0;
  double2* d_pk_in = nullptr;
  double2* d_pk_out = nullptr;
  int64_t pk_cap = 0;  // capacity in complex entries
  void* d_xb = nullptr;  // kBatch-RHS padded vectors (allocated on first use, FP64 size)
  void* d_yb = nullptr;
  float2* d_x32 = nullptr;
  float2* d_y32 = nullptr;
  double2* d_x64 = nullptr;
  double2* d_y64 = nullptr;
  float last_kernel_ms = 0.f;
};

void smh_op_free(smh_op* op);

namespace smh {
// right-hand sides per batched matvec / batched GMRES block
constexpr int kBatchRhs = 8;
constexpr int kBatch = kBatchRhs;
// Index of (particle p, multipole n, rhs r) in the FP32 kBatch-RHS input:
// slot layout [n_part][4][kBatch][S/4] (slot k = slot[n] sits at (k % 4, r,
// k / 4)) when a slot table is given, else the plain layout [n_part][S = B][kBatch].
__host__ __device__ inline int64_t b32_index(int64_t p, int n, int r, int S, const int* slot) {
  if (!slot) return (p * S + n) * kBatchRhs + r;
  const int k = slot[n];
  return p * S * kBatchRhs + ((k & 3) * kBatchRhs + r) * (S / 4) + (k >> 2);
}
// Index (in halves) of Re x for (particle p, multipole n, rhs r) in the FP16
// input of the FP16-operand kBatch-RHS kernel: [n_part][4 (q)][kBatch][S/16
// (c)][2 (re, im)][4 (j)] for slot k = slot[n] = 16 c + 4 q + j; Im x sits 4
// halves later. The values are x / xmax, xmax the input's largest |Re| + |Im|.
__host__ __device__ inline int64_t b16_index(int64_t p, int n, int r, int S, const int* slot) {
  const int k = slot[n];
  const int c = k >> 4, q = (k >> 2) & 3, j = k & 3;
  return p * S * kBatchRhs * 2 + ((int64_t)(q * kBatchRhs + r) * (S / 16) + c) * 8 + j;
}
// FP64 padded kBatch-RHS vectors [n_part][B][kBatch] -> the FP16 input above
// (with the scale written to op->d_xmax); padding slots of out must be zero.
int op_to_half_batch(smh_op* op, const double2* v, void* out);
// packed -> padded index map of the particles' multipole blocks
int op_index_map(smh_op* op, const int* lmax, const int* mmax);
// volume operator: matvec, table upload / download (smh_volume.hip)
int vol_apply(smh_op* op, int precision, const void* x_pad, void* y_pad, bool time_kernel);
int vol_apply_batch(smh_op* op, int precision, const void* x_pad, void* y_pad, bool time_kernel);
// particle ensembles of an operator: n_ens offsets ens_off[0] = 0 < ... < ens_off[n_ens] = n_part
int op_set_segments(smh_op* op, int n_ens, const int64_t* ens_off);
int vol_upload(smh_op* op, const double* plus, const double* minus);
int vol_download(smh_op* op, double* plus, double* minus);
// explicit operator: dense matvec and the padded matrix (smh_explicit.hip)
int explicit_apply(smh_op* op, int precision, const void* x_pad, void* y_pad, bool time_kernel);
int explicit_download(smh_op* op, double* W);
// solver workspace release, and the drop of its captured graphs when the
// operator's buffers change (smh_solver.hip)
void solver_release(smh_op* op);
void solver_drop_graphs(smh_op* op);
// raise a kernel's dynamic LDS limit once per operator (outside stream capture)
int op_set_lds(smh_op* op, void* fn, size_t lds);
// y = P x (+ y when accumulate) for kBatch padded vectors [n_part][B][kBatch]
int op_apply_precond_batch(smh_op* op, const double2* x, double2* y, bool accumulate);
// preconditioner storage for k1 blocks per particle (kept when the size matches)
int op_reserve_precond(smh_op* op, int k1);
// explicit operator: LU solve on the device (smh_lu.hip) and its release
int lu_solve(smh_op* op, const double* t_diag, const double* rhs, double* x, int64_t nrhs, double* stats);
void lu_release(smh_op* op);
// kBatch-RHS matvec (radial operators): FP64 input and every output on padded
// [n_part][B][kBatch] vectors; the FP32 input in the layout of b32_index()
bool op_has_batch(const smh_op* op, int precision);
int op_apply_batch(smh_op* op, int precision, const void* x_pad, void* y_pad, bool time_kernel);
int op_unpack_batch(smh_op* op, int precision, const double2* pk, int nrhs, void* pad);
int op_pack_batch(smh_op* op, int precision, const void* pad, int nrhs, double2* pk);
int op_reserve_batch(smh_op* op);
// Launch one coupling matvec on padded device vectors (x -> y, [n_part][B]).
int op_apply(smh_op* op, int precision, const void* x_pad, void* y_pad, bool time_kernel);
// Ensure packed host staging buffers for nrhs vectors.
int op_reserve_packed(smh_op* op, int64_t nrhs);
// packed (smuthi layout) <-> padded conversions on device
int op_unpack(smh_op* op, int precision, const double2* pk, void* pad);
int op_pack(smh_op* op, int precision, const void* pad, double2* pk);
// derive the tap rows (and spline moments) from the FP64 table
int op_make_taps(smh_op* op);
}  // namespace smh
