// Pieces shared by the radial and the volume lookup operators: work-item
// construction over bin entries, per-item fixed-point statistics, the chunk
// reduction and small helper kernels. Included by the translation unit of
// each operator kind and by the autotuner (everything here has internal
// linkage).
#pragma once

#include <algorithm>
#include <vector>

#include "smh_op.h"
#include "smh_radial_kernel.h"

namespace {

using namespace smh;

inline int blocks_for(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

struct KernelInfo {
  void* fn = nullptr;
  int threads = 0;
  int sy = 0;     // accumulator stride per receiver
  int slots = 0;  // kBatch-RHS kernels in slot order: slots per particle (0: plain layout)
  void (*slot_tables)(int*) = nullptr;  // host slot order: slot_of[B], then n_of[slots]
};

// Receiver chunk of a precision: as many receivers (at most 512, the pair
// record's receiver field) as the LDS accumulator [Rc][SY] and one counter
// int per receiver fit in one workgroup's LDS, less 1 KiB.
inline int chunk_size(const smh_op* op, size_t real_size, int sy, size_t& lds) {
  const size_t limit = std::min<size_t>(op->ctx->lds_per_block ? op->ctx->lds_per_block : 65536, 163840) - 1024;
  const size_t per = (size_t)sy * real_size + sizeof(int);
  int rc = (int)std::min<size_t>(512, (limit - 16) / per);
  if (rc < 16) {
    lds = 0;
    return 0;
  }
  lds = (size_t)rc * per + 16;
  return rc;
}

// Sum the work-item accumulators of each receiver chunk in fixed order
// (y = sum, or y += sum when accumulate is set).
template <typename Real>
__global__ void reduce_kernel(const Real* __restrict__ ypart, const int* __restrict__ chunk_items, int Rc, int B,
                              int64_t n_part, typename Vec2<Real>::t* y, int accumulate) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_part * B) return;
  const int64_t i1 = j / B;
  const int n1 = (int)(j - i1 * B);
  const int c = (int)(i1 / Rc), il = (int)(i1 - (int64_t)c * Rc);
  Real re = 0, im = 0;
  for (int w = chunk_items[c]; w < chunk_items[c + 1]; ++w) {
    const Real* p = ypart + ((size_t)w * Rc + il) * 2 * B;
    re += p[n1];
    im += p[B + n1];
  }
  if (accumulate) {
    re += y[j].x;
    im += y[j].y;
  }
  y[j] = typename Vec2<Real>::t{re, im};
}

// Sum the work items of each chunk for the kBatch-RHS kernels: partial
// [Rc][kBatch][2][S] (slot order when slot is given, else S = B) -> y
// [n_part][B][kBatch] (y = sum, or y += sum when accumulate is set).
template <typename Real>
__global__ void reduce_batch_kernel(const Real* __restrict__ ypart, const int* __restrict__ chunk_items, int Rc,
                                    int B, int S, const int* __restrict__ slot, int64_t n_part,
                                    typename Vec2<Real>::t* y, int accumulate) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_part * B * kBatchRhs) return;
  const int r = (int)(j % kBatchRhs);
  const int64_t in = j / kBatchRhs;
  const int64_t i1 = in / B;
  const int n1 = (int)(in - i1 * B);
  const int k1 = slot ? slot[n1] : n1;
  const int c = (int)(i1 / Rc), il = (int)(i1 - (int64_t)c * Rc);
  Real re = 0, im = 0;
  for (int w = chunk_items[c]; w < chunk_items[c + 1]; ++w) {
    const Real* p = ypart + ((size_t)w * Rc + il) * 2 * kBatchRhs * S + (size_t)(2 * r) * S;
    re += p[k1];
    im += p[S + k1];
  }
  if (accumulate) {
    re += y[j].x;
    im += y[j].y;
  }
  y[j] = typename Vec2<Real>::t{re, im};
}

// Row sums of tap rows: rowsum[row] = max_n1 sum_n2 (|Re| + |Im|) of row
// `row` of a [rows][B][B] table (with `interleave`, even rows come from
// table and odd rows from mh, both indexed row / 2), with a small margin for
// the FP32 rounding of the tap table.
template <typename Src>
__global__ void tap_rowsum_kernel(const Src* __restrict__ table, const Src* __restrict__ mh, int64_t rows, int B,
                                  int interleave, float* rowsum) {
  for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < rows;
       row += (int64_t)gridDim.x * blockDim.x) {
    const Src* src = table;
    int64_t j = row;
    if (interleave) {
      j = row >> 1;
      if (row & 1) src = mh;
    }
    double best = 0.0;
    for (int n1 = 0; n1 < B; ++n1) {
      double sum = 0.0;
      for (int n2 = 0; n2 < B; ++n2) {
        const Src v = src[(j * B + n1) * B + n2];
        sum += fabs((double)v.x) + fabs((double)v.y);
      }
      best = fmax(best, sum);
    }
    rowsum[row] = (float)(best * 1.001);
  }
}

// Per work item: largest row sum over the tap rows its bin entries use (tap t
// of an entry at row x sits at x + t for 1-D lookups, vmode 0; 2-D lookups:
// at x + vol_tap_offset(vmode, s, t, len_z) over all of their z slices s).
__global__ void item_lsum_kernel(const int4* __restrict__ entries, const int2* __restrict__ items, int n_items,
                                 int n_entries, int T, int vmode, int len_z, const float* __restrict__ rowsum,
                                 float* item_lsum) {
  const int k = blockIdx.x;
  const int e0 = items[k].y, e1 = k + 1 < n_items ? items[k + 1].y : n_entries;
  const int ns = vmode >= 2 ? 4 : 1;
  float m = 0.f;
  for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x)
    for (int sl = 0; sl < ns; ++sl)
      for (int t = 0; t < T; ++t) {
        const int off = vmode ? vol_tap_offset(vmode, sl, t, len_z) : t;
        m = fmaxf(m, rowsum[entries[e].x + off]);
      }
  __shared__ float red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) item_lsum[k] = red[0];
}

// Per work item: eN = ceil(log2) of the most pairs a single receiver has among
// the item's pairs (a contiguous range of the records).
template <typename Rec>
__global__ void item_count_kernel(const Rec* __restrict__ recs, const int4* __restrict__ entries,
                                  const int2* __restrict__ items, int n_items, int n_entries, int* item_en) {
  __shared__ int hist[512];
  __shared__ int red[256];
  const int k = blockIdx.x;
  const int e0 = items[k].y, e1 = (k + 1 < n_items ? items[k + 1].y : n_entries) - 1;
  for (int j = threadIdx.x; j < 512; j += blockDim.x) hist[j] = 0;
  __syncthreads();
  const int p0 = entries[e0].y, p1 = entries[e1].y + entries[e1].z;
  for (int p = p0 + (int)threadIdx.x; p < p1; p += blockDim.x) atomicAdd(&hist[recs[p].ij >> kPairShift], 1);
  __syncthreads();
  int m = 0;
  for (int j = threadIdx.x; j < 512; j += blockDim.x) m = max(m, hist[j]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int eN = 1;
    while ((1 << eN) < red[0]) ++eN;
    item_en[k] = eN;
  }
}

// Largest |Re x| + |Im x| of an input, in two passes: kXmaxBlocks workgroups
// write their maxima after xmax[0], one workgroup takes the largest (times a
// headroom factor) into xmax[0]. The buffer holds 1 + kXmaxBlocks floats.
constexpr int kXmaxBlocks = 256;
template <typename V2>
__global__ void __launch_bounds__(256) xmax_part_kernel(const V2* __restrict__ x, int64_t n, float* xmax) {
  __shared__ float red[256];
  float m = 0.f;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const V2 v = x[j];
    m = fmaxf(m, (float)(fabs((double)v.x) + fabs((double)v.y)));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) xmax[1 + blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(kXmaxBlocks) xmax_final_kernel(float* xmax, float headroom) {
  __shared__ float red[kXmaxBlocks];
  red[threadIdx.x] = xmax[1 + threadIdx.x];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) xmax[0] = red[0] * headroom;
}
// FP32 input: the bound of the fixed-point scales.
inline void launch_xmax(const float2* x, int64_t n, float* xmax, hipStream_t s) {
  hipLaunchKernelGGL(xmax_part_kernel<float2>, dim3(kXmaxBlocks), dim3(256), 0, s, x, n, xmax);
  hipLaunchKernelGGL(xmax_final_kernel, dim3(1), dim3(kXmaxBlocks), 0, s, xmax, 1.f);
}
// FP64 input (the FP16-operand kernel's input is scaled by it while being
// converted): one ulp of headroom, so the scaled inputs stay inside [-1, 1].
inline void launch_xmax(const double2* x, int64_t n, float* xmax, hipStream_t s) {
  hipLaunchKernelGGL(xmax_part_kernel<double2>, dim3(kXmaxBlocks), dim3(256), 0, s, x, n, xmax);
  hipLaunchKernelGGL(xmax_final_kernel, dim3(1), dim3(kXmaxBlocks), 0, s, xmax, 1.f + 1.2e-7f);
}

// FP64 padded [n_part][B][kBatchRhs] -> FP64 input of the FP64 slot-order
// kernel (b32_index layout)
__global__ void to_slots64_kernel(const double2* __restrict__ v, int64_t n8, int B, int S,
                                  const int* __restrict__ slot, double2* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n8) return;
  const int r = (int)(e % kBatchRhs);
  const int64_t ip = e / kBatchRhs, p = ip / B;
  out[b32_index(p, (int)(ip - p * B), r, S, slot)] = v[e];
}

// Zero-fill n 32-bit words with a kernel (callers may capture the launch in a
// graph, where a memset node is avoided).
__global__ void zero_words_kernel(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    p[j] = 0u;
}
inline void fill_zero(void* p, size_t bytes, hipStream_t s) {
  const int64_t n = (int64_t)(bytes / 4);
  hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, s,
                     (uint32_t*)p, n);
}

// Work items per compute unit: items of one receiver chunk run one after the
// other on a CU's single workgroup slot, so ~16 rounds keep the last,
// partly filled round short. The autotuner (smh_op_autotune) may regroup.
constexpr int kItemsPerCu = 16;

// Launch the item statistics that depend on the grouping: the most pairs per
// receiver (always) and the largest tap-row sum (once the table is set).
inline void item_stats(smh_op::Binning& bn, hipStream_t s) {
  if (bn.n_items <= 0) return;
  bn.count_items(bn, s);
  if (bn.lsum_rows)
    hipLaunchKernelGGL(item_lsum_kernel, dim3(bn.n_items), dim3(256), 0, s, bn.d_entries, bn.d_items, bn.n_items,
                       bn.n_entries, bn.lsum_T, bn.lsum_vmode, bn.lsum_len_z, bn.lsum_rows, bn.d_item_lsum);
}

// Record the tap-row sums the item bounds come from and compute them.
inline void set_item_lsum(smh_op::Binning& bn, const float* rowsum, int T, int vmode, int len_z, hipStream_t s) {
  bn.lsum_rows = rowsum;
  bn.lsum_T = T;
  bn.lsum_vmode = vmode;
  bn.lsum_len_z = len_z;
  if (bn.n_items > 0)
    hipLaunchKernelGGL(item_lsum_kernel, dim3(bn.n_items), dim3(256), 0, s, bn.d_entries, bn.d_items, bn.n_items,
                       bn.n_entries, T, vmode, len_z, rowsum, bn.d_item_lsum);
}

// Host: group the binning's entries (h_entries, in (chunk, bin) order) into
// work items of similar cost, about items_per_cu per compute unit and never
// spanning two receiver chunks, and upload them with their launch order and
// statistics (replacing an earlier grouping). Cost of an entry: single-vector
// kernels one unit per tile of 16 pairs plus 1/4 for loading the bin's taps;
// the kBatch-RHS kernels one unit per pair plus one for the taps.
// order: 0 chunk order, 1 sorted by first table row and dealt to the XCDs in
// contiguous shares, n >= 2 cut at the boundaries of n equal bands of table
// rows and launched band after band (kRowBandOrder).
constexpr int kRowBandOrder = 2;
inline int group_work_items(smh_op* op, smh_op::Binning& bn, int items_per_cu, int order_mode) {
  const bool batch = bn.batch;
  auto cost = [batch](int np) { return batch ? np + 1.0 : (np + 15) / 16 + 0.25; };
  smh_ctx* ctx = op->ctx;
  hipStream_t s = ctx->stream;
  const std::vector<int4>& entries = bn.h_entries;
  const std::vector<int>& ent_chunk = bn.h_ent_chunk;
  double total = 0;
  for (const int4& e : entries) total += cost(e.z);
  const int64_t target_items = std::max<int64_t>((int64_t)items_per_cu * ctx->num_cu, bn.n_chunks);
  const double budget = std::max(1.0, total / (double)target_items);
  // row bands: n_bands equal ranges of the entries' first tap rows
  const int n_bands = order_mode >= kRowBandOrder ? order_mode : 1;
  int max_row = 0;
  for (const int4& e : entries) max_row = std::max(max_row, e.x);
  const int band_rows = (int)(((int64_t)max_row + n_bands) / n_bands);
  auto band_of = [&](size_t e) { return entries[e].x / band_rows; };
  std::vector<int2> items;
  std::vector<int> chunk_items(bn.n_chunks + 1, 0);
  int cur_chunk = -1;
  double run = 0;
  for (size_t e = 0; e < entries.size(); ++e) {
    const int c = ent_chunk[e];
    if (c != cur_chunk || run >= budget || (n_bands > 1 && band_of(e) != band_of(e - 1))) {
      while (cur_chunk < c) chunk_items[++cur_chunk] = (int)items.size();
      items.push_back(make_int2(c, (int)e));
      run = 0;
    }
    run += cost(entries[e].z);
  }
  while (cur_chunk < bn.n_chunks) chunk_items[++cur_chunk] = (int)items.size();
  const int ni = (int)items.size();
  // Launch order (order_mode 1): items sorted by their first table row;
  // blocks are dealt to the XCDs round-robin (XCD = block mod kXcds), so XCD x
  // gets the x-th contiguous share of the sorted list and its CUs run items of
  // neighbouring bins (from different receiver chunks) at the same time.
  // Row bands: the items of band 0 of every chunk first, then band 1, ..., so
  // all XCDs read the same band's tap rows at a time and share them through
  // the Infinity Cache. Order 0: blocks run the items in (chunk, bin) order.
  constexpr int kXcds = 8;
  std::vector<int> order(ni);
  for (int k = 0; k < ni; ++k) order[k] = k;
  if (n_bands > 1) {
    std::stable_sort(order.begin(), order.end(),
                     [&](int p, int q) { return band_of(items[p].y) < band_of(items[q].y); });
  } else if (order_mode == 1) {
    std::vector<int> sorted(order);
    std::stable_sort(sorted.begin(), sorted.end(),
                     [&](int p, int q) { return entries[items[p].y].x < entries[items[q].y].x; });
    std::vector<int> xoff(kXcds + 1, 0);
    for (int x = 0; x < kXcds; ++x) xoff[x + 1] = xoff[x] + (ni - x + kXcds - 1) / kXcds;
    for (int b = 0; b < ni; ++b) order[b] = sorted[xoff[b % kXcds] + b / kXcds];
  }
  // the previous grouping's arrays may still be read by queued work
  if (hipStreamSynchronize(s) != hipSuccess) return smh_fail(ctx, SMH_ERR_HIP, "stream failed before regrouping");
  void* old[] = {bn.d_items, bn.d_chunk_items, bn.d_item_lsum, bn.d_item_en, bn.d_order};
  for (void* p : old)
    if (p) hipFree(p);
  bn.d_items = nullptr;
  bn.d_chunk_items = nullptr;
  bn.d_item_lsum = nullptr;
  bn.d_item_en = nullptr;
  bn.d_order = nullptr;
  bn.n_items = ni;
  bn.items_per_cu = items_per_cu;
  bn.order_mode = order_mode;
  if (hipMalloc(&bn.d_items, std::max<size_t>(items.size(), 1) * sizeof(int2)) != hipSuccess ||
      hipMalloc(&bn.d_chunk_items, chunk_items.size() * sizeof(int)) != hipSuccess ||
      hipMalloc(&bn.d_item_lsum, std::max(ni, 1) * sizeof(float)) != hipSuccess ||
      hipMalloc(&bn.d_item_en, std::max(ni, 1) * sizeof(int)) != hipSuccess ||
      hipMalloc(&bn.d_order, std::max(ni, 1) * sizeof(int)) != hipSuccess ||
      hipMemcpyAsync(bn.d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(bn.d_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(bn.d_chunk_items, chunk_items.data(), chunk_items.size() * sizeof(int), hipMemcpyHostToDevice,
                     s) != hipSuccess) {
    return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (work items)");
  }
  item_stats(bn, s);
  if (hipStreamSynchronize(s) != hipSuccess) return smh_fail(ctx, SMH_ERR_HIP, "work item statistics failed");
  return SMH_OK;
}

// Host: bin entries (at most kEntryPairs pairs each) of every non-empty
// (chunk, bin) key, uploaded once, then grouped into work items
// (group_work_items). cnt / off: pair count and first pair of each key
// (chunk-major, bins_per_chunk keys per chunk); row_of(bin) gives the bin's
// first tap row.
template <typename Rec, typename RowOf>
int build_work_items(smh_op* op, smh_op::Binning& bn, const std::vector<int>& cnt, const std::vector<int>& off,
                     int64_t bins_per_chunk, RowOf row_of, bool batch = false) {
  smh_ctx* ctx = op->ctx;
  hipStream_t s = ctx->stream;
  bn.h_entries.clear();
  bn.h_ent_chunk.clear();
  for (int c = 0; c < bn.n_chunks; ++c)
    for (int64_t r = 0; r < bins_per_chunk; ++r) {
      const int64_t key = (int64_t)c * bins_per_chunk + r;
      const int n = cnt[key];
      for (int p0 = 0; p0 < n; p0 += kEntryPairs) {
        const int np = std::min(kEntryPairs, n - p0);
        bn.h_entries.push_back(make_int4(row_of(r), off[key] + p0, np, 0));
        bn.h_ent_chunk.push_back(c);
      }
    }
  bn.n_entries = (int)bn.h_entries.size();
  bn.batch = batch;
  bn.count_items = [](smh_op::Binning& b, hipStream_t st) {
    hipLaunchKernelGGL(item_count_kernel<Rec>, dim3(b.n_items), dim3(256), 0, st, (const Rec*)b.d_rec, b.d_entries,
                       b.d_items, b.n_items, b.n_entries, b.d_item_en);
  };
  if (hipMalloc(&bn.d_entries, std::max<size_t>(bn.h_entries.size(), 1) * sizeof(int4)) != hipSuccess ||
      hipMemcpyAsync(bn.d_entries, bn.h_entries.data(), bn.h_entries.size() * sizeof(int4), hipMemcpyHostToDevice,
                     s) != hipSuccess)
    return smh_fail(ctx, SMH_ERR_HIP, "device allocation failed (bin entries)");
  return group_work_items(op, bn, kItemsPerCu, 1);
}

}  // namespace
