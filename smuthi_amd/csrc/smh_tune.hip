// Launch-shape autotuner of the lookup matvecs.
//
// A lookup matvec launches one workgroup per work item; a work item is a run
// of bin entries of one receiver chunk (the chunk's accumulator lives in the
// workgroup's LDS). How finely the entries are cut into items sets the
// balance between the tail of the launch (few large items leave CUs idle in
// the last round) and per-item overhead (every item loads its entries' taps
// and writes its own partial sums, which the reduction reads back). Where the
// items go in the launch order sets cache reuse: sorted by table row and
// dealt to the XCDs in contiguous shares, the CUs of one XCD run neighbouring
// bins of different chunks together and share tap rows in that XCD's L2; cut
// into bands of table rows and launched band after band, every XCD reads the
// same band's rows at a time and they share it through the Infinity Cache
// (volume tables, whose rows far exceed the L2s, and whose receiver chunks
// each hold about one pair per bin).
//
// smh_op_autotune regroups the operator's entries for each candidate (items
// per CU, launch order), times the matvec of the requested shape (its kernels
// and the reduction of their partial sums, whose size grows with the number
// of items) on a synthetic input and keeps the fastest grouping. The best choice
// depends on the pair distribution (ensemble size and density, table
// resolution) and on the kernel (its LDS footprint decides how many
// workgroups share a CU).
#include <algorithm>
#include <vector>

#include "smh_lookup.h"

namespace {

// The binnings one matvec of this shape reads (volume operators: one per table).
std::vector<smh_op::Binning*> tuned_binnings(smh_op* op, bool batch, int pi) {
  std::vector<smh_op::Binning*> out;
  if (op->kind == 0) {
    out.push_back(batch ? &op->bbin[pi] : &op->bin[pi]);
  } else if (op->kind == 1) {
    for (int t = 0; t < 2; ++t) out.push_back(batch ? &op->vbbin[t][pi] : &op->vbin[t][pi]);
  }
  return out;
}

// Regroup every binning and grow the partial-result buffer when the new
// grouping needs more than it holds.
int regroup(smh_op* op, const std::vector<smh_op::Binning*>& bins, int per_cu, int order_mode) {
  smh_ctx* ctx = op->ctx;
  size_t need = 0;
  for (smh_op::Binning* bn : bins) {
    if (int rc = group_work_items(op, *bn, per_cu, order_mode)) return rc;
    need = std::max(need, (size_t)bn->n_items * bn->part_bytes);
  }
  if (need > op->ypart_bytes) {
    SMH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SMH_HIP(ctx, hipFree(op->d_ypart));
    op->d_ypart = nullptr;
    op->ypart_bytes = 0;
    SMH_HIP(ctx, hipMalloc(&op->d_ypart, need));
    op->ypart_bytes = need;
  }
  return SMH_OK;
}

}  // namespace

extern "C" int smh_op_autotune(smh_op* op, int batch, int precision, int iters, int n_cand, const int* items_per_cu,
                               const int* launch_order, double* matvec_ms, int* best) {
  if (!op) return SMH_ERR_ARG;
  smh_ctx* ctx = op->ctx;
  SMH_CHECK(ctx, op->kind == 0 || op->kind == 1, SMH_ERR_UNSUPPORTED, "the autotuner covers lookup operators");
  SMH_CHECK(ctx, op->has_table, SMH_ERR_STATE, "coupling table not set");
  SMH_CHECK(ctx, iters >= 1 && n_cand >= 1 && n_cand <= 64 && items_per_cu && launch_order && matvec_ms, SMH_ERR_ARG,
            "invalid autotune arguments");
  SMH_CHECK(ctx, precision == SMH_FP32 || precision == SMH_FP64 || (batch && precision == SMH_FP16), SMH_ERR_ARG,
            "invalid precision %d for %s matvecs", precision, batch ? "batched" : "single-vector");
  for (int c = 0; c < n_cand; ++c) {
    SMH_CHECK(ctx, items_per_cu[c] >= 1 && items_per_cu[c] <= 1024, SMH_ERR_ARG, "items per CU %d out of [1, 1024]",
              items_per_cu[c]);
    SMH_CHECK(ctx, launch_order[c] >= 0 && launch_order[c] <= 4096, SMH_ERR_ARG, "launch order %d out of [0, 4096]",
              launch_order[c]);
  }
  const int pi = precision == SMH_FP64 ? 1 : 0;
  if (batch)
    SMH_CHECK(ctx, op_has_batch(op, precision), SMH_ERR_UNSUPPORTED, "no batched matvec for this operator");
  const std::vector<smh_op::Binning*> bins = tuned_binnings(op, batch != 0, pi);
  for (smh_op::Binning* bn : bins)
    SMH_CHECK(ctx, bn->Rc > 0 && bn->count_items, SMH_ERR_UNSUPPORTED, "no %s kernel for this operator",
              pi ? "FP64" : "FP32");
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  // captured solver graphs hold the item arrays' addresses and grid sizes
  smh::solver_drop_graphs(op);
  int pick = -1;
  for (int c = 0; c < n_cand; ++c) {
    if (int rc = regroup(op, bins, items_per_cu[c], launch_order[c])) return rc;
    double kms = 0, tms = 0;
    const int rc = batch ? smh_op_bench_matvec_batch(op, precision, iters, &kms, &tms)
                         : smh_op_bench_matvec(op, precision, iters, &kms, &tms);
    if (rc) return rc;
    matvec_ms[c] = tms;
    if (pick < 0 || tms < matvec_ms[pick]) pick = c;
  }
  if (int rc = regroup(op, bins, items_per_cu[pick], launch_order[pick])) return rc;
  if (best) *best = pick;
  return SMH_OK;
}
