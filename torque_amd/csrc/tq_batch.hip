// tq_batch.hip -- many independent contractions as one launch sequence, and
// HIP-graph capture of a plan's execute.
//
// A batch is a plan whose pair list is the union of several plans' pairs:
// their GEMM descriptors are rebased onto shared A, B and C buffers, their
// pack jobs onto shared panels, and the union is re-sorted by cost into one
// work list. Executing it packs all operands with one launch per side and
// multiplies every pair with one streaming launch and one grouped-GEMM
// launch. No reference counterpart: the reference runs one contraction at a
// time with per-block host loops (src/gpu/block_sparse.cuh:693-1011).
#include <algorithm>
#include <cstring>
#include <memory>

#include "tq_common.h"
#include "tq_plan.h"

namespace {

template <class T>
tq_status download(tq_handle_t h, std::vector<T> &dst, const DevBuf &src, int64_t n) {
  dst.resize(std::max<int64_t>(n, 0));
  if (n > 0)
    TQ_HIP(h, hipMemcpyAsync(dst.data(), src.p, sizeof(T) * n, hipMemcpyDeviceToHost,
                             h->stream));
  return TQ_SUCCESS;
}

template <class T>
tq_status upload(tq_handle_t h, DevBuf &dst, const std::vector<T> &src) {
  TQ_HIP(h, dst.alloc(sizeof(T) * std::max<size_t>(src.size(), 1)));
  if (!src.empty())
    TQ_HIP(h, hipMemcpyAsync(dst.p, src.data(), sizeof(T) * src.size(),
                             hipMemcpyHostToDevice, h->stream));
  return TQ_SUCCESS;
}

// Merges one operand side of the plans into the batch's side.
tq_status merge_side(tq_handle_t h, tq_plan_s &out, bool is_b, int32_t n,
                     const tq_plan_t *plans, const int64_t *base,
                     std::vector<int64_t> &panel_base) {
  PackSide &dst = is_b ? out.pack_b : out.pack_a;
  std::vector<CopyJob> jobs;
  int64_t panel = 0;
  bool any = false, split = false;
  panel_base.assign(n, 0);
  for (int32_t i = 0; i < n; ++i) {
    const PackSide &s = is_b ? plans[i]->pack_b : plans[i]->pack_a;
    if (s.mode == kModeDirect) continue;
    any = true;
    split = s.split;
    panel_base[i] = panel;
    std::vector<CopyJob> pj;
    TQ_CHECK(download(h, pj, s.jobs, s.n_jobs));
    TQ_HIP(h, hipStreamSynchronize(h->stream));
    for (CopyJob j : pj) {
      j.src_off += base[i];
      j.dst_off += panel;
      jobs.push_back(j);
    }
    // panels stay multiples of 8 elements so split rows keep 16-byte alignment
    panel += (s.panel_elems + 7) & ~int64_t(7);
  }
  dst.mode = any ? kModeBlockPanel : kModeDirect;
  dst.split = split;
  if (!any) return TQ_SUCCESS;
  std::vector<int32_t> wg(jobs.size() + 1, 0);
  int64_t total = 0;
  for (size_t q = 0; q < jobs.size(); ++q) {
    wg[q] = (int32_t)total;
    total += pack_job_wgs(jobs[q]);
  }
  if (total > 0x7fffffff)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "batch: pack too large");
  wg[jobs.size()] = (int32_t)total;
  TQ_CHECK(upload(h, dst.jobs, jobs));
  TQ_CHECK(upload(h, dst.wg32, wg));
  dst.n_jobs = (int64_t)jobs.size();
  dst.total_wg = total;
  dst.panel_elems = panel;
  if (split) {
    const size_t bytes = 3 * sizeof(uint16_t) * std::max<int64_t>(panel, 1);
    TQ_HIP(h, dst.panel.alloc(bytes));
    TQ_HIP(h, hipMemsetAsync(dst.panel.p, 0, bytes, h->stream));
  } else {
    TQ_HIP(h, dst.panel.alloc(tq_elem_size(out.dtype) * std::max<int64_t>(panel, 1)));
  }
  return TQ_SUCCESS;
}

}  // namespace

extern "C" {

tq_status tq_batch_create(tq_handle_t h, int32_t n, const tq_plan_t *plans,
                          const int64_t *a_offset, const int64_t *b_offset,
                          const int64_t *c_offset, tq_plan_t *batch) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (!batch || n < 1 || !plans || !a_offset || !b_offset || !c_offset)
    return TQ_ERROR_INVALID_VALUE;
  *batch = nullptr;
  for (int32_t i = 0; i < n; ++i) {
    if (!plans[i]) return TQ_ERROR_NOT_INITIALIZED;
    if (plans[i]->is_batch || plans[i]->ro == 0 || plans[i]->dtype != plans[0]->dtype ||
        plans[i]->prec != plans[0]->prec ||
        plans[i]->stream_max_k != plans[0]->stream_max_k)
      return tq_fail(h, TQ_ERROR_INVALID_VALUE,
                     "batch: plans must share dtype, precision and small-K threshold, "
                     "and have a tensor (not scalar) result");
    if (plans[i]->member_stream)
      return tq_fail(h, TQ_ERROR_INVALID_VALUE,
                     "batch: a merged plan that streams its groups (short pairs) "
                     "cannot join a batch");
    if (plans[i]->split_direct != plans[0]->split_direct)
      return tq_fail(h, TQ_ERROR_INVALID_VALUE,
                     "batch: split-BF16 plans must all read their operands in place "
                     "(plain matrices) or all from split panels");
  }
  std::unique_ptr<tq_plan_s> pl(new tq_plan_s());
  pl->h = h;
  pl->is_batch = true;
  pl->dtype = plans[0]->dtype;
  pl->prec = plans[0]->prec;
  pl->ro = plans[0]->ro;
  pl->stream_max_k = plans[0]->stream_max_k;
  pl->split_direct = plans[0]->split_direct;
  if (pl->prec == TQ_PREC_BF16X3) {
    TQ_HIP(h, pl->range_flag.alloc(sizeof(int32_t)));
    TQ_HIP(h, hipMemsetAsync(pl->range_flag.p, 0, sizeof(int32_t), h->stream));
  }
  std::vector<int64_t> pa_base, pb_base;
  TQ_CHECK(merge_side(h, *pl, false, n, plans, a_offset, pa_base));
  TQ_CHECK(merge_side(h, *pl, true, n, plans, b_offset, pb_base));

  std::vector<GemmDesc> desc;
  std::vector<i64> cost;
  int64_t out_end = 0;
  for (int32_t i = 0; i < n; ++i) {
    const tq_plan_s &p = *plans[i];
    std::vector<GemmDesc> pd;
    TQ_CHECK(download(h, pd, p.desc, p.n_pairs));
    TQ_HIP(h, hipStreamSynchronize(h->stream));
    for (GemmDesc d : pd) {
      d.a_off += (d.flags & kAPacked) ? pa_base[i] : a_offset[i];
      d.b_off += (d.flags & kBPacked) ? pb_base[i] : b_offset[i];
      d.c_off += c_offset[i];
      desc.push_back(d);
      cost.push_back(pair_order_key(d.m, d.n, d.k));
      pl->h_m.push_back(d.m);
      pl->h_n.push_back(d.n);
      pl->h_k.push_back(d.k);
      pl->max_ld = std::max<int64_t>(pl->max_ld, std::max(d.lda, d.ldb));
    }
    pl->flops += p.flops;
    pl->bytes += p.bytes;
    out_end = std::max<int64_t>(out_end, c_offset[i] + p.out_n_elem);
  }
  pl->n_pairs = (int64_t)desc.size();
  pl->out_n_elem = out_end;
  if (pl->n_pairs > 0x7fffffff)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED, "batch: too many block pairs");
  TQ_CHECK(upload(h, pl->desc, desc));
  pl->h_desc = desc;
  DevBuf dcost;
  TQ_CHECK(upload(h, dcost, cost));
  TQ_HIP(h, pl->order.alloc(sizeof(i64) * std::max<int64_t>(pl->n_pairs, 1)));
  TQ_CHECK(tq_sort_by_cost(h, dcost.as<i64>(), pl->n_pairs, pl->order.as<i64>()));
  pl->a_layout = pl->pack_a.mode == kModeDirect ? 0 : 2;
  pl->b_layout = pl->pack_b.mode == kModeDirect ? 0 : 2;
  plan_pick_defaults(pl.get());
  if (pl->n_pairs > 0) TQ_CHECK(plan_build_tiles(pl.get()));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  *batch = pl.release();
  return TQ_SUCCESS;
}

tq_status tq_plan_capture(tq_plan_t pl, const void *a, const void *b, void *c) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  tq_handle_t h = pl->h;
  if (pl->ro == 0 || pl->n_pairs == 0)
    return tq_fail(h, TQ_ERROR_UNSUPPORTED,
                   "capture: scalar and empty results are not captured");
  // everything the execute needs is allocated before capture starts
  TQ_CHECK(plan_build_tiles(pl));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  pl->drop_graph();
  if (!pl->capture_stream) TQ_HIP(h, hipStreamCreateWithFlags(&pl->capture_stream,
                                                              hipStreamNonBlocking));
  hipStream_t user = h->stream;
  h->stream = pl->capture_stream;
  tq_status st = TQ_SUCCESS;
  hipError_t e = hipStreamBeginCapture(pl->capture_stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    st = tq_bs_contract_execute(pl, a, b, c);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(pl->capture_stream, &g);
    if (st == TQ_SUCCESS && e2 != hipSuccess) e = e2;
    pl->graph = g;
  }
  h->stream = user;
  if (e != hipSuccess)
    return tq_fail(h, TQ_ERROR_HIP, std::string("capture: ") + hipGetErrorString(e));
  if (st != TQ_SUCCESS) return st;
  TQ_HIP(h, hipGraphInstantiate(&pl->graph_exec, pl->graph, nullptr, nullptr, 0));
  return TQ_SUCCESS;
}

struct tq_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

tq_status tq_capture_begin(tq_handle_t h) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (h->capturing) return tq_fail(h, TQ_ERROR_INVALID_VALUE, "capture: already capturing");
  if (!h->capture_stream)
    TQ_HIP(h, hipStreamCreateWithFlags(&h->capture_stream, hipStreamNonBlocking));
  TQ_HIP(h, hipStreamBeginCapture(h->capture_stream, hipStreamCaptureModeThreadLocal));
  h->saved_stream = h->stream;
  h->stream = h->capture_stream;
  h->capturing = true;
  return TQ_SUCCESS;
}

tq_status tq_capture_end(tq_handle_t h, tq_graph_t *graph) {
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (!h->capturing) return tq_fail(h, TQ_ERROR_INVALID_VALUE, "capture: not capturing");
  h->stream = h->saved_stream;
  h->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(h->capture_stream, &g);
  if (e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return tq_fail(h, TQ_ERROR_HIP,
                   std::string("capture: ") + hipGetErrorString(e) +
                       " (an entry point that allocates or synchronises ran while capturing)");
  }
  if (!graph) {
    (void)hipGraphDestroy(g);
    return TQ_ERROR_INVALID_VALUE;
  }
  std::unique_ptr<tq_graph_s> out(new tq_graph_s());
  out->graph = g;
  TQ_HIP(h, hipGraphInstantiate(&out->exec, g, nullptr, nullptr, 0));
  *graph = out.release();
  return TQ_SUCCESS;
}

tq_status tq_graph_launch(tq_graph_t graph, tq_handle_t h) {
  if (!graph || !h) return TQ_ERROR_NOT_INITIALIZED;
  if (h->capturing) return tq_fail(h, TQ_ERROR_INVALID_VALUE, "graph launch: capturing");
  TQ_HIP(h, hipGraphLaunch(graph->exec, h->stream));
  return TQ_SUCCESS;
}

tq_status tq_graph_destroy(tq_graph_t graph) {
  if (!graph) return TQ_ERROR_NOT_INITIALIZED;
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  delete graph;
  return TQ_SUCCESS;
}

tq_status tq_plan_graph_launch(tq_plan_t pl) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (!pl->graph_exec)
    return tq_fail(pl->h, TQ_ERROR_NOT_INITIALIZED, "graph launch: plan not captured");
  TQ_HIP(pl->h, hipGraphLaunch(pl->graph_exec, pl->h->stream));
  return TQ_SUCCESS;
}

}  // extern "C"
