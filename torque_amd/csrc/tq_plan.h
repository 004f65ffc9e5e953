// tq_plan.h -- the contraction plan object and the argument blocks of the
// planning kernels (internal).
#pragma once

#include "tq_common.h"

struct LegList {
  int32_t a[TQ_MAX_RANK];
  int32_t b[TQ_MAX_RANK];
};

struct GeomArgs {
  int64_t n_pairs;
  int32_t a_rank, b_rank, nl;
  const i64 *a_begin, *a_end, *a_str, *a_off;
  const i64 *b_begin, *b_end, *b_str, *b_off;
  LegList legs;
  int32_t a_free[TQ_MAX_RANK];
  int32_t b_free[TQ_MAX_RANK];
};

// Whole-block panels. Plain panels list legs [free..., contracted in pair
// order...]; split-BF16 panels list [contracted..., free...] for the decode
// and store the three planes k-chunked (CopyJob), kdim padded to kpad.
struct PanelArgs {
  int64_t n_blocks;
  int32_t rank;
  int32_t nk;     // contracted legs
  int32_t split;  // 1 = split-BF16 panel
  const i64 *begin, *end, *strides, *offsets;
  int32_t order[TQ_MAX_RANK];
};

struct BoxArgs {
  int64_t n_pairs;
  int32_t is_b;
  int32_t split;
  int32_t kfast;  // plain panel stored k-fast ([contracted..., free...])
  int32_t self_rank, other_rank, nl, n_free;
  const i64 *self_begin, *self_end, *other_begin, *other_end;
  int32_t free_legs[TQ_MAX_RANK];
  int32_t self_legs[TQ_MAX_RANK];
  int32_t other_legs[TQ_MAX_RANK];
};

// Flags OR-ed over all pairs by the geometry kernel.
enum : int32_t {
  kNeedPackA = 1,       // some pair's A view is not a plain matrix
  kNeedPackB = 2,
  kNeedPairPanelA = 4,  // some pair's A sub-box is not plain inside a block panel
  kNeedPairPanelB = 8,
  kTooLarge = 16,
  kNotFullA = 32,  // some pair uses only part of its A block's contracted extent
  kNotFullB = 64,
};

enum { kModeDirect = 0, kModeBlockPanel = 1, kModePairPanel = 2 };

struct PlanUpload {
  const i64 *a_begin = nullptr, *a_end = nullptr, *a_str = nullptr, *a_off = nullptr;
  const i64 *b_begin = nullptr, *b_end = nullptr, *b_str = nullptr, *b_off = nullptr;
};

struct PackSide {
  int mode = kModeDirect;        // active: kModeDirect skips the pack
  int built_mode = kModeDirect;  // what the jobs below pack
  bool split = false;  // panels hold three BF16 planes
  DevBuf panel;      // packed operand
  DevBuf panel_off;  // per block or per pair
  DevBuf jobs;       // CopyJob list
  DevBuf wg;         // int64 workgroup prefix (scratch)
  DevBuf wg32;       // int32 workgroup prefix used by the gather kernel
  int64_t n_jobs = 0;
  int64_t total_wg = 0;
  int64_t panel_elems = 0;  // elements per plane
};

// Smallest min(M, N) of a pair in the large MFMA tile class.
constexpr int kLargeMin = 128;

// One MFMA tile class of a plan: its configuration and tile list.
struct TileClass {
  DevBuf tile_pair, tile_first;
  DevBuf tile_ctr;  // the XCDs' tile counters of the wave-tile GEMM launch
  int64_t n_tiles = 0;
  int tiles_cfg = -1;
  bool ready = false;
  bool runs = false;  // the tile list indexes runs (wave-tile configurations)
  int cfg = -1;  // user-selected configuration (-1 = automatic)
  int auto_cfg = 0;
  int current() const { return cfg >= 0 ? cfg : auto_cfg; }
  // split-K: each pair's contracted range cut into up to `split` slices;
  // the slices' partial products go to ws (one output block per slice) and a
  // reduction adds them into C in slice order (deterministic)
  int split = 1;
  // the split was chosen by the caller, the autotuner or the cost model (the
  // analytic default no longer applies)
  bool split_user = false;
  int tiles_split = 1;  // the split the tile list was built for
  DevBuf vdesc;         // descriptors of the (pair, slice) GEMMs
  DevBuf ws;
  DevBuf red_jobs;
  int64_t n_red_jobs = 0;
};

// k extent of one split-K slice: ceil(K / split) rounded up to whole 32-value
// k steps (pairs with K <= 32 stay whole).
inline int64_t split_slice_k(int64_t k, int split) {
  if (split <= 1) return k;
  const int64_t ks = ceil_div(ceil_div(k, split), 32) * 32;
  return ks < k ? ks : k;
}

// One reduction job of split-K: C[c_off + e] = sum over slices q of
// ws[ws_off + q * stride + e] for e < n, slices added in order.
struct SplitReduceJob {
  i64 c_off, ws_off;
  int32_t n, stride, slices, pad;
};
constexpr int kSplitReduceElems = 4096;
constexpr int kMaxSplit = 64;
tq_status tq_launch_split_reduce(tq_handle_t h, tq_dtype out, const SplitReduceJob *jobs,
                                 int64_t n_jobs, const void *ws, void *c);

struct tq_plan_s {
  tq_handle_t h = nullptr;
  tq_dtype dtype = TQ_F64;
  tq_precision prec = TQ_PREC_NATIVE;
  int ra = 0, rb = 0, ro = 0, nl = 0;
  int64_t na = 0, nb = 0;
  int32_t legs_a[TQ_MAX_RANK] = {0}, legs_b[TQ_MAX_RANK] = {0};
  int32_t a_free[TQ_MAX_RANK] = {0}, b_free[TQ_MAX_RANK] = {0};
  // plain (not split) panels of an operand whose fastest leg in memory is a
  // contracted one are stored k-fast, so the pack reads along that leg
  bool kfast_panel_a = false, kfast_panel_b = false;
  std::vector<i64> out_dims;
  int64_t n_pairs = 0;
  int64_t out_n_elem = 1;
  double flops = 0.0, bytes = 0.0;
  int a_layout = 0, b_layout = 0;
  // split-BF16 tier with both operands plain matrices in place: the GEMM
  // splits FP64 values as it stages them, and no pack runs
  bool split_direct = false;
  // split-BF16: a device word or-ed with 1 by the pack and GEMM kernels when
  // an operand value lies outside the split's domain (zero from plan time)
  DevBuf range_flag;
  // split-BF16 operand form: the plan may read plain operands in place or
  // pack them into split panels (built on first use); split_form is the
  // caller's request (0 automatic: the autotuner times both, in place until
  // then; 1 panels; 2 in place)
  bool can_split_direct = false;
  int split_form = 0;
  bool panels_built = false;
  int64_t geom_flags = 0;
  DevBuf desc_inplace, desc_panels;
  std::vector<GemmDesc> h_desc_inplace, h_desc_panels;

  DevBuf meta;       // uploaded A/B metadata
  PlanUpload up;     // device pointers into meta
  DevBuf pair_a, pair_b;  // int32 per pair
  DevBuf sub_off;    // int64 [2*n_pairs]: A and B sub-block offsets
  DevBuf desc;       // GemmDesc per pair
  DevBuf out_be;     // out begin then out end (device)
  DevBuf c_off;      // int64 [n_pairs+1]
  DevBuf a_used, b_used;  // uint8 per block
  DevBuf order;      // int64 work list (pair ids, largest cost first)
  PackSide pack_a, pack_b;

  // Pairs with K <= stream_max_k go to the streaming small-K kernel (segment
  // list over column slices of the pairs); the others to MFMA tiles in two classes, each with its own tile
  // configuration: class 1 when min(M, N) >= large_min, class 0 otherwise.
  int stream_max_k = 0;
  DevBuf chunk_pair, chunk_first;
  DevBuf stream_desc;  // GemmDesc per column slice of a streamed pair
  int64_t n_chunks = 0;
  bool chunks_ready = false;
  int large_min = kLargeMin;
  int64_t max_ld = 0;  // largest lda / ldb of the GEMM descriptors
  bool wave_ok() const { return max_ld < kWaveMaxLd; }
  TileClass cls[2];
  int64_t n_tiles = 0;  // both classes

  // rank-0 reduction scratch
  DevBuf partial, red_tmp;

  // captured execute (tq_plan_capture)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipStream_t capture_stream = nullptr;
  bool is_batch = false;
  bool cached = false;  // owned by its handle's plan cache

  // Releases a captured execute. Every call that may rebuild the tile or
  // segment lists (configuration setters, autotune) drops the graph first:
  // it holds the old buffers' addresses and tile counts.
  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }
  ~tq_plan_s() {
    drop_graph();
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
  }

  // runs of pairs sharing an A operand with adjacent outputs (built on the
  // host from h_desc when a class first uses a wave-tile configuration)
  std::vector<GemmDesc> h_desc;
  std::vector<RunDesc> h_runs;
  DevBuf run_desc;
  bool runs_built = false;

  // merged output: GEMM units are groups of pairs sharing an output block;
  // n_pairs then counts groups and n_matched the matched pairs
  bool merged = false;
  int64_t n_matched = 0;
  std::vector<i64> h_pair_m, h_pair_n, h_pair_k;
  // streamed groups (every pair K <= kStreamMaxK, at most kGroupMaxMembers
  // pairs per group, not the split-BF16 tier): the operands stay packed as
  // for the unmerged plan and the streaming kernel sums each group's
  // members (member_desc, in group order, group g from h_member_first[g])
  // instead of reading K-concatenated group panels; every group streams
  bool member_stream = false;
  std::vector<GemmDesc> h_members;
  std::vector<int64_t> h_member_first;
  DevBuf member_first;  // h_member_first on the device
  DevBuf member_desc;
  DevBuf group_slices;  // GroupSlice per column slice of a streamed group

  // host copies
  std::vector<i64> h_m, h_n, h_k;
  std::vector<i64> h_out_begin, h_out_end, h_c_off;
};

// Builds the streaming segment list and both tile classes' lists for their
// current configurations (no work when already built).
tq_status plan_build_tiles(tq_plan_t pl);
// Switches a split-BF16 plan with plain operands between reading them in
// place (in_place) and packing them into split panels; drops the work lists
// and any captured graph when the form changes.
tq_status plan_set_split_form(tq_plan_t pl, bool in_place);
// Analytic configuration choice for one tile class.
int tq_plan_pick_config(tq_plan_t pl, int cls);
int tq_plan_pick_split(tq_plan_t pl, int cls, int cfg);
// The analytic defaults of both tile classes: configuration, and split-K
// where no caller, autotuner or cost model chose one.
void plan_pick_defaults(tq_plan_t pl);
// Whether configuration cfg can run this plan: wave tiles address operands
// with 32-bit offsets; FP16 data runs the workgroup-tile kernel with
// 16-value k steps only.
inline bool plan_config_ok(const tq_plan_s &pl, int cfg) {
  const GemmConfig g = tq_gemm_config(pl.prec, cfg);
  if (pl.dtype == TQ_F16) return !g.wave_tile && g.bk % 16 == 0;
  return !(g.wave_tile && !pl.wave_ok());
}
// 0, 1: MFMA tile class of a pair; -1: streaming.
inline int pair_class(const tq_plan_s &pl, int64_t m, int64_t n, int64_t k) {
  if (pl.member_stream || k <= pl.stream_max_k) return -1;
  return (m >= pl.large_min && n >= pl.large_min) ? 1 : 0;
}
tq_status tq_sort_by_cost(tq_handle_t h, const i64 *cost, int64_t n, i64 *order);
