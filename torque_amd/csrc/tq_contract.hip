// tq_contract.hip -- execution of a contraction plan, the rank-0 (network
// scalar) branch, tile-configuration choice, and the dense contraction.
//
// Execute = [pack A] -> [pack B] -> [streaming small-K launch] -> one
// grouped-GEMM launch per MFMA tile class (or, for a rank-0 output, one
// pair-dot launch and a deterministic FP64 sum). This
// replaces the per-A-block sequence of src/gpu/block_sparse.cuh:693-1011
// (reshape, cuTT, strided-batched GEMM, reshape, D2D copy, dot + add<<<1,1>>>).
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "tq_common.h"
#include "tq_plan.h"

namespace {

// One wavefront per pair: partial[p] = sum_k A_p[k] * B_p[k] in FP64, where
// the operands are the pair's (M = N = 1) views read with the same layout
// flags as the GEMM.
template <class T>
__global__ __launch_bounds__(256) void pair_dot_kernel(
    const GemmDesc *__restrict__ descs, int64_t n_pairs,
    const T *__restrict__ a, const T *__restrict__ a_packed,
    const T *__restrict__ b, const T *__restrict__ b_packed,
    double *__restrict__ partial) {
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= n_pairs) return;
  const int lane = threadIdx.x & 63;
  const GemmDesc d = descs[p];
  const T *pa = ((d.flags & kAPacked) ? a_packed : a) + d.a_off;
  const T *pb = ((d.flags & kBPacked) ? b_packed : b) + d.b_off;
  const i64 sa = (d.flags & kAKFast) ? 1 : d.lda;
  const i64 sb = (d.flags & kBKFast) ? 1 : d.ldb;
  double s = 0.0;
  for (int k = lane; k < d.k; k += 64) s += (double)pa[sa * k] * (double)pb[sb * k];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) partial[p] = s;
}

template <class T>
__global__ void store_scalar_kernel(const double *__restrict__ v, T *__restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = (T)v[0];
}

// Split-K reduction: one workgroup per job; every element adds its slices
// in slice order, so the result does not depend on the schedule.
template <class T, class TS>
__global__ __launch_bounds__(256) void split_reduce_kernel(const SplitReduceJob *__restrict__ jobs,
                                                           const T *__restrict__ ws,
                                                           T *__restrict__ c) {
  const SplitReduceJob j = jobs[blockIdx.x];
  for (int e = threadIdx.x; e < j.n; e += 256) {
    TS s = (TS)ws[j.ws_off + e];
    for (int q = 1; q < j.slices; ++q) s += (TS)ws[j.ws_off + (i64)q * j.stride + e];
    c[j.c_off + e] = (T)s;
  }
}

}  // namespace

tq_status tq_launch_split_reduce(tq_handle_t h, tq_dtype out, const SplitReduceJob *jobs,
                                 int64_t n_jobs, const void *ws, void *c) {
  if (n_jobs <= 0) return TQ_SUCCESS;
  if (n_jobs > 0x7fffffff) return tq_fail(h, TQ_ERROR_UNSUPPORTED, "split-K: too many jobs");
  if (out == TQ_F32)
    hipLaunchKernelGGL((split_reduce_kernel<float, float>), dim3((unsigned)n_jobs), dim3(256),
                       0, h->stream, jobs, (const float *)ws, (float *)c);
  else if (out == TQ_F16)
    hipLaunchKernelGGL((split_reduce_kernel<_Float16, float>), dim3((unsigned)n_jobs),
                       dim3(256), 0, h->stream, jobs, (const _Float16 *)ws, (_Float16 *)c);
  else
    hipLaunchKernelGGL((split_reduce_kernel<double, double>), dim3((unsigned)n_jobs),
                       dim3(256), 0, h->stream, jobs, (const double *)ws, (double *)c);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

// Analytic choice of the tile configuration for one MFMA tile class: models
// each tile as MFMA work on the padded tile plus its operand and output
// traffic, and picks the configuration with the least modelled time. The
// autotuner measures the candidates instead and overrides this.
int tq_plan_pick_config(tq_plan_t pl, int cls) {
  const int nc = tq_gemm_n_configs(pl->prec);
  const double es = (double)tq_elem_size(pl->dtype);
  const double peak = pl->dtype == TQ_F16                            ? 2.5e15
                      : pl->dtype == TQ_F32 || pl->prec == TQ_PREC_F32 ? 157e12
                      : pl->prec == TQ_PREC_BF16X3               ? 2.5e15 / 6
                                                                   : 78.6e12;
  const double bw = 8e12;
  const double n_cu = pl->h ? pl->h->n_cu : 256;
  int best = 0;
  double best_t = 1e300;
  for (int c = 0; c < nc; ++c) {
    const GemmConfig g = tq_gemm_config(pl->prec, c);
    if (!plan_config_ok(*pl, c)) continue;
    double flop = 0, bytes = 0, tiles = 0;
    for (int64_t p = 0; p < pl->n_pairs; ++p) {
      if (pair_class(*pl, pl->h_m[p], pl->h_n[p], pl->h_k[p]) != cls) continue;
      const double tm = std::ceil((double)pl->h_m[p] / g.bm);
      const double tn = std::ceil((double)pl->h_n[p] / g.bn);
      const double kp = std::ceil((double)pl->h_k[p] / 4.0) * 4.0;
      tiles += tm * tn;
      flop += tm * tn * 2.0 * g.bm * g.bn * kp;
      bytes += tm * tn * (g.bm + g.bn) * kp * es * 0.5 +
               (double)pl->h_m[p] * pl->h_n[p] * es;
    }
    // tail effect: the last round of tiles on 4 workgroups per CU
    const double waves = std::ceil(tiles / (n_cu * (g.wave_tile ? 16 : 4)));
    const double t = std::max(flop / peak, bytes / bw) + waves * 2e-6;
    if (t < best_t) {
      best_t = t;
      best = c;
    }
  }
  return best;
}

// Analytic split-K for one tile class under configuration cfg: a class whose
// tiles would occupy less than half the GPU's tile slots (one wavefront tile
// per SIMD for wave-tile configurations, two workgroup tiles per CU
// otherwise) leaves most of it idle while each tile walks its whole K; its
// contracted range is then cut into the fewest power-of-two slices that fill
// the slots, keeping at least 256 k values per slice (the partial products
// and their reduction cost M N per slice; at 256 k a slice does 64 flops per
// byte of partials). The autotuner measures splits instead.
int tq_plan_pick_split(tq_plan_t pl, int cls, int cfg) {
  const GemmConfig g = tq_gemm_config(pl->prec, cfg);
  const double n_cu = pl->h ? pl->h->n_cu : 256;
  const double slots = n_cu * (g.wave_tile ? 4.0 * g.min_waves : 2.0);
  double tiles = 0;
  int64_t max_k = 0;
  for (int64_t p = 0; p < pl->n_pairs; ++p) {
    if (pair_class(*pl, pl->h_m[p], pl->h_n[p], pl->h_k[p]) != cls) continue;
    tiles += std::ceil((double)pl->h_m[p] / g.bm) * std::ceil((double)pl->h_n[p] / g.bn);
    max_k = std::max<int64_t>(max_k, pl->h_k[p]);
  }
  if (tiles <= 0 || 2.0 * tiles >= slots) return 1;
  int s = 1;
  while (s < 16 && tiles * s < slots && max_k / (2 * s) >= 256) s *= 2;
  return s;
}

void plan_pick_defaults(tq_plan_t pl) {
  for (int c = 0; c < 2; ++c) {
    TileClass &tc = pl->cls[c];
    tc.auto_cfg = tq_plan_pick_config(pl, c);
    if (!tc.split_user) tc.split = tq_plan_pick_split(pl, c, tc.current());
  }
}

namespace {

// One grouped-GEMM launch over the tiles of one MFMA tile class (for a
// split-K class: the slices' GEMMs into the workspace, then the reduction).
tq_status launch_class(tq_plan_t pl, int cls, const void *a, const void *b, void *c) {
  tq_handle_t h = pl->h;
  const TileClass &tc = pl->cls[cls];
  if (tc.n_tiles == 0) return TQ_SUCCESS;
  const bool split = tc.tiles_split > 1;
  const GemmDesc *descs = split ? tc.vdesc.as<GemmDesc>() : pl->desc.as<GemmDesc>();
  void *out = split ? tc.ws.p : c;
  const RunDesc *runs = tc.runs ? pl->run_desc.as<RunDesc>() : nullptr;
  if (pl->prec == TQ_PREC_BF16X3)
    TQ_CHECK(tq_launch_bf16x3_gemm(h, tc.tiles_cfg, pl->split_direct, descs, runs,
                                   tc.tile_pair.as<int32_t>(), tc.tile_first.as<int32_t>(),
                                   tc.n_tiles, pl->split_direct ? a : pl->pack_a.panel.p,
                                   pl->split_direct ? b : pl->pack_b.panel.p,
                                   pl->pack_a.panel_elems, pl->pack_b.panel_elems,
                                   reinterpret_cast<double *>(out), pl->range_flag.as<int32_t>(),
                                   tc.tile_ctr.as<uint32_t>()));
  else
    TQ_CHECK(tq_launch_grouped_gemm(h, pl->dtype, pl->prec, tc.tiles_cfg, descs, runs,
                                    tc.tile_pair.as<int32_t>(), tc.tile_first.as<int32_t>(),
                                    tc.n_tiles, a, pl->pack_a.panel.p, b, pl->pack_b.panel.p,
                                    out, tc.tile_ctr.as<uint32_t>()));
  if (!split) return TQ_SUCCESS;
  return tq_launch_split_reduce(h, pl->dtype, tc.red_jobs.as<SplitReduceJob>(), tc.n_red_jobs,
                                tc.ws.p, c);
}

}  // namespace

extern "C" {

namespace {

tq_status check_config(tq_plan_t pl, int32_t config) {
  if (config < -1 || config >= tq_gemm_n_configs(pl->prec)) return TQ_ERROR_INVALID_VALUE;
  if (config >= 0 && pl->dtype == TQ_F16 && !plan_config_ok(*pl, config))
    return tq_fail(pl->h, TQ_ERROR_UNSUPPORTED,
                   "gemm: FP16 runs the workgroup-tile configurations with 16-value k steps");
  if (config >= 0 && !plan_config_ok(*pl, config))
    return tq_fail(pl->h, TQ_ERROR_UNSUPPORTED,
                   "gemm: wave-tile configurations need leading dimensions below 2^22");
  return TQ_SUCCESS;
}

}  // namespace

tq_status tq_plan_set_config(tq_plan_t pl, int32_t config) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  TQ_CHECK(check_config(pl, config));
  pl->drop_graph();
  pl->cls[0].cfg = pl->cls[1].cfg = config;
  return TQ_SUCCESS;
}

tq_status tq_plan_set_class_config(tq_plan_t pl, int32_t cls, int32_t config) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (cls < 0 || cls > 1) return TQ_ERROR_INVALID_VALUE;
  TQ_CHECK(check_config(pl, config));
  pl->drop_graph();
  pl->cls[cls].cfg = config;
  return TQ_SUCCESS;
}

tq_status tq_plan_split_range_check(tq_plan_t pl, int32_t *out_of_range) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (!out_of_range) return TQ_ERROR_INVALID_VALUE;
  *out_of_range = 0;
  if (pl->prec != TQ_PREC_BF16X3 || !pl->range_flag.p) return TQ_SUCCESS;
  tq_handle_t h = pl->h;
  int32_t v = 0;
  TQ_HIP(h, hipMemcpyAsync(&v, pl->range_flag.p, sizeof(int32_t), hipMemcpyDeviceToHost,
                           h->stream));
  TQ_HIP(h, hipMemsetAsync(pl->range_flag.p, 0, sizeof(int32_t), h->stream));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  *out_of_range = v != 0;
  return TQ_SUCCESS;
}

tq_status tq_plan_set_class_split(tq_plan_t pl, int32_t cls, int32_t split) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (cls < 0 || cls > 1 || split < 1 || split > kMaxSplit)
    return tq_fail(pl->h, TQ_ERROR_INVALID_VALUE, "split-K: class 0 or 1, split 1..64");
  pl->drop_graph();
  pl->cls[cls].split = split;
  pl->cls[cls].split_user = true;
  return TQ_SUCCESS;
}

tq_status tq_plan_get_split(tq_plan_t pl, int32_t *split) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (!split) return TQ_ERROR_INVALID_VALUE;
  for (int c = 0; c < 2; ++c) split[c] = pl->cls[c].split;
  return TQ_SUCCESS;
}

tq_status tq_plan_get_config(tq_plan_t pl, int32_t *config) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (!config) return TQ_ERROR_INVALID_VALUE;
  for (int c = 0; c < 2; ++c) config[c] = pl->cls[c].current();
  return TQ_SUCCESS;
}

tq_status tq_bs_contract_pack(tq_plan_t pl, const void *a, const void *b) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  tq_handle_t h = pl->h;
  if (pl->n_pairs == 0) return TQ_SUCCESS;
  if (!a || !b) return TQ_ERROR_INVALID_VALUE;
  const void *src[2] = {a, b};
  PackSide *sides[2] = {&pl->pack_a, &pl->pack_b};
  for (int q = 0; q < 2; ++q) {
    PackSide &sd = *sides[q];
    if (sd.mode == kModeDirect) continue;
    if (sd.split)
      TQ_CHECK(tq_launch_split_jobs(h, sd.jobs.as<CopyJob>(), sd.wg32.as<int32_t>(),
                                    sd.n_jobs, sd.total_wg,
                                    reinterpret_cast<const double *>(src[q]),
                                    sd.panel.as<uint16_t>(), sd.panel_elems,
                                    pl->range_flag.as<int32_t>()));
    else
      TQ_CHECK(tq_launch_gather_jobs(h, pl->dtype, sd.jobs.as<CopyJob>(),
                                     sd.wg32.as<int32_t>(), sd.n_jobs, sd.total_wg,
                                     src[q], sd.panel.p));
  }
  return TQ_SUCCESS;
}

tq_status tq_bs_contract_multiply(tq_plan_t pl, const void *a, const void *b,
                                  void *c) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  tq_handle_t h = pl->h;
  if (!c) return TQ_ERROR_INVALID_VALUE;
  const size_t es = tq_elem_size(pl->dtype);
  if (pl->n_pairs == 0) {
    if (pl->ro == 0) TQ_HIP(h, hipMemsetAsync(c, 0, es, h->stream));
    return TQ_SUCCESS;
  }
  if (!a || !b) return TQ_ERROR_INVALID_VALUE;
  if (pl->ro > 0) {
    TQ_CHECK(plan_build_tiles(pl));
    if (pl->n_chunks > 0 && pl->member_stream)
      TQ_CHECK(tq_launch_small_k_groups(
          h, pl->dtype, pl->prec, pl->group_slices.as<GroupSlice>(),
          pl->member_desc.as<GemmDesc>(), pl->chunk_pair.as<int32_t>(),
          pl->chunk_first.as<int32_t>(), pl->n_chunks, a, pl->pack_a.panel.p, b,
          pl->pack_b.panel.p, c));
    else if (pl->n_chunks > 0)
      TQ_CHECK(tq_launch_small_k(h, pl->dtype, pl->prec, pl->split_direct,
                                 pl->stream_desc.as<GemmDesc>(),
                                 pl->chunk_pair.as<int32_t>(),
                                 pl->chunk_first.as<int32_t>(), pl->n_chunks, a,
                                 pl->pack_a.panel.p, b, pl->pack_b.panel.p, c,
                                 pl->pack_a.panel_elems, pl->pack_b.panel_elems,
                                 pl->range_flag.as<int32_t>()));
    // large tiles first: the small-tile launch then fills the tail
    TQ_CHECK(launch_class(pl, 1, a, b, c));
    return launch_class(pl, 0, a, b, c);
  }
  // rank-0 output: the network scalar, reduced in FP64 on the GPU
  const int64_t np = pl->n_pairs;
  TQ_HIP(h, pl->partial.alloc(sizeof(double) * (np + 1)));
  TQ_HIP(h, pl->red_tmp.alloc(tq_sum_tmp_bytes(np)));
  const unsigned g = (unsigned)ceil_div(np, 4);
  if (pl->dtype == TQ_F64) {
    hipLaunchKernelGGL(pair_dot_kernel<double>, dim3(g), dim3(256), 0, h->stream,
                       pl->desc.as<GemmDesc>(), np, (const double *)a,
                       (const double *)pl->pack_a.panel.p, (const double *)b,
                       (const double *)pl->pack_b.panel.p, pl->partial.as<double>());
  } else if (pl->dtype == TQ_F16) {
    hipLaunchKernelGGL(pair_dot_kernel<_Float16>, dim3(g), dim3(256), 0, h->stream,
                       pl->desc.as<GemmDesc>(), np, (const _Float16 *)a,
                       (const _Float16 *)pl->pack_a.panel.p, (const _Float16 *)b,
                       (const _Float16 *)pl->pack_b.panel.p, pl->partial.as<double>());
  } else {
    hipLaunchKernelGGL(pair_dot_kernel<float>, dim3(g), dim3(256), 0, h->stream,
                       pl->desc.as<GemmDesc>(), np, (const float *)a,
                       (const float *)pl->pack_a.panel.p, (const float *)b,
                       (const float *)pl->pack_b.panel.p, pl->partial.as<double>());
  }
  TQ_HIP(h, hipGetLastError());
  double *sum = pl->partial.as<double>() + np;
  TQ_CHECK(tq_sum_f64_dev(h, pl->partial.as<double>(), np, sum, pl->red_tmp.p));
  if (pl->dtype == TQ_F64)
    hipLaunchKernelGGL(store_scalar_kernel<double>, dim3(1), dim3(64), 0, h->stream,
                       sum, (double *)c);
  else if (pl->dtype == TQ_F16)
    hipLaunchKernelGGL(store_scalar_kernel<_Float16>, dim3(1), dim3(64), 0, h->stream,
                       sum, (_Float16 *)c);
  else
    hipLaunchKernelGGL(store_scalar_kernel<float>, dim3(1), dim3(64), 0, h->stream,
                       sum, (float *)c);
  TQ_HIP(h, hipGetLastError());
  return TQ_SUCCESS;
}

tq_status tq_plan_set_small_k(tq_plan_t pl, int32_t max_k) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  if (max_k < 0 || max_k > kStreamMaxK)
    return tq_fail(pl->h, TQ_ERROR_INVALID_VALUE, "small-K threshold above 8");
  pl->drop_graph();
  pl->stream_max_k = max_k;
  pl->chunks_ready = false;
  pl->runs_built = false;  // runs hold pairs of one class
  for (int c = 0; c < 2; ++c) pl->cls[c].ready = false;
  plan_pick_defaults(pl);
  return TQ_SUCCESS;
}

namespace {

// One configuration sweep of both tile classes in the plan's current operand
// form: every configuration of a class is timed over `reps` launches (after a
// warm-up) and the fastest kept. times_ms as in tq_plan_autotune.
tq_status sweep_configs(tq_plan_t pl, const void *a, const void *b, void *c, int32_t reps,
                        hipEvent_t e0, hipEvent_t e1, float *times_ms) {
  tq_handle_t h = pl->h;
  const int nc = tq_gemm_n_configs(pl->prec);
  TQ_CHECK(tq_bs_contract_pack(pl, a, b));
  TQ_CHECK(plan_build_tiles(pl));
  // each class is timed alone: its launches do not depend on the other's
  for (int k = 0; k < 2; ++k) {
    TileClass &tc = pl->cls[k];
    if (tc.n_tiles == 0) continue;
    tc.split = 1;  // configurations are compared on whole pairs
    int best_cfg = tc.current();
    float best_t = 3.4e38f;
    for (int q = 0; q < nc; ++q) {
      if (!plan_config_ok(*pl, q)) {
        if (times_ms) times_ms[k * nc + q] = -1.f;  // not applicable
        continue;
      }
      tc.cfg = q;
      TQ_CHECK(plan_build_tiles(pl));
      TQ_CHECK(launch_class(pl, k, a, b, c));  // warm-up
      TQ_HIP(h, hipEventRecord(e0, h->stream));
      for (int r = 0; r < reps; ++r) TQ_CHECK(launch_class(pl, k, a, b, c));
      TQ_HIP(h, hipEventRecord(e1, h->stream));
      TQ_HIP(h, hipEventSynchronize(e1));
      float ms = 0.f;
      TQ_HIP(h, hipEventElapsedTime(&ms, e0, e1));
      ms /= reps;
      if (times_ms) times_ms[k * nc + q] = ms;
      if (ms < best_t) {
        best_t = ms;
        best_cfg = q;
      }
    }
    tc.cfg = best_cfg;
    // split-K over the chosen configuration: the contracted range cut into
    // 2..16 slices (partials plus a reduction), kept when faster
    int best_split = 1;
    float best_split_t = 3.4e38f;
    for (int sp : {1, 2, 4, 8, 16}) {
      tc.split = sp;
      TQ_CHECK(plan_build_tiles(pl));
      TQ_CHECK(launch_class(pl, k, a, b, c));  // warm-up
      TQ_HIP(h, hipEventRecord(e0, h->stream));
      for (int r = 0; r < reps; ++r) TQ_CHECK(launch_class(pl, k, a, b, c));
      TQ_HIP(h, hipEventRecord(e1, h->stream));
      TQ_HIP(h, hipEventSynchronize(e1));
      float ms = 0.f;
      TQ_HIP(h, hipEventElapsedTime(&ms, e0, e1));
      ms /= reps;
      if (ms < best_split_t) {
        best_split_t = ms;
        best_split = sp;
      }
    }
    tc.split = best_split;
    tc.split_user = true;
  }
  return plan_build_tiles(pl);
}

// Mean time of `reps` whole executes (pack and multiply) after a warm-up.
tq_status time_execute(tq_plan_t pl, const void *a, const void *b, void *c, int32_t reps,
                       hipEvent_t e0, hipEvent_t e1, float *ms) {
  tq_handle_t h = pl->h;
  TQ_CHECK(tq_bs_contract_execute(pl, a, b, c));
  TQ_HIP(h, hipEventRecord(e0, h->stream));
  for (int r = 0; r < reps; ++r) TQ_CHECK(tq_bs_contract_execute(pl, a, b, c));
  TQ_HIP(h, hipEventRecord(e1, h->stream));
  TQ_HIP(h, hipEventSynchronize(e1));
  TQ_HIP(h, hipEventElapsedTime(ms, e0, e1));
  *ms /= reps;
  return TQ_SUCCESS;
}

struct Events {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~Events() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

}  // namespace

tq_status tq_plan_autotune(tq_plan_t pl, const void *a, const void *b, void *c,
                           int32_t reps, int32_t *best, float *times_ms) {
  if (!pl) return TQ_ERROR_NOT_INITIALIZED;
  tq_handle_t h = pl->h;
  if (reps < 1) reps = 1;
  const int nc = tq_gemm_n_configs(pl->prec);
  if (times_ms)
    for (int q = 0; q < 2 * nc; ++q) times_ms[q] = 0.f;
  if (pl->n_pairs == 0 || pl->ro == 0) {
    if (best)
      for (int k = 0; k < 2; ++k) best[k] = pl->cls[k].current();
    return TQ_SUCCESS;
  }
  pl->drop_graph();
  Events ev;
  TQ_HIP(h, hipEventCreate(&ev.e0));
  TQ_HIP(h, hipEventCreate(&ev.e1));
  if (pl->can_split_direct && pl->split_form == 0) {
    // split-BF16 with plain operands: sweep both operand forms and keep the
    // one whose whole execute is faster (reading in place skips the pack
    // pass but splits each value once per tile that reads it)
    std::vector<float> t_in(2 * nc, 0.f), t_pan(2 * nc, 0.f);
    float ms_in = 0.f, ms_pan = 0.f;
    int cfg_in[2];
    TQ_CHECK(plan_set_split_form(pl, true));
    TQ_CHECK(sweep_configs(pl, a, b, c, reps, ev.e0, ev.e1, t_in.data()));
    TQ_CHECK(time_execute(pl, a, b, c, reps, ev.e0, ev.e1, &ms_in));
    for (int k = 0; k < 2; ++k) cfg_in[k] = pl->cls[k].cfg;
    TQ_CHECK(plan_set_split_form(pl, false));
    TQ_CHECK(sweep_configs(pl, a, b, c, reps, ev.e0, ev.e1, t_pan.data()));
    TQ_CHECK(time_execute(pl, a, b, c, reps, ev.e0, ev.e1, &ms_pan));
    const bool in_place = ms_in <= ms_pan;
    if (in_place) {
      TQ_CHECK(plan_set_split_form(pl, true));
      for (int k = 0; k < 2; ++k) pl->cls[k].cfg = cfg_in[k];
      TQ_CHECK(plan_build_tiles(pl));
    }
    if (times_ms) std::copy_n((in_place ? t_in : t_pan).data(), 2 * nc, times_ms);
  } else {
    TQ_CHECK(sweep_configs(pl, a, b, c, reps, ev.e0, ev.e1, times_ms));
  }
  if (best)
    for (int k = 0; k < 2; ++k) best[k] = pl->cls[k].current();
  return TQ_SUCCESS;
}

tq_status tq_bs_contract_execute(tq_plan_t pl, const void *a, const void *b,
                                 void *c) {
  TQ_CHECK(tq_bs_contract_pack(pl, a, b));
  return tq_bs_contract_multiply(pl, a, b, c);
}

tq_status tq_dense_contract(tq_handle_t h, tq_dtype dtype, int32_t rank_a,
                            const int64_t *dims_a, int32_t rank_b,
                            const int64_t *dims_b, int32_t n_legs,
                            const int32_t *legs, const void *a, const void *b,
                            void *c) {
  // A dense tensor is a block-sparse tensor with one block covering it all;
  // the single matched pair is the whole GEMM (src/gpu/dense.cuh:439-555).
  if (!h) return TQ_ERROR_NOT_INITIALIZED;
  if (rank_a < 0 || rank_b < 0 || rank_a > TQ_MAX_RANK || rank_b > TQ_MAX_RANK)
    return TQ_ERROR_UNSUPPORTED;
  i64 ba[TQ_MAX_RANK] = {0}, ea[TQ_MAX_RANK], ta[TQ_MAX_RANK];
  i64 bb[TQ_MAX_RANK] = {0}, eb[TQ_MAX_RANK], tb[TQ_MAX_RANK];
  i64 off0 = 0;
  i64 sa = 1, sb = 1;
  for (int l = 0; l < rank_a; ++l) {
    if (dims_a[l] <= 0) return TQ_ERROR_INVALID_VALUE;
    ea[l] = dims_a[l] - 1;
    ta[l] = sa;
    sa *= dims_a[l];
  }
  for (int l = 0; l < rank_b; ++l) {
    if (dims_b[l] <= 0) return TQ_ERROR_INVALID_VALUE;
    eb[l] = dims_b[l] - 1;
    tb[l] = sb;
    sb *= dims_b[l];
  }
  tq_bs_desc da{rank_a, 1, dims_a, ba, ea, ta, &off0};
  tq_bs_desc db{rank_b, 1, dims_b, bb, eb, tb, &off0};
  // the handle's plan cache: a repeated contraction of the same shapes plans
  // once (the reference re-plans nothing either: cuBLAS and cuTT take the
  // shapes per call)
  tq_plan_t pl = nullptr;
  TQ_CHECK(tq_bs_contract_plan_cached(h, &da, &db, n_legs, legs, dtype, TQ_PREC_NATIVE, 0,
                                      &pl));
  TQ_CHECK(tq_bs_contract_execute(pl, a, b, c));
  TQ_HIP(h, hipStreamSynchronize(h->stream));
  return TQ_SUCCESS;
}

}  // extern "C"
