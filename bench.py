"""Benchmark: block-sparse contraction GFLOP/s on one MI355X.

Workload (default): the PEPS CTMRG-style absorption step at D=8, chi=128
with U(1) charges -2..2 and entries N(0,1) * 10^U(-3,0) per block. One step
is the double layer a = A . conj(A) over the physical leg followed by
E(chi,D,D,chi) . a over (l, l'): two block-sparse contractions through the
C-ABI (operand blocks permuted to GEMM layout, then one grouped-GEMM launch
each). Plans are built once (the block structure is fixed across CTMRG
steps); inputs are resident in HBM when the timed region starts.

Prints one JSON line: FP64 GFLOP/s as the value, a live roofline of the
dominant kernel (the E . a grouped GEMM, timed with HIP events on the stream
it runs on), the FP32 and split-BF16 tiers with their relative Frobenius
error against the FP64 CPU oracle on a sample of output blocks, and the CPU
oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Block-sparse contract GFLOP/s + %peak: FP64 vs 3xBF16 at ≤1e-6 rel err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def committed_traffic(name):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/<name>_traffic.md),
    or None when no summary exists for this kernel and workload."""
    path = os.path.join(ROOT, "profiles", name + "_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Threads for the all-core CPU leg: every core the process may use
    (OMP_NUM_THREADS caps it where the host sets a share)."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() else n


def median_time(fn, runs=5):
    """Median wall time of `runs` calls after one warm-up call."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--D", type=int, default=8)
    p.add_argument("--chi", type=int, default=128)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--tiers", default="fp32,bf16x3")
    p.add_argument("--cpu-sample-blocks", type=int, default=0,
                   help="E blocks in the CPU-oracle sample (0 = all)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-permute", action="store_true",
                   help="skip the standalone 200M-element rank-6 permutation")
    p.add_argument("--no-batched", action="store_true",
                   help="skip the 512-contraction batched workload in the default line")
    p.add_argument("--no-large", action="store_true",
                   help="skip the D=10, chi=256 enlarged-corner product")
    p.add_argument("--workload", default="peps",
                   choices=["peps", "batched", "permute", "large", "shapes", "corner6",
                            "scaling"],
                   help="peps: the D, chi absorption step (default); batched: --n-batch "
                        "independent rank-4 x rank-4 contractions in one captured launch")
    p.add_argument("--n-batch", type=int, default=512)
    p.add_argument("--oracle-contractions", type=int, default=24,
                   help="batched: contractions checked against the oracle")
    p.add_argument("--copy-calibration", action="store_true",
                   help="finish with a plain 1 GiB device copy (reference for "
                        "FETCH_SIZE / WRITE_SIZE counters)")
    return p.parse_args()


def bench_batched(args):
    """Batched workload: n independent U(1) rank-4 x rank-4 contractions over two
    legs (leg dims U{8..96}, 3-7 Dirichlet(1) sectors), all pairs in one work list,
    autotuned, captured in a HIP graph and replayed."""
    import torch

    from torque_amd import BatchPlan, BlockSparseTensor, TorqueError, default_handle
    from torque_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t0 = time.perf_counter()
    pairs = W.batched(args.n_batch, seed=args.seed)
    a_all = torch.tensor(np.concatenate([a.data for a, _, _ in pairs]), device=dev)
    b_all = torch.tensor(np.concatenate([b.data for _, b, _ in pairs]), device=dev)
    a_off = np.cumsum([0] + [a.data.size for a, _, _ in pairs])[:-1]
    b_off = np.cumsum([0] + [b.data.size for _, b, _ in pairs])[:-1]
    gen_s = time.perf_counter() - t0

    def build(precision, split_form=None):
        plans = []
        for i, (a, b, legs) in enumerate(pairs):
            A = BlockSparseTensor(a.dims, a_all[a_off[i]:a_off[i] + a.data.size], a.begin,
                                  a.end, device=dev)
            B = BlockSparseTensor(b.dims, b_all[b_off[i]:b_off[i] + b.data.size], b.begin,
                                  b.end, device=dev)
            p = A.plan_contract(B, legs, precision)
            if split_form:
                p.set_split_operands(split_form)
            plans.append(p)
        c_off = np.cumsum([0] + [p.out_n_elem for p in plans])[:-1]
        batch = BatchPlan(default_handle(), plans, a_off, b_off, c_off)
        return plans, c_off, batch

    t0 = time.perf_counter()
    plans, c_off, batch = build("native")
    c = torch.empty(max(batch.out_n_elem, 1), dtype=torch.float64, device=dev)
    plan_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    best, times = batch.autotune(a_all, b_all, c, reps=2)
    sweep_s = time.perf_counter() - t0
    batch.capture(a_all, b_all, c)
    n_tiles, n_segs = batch.work()
    print("# batched: %d contractions, %d pairs, %.1f GFLOP, out %.1f M elem; gen %.1f s, "
          "plans %.1f s; configs (small, large tiles) %s; per config ms %s / %s"
          % (args.n_batch, batch.n_pairs, batch.flops / 1e9, batch.out_n_elem / 1e6, gen_s,
             plan_s, best, ["%.2f" % x for x in times[0]], ["%.2f" % x for x in times[1]]),
          flush=True)

    def time_graph(bp):
        for _ in range(args.warmup):
            bp.graph_launch()
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream(dev)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            bp.graph_launch()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / args.steps

    dt = time_graph(batch)
    value = batch.flops / dt / 1e9
    print("# fp64 graph replay %.3f ms, %.1f GFLOP/s (%.1f%% of FP64 MFMA peak)"
          % (dt * 1e3, value, 100 * value / (FP64_PEAK_TFLOPS * 1e3)), flush=True)

    from oracle import oracle as O
    n_chk = min(args.oracle_contractions, len(pairs))
    refs = []
    for i in range(n_chk):
        a, b, legs = pairs[i]
        ref, _ = O.contract(O.BS(a.dims, a.begin, a.end), a.data, O.BS(b.dims, b.begin, b.end),
                            b.data, legs, threads=min(16, os.cpu_count() or 1))
        refs.append(ref)

    def errors(cbuf):
        host = cbuf.cpu().numpy()
        num = den = worst = 0.0
        for i in range(n_chk):
            got = host[c_off[i]:c_off[i] + plans[i].out_n_elem]
            d = np.linalg.norm(got - refs[i]) ** 2
            r = np.linalg.norm(refs[i]) ** 2
            num += d
            den += r
            worst = max(worst, np.sqrt(d / r) if r > 0 else 0.0)
        return float(np.sqrt(num / den)), float(worst)

    err64, worst64 = errors(c)

    # the committed cost model's choice against the sweep: regret per class
    # (from the sweep's own per-configuration times) and on the whole replay,
    # and its time to schedule against the sweep's
    from torque_amd import costmodel
    cost_model = None
    model = costmodel.load_default()
    if model is not None:
        t0 = time.perf_counter()
        chosen = model.choose(batch)
        choose_s = time.perf_counter() - t0
        split = batch.split()
        regret = {}
        for cls in range(2):
            t = np.array(times[cls])
            if (t > 0).any() and split[cls] == 1:
                regret["class%d" % cls] = round(float(t[chosen[cls]] / t[t > 0].min() - 1), 4)
        batch.capture(a_all, b_all, c)
        mdt = time_graph(batch)
        regret["replay"] = round(mdt / dt - 1, 4)
        cost_model = {"chosen": chosen, "split": split, "regret_vs_sweep": regret,
                      "replay_ms": round(mdt * 1e3, 4),
                      "time_to_schedule_s": round(choose_s, 4),
                      "sweep_time_to_schedule_s": round(sweep_s, 3)}
        print("# cost model: chose %s split %s, replay %.3f ms, regret %s, %.4f s to schedule "
              "(sweep %.2f s)" % (chosen, split, mdt * 1e3, regret, choose_s, sweep_s), flush=True)
        for cls in range(2):
            batch.set_class_split(cls, 1)
            batch.set_class_config(cls, best[cls])
    tiers = {"fp64": {"gflops": round(value, 1), "rel_err": err64, "max_contraction_rel_err": worst64,
                      "frac_fp64_peak": round(value / (FP64_PEAK_TFLOPS * 1e3), 4)}}
    for tier in [t for t in args.tiers.split(",") if t]:
        # split-BF16: the batch reads its plain operands in place or packs them
        # into split panels; both forms are timed and the faster reported
        forms = ["in_place", "panels"] if tier == "bf16x3" else [None]
        for form in forms:
            try:
                _, _, bt = build(tier, form)
                ct = torch.empty_like(c)
                bt.autotune(a_all, b_all, ct, reps=2)
                bt.capture(a_all, b_all, ct)
            except TorqueError as e:
                tiers[tier] = {"error": str(e)}
                continue
            tdt = time_graph(bt)
            e, w = errors(ct)
            entry = {"gflops": round(bt.flops / tdt / 1e9, 1), "ms": round(tdt * 1e3, 4),
                     "rel_err": e, "max_contraction_rel_err": w,
                     "speedup_vs_fp64": round(dt / tdt, 3)}
            if form:
                entry["operands"] = form
            print("# %s%s graph replay %.3f ms, %.1f GFLOP/s, speedup %.2fx, rel err %.2e"
                  % (tier, " (%s)" % form if form else "", tdt * 1e3, bt.flops / tdt / 1e9,
                     dt / tdt, e), flush=True)
            prev = tiers.get(tier)
            if prev is None or "error" in prev or entry["ms"] < prev["ms"]:
                if prev is not None and "ms" in prev:
                    entry["other_form_ms"] = prev["ms"]
                tiers[tier] = entry
            else:
                prev["other_form_ms"] = entry["ms"]
            bt.destroy()
            del ct

    cpu_baseline = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        a, b, legs = pairs[0]
        t0 = time.perf_counter()
        ref, rp = O.contract(O.BS(a.dims, a.begin, a.end), a.data, O.BS(b.dims, b.begin, b.end),
                             b.data, legs, threads=threads)
        cs = time.perf_counter() - t0
        f = float((2 * rp["m"] * rp["n"] * rp["k"]).sum())
        cpu_baseline = {"value": round(f / cs / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
                        "kind": "port",
                        "sample": "contraction 0 of the batch (%d pairs, %.2f GFLOP), %.2f s"
                                  % (rp["m"].size, f / 1e9, cs)}
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched_%d_u1_rank4x4" % args.n_batch,
                   "pairs": int(batch.n_pairs), "gflop_per_step": round(batch.flops / 1e9, 3),
                   "out_elements": int(batch.out_n_elem), "gemm_config": best,
                   "mfma_tiles": int(n_tiles), "stream_segments": int(n_segs),
                   "autotune_sweep_s": round(sweep_s, 3), "cost_model": cost_model,
                   "launch": "one hipGraph replay (pack A, pack B, small-K, grouped GEMM)",
                   "seed": args.seed},
        "roofline": {"bound": "mfma", "achieved": round(value / 1e3, 3),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(value / (FP64_PEAK_TFLOPS * 1e3), 4), "traffic": None,
                     "kernel": "graph replay (grouped_gemm_kernel dominant)"},
        "cpu_baseline": cpu_baseline, "tiers": tiers,
    }
    return line


PERMUTE_DIMS = (10, 20, 25, 40, 50, 20)  # 200,000,000 elements
PERMUTE_PERMS = ((5, 4, 3, 2, 1, 0), (1, 0, 2, 3, 4, 5), (0, 2, 1, 3, 4, 5),
                 (2, 3, 4, 5, 0, 1), (3, 5, 0, 4, 2, 1))


def bench_permute(dev, reps=10, dims=PERMUTE_DIMS, perms=PERMUTE_PERMS):
    """Standalone rank-6 permutation of a 200M-element FP64 tensor
    (tq_permute), per permutation: kernel time from HIP events on the
    handle's stream, algorithmic bytes 2 * 8 * N per launch, and a bitwise
    round trip (permute, inverse permute) over the whole tensor."""
    import ctypes

    import torch

    from torque_amd import default_handle, lib

    n = int(np.prod(dims))
    x = torch.randn(n, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    h = default_handle(dev.index)
    stream = torch.cuda.current_stream(dev)
    out = []
    for perm in perms:
        d = np.ascontiguousarray(dims, np.int64)
        p = np.ascontiguousarray(perm, np.int32)
        inv = np.ascontiguousarray(np.argsort(perm), np.int32)
        d2 = np.ascontiguousarray(np.asarray(dims)[list(perm)], np.int64)

        def run(src, dst, dd, pp):
            h.check(lib().tq_permute(h.ptr, 1, len(dims), dd.ctypes.data_as(ctypes.POINTER(
                ctypes.c_int64)), None, pp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr())))

        run(x, y, d, p)
        run(y, z, d2, inv)
        torch.cuda.synchronize()
        exact = bool(torch.equal(x, z))
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            run(x, y, d, p)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gbs = 2 * 8 * n / (ms * 1e-3) / 1e9
        out.append({"perm": list(perm), "ms": round(ms, 4), "gbs": round(gbs, 1),
                    "frac_hbm": round(gbs / HBM_PEAK_GBS, 4), "round_trip_bitwise": exact})
        print("# permute %s: %.3f ms, %.0f GB/s (%.1f%% of HBM), round trip %s"
              % (list(perm), ms, gbs, 100 * gbs / HBM_PEAK_GBS, "exact" if exact else "DIFFERS"),
              flush=True)
    # calibration: a plain device copy of the same buffer (torch), the
    # practical ceiling of a read-once write-once kernel on this GPU
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    y.copy_(x)
    e0.record(stream)
    for _ in range(reps):
        y.copy_(x)
    e1.record(stream)
    torch.cuda.synchronize()
    copy_gbs = 2 * 8 * n / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9
    print("# plain device copy of the same buffer: %.0f GB/s" % copy_gbs, flush=True)
    del x, y, z
    torch.cuda.empty_cache()
    return {"dims": list(dims), "elements": n, "dtype": "f64",
            "algorithmic_bytes_per_launch": 16 * n, "perms": out,
            "copy_calibration_gbs": round(copy_gbs, 1)}


def bench_large(dev, steps=5, warmup=2, tiers=("fp32", "bf16x3"), oracle_max_block=1368,
                seed=0, D=10, chi=256):
    """The corner product of a CTMRG step in the fused-leg view (enlarged
    corners as block-diagonal matrices over the fused (chi, D, D) leg; at
    D=10, chi=256: 19 GEMMs of 4..3600, 392 GFLOP; at D=8, chi=128: 13 GEMMs
    of 8..1568, 23.9 GFLOP): FP64 and the tiers, each autotuned by a full
    sweep, timed with HIP events over whole executes (pack + multiply) and
    over the multiply alone; relative error against the oracle on every
    block up to oracle_max_block; and the committed cost model's choice
    against the sweep (regret, time to schedule)."""
    import torch

    from oracle import oracle as O
    from torque_amd import BlockSparseTensor, costmodel
    from torque_amd import workloads as W

    c1, c2, legs = W.enlarged_corners(D, chi, seed=seed)
    A = BlockSparseTensor(c1.dims, c1.data, c1.begin, c1.end, device=dev)
    B = BlockSparseTensor(c2.dims, c2.data, c2.begin, c2.end, device=dev)
    stream = torch.cuda.current_stream(dev)
    sizes = (c1.end[:, 0] - c1.begin[:, 0] + 1).astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes * sizes)])
    sel = np.nonzero(sizes <= oracle_max_block)[0]

    def sub(t):
        data = np.concatenate([t.data[offs[i]:offs[i + 1]] for i in sel])
        return O.BS(t.dims, t.begin[sel], t.end[sel]), data

    TA, dA = sub(c1)
    TB, dB = sub(c2)
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    ref, _ = O.contract(TA, dA, TB, dB, legs, threads=threads)
    cpu_s = time.perf_counter() - t0
    sample_flops = float(2 * (sizes[sel].astype(np.float64) ** 3).sum())

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    res = {}
    fp64 = None
    for prec in ("native",) + tuple(tiers):
        p = A.plan_contract(B, legs, precision=prec)
        out = torch.empty(max(p.out_n_elem, 1), dtype=torch.float64, device=dev)
        t0 = time.perf_counter()
        best, times = p.autotune(A.data, B.data, out, reps=2)
        torch.cuda.synchronize()
        sweep_s = time.perf_counter() - t0
        ms = timed(lambda: p.execute(A.data, B.data, out))
        p.pack(A.data, B.data)
        gemm_ms = timed(lambda: p.multiply(A.data, B.data, out))
        host = out.cpu().numpy()
        got = np.concatenate([host[p.out_offsets[i]:p.out_offsets[i] + sizes[i] ** 2]
                              for i in sel])
        err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        worst, off = 0.0, 0
        for i in sel:
            n2 = int(sizes[i]) ** 2
            r = ref[off:off + n2]
            worst = max(worst, float(np.linalg.norm(got[off:off + n2] - r) / np.linalg.norm(r)))
            off += n2
        name = "fp64" if prec == "native" else prec
        e = {"ms": round(ms, 4), "gflops": round(p.flops / ms / 1e6, 1),
             "gemm_ms": round(gemm_ms, 4), "gemm_gflops": round(p.flops / gemm_ms / 1e6, 1),
             "rel_err_vs_oracle": err, "max_block_rel_err": worst, "gemm_config": best,
             "autotune_sweep_s": round(sweep_s, 3)}
        if prec == "bf16x3":
            e["operands"] = "split panels + pack" if p.a_layout == 2 else "in place"
            e["executed_bf16_mfma_tflops"] = round(6 * p.flops / gemm_ms / 1e9, 1)
        if prec == "native":
            fp64 = (ms, gemm_ms)
            e["frac_fp64_mfma_peak"] = round(p.flops / ms / 1e9 / FP64_PEAK_TFLOPS, 4)
            e["gemm_frac_fp64_mfma_peak"] = round(p.flops / gemm_ms / 1e9 / FP64_PEAK_TFLOPS, 4)
            model = costmodel.load_default()
            if model is not None:
                t0 = time.perf_counter()
                chosen = model.choose(p)
                choose_s = time.perf_counter() - t0
                split = p.split()
                regret = {}
                for cls in range(2):
                    t = np.array(times[cls])
                    # the sweep's per-configuration times are whole-pair GEMMs: a
                    # class the model split along K is judged by the measured
                    # multiply below only
                    if (t > 0).any() and split[cls] == 1:
                        regret["class%d" % cls] = round(float(t[chosen[cls]] / t[t > 0].min() - 1),
                                                        4)
                # both launches together, measured under the model's choice against
                # the autotuned multiply: a class of microsecond launches can show a
                # large relative regret that costs nothing next to the other class
                model_ms = timed(lambda: p.multiply(A.data, B.data, out))
                regret["multiply"] = round(model_ms / gemm_ms - 1, 4)
                e["cost_model"] = {"chosen": chosen, "split": split, "regret_vs_sweep": regret,
                                   "multiply_ms": round(model_ms, 4),
                                   "time_to_schedule_s": round(choose_s, 4),
                                   "sweep_time_to_schedule_s": round(sweep_s, 3)}
        else:
            e["speedup_vs_fp64"] = round(fp64[0] / ms, 3)
            e["gemm_speedup_vs_fp64"] = round(fp64[1] / gemm_ms, 3)
        res[name] = e
        print("# corner product D=%d chi=%d %s: %.3f ms (%.1f GFLOP/s), multiply alone %.3f ms "
              "(%.1f GFLOP/s), rel err %.2e (block max %.2e), configs %s"
              % (D, chi, name, ms, p.flops / ms / 1e6, gemm_ms, p.flops / gemm_ms / 1e6, err,
                 worst, best), flush=True)
        p.destroy()
        del out
    if "cost_model" in res.get("fp64", {}):
        print("# cost model: chose %s split %s, regret %s, %.4f s to schedule (sweep %.2f s)"
              % (res["fp64"]["cost_model"]["chosen"], res["fp64"]["cost_model"]["split"],
                 res["fp64"]["cost_model"]["regret_vs_sweep"],
                 res["fp64"]["cost_model"]["time_to_schedule_s"],
                 res["fp64"]["cost_model"]["sweep_time_to_schedule_s"]), flush=True)
    del A, B
    torch.cuda.empty_cache()
    return {"workload": "enlarged_corner_product_D%d_chi%d_fused" % (D, chi),
            "blocks": sizes.tolist(), "gflop": round(float(2 * (sizes.astype(np.float64) ** 3).sum())
                                                     / 1e9, 3),
            "tiers": res,
            "cpu_oracle": {"gflops": round(sample_flops / cpu_s / 1e9, 3), "cores": threads,
                           "sample": "blocks up to %d (%.2f GFLOP), %.2f s"
                                     % (oracle_max_block, sample_flops / 1e9, cpu_s)}}


def bench_corner_rank6(dev, steps=5, warmup=2, D=10, chi=256, seed=0, oracle_max_block=1368):
    """The corner product on enlarged corners held as rank-6 U(1) tensors
    C(chi, D ket, D bra | chi, D ket, D bra) (charges -3..3 at chi > 128),
    contracted over the three legs of the shared side, two ways:
      * fused: both corners fused into block-diagonal matrices over the fused
        (chi, D, D) legs (U1Tensor.fuse, one box-copy launch each), one GEMM
        per fused charge (19 GEMMs of 4..3600 at D=10, chi=256), and the
        product split back into rank 6; FP64 and split-BF16;
      * unfused: the merged-output contraction of the rank-6 tensors (one
        K-concatenated GEMM per output block), FP64.
    Times each step with HIP events (whole chains, and fuse / product /
    split alone), compares the two paths' FP64 results, and checks the fused
    product's blocks up to oracle_max_block against the oracle."""
    import torch

    from oracle import oracle as O
    from torque_amd import CopyPlan, U1Leg, U1Tensor, default_handle
    from torque_amd import workloads as W
    from torque_amd.tensor import BOX_DTYPE

    lchi, lD, _ = W.peps_legs(D, chi)
    lc = U1Leg([q for q, _, _ in lchi], [n for _, _, n in lchi])
    ld = U1Leg([q for q, _, _ in lD], [n for _, _, n in lD])
    legs = [lc, ld, ld.dual(), lc.dual(), ld.dual(), ld]
    g = torch.Generator(device=dev).manual_seed(seed)
    C1 = U1Tensor.random(legs, 0, g, device=dev)
    C2 = U1Tensor.random(legs, 0, g, device=dev)
    stream = torch.cuda.current_stream(dev)
    groups = [[0, 1, 2], [3, 4, 5]]

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    t0 = time.perf_counter()
    fplan = C1.fusion_plan(groups)  # C2 has the same structure: the same plan
    F1, F2 = fplan.run(C1), fplan.run(C2)
    torch.cuda.synchronize()
    fuse_plan_s = time.perf_counter() - t0
    n_elem = int(C1.bs.block_n_elem.sum())
    n_blocks6 = C1.n_blocks

    def fuse_both():
        fplan.execute(C1.bs.data, F1.bs.data)
        fplan.execute(C2.bs.data, F2.bs.data)

    fuse_ms = timed(fuse_both)
    sizes = [int(n) for n in F1.legs[0].sizes]

    # oracle on the fused blocks up to oracle_max_block (block-diagonal: key (q, q))
    sel = [i for i, k in enumerate(F1.keys) if F1.legs[0].sizes[F1.legs[0].sector(k[0])]
           <= oracle_max_block]

    def sub(F):
        host = F.bs.data.cpu().numpy()
        offs, nel = F.bs.block_offsets, F.bs.block_n_elem
        data = np.concatenate([host[offs[i]:offs[i] + nel[i]] for i in sel])
        return O.BS(F.bs.dimension, F.bs.begin_points[sel], F.bs.end_points[sel]), data

    TA, dA = sub(F1)
    TB, dB = sub(F2)
    threads = min(16, os.cpu_count() or 1)
    ref, ref_plan = O.contract(TA, dA, TB, dB, [[1, 0]], threads=threads)
    ref_off = {tuple(b): o for b, o in zip(ref_plan["out_begin"].tolist(),
                                            ref_plan["out_offsets"].tolist())}

    res = {"fuse_both_ms": round(fuse_ms, 4),
           "fuse_gbs": round(2 * 2 * n_elem * 8 / fuse_ms / 1e6, 1),
           "fuse_plan_s": round(fuse_plan_s, 3)}
    split_out = None
    flops = None
    for prec in ("native", "bf16x3"):
        p = F1.bs.plan_contract(F2.bs, [[1, 0]], precision=prec, merged=True)
        flops = p.flops
        out = torch.empty(max(p.out_n_elem, 1), dtype=torch.float64, device=dev)
        best, _ = p.autotune(F1.bs.data, F2.bs.data, out, reps=2)
        Fc = U1Tensor._wrap(p.output_tensor(out), [F1.legs[0], F2.legs[1]], C1.flux + C2.flux)
        t0 = time.perf_counter()
        splan = Fc.split_plan()
        split_plan_s = time.perf_counter() - t0
        cout = splan.new_output(torch.float64, dev)
        prod_ms = timed(lambda: p.execute(F1.bs.data, F2.bs.data, out))
        split_ms = timed(lambda: splan.execute(out, cout))

        def chain():
            fuse_both()
            p.execute(F1.bs.data, F2.bs.data, out)
            splan.execute(out, cout)

        chain_ms = timed(chain)
        host = out.cpu().numpy()
        got = np.concatenate([host[int(o):int(o) + int(n)] for b, o, n in zip(
            Fc.bs.begin_points.tolist(), Fc.bs.block_offsets, Fc.bs.block_n_elem)
            if tuple(b) in ref_off])
        want = np.concatenate([ref[ref_off[tuple(b)]:ref_off[tuple(b)] + int(n)] for b, n in zip(
            Fc.bs.begin_points.tolist(), Fc.bs.block_n_elem) if tuple(b) in ref_off])
        err = float(np.linalg.norm(got - want) / np.linalg.norm(want))
        name = "fp64" if prec == "native" else prec
        res[name] = {"chain_ms": round(chain_ms, 4), "chain_gflops": round(flops / chain_ms / 1e6, 1),
                     "product_ms": round(prod_ms, 4),
                     "product_gflops": round(flops / prod_ms / 1e6, 1),
                     "split_ms": round(split_ms, 4), "split_plan_s": round(split_plan_s, 3),
                     "gemm_config": best, "rel_err_vs_oracle": err}
        print("# rank-6 corner product D=%d chi=%d %s, fused: chain %.3f ms (%.1f GFLOP/s) = fuse "
              "both %.3f + product %.3f + split %.3f ms; rel err vs oracle %.2e (fused blocks "
              "up to %d)" % (D, chi, name, chain_ms, flops / chain_ms / 1e6, fuse_ms, prod_ms,
                              split_ms, err, oracle_max_block), flush=True)
        if prec == "native":
            split_out = splan.run(Fc)
        p.destroy()
        del out, cout

    # the unfused path: merged-output contraction of the rank-6 tensors
    t0 = time.perf_counter()
    pd = C1.bs.plan_contract(C2.bs, [[3, 0], [4, 1], [5, 2]], merged=True)
    outd = torch.empty(max(pd.out_n_elem, 1), dtype=torch.float64, device=dev)
    pd.autotune(C1.bs.data, C2.bs.data, outd, reps=2)
    torch.cuda.synchronize()
    direct_plan_s = time.perf_counter() - t0
    direct_ms = timed(lambda: pd.execute(C1.bs.data, C2.bs.data, outd))
    Cd = U1Tensor._wrap(pd.output_tensor(outd), split_out.legs, 0)
    # both results block by block: the direct blocks moved into the split
    # result's block order by one copy launch
    boxes = np.zeros(split_out.n_blocks, BOX_DTYPE)
    for i, k in enumerate(split_out.keys):
        j = Cd._index[k]
        boxes[i]["src_off"] = Cd.bs.block_offsets[j]
        boxes[i]["dst_off"] = split_out.bs.block_offsets[i]
        boxes[i]["rank"] = 1
        boxes[i]["dims"][0] = split_out.bs.block_n_elem[i]
        boxes[i]["src_strides"][0] = 1
        boxes[i]["dst_strides"][0] = 1
    h = default_handle(dev.index)
    reorder = CopyPlan(h, torch.float64, boxes, int(Cd.bs.block_n_elem.sum()), n_elem)
    moved = torch.zeros_like(split_out.bs.data)
    reorder.execute(Cd.bs.data, moved)
    diff = float(torch.linalg.vector_norm(moved[:n_elem] - split_out.bs.data[:n_elem]) /
                 torch.linalg.vector_norm(moved[:n_elem]))
    reorder.destroy()
    res["unfused_fp64"] = {"ms": round(direct_ms, 4), "gflops": round(pd.flops / direct_ms / 1e6, 1),
                           "pairs": int(pd.n_matched), "out_blocks": int(pd.n_pairs),
                           "plan_s": round(direct_plan_s, 3),
                           "rel_diff_vs_fused": diff}
    print("# rank-6 corner product D=%d chi=%d fp64, unfused merged contraction: %.3f ms "
          "(%.1f GFLOP/s), %d pairs into %d blocks; rel diff vs the fused chain %.2e"
          % (D, chi, direct_ms, pd.flops / direct_ms / 1e6, pd.n_matched, pd.n_pairs, diff),
          flush=True)
    res["fp64"]["speedup_vs_unfused"] = round(direct_ms / res["fp64"]["chain_ms"], 3)
    res["bf16x3"]["chain_speedup_vs_fp64"] = round(res["fp64"]["chain_ms"] /
                                                   res["bf16x3"]["chain_ms"], 3)
    pd.destroy()
    del outd, Cd, split_out, moved, F1, F2, C1, C2
    torch.cuda.empty_cache()
    return dict({"workload": "enlarged_corner_product_D%d_chi%d_rank6" % (D, chi),
                 "blocks_per_corner": n_blocks6,
                 "fused_blocks": sizes, "gflop": round(flops / 1e9, 3),
                 "elements_per_corner": n_elem}, **res)


def bench_split_k(dev, reps=20):
    """Split-K on a contraction with few output tiles and long contracted
    ranges (block-diagonal, 8 blocks of 64 x 16384 times 16384 x 64, FP64):
    the multiply with whole pairs against the autotuned split."""
    import torch

    from torque_amd import BlockSparseTensor

    n_blk, m, k = 8, 64, 16384
    off_m = np.arange(n_blk, dtype=np.int64) * m
    off_k = np.arange(n_blk, dtype=np.int64) * k
    rng = np.random.default_rng(3)
    ba = np.stack([off_m, off_k], 1)
    ea = ba + np.array([m - 1, k - 1])
    bb = np.stack([off_k, off_m], 1)
    eb = bb + np.array([k - 1, m - 1])
    A = BlockSparseTensor([n_blk * m, n_blk * k], rng.standard_normal(n_blk * m * k), ba, ea,
                          device=dev)
    B = BlockSparseTensor([n_blk * k, n_blk * m], rng.standard_normal(n_blk * m * k), bb, eb,
                          device=dev)
    p = A.plan_contract(B, [[1, 0]])
    out = torch.empty(p.out_n_elem, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)

    def timed():
        p.execute(A.data, B.data, out)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            p.execute(A.data, B.data, out)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    default_split = p.split()  # the analytic choice, before any tuning
    ms_default = timed()
    best, _ = p.autotune(A.data, B.data, out, reps=3)
    split = p.split()
    ms_split = timed()
    p.set_class_split(0, 1)
    p.set_class_split(1, 1)
    ms_whole = timed()
    res = {"workload": "block_diagonal_8x_64x16384x64", "gflop": round(p.flops / 1e9, 3),
           "whole_pairs_ms": round(ms_whole, 4), "split_ms": round(ms_split, 4),
           "split": split, "gemm_config": best, "speedup": round(ms_whole / ms_split, 2),
           "default_split": default_split, "default_ms": round(ms_default, 4),
           "split_gflops": round(p.flops / ms_split / 1e6, 1)}
    print("# split-K (8 blocks 64 x 16384 x 64): whole pairs %.3f ms, split %s %.3f ms (%.2fx); "
          "analytic default split %s %.3f ms"
          % (ms_whole, split, ms_split, ms_whole / ms_split, default_split, ms_default),
          flush=True)
    p.destroy()
    return res


def bench_ctmrg(dev, D, chi, steps, warmup):
    """One CTMRG edge update (double layer, absorption, two projector
    contractions, all merged-output plans, then the FP64 normalisation on the
    device) run eagerly and as one captured HIP graph; both timed over whole
    steps."""
    import torch

    from torque_amd import BlockSparseTensor, default_handle, normalize_f64_noalloc
    from torque_amd import workloads as W

    E, A, Ac, P, Q, legs = W.ctmrg_step(D, chi)
    t = [BlockSparseTensor(x.dims, x.data, x.begin, x.end, device=dev) for x in (E, A, Ac, P, Q)]
    gE, gA, gAc, gP, gQ = t
    plans, bufs = [], []
    prev = None
    for i, (lhs, rhs) in enumerate([(gA, gAc), (gE, None), (None, gP), (None, gQ)]):
        a = lhs if lhs is not None else prev
        b = rhs if rhs is not None else prev
        p = a.plan_contract(b, legs[i], merged=True)
        buf = torch.empty(max(p.out_n_elem, 1), dtype=torch.float64, device=dev)
        p.autotune(a.data, b.data, buf)
        plans.append((p, a, b))
        bufs.append(buf)
        prev = p.output_tensor(buf)
    norm = torch.empty(1, dtype=torch.float64, device=dev)
    g = bufs[-1][:plans[-1][0].out_n_elem]

    def step():
        for (p, a, b), buf in zip(plans, bufs):
            p.execute(a.data, b.data, buf)
        normalize_f64_noalloc(g, norm)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    eager_s = timed(step)
    h = default_handle(dev.index)
    with h.capture() as cap:
        step()
    graph_s = timed(cap.graph.launch)
    flops = sum(p.flops for p, _, _ in plans)
    res = {"workload": "ctmrg_edge_update_D%d_chi%d" % (D, chi),
           "blocks_out": [int(p.n_pairs) for p, _, _ in plans],
           "pairs": [int(p.n_matched) for p, _, _ in plans],
           "gflop": round(flops / 1e9, 4),
           "eager_ms": round(eager_s * 1e3, 4), "graph_ms": round(graph_s * 1e3, 4),
           "graph_gflops": round(flops / graph_s / 1e9, 1), "norm": float(norm[0])}
    print("# CTMRG edge update D=%d chi=%d (4 merged contractions + FP64 normalisation, %.3f "
          "GFLOP): eager %.3f ms, one graph %.3f ms" % (D, chi, flops / 1e9, eager_s * 1e3,
                                                       graph_s * 1e3), flush=True)
    del bufs
    torch.cuda.empty_cache()
    return res


SCALING_LENGTHS = (2, 4, 8, 16, 32, 64)
SCALING_SLICES = (0, 2, 3, 4, 4, 4)


def bench_scaling(dev, reps=5):
    """The reference's large-tensor benchmark (test/gpu/large_tensor_benchmark.cu:28-760)
    through the drop-in classes: per-call contract time of a rank-2, -3 and -4
    tensor with a matrix (legs {1,0}, or {1,1},{2,0}) at leg lengths 2..64,
    block-sparse with the blocks sliced j times by 2 on every leg (4^j, 8^j,
    16^j blocks), dense, and element-sparse with every index stored. Block
    sparse calls are timed cold (a new plan, as the reference plans every
    call) and warm (the handle's cached plan); each time is the median of
    `reps` calls, host to host around a device synchronisation."""
    import torch

    from torque_amd import BlockSparseTensor, DenseTensor, SparseTensor, default_handle
    from torque_amd.tensor import ContractionPlan

    rng = np.random.default_rng(0)
    h = default_handle(dev.index)

    def median_call(fn):
        ts = []
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts[1:])) * 1e3

    def legs_for(rank):
        return np.array([[1, 0]] if rank == 2 else [[1, 1], [2, 0]], np.int32)

    bs, dense, sparse = [], [], []
    for rank in (2, 3, 4):
        legs = legs_for(rank)
        for length, jmax in zip(SCALING_LENGTHS, SCALING_SLICES):
            dims_t, dims_m = [length] * rank, [length] * 2
            xt = rng.random(length ** rank)
            xm = rng.random(length ** 2)
            t = BlockSparseTensor(dims_t, xt, np.zeros((1, rank), np.int64),
                                  np.full((1, rank), length - 1, np.int64), device=dev)
            m = BlockSparseTensor(dims_m, xm, np.zeros((1, 2), np.int64),
                                  np.full((1, 2), length - 1, np.int64), device=dev)
            for j in range(jmax + 1):
                if j:
                    t = t.slice([2] * rank)
                    m = m.slice([2, 2])

                def cold():
                    p = ContractionPlan(h, t, m, legs, "native", False)
                    out = torch.empty(max(p.out_n_elem, 1), dtype=torch.float64, device=dev)
                    p.execute(t.data, m.data, out)
                    p.output_tensor(out)
                    p.destroy()

                cold_ms = median_call(cold)
                t.contract(m, legs)
                warm_ms = median_call(lambda: t.contract(m, legs))
                p = t.plan_contract(m, legs)
                bs.append({"rank": rank, "length": length, "slices": j,
                           "blocks": int(t.n_blocks), "pairs": int(p.n_pairs),
                           "cold_ms": round(cold_ms, 3), "warm_ms": round(warm_ms, 3)})
            dt_ = DenseTensor(dims_t, xt, device=dev)
            dm = DenseTensor(dims_m, xm, device=dev)
            d_ms = median_call(lambda: dt_.contract(dm, legs))
            flop = 2.0 * length ** (rank + (1 if rank == 2 else 0))
            dense.append({"rank": rank, "length": length, "ms": round(d_ms, 3),
                          "gflops": round(flop / d_ms / 1e6, 3)})
            st = SparseTensor(dims_t, xt, np.arange(length ** rank), device=dev)
            sm = SparseTensor(dims_m, xm, np.arange(length ** 2), device=dev)
            s_ms = median_call(lambda: st.contract(sm, legs))
            sparse.append({"rank": rank, "length": length, "nnz": length ** rank,
                           "ms": round(s_ms, 3)})
        big = [e for e in bs if e["rank"] == rank][-1]
        print("# scaling rank %d: block-sparse %d blocks (%d pairs) cold %.2f ms warm %.3f ms; "
              "dense length 64 %.3f ms; sparse %d nnz %.3f ms"
              % (rank, big["blocks"], big["pairs"], big["cold_ms"], big["warm_ms"],
                 dense[-1]["ms"], sparse[-1]["nnz"], sparse[-1]["ms"]), flush=True)
    return {"source": "test/gpu/large_tensor_benchmark.cu", "lengths": list(SCALING_LENGTHS),
            "block_sparse": bs, "dense": dense, "sparse": sparse}


SHAPES = ((128, 128, 128), (64, 64, 256), (64, 64, 64), (40, 36, 35), (32, 32, 32),
          (16, 16, 64))


def bench_shapes(dev, gflop=20.0):
    """FP64 grouped GEMM on uniform block-diagonal products (nb identical
    M x K . K x N pairs, A stored row index fastest, B k fastest, as the batched
    workload's operands), autotuned, against the vendor library's FP64 batched
    GEMM on the same shapes (torch.bmm: rocBLAS / hipBLASLt): the rates the
    batched workload's pair shapes allow."""
    import torch

    from torque_amd import BlockSparseTensor

    def timed(f, n=5):
        f()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    out = []
    for (M, N, K) in SHAPES:
        nb = int(min(gflop * 1e9 / (2 * M * N * K), 3e8 / max(M * K, N * K, M * N)))
        ia = np.arange(nb, dtype=np.int64)
        ba = np.stack([ia * M, ia * K], 1)
        ea = np.stack([ia * M + M - 1, ia * K + K - 1], 1)
        bb = np.stack([ia * K, ia * N], 1)
        eb = np.stack([ia * K + K - 1, ia * N + N - 1], 1)
        da = torch.randn(nb * M * K, dtype=torch.float64, device=dev)
        db = torch.randn(nb * K * N, dtype=torch.float64, device=dev)
        A = BlockSparseTensor([nb * M, nb * K], da, ba, ea, device=dev)
        B = BlockSparseTensor([nb * K, nb * N], db, bb, eb, device=dev)
        p = A.plan_contract(B, [[1, 0]], "native")
        c = torch.empty(p.out_n_elem, dtype=torch.float64, device=dev)
        best, _ = p.autotune(A.data, B.data, c, reps=2)
        ms = timed(lambda: p.execute(A.data, B.data, c))
        # the same products as one strided batch of contiguous A_i (M x K)
        # and B_i (K x N)
        ta = da.view(nb, K, M).transpose(1, 2).contiguous()
        tb = db.view(nb, N, K).transpose(1, 2).contiguous()
        ms_lib = timed(lambda: torch.bmm(ta, tb))
        ref = torch.bmm(ta, tb)
        got = c.view(nb, N, M).transpose(1, 2)
        err = float((got - ref).norm() / ref.norm())
        del ta, tb, ref, got
        r = {"m": M, "n": N, "k": K, "pairs": nb, "gflop": round(p.flops / 1e9, 2),
             "tflops": round(p.flops / ms / 1e9, 1),
             "library_tflops": round(p.flops / ms_lib / 1e9, 1), "config": best,
             "rel_err_vs_library": err}
        out.append(r)
        print("# %d x (%d, %d, %d): grouped GEMM %.1f TFLOP/s (configs %s), library batched "
              "GEMM %.1f TFLOP/s, rel diff %.1e" % (nb, M, N, K, r["tflops"], best,
                                                    r["library_tflops"], err), flush=True)
        del A, B, da, db, c, p
        torch.cuda.empty_cache()
    return {"workload": "uniform_block_diagonal_fp64", "shapes": out}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload in ("shapes", "scaling"):
        import torch
        dev = torch.device("cuda:0")
        torch.cuda.set_device(dev)
        fn = bench_shapes if args.workload == "shapes" else bench_scaling
        print(json.dumps(fn(dev)), flush=True)
        return
    if args.workload in ("permute", "large", "corner6"):
        import torch
        dev = torch.device("cuda:0")
        torch.cuda.set_device(dev)
        fn = {"permute": bench_permute, "large": bench_large,
              "corner6": bench_corner_rank6}[args.workload]
        print(json.dumps(fn(dev)), flush=True)
        return
    import torch

    from torque_amd import BlockSparseTensor, TorqueError
    from torque_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    E, A, Ac, l1, l2 = W.peps(args.D, args.chi, seed=args.seed, log10_scale_range=(-3.0, 0.0))
    gE = BlockSparseTensor(E.dims, E.data, E.begin, E.end, device=dev)
    gA = BlockSparseTensor(A.dims, A.data, A.begin, A.end, device=dev)
    gAc = BlockSparseTensor(Ac.dims, Ac.data, Ac.begin, Ac.end, device=dev)

    t0 = time.perf_counter()
    p1 = gA.plan_contract(gAc, l1)
    aa = torch.empty(max(p1.out_n_elem, 1), dtype=torch.float64, device=dev)
    p1.execute(gA.data, gAc.data, aa)
    ga = p1.output_tensor(aa)
    p2 = gE.plan_contract(ga, l2)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    out = torch.empty(max(p2.out_n_elem, 1), dtype=torch.float64, device=dev)
    t0 = time.perf_counter()
    best1, _ = p1.autotune(gA.data, gAc.data, aa)
    best2, times2 = p2.autotune(gE.data, aa, out)
    tune_s = time.perf_counter() - t0
    print("# autotune %.2f s: A.A* configs %s, E.a configs %s; E.a ms per config %s / %s"
          % (tune_s, best1, best2, ["%.3f" % x for x in times2[0]],
             ["%.3f" % x for x in times2[1]]), flush=True)
    print("# workload D=%d chi=%d: A.A* %d pairs, E.a %d pairs, %.3f GFLOP, out %.1f M elem, "
          "plans %.3f s" % (args.D, args.chi, p1.n_pairs, p2.n_pairs,
                            (p1.flops + p2.flops) / 1e9, p2.out_n_elem / 1e6, plan_s),
          flush=True)

    def step():
        p1.execute(gA.data, gAc.data, aa)
        p2.execute(gE.data, aa, out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    flops = p1.flops + p2.flops
    value = flops / dt / 1e9
    print("# fp64 step %.3f ms, %.1f GFLOP/s" % (dt * 1e3, value), flush=True)

    # dominant kernel: the E.a grouped GEMM, events on the handle's stream
    stream = torch.cuda.current_stream(dev)
    p2.pack(gE.data, aa)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    n_rep = max(args.steps, 5)
    ev0.record(stream)
    for _ in range(n_rep):
        p2.multiply(gE.data, aa, out)
    ev1.record(stream)
    torch.cuda.synchronize()
    gemm_s = ev0.elapsed_time(ev1) / 1e3 / n_rep
    alg_bytes = p2.bytes
    achieved = alg_bytes / gemm_s / 1e9
    n_tiles, n_segs = p2.work()
    kernel = ("small_k_kernel" if n_tiles == 0 else
              "grouped_gemm_kernel" if n_segs == 0 else "small_k_kernel + grouped_gemm_kernel")
    traffic = committed_traffic("%s_peps_absorption_D%d_chi%d" % (kernel, args.D, args.chi))
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kernel + " (E.a)", "kernel_ms": round(gemm_s * 1e3, 4),
                "kernel_gflops": round(p2.flops / gemm_s / 1e9, 1),
                "algorithmic_bytes": alg_bytes}
    print("# E.a multiply (%s; %d tiles, %d segments) %.3f ms: %.0f GB/s algorithmic "
          "(%.1f%% of HBM), %.0f GFLOP/s"
          % (kernel, n_tiles, n_segs, gemm_s * 1e3, achieved, 100 * achieved / HBM_PEAK_GBS,
             p2.flops / gemm_s / 1e9), flush=True)

    # ---- oracle sample: E blocks [0, s) x all of a, pairs contiguous in out
    from oracle import oracle as O
    s_blocks = E.n_blocks if args.cpu_sample_blocks <= 0 else min(args.cpu_sample_blocks,
                                                                 E.n_blocks)
    n_sub = int(np.prod(E.end[:s_blocks] - E.begin[:s_blocks] + 1, axis=1).sum())
    TEs = O.BS(E.dims, E.begin[:s_blocks], E.end[:s_blocks])
    Ta = O.BS(ga.dimension, ga.begin_points, ga.end_points)
    aa_host = aa.cpu().numpy()
    threads = host_threads()
    cpu_baseline = None
    ref_c, ref_p = O.contract(TEs, E.data[:n_sub], Ta, aa_host, l2, threads=threads)
    sample_flops = float((2 * ref_p["m"] * ref_p["n"] * ref_p["k"]).sum())
    if not args.no_cpu:
        # the oracle (the reference's per-pair pack + GEMM + append, restated)
        # on this host: all threads over the whole E.a, one thread over its
        # first tenth of the E blocks; median of 5 runs after a warm-up each
        cpu_s = median_time(lambda: O.contract(TEs, E.data[:n_sub], Ta, aa_host, l2,
                                               threads=threads))
        s1 = max(1, s_blocks // 10)
        n1 = int(np.prod(E.end[:s1] - E.begin[:s1] + 1, axis=1).sum())
        TE1 = O.BS(E.dims, E.begin[:s1], E.end[:s1])
        _, p1s = O.contract(TE1, E.data[:n1], Ta, aa_host, l2, threads=threads)
        f1 = float((2 * p1s["m"] * p1s["n"] * p1s["k"]).sum())
        cpu1_s = median_time(lambda: O.contract(TE1, E.data[:n1], Ta, aa_host, l2, threads=1))
        cpu_baseline = {"value": round(sample_flops / cpu_s / 1e9, 3), "unit": "GFLOP/s",
                        "cores": threads, "kind": "port",
                        "sample": "E.a restricted to the first %d E blocks (%d pairs, %.3f "
                                  "GFLOP), oracle per-pair pack + GEMM, median %.3f s of 5"
                                  % (s_blocks, ref_p["m"].size, sample_flops / 1e9, cpu_s),
                        "single_thread": {"value": round(f1 / cpu1_s / 1e9, 3),
                                          "sample": "first %d E blocks (%d pairs, %.3f GFLOP), "
                                                    "median %.3f s of 5"
                                                    % (s1, p1s["m"].size, f1 / 1e9, cpu1_s)},
                        "nproc": os.cpu_count(), "cpu_model": cpu_model()}
        print("# CPU oracle (%s, nproc %d): %d threads %.2f GFLOP/s, 1 thread %.2f GFLOP/s"
              % (cpu_model(), os.cpu_count() or 0, threads, sample_flops / cpu_s / 1e9,
                 f1 / cpu1_s / 1e9), flush=True)
    n_ref = ref_c.size

    def rel(x):
        return float(np.linalg.norm(x[:n_ref] - ref_c) / np.linalg.norm(ref_c))

    def per_block_max(x):
        worst = 0.0
        offs = ref_p["out_offsets"]
        sizes = ref_p["m"] * ref_p["n"]
        for o, n in zip(offs, sizes):
            r = ref_c[o:o + n]
            nr = np.linalg.norm(r)
            if nr > 0:
                worst = max(worst, float(np.linalg.norm(x[o:o + n] - r) / nr))
        return worst

    got64 = out[:n_ref].cpu().numpy()
    err64 = rel(got64)
    tiers = {"fp64": {"gflops": round(value, 1), "rel_err": err64,
                      "max_block_rel_err": per_block_max(got64)}}

    for tier in [t for t in args.tiers.split(",") if t]:
        try:
            pt1 = gA.plan_contract(gAc, l1, precision=tier)
            pt2 = gE.plan_contract(ga, l2, precision=tier)
        except TorqueError as e:
            tiers[tier] = {"error": str(e)}
            continue
        aa_t = torch.empty_like(aa)
        out_t = torch.empty_like(out)
        try:
            pt1.autotune(gA.data, gAc.data, aa_t)
            pt2.autotune(gE.data, aa, out_t)
        except TorqueError as e:
            tiers[tier] = {"error": str(e)}
            continue

        def tstep():
            pt1.execute(gA.data, gAc.data, aa_t)
            pt2.execute(gE.data, aa_t, out_t)

        try:
            for _ in range(args.warmup):
                tstep()
        except TorqueError as e:
            tiers[tier] = {"error": str(e)}
            continue
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            tstep()
        torch.cuda.synchronize()
        tdt = (time.perf_counter() - t0) / args.steps
        # error of the tier's E.a on the FP64 double layer, vs the oracle
        pt2.execute(gE.data, aa, out_t)
        got = out_t[:n_ref].cpu().numpy()
        tiers[tier] = {"gflops": round(flops / tdt / 1e9, 1), "ms_per_step": round(tdt * 1e3, 3),
                       "rel_err": rel(got), "max_block_rel_err": per_block_max(got),
                       "speedup_vs_fp64": round(dt / tdt, 3)}
        print("# %s step %.3f ms, %.1f GFLOP/s, rel err %.2e" %
              (tier, tdt * 1e3, flops / tdt / 1e9, rel(got)), flush=True)

    # merged output: E.a writes one block per distinct output block (the sum
    # of its pairs) instead of one per matched pair; same dense result
    pm = gE.plan_contract(ga, l2, merged=True)
    outm = torch.empty(max(pm.out_n_elem, 1), dtype=torch.float64, device=dev)
    bestm, _ = pm.autotune(gE.data, aa, outm)

    def mstep():
        p1.execute(gA.data, gAc.data, aa)
        pm.execute(gE.data, aa, outm)

    for _ in range(args.warmup):
        mstep()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mstep()
    torch.cuda.synchronize()
    mdt = (time.perf_counter() - t0) / args.steps
    merged = {"ms_per_step": round(mdt * 1e3, 4), "gflops": round(flops / mdt / 1e9, 1),
              "out_blocks": int(pm.n_pairs), "matched_pairs": int(pm.n_matched),
              "out_elements": int(pm.out_n_elem), "speedup_vs_unmerged": round(dt / mdt, 3),
              "gemm_config": bestm}
    print("# merged output: E.a %d blocks for %d pairs, %.1f M elements; step %.3f ms, "
          "%.1f GFLOP/s (%.2fx the unmerged step)"
          % (pm.n_pairs, pm.n_matched, pm.out_n_elem / 1e6, mdt * 1e3, flops / mdt / 1e9,
             dt / mdt), flush=True)
    del outm
    pm.destroy()

    # per-call cost of the drop-in contract(): planning included (cold, a
    # fresh plan), a repeated call on the cached plan (warm), and the execute
    # alone on the same plan
    from torque_amd import default_handle
    default_handle().plans.clear()
    del out
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    cold = gE.contract(ga, l2)
    torch.cuda.synchronize()
    cold_s = time.perf_counter() - t0
    del cold

    def warm_call():
        c = gE.contract(ga, l2)
        torch.cuda.synchronize()
        return c

    warm_s = median_time(warm_call)
    pw = gE.plan_contract(ga, l2)
    outw = torch.empty(max(pw.out_n_elem, 1), dtype=torch.float64, device=dev)

    def exec_call():
        pw.execute(gE.data, ga.data, outw)
        torch.cuda.synchronize()

    exec_s = median_time(exec_call)
    per_call = {"workload": "E.a of the D=%d chi=%d step" % (args.D, args.chi),
                "cold_contract_ms": round(cold_s * 1e3, 3),
                "warm_contract_ms": round(warm_s * 1e3, 3),
                "execute_ms": round(exec_s * 1e3, 3),
                "warm_over_execute": round(warm_s / exec_s, 3)}
    print("# contract() per call: cold %.2f ms (plan + execute), warm %.3f ms, execute %.3f ms"
          % (cold_s * 1e3, warm_s * 1e3, exec_s * 1e3), flush=True)
    del outw

    ctmrg = bench_ctmrg(dev, args.D, args.chi, args.steps, args.warmup)
    split_k = bench_split_k(dev)
    permute = None if args.no_permute else bench_permute(dev)
    large = None if args.no_large else bench_large(dev)
    corner128 = None if args.no_large else bench_large(dev, D=8, chi=128, oracle_max_block=4096)
    corner6 = None if args.no_large else bench_corner_rank6(dev)
    scaling = bench_scaling(dev)
    batched = None
    if not args.no_batched:
        bl = bench_batched(args)
        batched = {"workload": bl["config"]["workload"], "pairs": bl["config"]["pairs"],
                   "gflop": bl["config"]["gflop_per_step"], "ms": bl["ms_per_step"],
                   "gflops": bl["value"], "frac_fp64_mfma_peak": bl["roofline"]["frac"],
                   "gemm_config": bl["config"]["gemm_config"], "tiers": bl["tiers"],
                   "cost_model": bl["config"]["cost_model"],
                   "autotune_sweep_s": bl["config"]["autotune_sweep_s"],
                   "cpu_baseline": bl["cpu_baseline"]}
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "peps_absorption_D%d_chi%d" % (args.D, args.chi),
                   "charges": "-2..2" if args.chi <= 128 else "-3..3",
                   "pairs": [int(p1.n_pairs), int(p2.n_pairs)],
                   "gflop_per_step": round(flops / 1e9, 4),
                   "out_elements": int(p2.out_n_elem), "plan_seconds": round(plan_s, 3),
                   "gemm_config": {"A.A*": best1, "E.a": best2},
                   "entries": "N(0,1)*10^U(-3,0) per block", "seed": args.seed},
        "roofline": roofline,
        "cpu_baseline": cpu_baseline,
        "tiers": tiers,
        "fp64_rel_err_vs_oracle": err64,
        "merged_output": merged,
        "ctmrg_step_graph": ctmrg,
        "split_k": split_k,
        "per_call": per_call,
        "enlarged_corner_D8_chi128": corner128,
        "permute_rank6_200M": permute,
        "enlarged_corner_D10_chi256": large,
        "enlarged_corner_D10_chi256_rank6": corner6,
        "batched_512_u1_rank4x4": batched,
        "reference_scaling": scaling,
    }
    print(json.dumps(line), flush=True)
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("# calibration copy: 1 GiB read, 1 GiB written", flush=True)


if __name__ == "__main__":
    main()
