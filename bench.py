"""Benchmark of the MI355X lazy equi-join path.  Prints one JSON line.

Default workload (`--workload pro`): PRO on |R| = |S| = 2^27 tuples of 8 bytes
{payloadID, key}.  R keys are a seeded permutation of 1..2^27, S keys uniform on
1..2^27, payload = row id.  A step is one PRO join of the device-resident
relations: partition passes, LDS build/probe, count/checksum reduction.

Roofline: per-kernel device times come from HIP events recorded around every
launch on the join stream (library option kernel_timing) during the timed
steps.  The dominant kernel's algorithmic bytes per launch (hist 8 B/tuple,
scatter 16 B/tuple, build/probe 8 B/tuple) over its mean duration gives
`achieved`.

Other workloads (each prints its own line):
  --workload skew     |R| = 2^24 unique keys, |S| = 2^28 Zipf(--theta) over R's keys
  --workload windows  1024 tumbling windows of two Poisson streams (1.6 tuples/us),
                      per-side sizes log-uniform in [4K, 4M], duplicate factor 4;
                      one graph-captured batched launch vs per-window launches
  --workload serve    64 stream pairs of window-join requests (Poisson arrivals,
                      lognormal sizes median 256K sigma 1 capped 16M per side, R keys
                      uniform over 2^24, S keys uniform / Zipf 1.0 half-half) served
                      under a 200 GB HBM budget at offered loads 50..120% of the
                      batched peak (ajbench -a SERVE); p50/p99/p999 latency
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PRO_BYTES_PER_TUPLE = 32  # histogram read 8 + scatter read 8 + write 8 + build/probe read 8
# NPO per R tuple: read 8 + 32 of table slots written (16-byte slots at half
# load); per S tuple: read 8 + one 16-byte slot read.  Average over |R|+|S|.
NPO_BYTES_PER_TUPLE = 32
# SHJ per tuple: read 8 + its own table's slot 16 + a slot of the other table
# 16 + 32 of its own table's slots initialised
SHJ_BYTES_PER_TUPLE = 72
# window batches, credited per tuple whatever passes run (as PRO's 32): 12 read
# for the window's histogram + 12 re-read and 8 written by the scatter + 8 read
# by the build/probe.  The staged pass moves 28 of them (no histogram read).
WINDOW_BYTES_PER_TUPLE = 40
# algorithmic bytes per input tuple of each kernel (8-byte tuples), credited for
# the tuples that kernel reads: NPO's build reads R only, its probe S only.  With
# the histogram passes pass 1's histogram kernel (k_hist2d) also counts the
# pass-2 digits, and the "hist2" phase is only the pass-2 unit planning.
KERNEL_BYTES = {"hist1": 8, "scatter1": 16, "hist2": 0, "scatter2": 16, "join": 8,
                "npo_build": 40, "npo_probe": 24, "shj_step": 72}
KERNEL_SIDE = {"npo_build": "R", "npo_probe": "S"}  # others: |R| + |S|
KERNEL_NAMES = {"hist1": "k_hist2d (pass 1)", "scatter1": "k_scatter_pipe (pass 1)",
                "hist2": "pass-2 planning (k_p2_units, k_h2_usum, k_h2_base)",
                "scatter2": "k_scatter_p2 (pass 2)",
                "join": "k_join", "npo_build": "NPO build (k_stage1 + k_stage2 hashing R into "
                "the table's regions, k_npo_region)", "npo_probe": "k_npo_probe",
                "shj_step": "k_shj_step (all eager steps)",
                "heavy_split": "k_heavy_split (R and S read once: the hottest keys of S summed, "
                               "the rest of S written out; 8 B per tuple read + 8 B per S "
                               "tuple written)"}
# the staged partitioning (aj_join_out_t.flags & AJ_OUT_STAGED): one pass-1 and one
# pass-2 launch per super-chunk of STAGE_CHUNK tuples (the library's default)
STAGED_NAMES = {"scatter1": "k_stage1 (pass 1)", "scatter2": "k_stage2 (pass 2)"}
STAGE_CHUNK = 1 << 25


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["pro", "skew", "windows", "serve"], default="pro")
    ap.add_argument("--peak", type=float, default=0.0,
                    help="serve: the M tuples/s the offered load is relative to (0 = the "
                         "batched peak measured in the same run)")
    ap.add_argument("--calib-windows", type=int, default=256,
                    help="serve: resident windows of the reference-rate measurement")
    ap.add_argument("--loads", default="0.5,0.7,0.9,1.0,1.1,1.2")
    ap.add_argument("--seconds", type=float, default=2.0)
    ap.add_argument("--algo", choices=["PRO", "NPO", "SHJ"], default="PRO",
                    help="pro workload: lazy PRO / NPO, or the eager symmetric hash join")
    ap.add_argument("--batch", type=int, default=0,
                    help="SHJ: tuples per side per eager step (0 = library default)")
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--theta", type=float, default=1.0)
    ap.add_argument("--nwin", type=int, default=1024)
    ap.add_argument("--no-graph", action="store_true",
                    help="windows workload, for counter runs: enqueue the batch's kernels "
                         "directly instead of replaying a captured hipGraph, and wait for "
                         "every launch (rocprofv3 --pmc ended the process with launches "
                         "queued back to back); the times include those waits")
    ap.add_argument("--radix-bits", type=int, default=0)
    ap.add_argument("--passes", type=int, default=0)
    ap.add_argument("--cap", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-staged", action="store_true",
                    help="PRO: histogram-based partition passes instead of the staged ones")
    ap.add_argument("--no-tune", action="store_true",
                    help="pro/skew: use the size heuristic instead of autotuning the geometry")
    ap.add_argument("--cpu-log2n", type=int, default=27,
                    help="SHJ's CPU baseline sample and the cap of the reference's RJ_st sample")
    ap.add_argument("--calib", action="store_true",
                    help="also run aj_copy_kernel over 1 GiB (traffic-counter calibration)")
    return ap.parse_args()


def cpu_baseline_shj(aj, log2n: int):
    """The reference's own single-thread symmetric hash join (shj, compiled from
    its sources into oracle/_ref) on a bounded sample of the same workload."""
    import oracle
    if not oracle.ref_available():
        return None
    n = 1 << log2n
    R = aj.gen_unique(n, 12345)
    S = aj.gen_uniform(n, 1, n, 54321)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        assert oracle.ref_shj(R, S) == n
        times.append(time.perf_counter() - t0)
    sec = sorted(times)[1]
    return {"value": round(2 * n / sec / 1e6, 2), "unit": "M tuples/s", "cores": 1,
            "kind": "reference",
            "sample": f"reference shj (helper/localjoiner.cpp:33-64, one thread) "
                      f"|R|=|S|=2^{log2n}, same key distribution, median of 3 "
                      f"(input copies included)"}


def host_info():
    """The host the CPU baselines run on: model, sockets, nproc and the CPUs this
    process may use (the box's share)."""
    model, sockets = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("physical id"):
                    sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "sockets": max(len(sockets), 1), "nproc": os.cpu_count(),
            "usable_cpus": usable}


def cpu_pro_timing(R, S, threads: int, reps: int):
    """The CPU oracle's PRO (clean-room restatement of the reference's multithreaded
    PRO, 8 radix bits in 2 passes: its defaults) on host relations: median of `reps`
    runs after one warm-up, temporaries included."""
    import oracle
    times = []
    for i in range(reps + 1):
        r = oracle.pro(R, S, nthreads=threads, radix_bits=8, passes=2)
        assert r.count == len(S)
        if i:
            times.append(r.total_ms)
    return float(np.median(times))


def cpu_baseline(aj, R, S, host):
    """The CPU baseline of the 2^27 x 2^27 join: the oracle's PRO on the same relations
    the GPU joined, on the host CPUs this process may use (at most 16, the box's
    share of its host)."""
    threads = min(host["usable_cpus"] or 1, 16)
    ms = cpu_pro_timing(R, S, threads, 3)
    n = len(R) + len(S)
    return {"value": round(n / (ms * 1e-3) / 1e6, 2), "unit": "M tuples/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle PRO (the reference's multithreaded PRO restated) on the same "
                      f"|R|=|S|=2^{len(R).bit_length() - 1} relations, {threads} threads, 8 radix "
                      f"bits / 2 passes (reference defaults), median of 3 after a warm-up, "
                      f"temporaries included; host {host['model']}, {host['sockets']} socket(s), "
                      f"nproc {host['nproc']}, {host['usable_cpus']} CPUs usable"}


def cpu_baseline_small(aj, host):
    """The reference's CPU-sized case: 2^20 x 2^20, 8 threads, median of 5 after a warm-up."""
    n = 1 << 20
    R = aj.gen_unique(n, 12345)
    S = aj.gen_uniform(n, 1, n, 54321)
    ms = cpu_pro_timing(R, S, 8, 5)
    return {"value": round(2 * n / (ms * 1e-3) / 1e6, 2), "unit": "M tuples/s", "cores": 8,
            "kind": "port", "ms": round(ms, 3),
            "sample": "oracle PRO |R|=|S|=2^20, R a permutation, S uniform, 8 threads, 8 radix "
                      "bits / 2 passes, median of 5 after a warm-up"}


def cpu_reference_pro(aj, log2n: int):
    """The reference's own single-thread radix join (RJ_st, compiled from its sources
    into oracle/_ref: the same radix_cluster / bucket_chaining_join code as its
    multithreaded PRO, which cannot run off its authors' machine) on a bounded sample
    of the same workload; reported beside the multithreaded port, not as the target."""
    import oracle
    if not oracle.ref_available():
        return None
    n = 1 << log2n
    R = aj.gen_unique(n, 12345)
    S = aj.gen_uniform(n, 1, n, 54321)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        assert oracle.ref_join("RJ_st", R, S)[0] == n
        times.append(time.perf_counter() - t0)
    sec = sorted(times)[1]
    return {"value": round(2 * n / sec / 1e6, 2), "unit": "M tuples/s", "cores": 1,
            "kind": "reference",
            "sample": f"reference RJ_st (parallel_radix_join.cpp:1719-1844, one thread) "
                      f"|R|=|S|=2^{log2n}, same key distribution, median of 3 "
                      f"(input copies included)"}


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 traffic summary of
    this workload (profiles/*_traffic.json), or None when none was measured for it."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.json"))):
        try:
            with open(path) as f:
                d = json.load(f)
        except (OSError, ValueError):
            continue
        if d.get("workload") == workload:
            return d["hbm_bytes_per_launch"].get(kernel)
    return None


def closed_form_fk(R: np.ndarray, S: np.ndarray):
    """count/checksum of a foreign-key join where R's keys are a permutation of
    1..|R| and every S key is one of them: count = |S| and the checksum follows
    from R's inverse permutation."""
    inv = np.empty(len(R) + 1, np.uint64)
    inv[R["key"]] = R["payloadID"].astype(np.uint64)
    x = (inv[S["key"]] << np.uint64(32)).sum(dtype=np.uint64)
    x = (int(x) + int(S["payloadID"].astype(np.uint64).sum(dtype=np.uint64))) % (1 << 64)
    return len(S), x


def autotune(aj, ctx, dR, nR, dS, nS, algo, kw, args):
    """Autotuned geometry for this size bucket (untimed, before the warm-up); the
    timed joins then run with the table's entry (their options leave it open)."""
    if args.no_tune or algo != aj.AJ_ALGO_PRO or any(kw.values()):
        return None
    t0 = time.time()
    best = aj.tune_pro(ctx, dR.ptr, nR, dS.ptr, nS, reps=3)
    print(f"[bench] tuned in {time.time() - t0:.1f}s: {best}", file=sys.stderr, flush=True)
    return best


def timed_joins(aj, ctx, dR, nR, dS, nS, algo, kw, steps, warmup, want):
    """W warm-up joins, then exactly K timed joins (host clock around them, a device
    synchronisation on both sides), then K more joins with HIP events around every
    kernel launch on the join stream for the per-kernel breakdown.  The events cost
    about 4% of a join, so the timed region runs without them."""
    for _ in range(warmup):
        out = ctx.join(dR.ptr, nR, dS.ptr, nS, algo, **kw)
        assert (out.count, out.checksum) == want, (out.count, out.checksum, want)
    aj.device_sync()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        out = ctx.join(dR.ptr, nR, dS.ptr, nS, algo, **kw)
        dev_ms.append(out.device_ms)
        assert (out.count, out.checksum) == want
    aj.device_sync()
    wall = time.perf_counter() - t0
    aj.set_option("kernel_timing", 1)
    phases, ev_ms = [], []
    try:
        for _ in range(steps):
            o2 = ctx.join(dR.ptr, nR, dS.ptr, nS, algo, **kw)
            phases.append({**o2.phases(), "heavy_split": float(o2.heavy_ms)})
            ev_ms.append(o2.device_ms)
            assert (o2.count, o2.checksum) == want
    finally:
        aj.set_option("kernel_timing", 0)
    return out, float(np.mean(dev_ms)), wall * 1e3 / steps, phases, float(np.mean(ev_ms))


def join_line(aj, args, workload_name, nR, nS, algo, out, dev, ms_step, phases, metric,
              ev_dev=None):
    tuples = nR + nS
    mean_ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    if algo == aj.AJ_ALGO_SHJ:  # one kernel type: the steps are the whole device time
        mean_ph = {"shj_step": dev}
    # With the hottest keys of S joined by their aggregates (AJ_OUT_HEAVY), one
    # pass reads R and S and writes out the rest of S (out.light tuples); the
    # partitioned join then takes R and that rest only.
    heavy = bool(getattr(out, "flags", 0) & aj.AJ_OUT_HEAVY)
    light = int(out.light) if heavy else nS

    def kernel_bytes(k):
        if k == "heavy_split":
            return 8 * (nR + nS) + 8 * light
        side = KERNEL_SIDE.get(k)
        return KERNEL_BYTES[k] * (nR if side == "R" else nS if side == "S" else nR + light)

    cands = list(KERNEL_BYTES) + (["heavy_split"] if heavy else [])
    dom = max(cands, key=lambda k: mean_ph.get(k, 0.0))
    dom_bytes = kernel_bytes(dom)
    achieved = dom_bytes / (mean_ph[dom] * 1e-3) / 1e9
    staged = bool(getattr(out, "flags", 0) & aj.AJ_OUT_STAGED)
    kname = (STAGED_NAMES.get(dom) if staged else None) or KERNEL_NAMES[dom]
    launches = 1
    if staged and dom in STAGED_NAMES:
        chunk = aj.get_option("stage_chunk") or STAGE_CHUNK
        launches = -(-nR // chunk) + -(-nS // chunk)
    bpt = {aj.AJ_ALGO_PRO: PRO_BYTES_PER_TUPLE, aj.AJ_ALGO_NPO: NPO_BYTES_PER_TUPLE,
           aj.AJ_ALGO_SHJ: SHJ_BYTES_PER_TUPLE}[algo]
    whole_gbs = bpt * tuples / (ms_step * 1e-3) / 1e9
    dev_gbs = bpt * tuples / (dev * 1e-3) / 1e9
    return {
        "metric": metric,
        "value": round(tuples / (ms_step * 1e-3) / 1e6, 1),
        "unit": "M tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": workload_name, "n_r": nR, "n_s": nS},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": kname, "launches_per_join": launches,
                     "bytes_per_tuple": KERNEL_BYTES.get(dom, 8),
                     "bytes_per_launch": dom_bytes // launches,
                     "mean_launch_ms": round(mean_ph[dom] / launches, 4),
                     "whole_join": {"achieved": round(whole_gbs, 1),
                                    "frac": round(whole_gbs / HBM_PEAK_GBS, 4),
                                    "bytes_per_tuple": bpt, "ms": round(ms_step, 3)}},
        "pipeline": {"device_ms": round(dev, 3), "alg_bytes_per_tuple": bpt,
                     "alg_GBps": round(dev_gbs, 1),
                     "frac_of_peak": round(dev_gbs / HBM_PEAK_GBS, 4),
                     "partitioning": ("region build" if staged else "flat build")
                     if algo == aj.AJ_ALGO_NPO else "staged" if staged else "histogram passes",
                     "heavy_keys": {"S_by_aggregates": nS - light, "S_partitioned": light}
                     if heavy else None,
                     "kernel_ms": {k: round(v, 4) for k, v in mean_ph.items() if v > 0.0005},
                     "kernel_ms_from": f"{args.steps} further joins with HIP events around "
                                       f"every launch (device {ev_dev:.3f} ms per join with "
                                       f"the events)" if ev_dev else None},
        "checksum": str(out.checksum),
    }


def random_access_ceiling(aj, line, algo, nR, nS):
    """The hash-table joins are bound by random slot accesses, not streaming
    bytes: the same accesses timed alone (aj_access_kernel, a table of 2^28
    16-byte slots as the joins' at 2^27, 2^27 accesses), and the join's kernel
    time against that ceiling."""
    slots, ops = 1 << 28, 1 << 27
    tab = aj.DeviceArray(16 * slots)
    try:
        rate = {k: aj.access_rate(kind, tab.ptr, slots, ops)
                for k, kind in (("load", aj.AJ_ACCESS_LOAD), ("cas", aj.AJ_ACCESS_CAS),
                                ("load_cas", aj.AJ_ACCESS_LOAD_CAS))}
    finally:
        tab.free()
    out = {"table_bytes": 16 * slots, "accesses": ops,
           "G_per_s": {k: round(v, 2) for k, v in rate.items()}}
    km = line["pipeline"]["kernel_ms"]
    if algo == aj.AJ_ALGO_NPO and km.get("npo_probe"):
        # one random slot load per S tuple
        t = nS / (rate["load"] * 1e9) * 1e3
        out["probe"] = {"ceiling_ms": round(t, 4), "kernel_ms": km["npo_probe"],
                        "frac": round(t / km["npo_probe"], 3)}
    if algo == aj.AJ_ALGO_SHJ:
        # a probe load and an independent claim per tuple of either side
        t = (nR + nS) / (rate["load_cas"] * 1e9) * 1e3
        out["steps"] = {"ceiling_ms": round(t, 3), "join_ms": line["pipeline"]["device_ms"],
                        "frac": round(t / line["pipeline"]["device_ms"], 3)}
    return out


def calib_copy(aj):
    """One aj_copy_kernel over 1 GiB: the traffic counters' reading for a copy of
    known size, which scales the counters of the other kernels."""
    src, dst = aj.DeviceArray(1 << 30), aj.DeviceArray(1 << 30)
    aj.check(aj.lib.aj_copy_d2d(dst.ptr, src.ptr, 1 << 30))
    aj.device_sync()
    src.free()
    dst.free()


def run_pro(aj, args):
    n = 1 << args.log2n
    t0 = time.time()
    R = aj.gen_unique(n, 12345)
    S = aj.gen_uniform(n, 1, n, 54321)
    want = closed_form_fk(R, S)
    print(f"[bench] generated 2x{n} tuples in {time.time() - t0:.1f}s", file=sys.stderr,
          flush=True)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    aj.device_sync()
    if args.calib:
        calib_copy(aj)
    algo = {"PRO": aj.AJ_ALGO_PRO, "NPO": aj.AJ_ALGO_NPO, "SHJ": aj.AJ_ALGO_SHJ}[args.algo]
    kw = dict(radix_bits=args.radix_bits, passes=args.passes, cap=args.cap)
    if algo == aj.AJ_ALGO_SHJ:
        kw = dict(batch=args.batch)
    with aj.Context(0) as ctx:
        tuned = autotune(aj, ctx, dR, n, dS, n, algo, kw, args)
        out, dev, ms_step, phases, ev_dev = timed_joins(aj, ctx, dR, n, dS, n, algo, kw,
                                                        args.steps, args.warmup, want)
    dR.free()
    dS.free()
    name = (f"{args.algo} |R|=|S|=2^{args.log2n} 8-byte tuples, "
            f"unique R keys x uniform S keys")
    line = join_line(aj, args, name, n, n, algo, out, dev, ms_step, phases,
                     "join throughput M tuples/s (|R|+|S|)", ev_dev)
    if tuned:
        line["config"]["geometry"] = {k: tuned[k] for k in ("radix_bits", "passes", "cap")}
        line["config"]["geometry"]["chosen_by"] = "autotuner (aj_tune_pro)"
    # the traffic summary covers the default geometry (14 bits in 7 + 7)
    default_geom = algo == aj.AJ_ALGO_PRO and not (args.radix_bits or args.passes) and (
        not tuned or (tuned["radix_bits"], tuned["passes"]) == (14, 2))
    if default_geom or (algo != aj.AJ_ALGO_PRO and not args.batch):
        line["roofline"]["traffic"] = measured_traffic(line["roofline"]["kernel"], name)
    if algo == aj.AJ_ALGO_SHJ:
        line["metric"] = "eager join throughput M tuples/s (|R|+|S|)"
        line["config"]["batch"] = args.batch or max(65536, -(-n // 256))
    if algo != aj.AJ_ALGO_PRO:
        line["roofline"]["random_access"] = random_access_ceiling(aj, line, algo, n, n)
    if not args.no_cpu:
        host = host_info()
        if algo == aj.AJ_ALGO_SHJ:
            line["cpu_baseline"] = cpu_baseline_shj(aj, min(args.cpu_log2n, 23))
        else:
            line["cpu_baseline"] = cpu_baseline(aj, R, S, host)
            line["cpu_baseline_2e20_8threads"] = cpu_baseline_small(aj, host)
        line["host"] = host
        if algo == aj.AJ_ALGO_PRO:
            ref = cpu_reference_pro(aj, min(args.cpu_log2n, 25))
            if ref:
                line["cpu_reference"] = ref
    return line


def run_skew(aj, args):
    nR, nS = 1 << 24, 1 << 28
    R = aj.gen_unique(nR, 12345)
    S = aj.gen_zipf(nS, nR, args.theta, 54321)
    want = closed_form_fk(R, S)
    print(f"[bench] generated skewed relations (theta={args.theta})", file=sys.stderr,
          flush=True)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    aj.device_sync()
    del R, S
    algo = aj.AJ_ALGO_PRO if args.algo == "PRO" else aj.AJ_ALGO_NPO
    kw = dict(radix_bits=args.radix_bits, passes=args.passes, cap=args.cap)
    with aj.Context(0) as ctx:
        tuned = autotune(aj, ctx, dR, nR, dS, nS, algo, kw, args)
        out, dev, ms_step, phases, ev_dev = timed_joins(aj, ctx, dR, nR, dS, nS, algo, kw,
                                                        args.steps, args.warmup, want)
    name = (f"{args.algo} skewed foreign-key join |R|=2^24 unique x |S|=2^28 "
            f"Zipf(theta={args.theta})")
    line = join_line(aj, args, name, nR, nS, algo, out, dev, ms_step, phases,
                     "join throughput M tuples/s (|R|+|S|)", ev_dev)
    if tuned:
        line["config"]["geometry"] = {k: tuned[k] for k in ("radix_bits", "passes", "cap",
                                                            "part_block")}
        relabel_scatter(line, tuned)
    return line


def relabel_scatter(line, geom):
    """Passes of more than 10 radix bits run the non-pipelined k_scatter."""
    per_pass = -(-geom["radix_bits"] // max(geom["passes"], 1))
    if per_pass > 10 and line["roofline"]["kernel"].startswith("k_scatter_pipe"):
        line["roofline"]["kernel"] = line["roofline"]["kernel"].replace("k_scatter_pipe",
                                                                        "k_scatter")


def run_windows(aj, args):
    nwin = args.nwin
    rng = np.random.default_rng(2024)
    sizes = np.exp(rng.uniform(np.log(4096), np.log(4194304), nwin)).astype(np.uint32)
    cap = int(sizes.sum() * 1.001 + 8 * np.sqrt(sizes.sum()) + 4096)
    if args.calib:
        calib_copy(aj)
    with aj.Context(0) as ctx:
        streams = {}
        for side, seed in (("R", 101), ("S", 202)):
            d = aj.DeviceArray(12 * cap)
            b = aj.DeviceArray(4 * (nwin + 1))
            n = aj.gen_window_streams(ctx, seed, sizes, 1.6, 4, d.ptr, cap, b.ptr)
            streams[side] = (d, n, b)
        (dR, nR, bR), (dS, nS, _) = streams["R"], streams["S"]
        bounds = bR.to_host(np.uint32, nwin + 1)
        print(f"[bench] streams: {nR} + {nS} tuples, {nwin} windows", file=sys.stderr,
              flush=True)
        cnt, chk = aj.DeviceArray(8 * nwin), aj.DeviceArray(8 * nwin)
        geom = {}
        if not args.no_tune:  # the batch geometry chosen on these windows
            t0 = time.perf_counter()
            geom, tuned_ms = aj.tune_batch(ctx, dR.ptr, nR, dS.ptr, nS, bR.ptr, nwin,
                                           int(sizes.max() * 1.05), reps=2)
            print(f"[bench] batch tuned in {time.perf_counter() - t0:.1f}s: {geom} "
                  f"({tuned_ms:.2f} ms)", file=sys.stderr, flush=True)
        with aj.BatchPlan(ctx, nR, nS, nwin, int(sizes.max() * 1.05), graph=not args.no_graph,
                          **geom) as plan:
            def batch():
                plan.run(dR.ptr, nR, dS.ptr, nS, bR.ptr, nwin, cnt.ptr, chk.ptr)
                if args.no_graph:
                    aj.device_sync()
            for _ in range(args.warmup):
                batch()
            aj.device_sync()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                batch()
            aj.device_sync()
            t_batch = (time.perf_counter() - t0) / args.steps
        batch_res = np.stack([cnt.to_host(np.uint64, nwin), chk.to_host(np.uint64, nwin)])
        hR, hS = dR.to_host(aj.STREAM_DTYPE, nR), dS.to_host(aj.STREAM_DTYPE, nS)
        # each window's tuple range per side (the streams are in timestamp order)
        rng_r = np.searchsorted(hR["ts"], bounds, side="left")
        rng_s = np.searchsorted(hS["ts"], bounds, side="left")
        per_r, per_s = np.diff(rng_r), np.diff(rng_s)
        maxw = int(max(per_r.max(), per_s.max(), 1))
        # per-window launches: a plan sized to the largest window, one plain (not
        # batched, no graph) launch per window on that window's tuples, back to back
        # on one stream, one sync at the end
        cnt1, chk1 = aj.DeviceArray(8 * nwin), aj.DeviceArray(8 * nwin)
        with aj.BatchPlan(ctx, maxw, maxw, 1, maxw, graph=False) as one:
            def launch(w):
                one.run(dR.ptr + 12 * int(rng_r[w]), int(per_r[w]), dS.ptr + 12 * int(rng_s[w]),
                        int(per_s[w]), bR.ptr + 4 * w, 1, cnt1.ptr + 8 * w, chk1.ptr + 8 * w)
                if args.no_graph:
                    aj.device_sync()
            for w in range(min(nwin, 8)):  # warm
                launch(w)
            aj.device_sync()
            t0 = time.perf_counter()
            for w in range(nwin):
                launch(w)
            aj.device_sync()
            t_single = time.perf_counter() - t0
        single_res = np.stack([cnt1.to_host(np.uint64, nwin), chk1.to_host(np.uint64, nwin)])
        assert np.array_equal(batch_res, single_res), "batched != per-window results"
        in_r = int(rng_r[-1] - rng_r[0])
        in_s = int(rng_s[-1] - rng_s[0])
    tuples = in_r + in_s
    bpt = WINDOW_BYTES_PER_TUPLE
    gbs = bpt * tuples / t_batch / 1e9
    return {
        "metric": "batched window join throughput M tuples/s (|R|+|S| over all windows)",
        "value": round(tuples / t_batch / 1e6, 1),
        "unit": "M tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_batch * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{nwin} tumbling windows, two Poisson streams at 1.6 tuples/us, "
                               f"per-side sizes log-uniform [4K, 4M], duplicate factor 4, "
                               f"12-byte tuples; one hipGraph launch per step",
                   "windows": nwin, "tuples_in_windows": tuples,
                   "geometry": dict(geom, chosen_by="autotuner (aj_tune_batch)") if geom
                   else "plan defaults"},
        "per_window_launches": {"value": round(tuples / t_single / 1e6, 1),
                                "unit": "M tuples/s", "ms_total": round(t_single * 1e3, 2),
                                "speedup_of_batched": round(t_single / t_batch, 2),
                                "how": f"a plan sized to the largest window ({maxw} tuples), "
                                       f"one launch per window on its tuples, back to back, "
                                       f"one sync at the end"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole batched pipeline", "bytes_per_tuple": bpt},
        "parity": {"batched_equals_per_window": True},
    }


def run_serve(args):
    """Served windows through the C++ driver (ajbench -a SERVE).  First the two
    reference rates on 256 resident windows of the same distribution (plain
    one-window launches back to back, and the server's batched peak), then one
    run per load level, the load relative to that batched peak."""
    import subprocess
    exe = os.path.join(ROOT, "alliancedb_amd", "ajbench")

    def drive(extra, what):
        cmd = [exe, "-a", "SERVE", "--budget", "200e9"] + extra
        print(f"[bench] serve {what}", file=sys.stderr, flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise SystemExit(f"ajbench SERVE failed ({r.returncode}): {r.stderr[-2000:]}")
        return json.loads(r.stdout.strip().splitlines()[-1])

    cal = drive(["--calibrate", str(args.calib_windows)], "calibration")
    peak = args.peak or cal["batched_Mtps"]
    sweep = [drive(["--load", str(load), "--peak", str(peak), "--seconds", str(args.seconds),
                    "--verify", "64"], f"load={load}")
             for load in [float(x) for x in args.loads.split(",")]]
    at = min(sweep, key=lambda d: abs(d["load"] - 1.0))
    best = max(d["achieved_Mtps"] for d in sweep)
    return {
        "metric": "served window join throughput M tuples/s (|R|+|S|) at 100% offered load "
                  "(the batched peak of this window distribution, measured in the same run); "
                  "p99 window-to-result latency",
        "value": at["achieved_Mtps"],
        "unit": "M tuples/s",
        "n_gpus": 1,
        "steps": len(sweep),
        "warmup": 0,
        "ms_per_step": round(at["span_ms"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "64 stream pairs, Poisson window-join requests, per-side sizes "
                               "lognormal(median 256K, sigma 1) capped 16M, R keys uniform over "
                               "2^24, S keys uniform (streams 0-31) or Zipf 1.0 (32-63), 12-byte "
                               "tuples, HBM budget 200 GB, 3 batch slots; windows generated on "
                               "the GPU as they are admitted",
                   "peak_Mtps": round(peak, 1), "seconds_per_load": args.seconds},
        "reference_rates": {
            "windows": cal["windows"], "tuples": cal["tuples"],
            "plain_per_window_Mtps": cal["plain_Mtps"],
            "batched_peak_Mtps": cal["batched_Mtps"],
            "how": "the same windows resident in HBM: plain = a plan sized to the largest "
                   "window, one launch per window, back to back, one sync at the end; batched "
                   "peak = every window submitted to the server at once",
            "parity": {"results": cal["results"], "mismatches": cal["mismatches"]}},
        "served_vs_plain": {"at_100pct": round(at["achieved_Mtps"] / cal["plain_Mtps"], 2),
                            "best": round(best / cal["plain_Mtps"], 2),
                            "best_achieved_Mtps": best},
        "latency_ms": {"p50": at["p50_ms"], "p99": at["p99_ms"], "p999": at["p999_ms"]},
        "sweep": sweep,
        "parity": {"verified_windows": sum(d["verified"] for d in sweep),
                   "mismatches": sum(d["mismatches"] for d in sweep)},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(this path does not shard across GPUs)", file=sys.stderr)
        sys.exit(2)
    import alliancedb_amd as aj
    if args.no_staged:
        aj.set_option("staged", 0)
    if args.workload == "pro":
        line = run_pro(aj, args)
    elif args.workload == "skew":
        line = run_skew(aj, args)
    elif args.workload == "serve":
        line = run_serve(args)
    else:
        line = run_windows(aj, args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
