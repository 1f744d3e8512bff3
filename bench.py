"""Benchmark: GPU PRO join throughput on 1 MI355X.

Workload (default): |R| = |S| = 2^27 tuples of 8 bytes {payloadID, key};
R keys a seeded permutation of 1..2^27, S keys uniform on 1..2^27, payload = row
id.  A step is one PRO join of the device-resident relations (partition passes
+ LDS build/probe + count/checksum reduction).  Prints one JSON line.

Roofline: per-kernel device times come from HIP events recorded around every
launch on the join stream (library option kernel_timing) during the timed
steps; the dominant kernel's algorithmic bytes per launch (hist 8 B/tuple,
scatter 16 B/tuple, build/probe 8 B/tuple) over its mean duration gives
`achieved`.
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
NPO_BYTES_PER_TUPLE = 16  # R read 8 + slot write 8 / S read 8 + slot read 8
# algorithmic bytes per input tuple of each kernel
KERNEL_BYTES = {"hist1": 8, "scatter1": 16, "hist2": 8, "scatter2": 16, "join": 8,
                "npo_build": 16, "npo_probe": 16}
KERNEL_NAMES = {"hist1": "k_hist (pass 1)", "scatter1": "k_scatter (pass 1)",
                "hist2": "k_hist (pass 2)", "scatter2": "k_scatter (pass 2)",
                "join": "k_join", "npo_build": "k_npo_build", "npo_probe": "k_npo_probe"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--algo", choices=["PRO", "NPO"], default="PRO")
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--radix-bits", type=int, default=0)
    ap.add_argument("--passes", type=int, default=0)
    ap.add_argument("--cap", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-log2n", type=int, default=24)
    ap.add_argument("--calib", action="store_true",
                    help="also run aj_copy_kernel over 1 GiB (traffic-counter calibration)")
    return ap.parse_args()


def cpu_baseline(aj, log2n: int):
    """The CPU oracle's PRO (clean-room restatement of the reference's
    multithreaded PRO) on a bounded sample of the same workload."""
    import oracle
    n = 1 << log2n
    R = aj.gen_unique(n, 12345)
    S = aj.gen_uniform(n, 1, n, 54321)
    cores = min(os.cpu_count() or 1, 16)
    times = []
    for _ in range(3):
        r = oracle.pro(R, S, nthreads=cores, radix_bits=8, passes=2)
        assert r.count == n
        times.append(r.total_ms)
    ms = sorted(times)[len(times) // 2]
    return {"value": round(2 * n / (ms * 1e-3) / 1e6, 2), "unit": "M tuples/s",
            "cores": cores, "kind": "port",
            "sample": f"PRO_CPU |R|=|S|=2^{log2n}, same key distribution, 8 radix bits / "
                      f"2 passes (reference defaults), median of 3 (temporaries included)"}


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 traffic summary of
    this workload, or None when none was measured for it."""
    path = os.path.join(ROOT, "profiles", "pro_2e27_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except OSError:
        return None
    if d.get("workload") != workload:
        return None
    return d["hbm_bytes_per_launch"].get(kernel)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(this path does not shard across GPUs)", file=sys.stderr)
        sys.exit(2)
    import alliancedb_amd as aj
    n = 1 << args.log2n
    t0 = time.time()
    R = aj.gen_unique(n, 12345)
    S = aj.gen_uniform(n, 1, n, 54321)
    print(f"[bench] generated 2x{n} tuples in {time.time() - t0:.1f}s", file=sys.stderr,
          flush=True)
    dR, dS = aj.DeviceArray.from_host(R), aj.DeviceArray.from_host(S)
    aj.device_sync()
    del R, S
    if args.calib:
        dst = aj.DeviceArray(dR.nbytes)
        aj.check(aj.lib.aj_copy_d2d(dst.ptr, dR.ptr, dR.nbytes))
        dst.free()
    algo = aj.AJ_ALGO_PRO if args.algo == "PRO" else aj.AJ_ALGO_NPO
    ctx = aj.Context(0)
    kw = dict(radix_bits=args.radix_bits, passes=args.passes, cap=args.cap)
    for _ in range(args.warmup):
        out = ctx.join(dR.ptr, n, dS.ptr, n, algo, **kw)
    assert out.count == n, out.count
    aj.device_sync()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    aj.set_option("kernel_timing", 1)
    dev_ms, phases = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = ctx.join(dR.ptr, n, dS.ptr, n, algo, **kw)
        dev_ms.append(out.device_ms)
        phases.append(out.phases())
        assert out.count == n, out.count
    aj.device_sync()
    wall = time.perf_counter() - t0
    aj.set_option("kernel_timing", 0)
    ms_step = wall * 1e3 / args.steps
    dev = float(np.mean(dev_ms))
    tuples = 2 * n
    mean_ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    dom = max(KERNEL_BYTES, key=lambda k: mean_ph.get(k, 0.0))
    achieved = KERNEL_BYTES[dom] * tuples / (mean_ph[dom] * 1e-3) / 1e9
    bpt = PRO_BYTES_PER_TUPLE if algo == aj.AJ_ALGO_PRO else NPO_BYTES_PER_TUPLE
    pipeline_gbs = bpt * tuples / (dev * 1e-3) / 1e9
    line = {
        "metric": "join throughput M tuples/s (|R|+|S|)",
        "value": round(tuples / (dev * 1e-3) / 1e6, 1),
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
        "config": {"workload": f"{args.algo} |R|=|S|=2^{args.log2n} 8-byte tuples, "
                               f"unique R keys x uniform S keys",
                   "n_r": n, "n_s": n},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": KERNEL_NAMES[dom],
                     "bytes_per_tuple": KERNEL_BYTES[dom]},
        "pipeline": {"device_ms": round(dev, 3), "alg_bytes_per_tuple": bpt,
                     "alg_GBps": round(pipeline_gbs, 1),
                     "frac_of_peak": round(pipeline_gbs / HBM_PEAK_GBS, 4),
                     "kernel_ms": {k: round(v, 4) for k, v in mean_ph.items() if v > 0}},
        "checksum": str(out.checksum),
    }
    default_geom = not (args.radix_bits or args.passes or args.cap)
    if default_geom:
        line["roofline"]["traffic"] = measured_traffic(KERNEL_NAMES[dom], line["config"]["workload"])
    ctx.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(aj, args.cpu_log2n)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
