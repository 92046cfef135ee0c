"""Benchmark of the gfx950 radix sort: keys/s and HBM roofline on one GPU.

Headline workload: plain GPU radix sort of 2^28 u32 keys uniform over
[0, 2^32) (high halves of SplitMix64 with state 2), device-resident.  Each
timed step sorts its own copy of those keys, prepared in HBM before the timed
region; the K steps run back to back between two synchronizations.

Printed: one JSON line with the driver's fields plus ``roofline`` (the
one-sweep scatter kernel, timed with HIP events on its own stream) and
``cpu_baseline`` (the C++ restatement of turbosort::sort_parallel on this
host's cores, on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

WORKLOADS = {
    # name: (dtype, generator kind, seed, live digits, description)
    "u32": ("uint32", 0, 2, 4, "plain GPU radix, 2^28 u32 keys uniform over [0, 2^32) (SplitMix64 state 2)"),
    "u64": ("uint64", 1, 3, 8, "plain GPU radix, 2^28 u64 keys uniform over [0, 2^64) (SplitMix64 state 3)"),
    "u64z": ("uint64", 2, 4, 4, "plain GPU radix, 2^28 u64 keys = zero-extended u32 (SplitMix64 state 4)"),
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="u32", choices=sorted(WORKLOADS))
    p.add_argument("--log2n", type=int, default=28)
    p.add_argument("--also", default="u64,u64z", help="extra workloads reported beside the headline")
    p.add_argument("--cpu-seconds", type=float, default=8.0, help="wall budget of the CPU baseline")
    p.add_argument("--no-cpu-baseline", action="store_true")
    return p.parse_args()


def run_workload(name: str, log2n: int, steps: int, warmup: int, ctx):
    import torch

    import turbosort_amd as ts

    dt, kind, seed, live, desc = WORKLOADS[name]
    tdt = getattr(torch, dt)
    n = 1 << log2n
    w = 4 if dt == "uint32" else 8
    dev = torch.device("cuda", 0)
    src = torch.empty(n, dtype=tdt, device=dev)
    if kind == 2:
        ts.fill_splitmix64_zero_extended(src, seed, ctx=ctx)
    else:
        ts.fill_splitmix64(src, seed, ctx=ctx)
    bufs = [torch.empty_like(src) for _ in range(steps + warmup)]
    for b in bufs:
        b.copy_(src)
    ctx.reserve(tdt, n)
    before = ts.checksum(src, ctx=ctx)

    stream = torch.cuda.current_stream(dev)
    for i in range(warmup):
        ts.sort(bufs[i], ctx=ctx)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        ev[i][0].record(stream)
        ts.sort(bufs[warmup + i], ctx=ctx)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = sorted(a.elapsed_time(b) for a, b in ev)

    # The result must be the sorted permutation of the input.
    after = ts.checksum(bufs[-1], ctx=ctx)
    ok = after[0] == before[0] and after[1] == before[1] and after[2] == 0

    # Per-kernel durations from HIP events on the sort's stream.
    ctx.set_profiling(True)
    kt = []
    for i in range(3):
        bufs[0].copy_(src)
        torch.cuda.synchronize()
        ts.sort(bufs[0], ctx=ctx)
        torch.cuda.synchronize()
        kt.append(ctx.kernel_times())
    ctx.set_profiling(False)
    sweep = [t for run in kt for (nm, t, _b) in run if nm == "k_sweep"]
    sweep.sort(reverse=True)
    sweep_live = sweep[: live * len(kt)]
    hist = [t for run in kt for (nm, t, _b) in run if nm == "k_hist"]
    sweep_ms = sum(sweep_live) / len(sweep_live)
    hist_ms = sum(hist) / len(hist)
    sweep_bytes = 2.0 * w * n  # one read and one write of every key
    algo_bytes = w * (1 + 2 * live) * n  # canonical LSD: one histogram read + P read/write passes

    del bufs
    torch.cuda.empty_cache()
    ms_step = wall * 1e3 / steps
    return dict(
        name=name, desc=desc, n=n, w=w, live=live, ok=ok, ms_per_step=ms_step,
        event_ms_median=per[len(per) // 2], event_ms_best=per[0],
        keys_per_s=n / (wall / steps), sweep_ms=sweep_ms, hist_ms=hist_ms,
        sweep_gbs=sweep_bytes / (sweep_ms * 1e-3) / 1e9, hist_gbs=w * n / (hist_ms * 1e-3) / 1e9,
        algo_bytes_per_key=w * (1 + 2 * live), total_gbs=algo_bytes / (wall / steps) / 1e9,
    )


def pmc_traffic(workload: str, log2n: int):
    """HBM bytes per k_sweep launch from the committed rocprofv3 counter runs.

    ``profiles/k_sweep_<workload>_2p<log2n>.md`` holds FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by what the same counters read for a copy
    of known size, summed into a ``traffic_bytes_per_launch:`` line.  None when
    no such summary exists for this workload and size.
    """
    path = os.path.join(HERE, "profiles", f"k_sweep_{workload}_2p{log2n}.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(budget_s: float):
    """The C++ restatement of turbosort::sort_parallel on this host's cores."""
    import numpy as np

    from oracle import closed_form as cf
    from oracle import native as nv

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(cores, 16))
    n = 1 << 25
    src = cf.u32_keys(2, n)
    work = src.copy()
    nv.sort_parallel_inplace(work, threads)  # warm-up
    times = []
    t_end = time.perf_counter() + budget_s
    while time.perf_counter() < t_end or len(times) < 3:
        np.copyto(work, src)
        t0 = time.perf_counter()
        nv.sort_parallel_inplace(work, threads)
        times.append(time.perf_counter() - t0)
        if len(times) >= 50:
            break
    times.sort()
    med = times[len(times) // 2]
    return dict(value=n / med, unit="keys/s", cores=threads, kind="port",
                sample=f"2^25 u32 keys (SplitMix64 state 2), median of {len(times)} sorts by the C++ restatement "
                       f"of turbosort::sort_parallel with {threads} threads")


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); every figure is per single MI355X.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    import turbosort_amd as ts

    ctx = ts.Context(0)
    head = run_workload(a.workload, a.log2n, a.steps, a.warmup, ctx)
    also = {}
    for name in [x for x in a.also.split(",") if x and x != a.workload]:
        r = run_workload(name, a.log2n, max(3, a.steps // 3), 1, ctx)
        also[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items() if k != "desc"}
        also[name]["roofline_frac"] = round(r["total_gbs"] / HBM_PEAK_GBS, 4)
    cpu = None if a.no_cpu_baseline else cpu_baseline(a.cpu_seconds)

    sweep_gbs = head["sweep_gbs"]
    line = {
        "metric": "keys/s & % HBM roofline, 256M u32/u64; learned/segmented vs plain GPU radix",
        "value": head["keys_per_s"],
        "unit": "keys/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32" if a.workload == "u32" else "u64",
        "data": "synthetic: SplitMix64 keys generated on the device",
        "config": {"workload": head["desc"], "n": head["n"], "key_bytes": head["w"],
                   "algorithmic_bytes_per_key": head["algo_bytes_per_key"], "device_resident": True},
        "roofline": {"bound": "hbm", "achieved": sweep_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": sweep_gbs / HBM_PEAK_GBS, "traffic": pmc_traffic(a.workload, a.log2n),
                     "kernel": "k_sweep (one-sweep scatter pass; 2*w bytes per key per launch)",
                     "kernel_ms": head["sweep_ms"]},
        "whole_sort": {"algorithmic_GBps": head["total_gbs"], "frac": head["total_gbs"] / HBM_PEAK_GBS,
                       "event_ms_median": head["event_ms_median"], "hist_ms": head["hist_ms"],
                       "verified": head["ok"]},
        "cpu_baseline": cpu,
        "also": also,
    }
    print(json.dumps(line), flush=True)
    if not head["ok"] or any(not v["ok"] for v in also.values()):
        sys.exit(1)


if __name__ == "__main__":
    main()
