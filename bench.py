"""Benchmark of the gfx950 radix sort: keys/s and HBM roofline on one GPU.

Headline workload: plain GPU radix sort of 2^28 u32 keys uniform over
[0, 2^32) (high halves of SplitMix64 with state 2), device-resident.  Each
timed step sorts its own copy of those keys, prepared in HBM before the timed
region; the K steps run back to back between two synchronizations.

Printed: one JSON line with the driver's fields plus ``roofline`` (the
one-sweep scatter kernel, timed with HIP events on its own stream) and
``cpu_baseline`` (the C++ restatement of turbosort::sort_parallel on this
host's CPU share, on a bounded sample).  ``also`` holds every other workload
the same way: device-resident rate, rate with the copies from and to pinned
host memory, roofline fractions, and a CPU baseline on a sample of the same
input.
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
    p.add_argument("--also", default="u64,u64z,segmented,by_key,learned",
                   help="extra workloads reported beside the headline (u32, u64, u64z, segmented, by_key, learned)")
    p.add_argument("--cpu-seconds", type=float, default=2.5,
                   help="wall budget of each CPU baseline (one per workload)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-host", action="store_true", help="skip the rates that include host copies")
    p.add_argument("--learned-n", type=int, default=1_000_000_000)
    p.add_argument("--no-tune", action="store_true",
                   help="segmented: keep the context's default kernel shapes instead of autotuning them first")
    return p.parse_args()


def median(xs):
    xs = sorted(xs)
    return xs[len(xs) // 2]


# ------------------------------------------------------------ CPU baseline
def cpu_share():
    """Threads the CPU baselines use: the CPU share this process is given
    (OMP_NUM_THREADS where the host sets it, else the CPUs it may run on),
    with the host's CPU count and model for the record."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else affinity
    threads = max(1, min(threads, affinity))
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(threads=threads, nproc=os.cpu_count(), affinity=affinity, model=model)


def cpu_timed(prepare, run, budget_s: float, min_runs: int = 3, max_runs: int = 40):
    """Median wall time of run() after prepare() (untimed), repeated within a budget."""
    prepare()
    run()  # warm-up
    times = []
    t_end = time.perf_counter() + budget_s
    while (time.perf_counter() < t_end or len(times) < min_runs) and len(times) < max_runs:
        prepare()
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    return median(times), len(times)


def cpu_sort_baseline(src, threads: int, budget_s: float, what: str):
    """The C++ restatement of turbosort::sort_parallel on a copy of src."""
    import numpy as np

    from oracle import native as nv

    work = np.empty_like(src)
    t, runs = cpu_timed(lambda: np.copyto(work, src), lambda: nv.sort_parallel_inplace(work, threads), budget_s)
    return dict(value=src.size / t, unit="keys/s", cores=threads, kind="port",
                sample=f"{what}: {src.size} keys, median of {runs} sorts by the C++ restatement of "
                       f"turbosort::sort_parallel (parallel.rs:46-291) with {threads} threads")


# ------------------------------------------------------------ plain sorts
def run_workload(name: str, log2n: int, steps: int, warmup: int, ctx, host: bool = True, cpu=None):
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
    info = ctx.last_sort_info()

    # The result must be the sorted permutation of the input.
    after = ts.checksum(bufs[-1], ctx=ctx)
    ok = after[0] == before[0] and after[1] == before[1] and after[2] == 0

    # Per-kernel durations from HIP events on the sort's stream; the executed
    # scatter passes are the first ones launched.
    ctx.set_profiling(True)
    kt = []
    for i in range(3):
        bufs[0].copy_(src)
        torch.cuda.synchronize()
        ts.sort(bufs[0], ctx=ctx)
        torch.cuda.synchronize()
        kt.append(ctx.kernel_times())
    ctx.set_profiling(False)
    n_exec = info["executed_passes"]
    sweep_live = [t for run in kt for (_nm, t, _b) in [x for x in run if x[0] == "k_sweep"][:n_exec]]
    hist = [t for run in kt for (nm, t, _b) in run if nm == "k_hist"]
    finish = [sum(t for (nm, t, _b) in run if nm.startswith("k_finish")) for run in kt]
    sweep_ms = sum(sweep_live) / len(sweep_live)
    hist_ms = sum(hist) / len(hist)
    sweep_bytes = 2.0 * w * n  # one read and one write of every key
    algo_bytes = w * (1 + 2 * live) * n  # canonical LSD: one histogram read + P read/write passes
    # What this sort reads and writes as executed: the histogram read, the
    # executed passes, and after a divert the finish's read of every key.
    exec_bytes = w * (1 + 2 * n_exec + (1 if info["skipped_digits"] and not info["fallback"] else 0)) * n
    del bufs
    torch.cuda.empty_cache()
    ms_step = wall * 1e3 / steps
    out = dict(
        name=name, desc=desc, n=n, w=w, live=live, ok=ok, ms_per_step=ms_step,
        event_ms_median=per[len(per) // 2], event_ms_best=per[0],
        keys_per_s=n / (wall / steps), sweep_ms=sweep_ms, hist_ms=hist_ms,
        finish_ms=median(finish) if info["skipped_digits"] else 0.0,
        sweep_gbs=sweep_bytes / (sweep_ms * 1e-3) / 1e9, hist_gbs=w * n / (hist_ms * 1e-3) / 1e9,
        algo_bytes_per_key=w * (1 + 2 * live), total_gbs=algo_bytes / (wall / steps) / 1e9,
        roofline_frac=algo_bytes / (wall / steps) / 1e9 / HBM_PEAK_GBS,
        executed=dict(passes=n_exec, skipped_digits=info["skipped_digits"], second_sort=bool(info["fallback"]),
                      bytes_per_key=exec_bytes / n,
                      roofline_frac=exec_bytes / (wall / steps) / 1e9 / HBM_PEAK_GBS),
    )
    if host:
        out["with_copies"] = run_host(name, log2n, max(3, steps // 3), ctx)
    if cpu is not None:
        import numpy as np

        # The same keys, copied to the host: the whole workload, not a sample.
        signed = torch.int32 if w == 4 else torch.int64
        keys = src.view(signed).cpu().numpy().view(np.uint32 if w == 4 else np.uint64)
        out["cpu_baseline"] = cpu_sort_baseline(keys, cpu["threads"], cpu["budget"],
                                                f"the whole workload ({desc.split(', ', 1)[1]})")
        del keys
    del src
    return out


def run_host(name: str, log2n: int, steps: int, ctx):
    """sort_host on the workload's keys in pinned host memory: copy to the
    device, sort, copy back; the end-to-end rate for a caller whose slice
    lives in host memory (wall time, median)."""
    import torch

    import turbosort_amd as ts

    dt, kind, seed, _, _ = WORKLOADS[name]
    tdt = getattr(torch, dt)
    n = 1 << log2n
    dev = torch.device("cuda", 0)
    d = torch.empty(n, dtype=tdt, device=dev)
    if kind == 2:
        ts.fill_splitmix64_zero_extended(d, seed, ctx=ctx)
    else:
        ts.fill_splitmix64(d, seed, ctx=ctx)
    src = torch.empty(n, dtype=tdt, pin_memory=True)
    src.copy_(d)
    before = ts.checksum(d, ctx=ctx)
    h = torch.empty_like(src).pin_memory()
    h.copy_(src)
    ts.sort_host(h, ctx=ctx)  # warm: the context's device buffer grows once
    times = []
    for _ in range(steps):
        h.copy_(src)
        t0 = time.perf_counter()
        ts.sort_host(h, ctx=ctx)
        times.append(time.perf_counter() - t0)
    d.copy_(h)
    after = ts.checksum(d, ctx=ctx)
    ok = after[0] == before[0] and after[1] == before[1] and after[2] == 0
    t = median(times)
    w = d.element_size()
    del d, src, h
    torch.cuda.empty_cache()
    return dict(desc="sort_host: the keys in pinned host memory, copied to the device, sorted, copied back "
                     "(wall time, median)", ms_per_step=t * 1e3, keys_per_s=n / t,
                link_GBps_both_ways=2 * w * n / t / 1e9, ok=ok)


def host_roundtrip(pairs, fn, steps: int):
    """Median wall time of: copy each pinned host tensor to its device tensor,
    fn(), copy every device tensor back, wait.  pairs: (host, device)."""
    import torch

    times = []
    for _ in range(steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for h, d in pairs:
            d.copy_(h, non_blocking=True)
        fn()
        for h, d in pairs:
            h.copy_(d, non_blocking=True)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    return median(times[1:])


# ------------------------------------------------------------ generators
# SplitMix64 (Vigna's splitmix64.c): output i of the stream with state s is
# mix(s + (i + 1) * 0x9E3779B97F4A7C15); the device fill kernel computes the same.
def np_splitmix64(seed: int, n: int):
    import numpy as np

    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def np_unit(bits):
    import numpy as np

    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def segment_offsets(seed: int, n_segments: int):
    """Offsets of n_segments log-uniform lengths clamp(floor(exp(U(ln 2, ln 16385))), 2, 16384)."""
    import numpy as np

    u = np_unit(np_splitmix64(seed, n_segments))
    a, b = np.log(2.0), np.log(16385.0)
    lens = np.clip(np.floor(np.exp(a + u * (b - a))), 2, 16384).astype(np.int64)
    off = np.zeros(n_segments + 1, dtype=np.int64)
    off[1:] = np.cumsum(lens)
    return off


def score_records(n: int, seed: int):
    """n 16-byte records {u32 id = i, f32 score, u64 payload}.

    score = f32 of a standard normal (Box-Muller on two SplitMix64 streams,
    states seed and seed + 2^32); 1 % of the records (stream seed + 2^33
    below 0.01) get one of {+NaN, -NaN, +inf, -inf, -0.0} in equal shares;
    then 20 % (stream seed + 3*2^32 below 0.2) copy the score of the record
    whose index is stream seed + 2^34 modulo n; payload = stream seed + 5*2^32.
    """
    import numpy as np

    s = lambda k: seed + k * (1 << 32)  # noqa: E731
    u1 = np_unit(np_splitmix64(s(0), n))
    u2 = np_unit(np_splitmix64(s(1), n))
    score = (np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)).astype(np.float32)
    del u1, u2
    bits = score.view(np.uint32)
    r = np_unit(np_splitmix64(s(2), n))
    sp = r < 0.01
    specials = np.array([0x7FC00000, 0xFFC00000, 0x7F800000, 0xFF800000, 0x80000000], dtype=np.uint32)
    bits[sp] = specials[np.minimum((r[sp] * 500).astype(np.int64), 4)]
    dup = np_unit(np_splitmix64(s(3), n)) < 0.2
    src = (np_splitmix64(s(4), n) % np.uint64(n)).astype(np.int64)
    bits[dup] = bits[src[dup]]
    rec = np.zeros(n, dtype=np.dtype([("id", "<u4"), ("score", "<u4"), ("payload", "<u8")]))
    rec["id"] = np.arange(n, dtype=np.uint32)
    rec["score"] = bits
    rec["payload"] = np_splitmix64(s(5), n)
    return rec


def take(t, idx, chunk: int = 1 << 24):
    """t[idx] in chunks of 2^24 indices (torch's single-call gather has returned
    zeros past 2^30 output bytes on this stack)."""
    import torch

    return torch.cat([t[idx[c:c + chunk]] for c in range(0, idx.numel(), chunk)])


def timed(fn, steps: int):
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        fn(i)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def run_segmented(steps: int, ctx, n_segments: int = 100_000, seed: int = 7, host: bool = True, cpu=None,
                  tune: bool = True):
    """Batched segmented key-value sort: 10^5 log-uniform segments in [2, 16384],
    u32 keys uniform (SplitMix64 state seed + 2^32), u32 values = global index."""
    import torch

    import turbosort_amd as ts
    from turbosort_amd import autotune

    dev = torch.device("cuda", 0)
    # Kernel shape per size class, picked on this GPU before the timed runs
    # (or the context's defaults: the same kernels in every run, as the
    # counter passes of a profile need).
    if tune:
        tuning = autotune.tune_segmented(torch.uint32, True, ctx=ctx)
        chosen = {str(c): tuning["names"][c][v] for c, v in tuning["choice"].items()}
    else:
        chosen = {str(c): autotune.variants(torch.uint32, True, c)[autotune.get_variant(ctx, torch.uint32, True, c)]
                  for c in range(len(autotune.CLASS_CAPS))}
    off_np = segment_offsets(seed, n_segments)
    off = torch.from_numpy(off_np).to(dev)
    n = int(off[-1])
    src = torch.empty(n, dtype=torch.uint32, device=dev)
    ts.fill_splitmix64(src, seed + (1 << 32), ctx=ctx)
    warm = 1
    ks = [src.clone() for _ in range(steps + warm)]
    vs = [torch.arange(n, dtype=torch.int32, device=dev) for _ in range(steps + warm)]
    run = lambda i: ts.segmented_sort(ks[i], off, vs[i], max_segment_len=16384, ctx=ctx)  # noqa: E731
    for i in range(warm):
        run(i)
    dt = timed(lambda i: run(warm + i), steps)
    # Correct iff values stay in their segments as a permutation, each key is
    # the input key at its value, keys do not decrease and ties keep order.
    k, v = ks[-1].view(torch.int32), vs[-1].long()
    seg = torch.repeat_interleave(torch.arange(n_segments, device=dev), off[1:] - off[:-1])
    seen = torch.zeros(n, dtype=torch.bool, device=dev)
    seen[v] = True
    kk = k.long() & 0xFFFFFFFF
    same = seg[1:] == seg[:-1]
    ok = (bool(torch.equal(take(seg, v), seg)) and bool(seen.all()) and bool(torch.equal(take(src.view(torch.int32), v), k))
          and not bool((same & (kk[1:] < kk[:-1])).any())
          and not bool((same & (kk[1:] == kk[:-1]) & (v[1:] <= v[:-1])).any()))
    del ks, vs, seg, seen, kk, same
    # The same sort captured once in a graph and replayed on restored input.
    gk, gv = src.clone(), torch.arange(n, dtype=torch.int32, device=dev)
    graph = torch.cuda.CUDAGraph()
    cs = torch.cuda.Stream()
    cs.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cs):
        with torch.cuda.graph(graph, stream=cs):
            ts.segmented_sort(gk, off, gv, max_segment_len=16384, ctx=ctx)
    torch.cuda.current_stream().wait_stream(cs)
    times = []
    for _ in range(steps):
        gk.copy_(src)
        gv.copy_(torch.arange(n, dtype=torch.int32, device=dev))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        graph.replay()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    times.sort()
    del graph
    out = dict(desc="batched segmented key-value sort, 10^5 log-uniform segments in [2, 16384] (SplitMix64 "
                    "state 7), u32 keys uniform, u32 values = global index", n=n, segments=n_segments,
               ms_per_step=dt * 1e3, pairs_per_s=n / dt, algo_bytes_per_pair=16,
               algorithmic_GBps=16 * n / dt / 1e9, roofline_frac=16 * n / dt / 1e9 / HBM_PEAK_GBS,
               graph_replay_ms_median=times[len(times) // 2],
               **{"autotuned" if tune else "default_shapes": chosen}, ok=ok)
    if host:
        hk = torch.empty(n, dtype=torch.uint32, pin_memory=True)
        hv = torch.empty(n, dtype=torch.int32, pin_memory=True)
        hk.copy_(src)
        hv.copy_(torch.arange(n, dtype=torch.int32, device=dev))
        t = host_roundtrip([(hk, gk), (hv, gv)],
                           lambda: ts.segmented_sort(gk, off, gv, max_segment_len=16384, ctx=ctx), max(3, steps // 2))
        out["with_copies"] = dict(desc="keys and values in pinned host memory: copied in, sorted, copied out "
                                       "(offsets resident; wall time, median)", ms_per_step=t * 1e3,
                                  pairs_per_s=n / t)
        del hk, hv
    del gk, gv
    torch.cuda.empty_cache()
    if cpu is not None:
        import numpy as np

        from oracle import native as nv

        m = 10_000  # the first 10^4 segments of the same workload
        sub = off_np[: m + 1].astype(np.uint64)
        keys0 = src[: int(sub[-1])].view(torch.int32).cpu().numpy().view(np.uint32)
        kw, vw = np.empty_like(keys0), np.empty(keys0.size, dtype=np.uint32)

        def prep():
            np.copyto(kw, keys0)
            vw[:] = np.arange(keys0.size, dtype=np.uint32)

        t, runs = cpu_timed(prep, lambda: nv.segmented_sort_pairs_inplace(kw, sub, vw, cpu["threads"]), cpu["budget"])
        out["cpu_baseline"] = dict(value=keys0.size / t, unit="pairs/s", cores=cpu["threads"], kind="port",
                                   sample=f"the first {m} segments ({keys0.size} pairs), median of {runs} runs of the "
                                          "restated per-segment sort_by_key_with_buffer + move (by_key.rs:175-235, "
                                          f"315-334) with {cpu['threads']} threads taking segments in turn")
    return out


def run_by_key(steps: int, ctx, n: int = 100_000_000, seed: int = 5, host: bool = True, cpu=None):
    """Stable sort_by_key of 10^8 16-byte records keyed on an f32 score at
    offset 4, then apply_order on a second u64 slice (SplitMix64 state 6)."""
    import numpy as np
    import torch

    import turbosort_amd as ts

    dev = torch.device("cuda", 0)
    rec = score_records(n, seed)
    src = torch.from_numpy(rec.view(np.uint8).reshape(n, 16)).to(dev)
    scores_sample = rec["score"][: 10_000_000].copy().view(np.float32) if cpu is not None else None
    del rec
    other_src = torch.empty(n, dtype=torch.uint64, device=dev)
    ts.fill_splitmix64(other_src, 6, ctx=ctx)
    buf = ts.SortByKeyBuffer.with_capacity(n, device=dev, key_dtype=torch.float32, record_bytes=16)
    key = ts.KeyField(4, torch.float32)
    warm = 1
    rs = [src.clone() for _ in range(steps + warm)]
    for i in range(warm):
        ts.sort_by_key_with_buffer(rs[i], buf, key)
    dt = timed(lambda i: ts.sort_by_key_with_buffer(rs[warm + i], buf, key), steps)
    passes = int(buf._ctx.last_sort_info()["executed_passes"])
    # The 16-byte records move through the passes with their u32 positions:
    # the histogram reads each record once (16 B), the first pass reads it and
    # writes record + position (36 B), every later pass reads and writes both
    # (40 B; the last writes the position into order()).
    executed = 16 + (36 + 40 * (passes - 1) if passes else 0)
    others = [other_src.clone() for _ in range(steps)]
    dt_apply = timed(lambda i: buf.apply_order(others[i]), steps)
    # The same order applied to a copy of the 16-byte records: the gather a
    # by-key sort of records that are not carried would run.
    recs = [src.clone() for _ in range(steps)]
    dt_apply16 = timed(lambda i: buf.apply_order(recs[i]), steps)
    del recs
    # order() is the stable argsort iff it is a permutation along which the
    # scores' radix keys do not decrease and equal keys keep increasing order.
    order = buf.order().view(torch.int32).long()
    sbits = src[:, 4:8].contiguous().view(torch.int32).reshape(-1).long() & 0xFFFFFFFF
    radix = torch.where(sbits >= 0x80000000, sbits ^ 0xFFFFFFFF, sbits | 0x80000000)
    rk = take(radix, order)
    seen = torch.zeros(n, dtype=torch.bool, device=dev)
    seen[order] = True
    ok = (bool(seen.all()) and not bool((rk[1:] < rk[:-1]).any())
          and not bool(((rk[1:] == rk[:-1]) & (order[1:] <= order[:-1])).any())
          and bool(torch.equal(rs[-1], take(src, order)))
          and bool(torch.equal(others[-1].view(torch.int64), take(other_src.view(torch.int64), order))))
    del rs, others, order, sbits, radix, rk, seen
    out = dict(desc="stable sort_by_key, 10^8 16-byte records {u32 id, f32 score, u64 payload} keyed on the "
                    "score (N(0,1), 1 % NaN/inf/-0.0, 20 % duplicated), records carried through the key passes; "
                    "apply_order on u64[10^8] and on a copy of the records",
               n=n, ms_per_step=dt * 1e3, records_per_s=n / dt, algo_bytes_per_record=120,
               algorithmic_GBps=120 * n / dt / 1e9, roofline_frac=120 * n / dt / 1e9 / HBM_PEAK_GBS,
               executed=dict(passes=passes, bytes_per_record=executed,
                             roofline_frac=executed * n / dt / 1e9 / HBM_PEAK_GBS),
               apply_order_ms=dt_apply * 1e3, apply_order_GBps=20 * n / dt_apply / 1e9,
               apply_order_roofline_frac=20 * n / dt_apply / 1e9 / HBM_PEAK_GBS,
               apply_order_records_ms=dt_apply16 * 1e3,
               apply_order_records_roofline_frac=36 * n / dt_apply16 / 1e9 / HBM_PEAK_GBS, ok=ok)
    if host:
        hr = torch.empty((n, 16), dtype=torch.uint8, pin_memory=True)
        hr.copy_(src)
        d = src.clone()
        ho = torch.empty(n, dtype=torch.int32, pin_memory=True)

        def fn():
            ts.sort_by_key_with_buffer(d, buf, key)
            ho.copy_(buf.order().view(torch.int32), non_blocking=True)

        t = host_roundtrip([(hr, d)], fn, max(3, steps))
        out["with_copies"] = dict(desc="records in pinned host memory: copied in, sorted, records and order() "
                                       "copied out (wall time, median)", ms_per_step=t * 1e3,
                                  records_per_s=n / t)
        del hr, ho, d
    del src, other_src
    torch.cuda.empty_cache()
    if cpu is not None:
        from oracle import native as nv

        m = scores_sample.size
        order_out = np.empty(m, dtype=np.uint32)
        recs = np.zeros((m, 16), dtype=np.uint8)
        t, runs = cpu_timed(lambda: None, lambda: (nv.lib().tso_sort_by_key_order(
            8, scores_sample.ctypes.data, m, order_out.ctypes.data), nv.lib().tso_apply_order(
            order_out.ctypes.data, m, recs.ctypes.data, m, 16)), cpu["budget"])
        out["cpu_baseline"] = dict(value=m / t, unit="records/s", cores=1, kind="port",
                                   sample=f"the first {m} scores of the same records: median of {runs} runs of the "
                                          "restated order() of sort_by_key_with_buffer and move_into_order "
                                          "(by_key.rs:175-235, 315-334; the reference's by-key path is serial)")
    return out


GAMMA = 0x9E3779B97F4A7C15


def dev_unit(n: int, seed: int, start: int, ctx):
    """Uniform doubles in [0, 1) from SplitMix64 outputs start .. start+n-1 of
    the stream with state seed (generated on the device)."""
    import torch

    import turbosort_amd as ts

    bits = torch.empty(n, dtype=torch.uint64, device="cuda")
    ts.fill_splitmix64(bits, (seed + start * GAMMA) % (1 << 64), ctx=ctx)
    b = bits.view(torch.int64)
    return ((b >> 11) & ((1 << 53) - 1)).to(torch.float64) * (1.0 / 9007199254740992.0)


def lognormal_keys_dev(n: int, seed: int, ctx, as_f64: bool = False, hot_share: float = 0.05, n_hot: int = 10_000,
                       chunk: int = 1 << 27):
    """The learned-partitioner workload on the device: a mixture of lognormals
    (mu 0, sigma 0.5 / 1 / 2, weights 0.5 / 0.3 / 0.2) with 5 % of the
    positions replaced by Zipf(1.1) draws over 10^4 hot values from the same
    mixture; u64 key = floor(min(x / e^12, 1 - 2^-53) * 2^64), or x as f64.
    Streams: SplitMix64 states seed + k * 2^32 (the formulas of the oracle's
    generator, evaluated in float64 on the device)."""
    import math

    import torch

    st = lambda k: seed + k * (1 << 32)  # noqa: E731

    def mixture(m, start, k0):
        u1 = dev_unit(m, st(k0), start, ctx)
        u2 = dev_unit(m, st(k0 + 1), start, ctx)
        pick = dev_unit(m, st(k0 + 2), start, ctx)
        z = torch.sqrt(-2.0 * torch.log1p(-u1)) * torch.cos(2.0 * math.pi * u2)
        sigma = torch.where(pick < 0.5, 0.5, torch.where(pick < 0.8, 1.0, 2.0))
        return torch.exp(sigma * z)

    hot = mixture(n_hot, 0, 3)
    w = torch.arange(1, n_hot + 1, dtype=torch.float64, device="cuda") ** -1.1
    cdf = torch.cumsum(w, 0) / w.sum()
    out = torch.empty(n, dtype=torch.float64 if as_f64 else torch.uint64, device="cuda")
    for c0 in range(0, n, chunk):
        m = min(chunk, n - c0)
        x = mixture(m, c0, 0)
        rank = torch.clamp(torch.searchsorted(cdf, dev_unit(m, st(6), c0, ctx), right=True), max=n_hot - 1)
        x = torch.where(dev_unit(m, st(7), c0, ctx) < hot_share, hot[rank], x)
        if as_f64:
            out[c0:c0 + m] = x
        else:
            v = torch.clamp(x / math.exp(12.0), max=1.0 - 2.0 ** -53)
            hi = torch.floor(v * 4294967296.0)
            lo = torch.floor((v * 4294967296.0 - hi) * 4294967296.0)
            out.view(torch.int64)[c0:c0 + m] = (hi.to(torch.int64) << 32) | lo.to(torch.int64)
        del x
    return out


def equal_chunked(a, b, chunk: int = 1 << 26) -> bool:
    import torch

    a, b = a.view(torch.int64) if a.element_size() == 8 else a, b.view(torch.int64) if b.element_size() == 8 else b
    return all(bool(torch.equal(a[i:i + chunk], b[i:i + chunk])) for i in range(0, a.numel(), chunk))


def run_learned(steps: int, ctx, n: int = 1_000_000_000, seed: int = 8, host: bool = True, cpu=None):
    """Learned partitioner against the plain device radix on 10^9 skewed u64
    keys (and the same x as f64), with the model fitted on a 1 % sample (the
    workload's) and the output judged before the call returns; the 0.5 %
    sample and the untrained model (the ablation) beside it."""
    import torch

    import turbosort_amd as ts

    res = {}
    for name, as_f64 in (("u64", False), ("f64", True)):
        src = lognormal_keys_dev(n, seed, ctx, as_f64=as_f64)
        torch.cuda.synchronize()
        ctx.reserve(src.dtype, n)
        warm = 1

        def timed_copies(fn):
            bufs = [src.clone() for _ in range(steps + warm)]
            for i in range(warm):
                fn(bufs[i])
            dt = timed(lambda i: fn(bufs[warm + i]), steps)
            out = bufs[-1]
            del bufs[:-1]
            return dt, out

        dt_plain, ref = timed_copies(lambda b: ts.sort(b, ctx=ctx))
        plain_info = ctx.last_sort_info()
        cfg = ts.LearnedConfig(sample_per_mille=10)
        rep = {}
        dt_learn, out = timed_copies(lambda b: rep.update(ts.learned_sort(b, config=cfg, ctx=ctx)))
        ok = equal_chunked(out, ref) and rep["verified"] == 1
        del out
        rep_h = {}
        dt_half, out = timed_copies(lambda b: rep_h.update(ts.learned_sort(b, config=ts.LearnedConfig(sample_per_mille=5),
                                                                           ctx=ctx)))
        ok = ok and equal_chunked(out, ref)
        del out
        rep_nv = {}
        dt_nv, out = timed_copies(lambda b: rep_nv.update(ts.learned_sort(
            b, config=ts.LearnedConfig(sample_per_mille=10, verify=False), ctx=ctx)))
        ok = ok and equal_chunked(out, ref)
        del out
        rep_u = {}
        dt_untr, out = timed_copies(lambda b: rep_u.update(ts.learned_sort(
            b, config=ts.LearnedConfig(sample_per_mille=10, train=False), ctx=ctx)))
        ok = ok and equal_chunked(out, ref)
        del out
        r = dict(n=n, plain_ms=dt_plain * 1e3, learned_ms=dt_learn * 1e3, untrained_ms=dt_untr * 1e3,
                 plain_keys_per_s=n / dt_plain, learned_keys_per_s=n / dt_learn,
                 speedup_vs_plain=dt_plain / dt_learn, untrained_speedup_vs_plain=dt_plain / dt_untr,
                 learned_roofline_frac=40 * n / dt_learn / 1e9 / HBM_PEAK_GBS,
                 plain_roofline_frac=136 * n / dt_plain / 1e9 / HBM_PEAK_GBS,
                 plain_executed_passes=plain_info["executed_passes"],
                 overflow_rate=rep["overflow_keys"] / n, report=rep, untrained_report=rep_u,
                 sample_half_percent=dict(learned_ms=dt_half * 1e3, speedup_vs_plain=dt_plain / dt_half,
                                          overflow_rate=rep_h["overflow_keys"] / n, report=rep_h),
                 without_judge=dict(learned_ms=dt_nv * 1e3, speedup_vs_plain=dt_plain / dt_nv), ok=ok)
        if host and name == "u64":
            hsrc = torch.empty(n, dtype=src.dtype, pin_memory=True)
            hsrc.copy_(src)
            d = ref  # device buffer of the same size
            t = host_roundtrip([(hsrc, d)], lambda: ts.learned_sort(d, config=cfg, ctx=ctx), 2)
            r["with_copies"] = dict(desc="keys in pinned host memory: copied in, learned sort (1 % sample, judged), "
                                         "copied out (wall time, median)", ms_per_step=t * 1e3, keys_per_s=n / t)
            del hsrc
        if cpu is not None:
            from oracle import closed_form as cf

            sample = cf.lognormal_mixture_keys(seed, 1 << 25, as_f64=as_f64)
            r["cpu_baseline"] = cpu_sort_baseline(sample, cpu["threads"], cpu["budget"],
                                                  f"2^25 keys of the same mixture as {name}")
        del ref, src
        torch.cuda.empty_cache()
        res[name] = r
    return dict(desc="learned partitioner vs plain GPU radix, 10^9 keys: lognormal mixture (sigma 0.5/1/2) with "
                     "5 % Zipf(1.1) hot values, as u64 and as f64; model trained on a 1 % sample, output judged "
                     "(sortedness + multiset fingerprint); untrained (uniform-knot) model as the ablation",
                ok=all(v["ok"] for v in res.values()), **res)


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


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); every figure is per single MI355X.",
              file=sys.stderr)
        sys.exit(2)
    import torch  # noqa: F401

    import turbosort_amd as ts

    ctx = ts.Context(0)
    share = cpu_share()
    cpu = None if a.no_cpu_baseline else dict(threads=share["threads"], budget=a.cpu_seconds)
    host = not a.no_host
    t0 = time.perf_counter()
    head = run_workload(a.workload, a.log2n, a.steps, a.warmup, ctx, host=host, cpu=cpu)
    print(f"bench: {a.workload} done ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    also = {}
    for name in [x for x in a.also.split(",") if x and x != a.workload]:
        if name == "segmented":
            r = run_segmented(max(3, a.steps // 2), ctx, host=host, cpu=cpu, tune=not a.no_tune)
        elif name == "by_key":
            r = run_by_key(max(3, a.steps // 3), ctx, host=host, cpu=cpu)
        elif name == "learned":
            r = run_learned(max(2, a.steps // 5), ctx, n=a.learned_n, host=host, cpu=cpu)
        else:
            r = run_workload(name, a.log2n, max(3, a.steps // 3), 1, ctx, host=host, cpu=cpu)
        also[name] = r
        print(f"bench: {name} done ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)

    def rounded(x):
        if isinstance(x, float):
            return round(x, 4)
        if isinstance(x, dict):
            return {k: rounded(v) for k, v in x.items()}
        return x

    sweep_gbs = head["sweep_gbs"]
    cpu_head = head.get("cpu_baseline")
    if cpu_head is not None:
        cpu_head = dict(cpu_head, host_cpus=share["nproc"], cpu_share=share["affinity"], cpu_model=share["model"])
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
                       "executed": rounded(head["executed"]), "verified": head["ok"]},
        "with_copies": rounded(head.get("with_copies")),
        "cpu_baseline": rounded(cpu_head),
        "also": {k: rounded(v) for k, v in also.items()},
    }
    print(json.dumps(line), flush=True)
    if not head["ok"] or any(not v["ok"] for v in also.values()) or (
            head.get("with_copies") and not head["with_copies"]["ok"]):
        sys.exit(1)


if __name__ == "__main__":
    main()
