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
    p.add_argument("--also", default="u64,u64z,host,segmented,by_key,learned",
                   help="extra workloads reported beside the headline (u32, u64, u64z, host, segmented, by_key, learned)")
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


def run_segmented(steps: int, ctx, n_segments: int = 100_000, seed: int = 7):
    """Batched segmented key-value sort: 10^5 log-uniform segments in [2, 16384],
    u32 keys uniform (SplitMix64 state seed + 2^32), u32 values = global index."""
    import torch

    import turbosort_amd as ts
    from turbosort_amd import autotune

    dev = torch.device("cuda", 0)
    # Kernel shape per size class, picked on this GPU before the timed runs.
    tuning = autotune.tune_segmented(torch.uint32, True, ctx=ctx)
    off = torch.from_numpy(segment_offsets(seed, n_segments)).to(dev)
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
    del gk, gv, graph
    torch.cuda.empty_cache()
    return dict(desc="batched segmented key-value sort, 10^5 log-uniform segments in [2, 16384] (SplitMix64 "
                     "state 7), u32 keys uniform, u32 values = global index", n=n, segments=n_segments,
                ms_per_step=dt * 1e3, pairs_per_s=n / dt, algo_bytes_per_pair=16,
                algorithmic_GBps=16 * n / dt / 1e9, roofline_frac=16 * n / dt / 1e9 / HBM_PEAK_GBS,
                graph_replay_ms_median=times[len(times) // 2],
                autotuned={str(c): tuning["names"][c][v] for c, v in tuning["choice"].items()}, ok=ok)


def run_by_key(steps: int, ctx, n: int = 100_000_000, seed: int = 5):
    """Stable sort_by_key of 10^8 16-byte records keyed on an f32 score at
    offset 4, then apply_order on a second u64 slice (SplitMix64 state 6)."""
    import numpy as np
    import torch

    import turbosort_amd as ts

    dev = torch.device("cuda", 0)
    rec = score_records(n, seed)
    src = torch.from_numpy(rec.view(np.uint8).reshape(n, 16)).to(dev)
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
    others = [other_src.clone() for _ in range(steps)]
    dt_apply = timed(lambda i: buf.apply_order(others[i]), steps)
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
    del rs, others, order, sbits, radix, rk, seen, src
    torch.cuda.empty_cache()
    return dict(desc="stable sort_by_key, 10^8 16-byte records {u32 id, f32 score, u64 payload} keyed on the "
                     "score (N(0,1), 1 % NaN/inf/-0.0, 20 % duplicated), records gathered; apply_order on u64[10^8]",
                n=n, ms_per_step=dt * 1e3, records_per_s=n / dt, algo_bytes_per_record=120,
                algorithmic_GBps=120 * n / dt / 1e9, roofline_frac=120 * n / dt / 1e9 / HBM_PEAK_GBS,
                apply_order_ms=dt_apply * 1e3, apply_order_GBps=20 * n / dt_apply / 1e9, ok=ok)


def run_host(steps: int, ctx, log2n: int = 28, seed: int = 2):
    """sort_host on 2^28 u32 keys in pinned host memory: H2D + sort + D2H, the
    end-to-end rate for a caller whose slice lives in host memory."""
    import torch

    import turbosort_amd as ts

    n = 1 << log2n
    dev = torch.device("cuda", 0)
    d = torch.empty(n, dtype=torch.uint32, device=dev)
    ts.fill_splitmix64(d, seed, ctx=ctx)
    src = torch.empty(n, dtype=torch.uint32, pin_memory=True)
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
    t = sorted(times)[len(times) // 2]
    del d, src, h
    torch.cuda.empty_cache()
    return dict(desc="sort_host: 2^28 u32 keys (SplitMix64 state 2) in pinned host memory; copy to the device, "
                     "sort, copy back (wall time, median)", n=n, ms_per_step=t * 1e3, keys_per_s=n / t,
                link_GBps_both_ways=2 * 4 * n / t / 1e9, ok=ok)


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


def run_learned(steps: int, ctx, n: int = 1_000_000_000, seed: int = 8):
    """Learned partitioner against the plain device radix on 10^9 skewed u64
    keys (and the same x as f64), with the untrained model as the ablation."""
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
        rep = {}
        dt_learn, out = timed_copies(lambda b: rep.update(ts.learned_sort(b, ctx=ctx)))
        ok = equal_chunked(out, ref)
        del out
        rep_u = {}
        dt_untr, out = timed_copies(lambda b: rep_u.update(ts.learned_sort(b, config=ts.LearnedConfig(train=False),
                                                                            ctx=ctx)))
        ok = ok and equal_chunked(out, ref)
        del out, ref, src
        torch.cuda.empty_cache()
        res[name] = dict(n=n, plain_ms=dt_plain * 1e3, learned_ms=dt_learn * 1e3, untrained_ms=dt_untr * 1e3,
                         plain_keys_per_s=n / dt_plain, learned_keys_per_s=n / dt_learn,
                         speedup_vs_plain=dt_plain / dt_learn, untrained_speedup_vs_plain=dt_plain / dt_untr,
                         learned_roofline_frac=40 * n / dt_learn / 1e9 / HBM_PEAK_GBS,
                         plain_roofline_frac=136 * n / dt_plain / 1e9 / HBM_PEAK_GBS,
                         overflow_rate=rep["overflow_keys"] / n, report=rep, untrained_report=rep_u, ok=ok)
    return dict(desc="learned partitioner vs plain GPU radix, 10^9 keys: lognormal mixture (sigma 0.5/1/2) with "
                     "5 % Zipf(1.1) hot values, as u64 and as f64; untrained (uniform-knot) model as the ablation",
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
        if name == "segmented":
            r = run_segmented(max(3, a.steps // 2), ctx)
        elif name == "by_key":
            r = run_by_key(max(3, a.steps // 3), ctx)
        elif name == "learned":
            r = run_learned(max(2, a.steps // 5), ctx)
        elif name == "host":
            r = run_host(max(3, a.steps // 3), ctx)
        else:
            r = run_workload(name, a.log2n, max(3, a.steps // 3), 1, ctx)
            r["roofline_frac"] = r["total_gbs"] / HBM_PEAK_GBS
        also[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
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
