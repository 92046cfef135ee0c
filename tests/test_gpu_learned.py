"""Learned partitioner on the GPU: exact output, training kernel, ablation.

learned_sort must give exactly sort()'s output (the unique ascending order of
the radix keys, lib.rs:99-101) for every input, whatever the model.  The
training kernel's gradient is checked against the float64 oracle
(oracle/pwl.py, itself pinned by finite differences), and the fitted model
must balance buckets better than the untrained (uniform-knot) model.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import turbosort_amd as ts
from gpu_util import bits, to_dev, to_host
from oracle import closed_form as cf
from oracle import pwl
from shapes import float_specials_f32, random_of, rng, structured_u64

pytestmark = pytest.mark.gpu


def check_sorted(x: np.ndarray, dev, config=None) -> dict:
    t = to_dev(x, dev)
    rep = ts.learned_sort(t, config=config)
    torch.cuda.synchronize()
    assert np.array_equal(bits(to_host(t, x.dtype)), bits(cf.sort(x)))
    if x.size > 1 and (config is None or config.verify):
        assert rep["verified"] == 1, rep  # the output judge ran and passed
    return rep


@pytest.mark.parametrize("n", [2, 3, 100, 1025, 5000, 100_003, 1_000_000])
@pytest.mark.parametrize("dtype", ["u64", "f64", "u32", "i64", "i32", "f32"])
def test_uniform_bits(gpu, n, dtype):
    check_sorted(random_of(dtype, n, seed=n + len(dtype)), gpu)


@pytest.mark.parametrize("n", [1000, 300_000, 2_000_000])
def test_lognormal_mixture(gpu, n):
    rep = check_sorted(cf.lognormal_mixture_keys(8, n), gpu)
    assert rep["loss_after"] < rep["loss_before"]
    check_sorted(cf.lognormal_mixture_keys(8, n, as_f64=True), gpu)


def test_structured_and_degenerate(gpu):
    for shape, x in structured_u64(200_000, seed=31).items():
        check_sorted(x, gpu)
    n = 150_000
    i = np.arange(n, dtype=np.uint64)
    for x in (np.full(n, 7, dtype=np.uint64), i, i[::-1].copy(), (i % 3).astype(np.uint64),
              np.concatenate([np.zeros(n // 2, np.uint64), np.full(n - n // 2, 2**63, np.uint64)]),
              (i % 2).astype(np.uint64) * np.uint64(2**64 - 1)):
        check_sorted(x, gpu)
    check_sorted(float_specials_f32(n, seed=4), gpu)
    check_sorted(float_specials_f32(n, seed=5).astype(np.float64), gpu)


@pytest.mark.parametrize("buckets_log2", [1, 8, 12, 20])
@pytest.mark.parametrize("train", [True, False])
def test_configurations(gpu, buckets_log2, train):
    x = cf.lognormal_mixture_keys(9, 400_000)
    cfg = ts.LearnedConfig(buckets_log2=buckets_log2, train=train, iters=50)
    rep = check_sorted(x, gpu, cfg)
    assert rep["buckets"] == 1 << buckets_log2
    assert rep["sample"] == 400_000 * cfg.sample_per_mille // 1000


def test_gradient_matches_oracle(gpu):
    g = rng(12)
    T = 40_000
    u = np.sort(g.uniform(0, 1024, T)).astype(np.float32)
    t = np.sort(g.uniform(0, 1, T)).astype(np.float32)
    theta = g.normal(0.5, 0.7, 1024).astype(np.float32)
    th_dev = torch.from_numpy(theta.copy()).to(gpu)
    grad, l0, l1 = ts.pwl_fit(torch.from_numpy(u).to(gpu), torch.from_numpy(t).to(gpu), th_dev, iters=0)
    want_l, want_g = pwl.loss_grad(theta, u, t)
    got = grad.cpu().numpy().astype(np.float64)
    assert abs(l0 - want_l) <= 1e-5 * max(1.0, want_l) + 1e-7
    assert l1 == l0  # no step taken
    # float32 accumulation over 40k points against float64
    scale = np.abs(want_g).max()
    assert np.max(np.abs(got - want_g)) <= 2e-3 * scale, (np.max(np.abs(got - want_g)), scale)
    assert np.array_equal(th_dev.cpu().numpy(), theta)


def test_training_lowers_the_loss(gpu):
    x = np.sort(cf.to_radix(cf.lognormal_mixture_keys(8, 200_000)))
    # the points the sort trains on: log coordinate over the sample's range
    lg = np.log2(np.maximum(x.astype(np.float64), 1.0))
    u = ((lg - lg[0]) / (lg[-1] - lg[0]) * 1023.999).astype(np.float32)
    t = ((np.arange(x.size) + 0.5) / x.size).astype(np.float32)
    theta = torch.full((1024,), 0.54132485, dtype=torch.float32, device=gpu)
    _, l0, l1 = ts.pwl_fit(torch.from_numpy(u).to(gpu), torch.from_numpy(t).to(gpu), theta, iters=300)
    assert l1 < 0.05 * l0, (l0, l1)
    # the trained theta gives the oracle the same loss
    assert abs(pwl.loss(theta.cpu().numpy(), u, t) - l1) < 1e-3 * l0


def test_trained_model_balances_buckets_better_than_uniform_knots(gpu):
    """Ablation: the fitted CDF against the untrained model (equal-width
    buckets over the sample's log-key range) on the skewed mixture."""
    x = cf.lognormal_mixture_keys(8, 2_000_000, hot_share=0.0)
    base = ts.LearnedConfig(buckets_log2=12)
    trained = check_sorted(x, gpu, base)
    plain = check_sorted(x, gpu, ts.LearnedConfig(buckets_log2=12, train=False))
    mean = x.size / 4096
    assert trained["loss_after"] <= trained["loss_before"] < trained["loss_uniform"], trained
    assert trained["largest_bucket"] < 4 * mean, trained
    assert plain["largest_bucket"] > 2 * trained["largest_bucket"], (plain, trained)


@pytest.mark.parametrize("rounds", [0, 1])
def test_overflow_without_split_rounds(gpu, rounds):
    """Overflowing buckets that the split rounds leave take the device radix
    sort (on the learned sort's second stream); the output stays exact."""
    n = 2_000_000
    g = rng(81)
    x = random_of("u64", n, seed=82)
    hot = random_of("u64", 2, seed=83)
    x = np.where(g.random(n) < 0.4, hot[0], np.where(g.random(n) < 0.2, hot[1], x)).astype(np.uint64)
    ctx = ts.Context(gpu)
    ctx.set_split_rounds(rounds)
    t = to_dev(x, gpu)
    rep = ts.learned_sort(t, ctx=ctx)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(t, np.uint64), np.sort(x))
    assert rep["overflow_buckets"] >= 1 and rep["split_rounds"] <= rounds, rep


def test_rejects_other_widths(gpu):
    t = torch.zeros(10, dtype=torch.uint16, device=gpu)
    with pytest.raises(ValueError, match="4- and 8-byte"):
        ts.learned_sort(t)


@pytest.mark.parametrize("dtype", ["u64", "f64", "u32"])
def test_heavy_hitters_take_the_split_path(gpu, dtype):
    """Buckets far beyond one workgroup (a few values repeated millions of
    times) go through three-way splits; the output stays exact."""
    n = 3_000_000
    g = rng(77)
    x = random_of(dtype, n, seed=78)
    hot = random_of(dtype, 3, seed=79)
    pick = g.random(n)
    x = np.where(pick < 0.5, hot[0], np.where(pick < 0.7, hot[1], np.where(pick < 0.75, hot[2], x)))
    rep = check_sorted(x.astype(cf.NP_DTYPES[dtype]), gpu)
    assert rep["overflow_buckets"] >= 1 and rep["split_rounds"] >= 1, rep


@pytest.mark.parametrize("per_mille", [10, 600])
def test_config_sample_sizes(gpu, per_mille):
    """The workload's 1 % sample, and a sample above half the keys (its own
    radix sort then runs tiles the partition's status rows were not sized
    for alone)."""
    cfg = ts.LearnedConfig(sample_per_mille=per_mille)
    for n in (1_000_000, 3_000_017):
        rep = check_sorted(cf.lognormal_mixture_keys(8, n), gpu, cfg)
        assert rep["sample"] == n * per_mille // 1000


def test_judge_rejects_a_corrupted_output(gpu):
    """ts_verify_sorted, the learned sort's judge, on outputs corrupted from
    the host: a swapped adjacent pair fails the order check, a changed key
    the multiset fingerprint; the true output passes both."""
    x = cf.lognormal_mixture_keys(8, 1_000_000)
    want = cf.sort(x)
    fp = ts.checksum(to_dev(x, gpu))[:2]
    ts.verify_sorted(to_dev(want, gpu), fp)
    bad = want.copy()
    bad[[5000, 5001]] = bad[[5001, 5000]]
    if bad[5000] != bad[5001]:
        with pytest.raises(RuntimeError, match="out of order"):
            ts.verify_sorted(to_dev(bad, gpu), fp)
    changed = want.copy()
    changed[-1] = changed[-1] + np.uint64(1)
    ts.verify_sorted(to_dev(changed, gpu))  # still sorted
    with pytest.raises(RuntimeError, match="multiset"):
        ts.verify_sorted(to_dev(changed, gpu), fp)


def test_learned_sort_without_the_judge(gpu):
    rep = check_sorted(cf.lognormal_mixture_keys(9, 500_000), gpu, ts.LearnedConfig(verify=False))
    assert rep["verified"] == 0
