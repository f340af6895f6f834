"""Statistical uncertainty and the stopping criterion on the GPU.

mqi_stat restates calculate_standard_deviation + calculate_stat
(moqui/kernel_functions/mqi_variables.hpp:21-48,
moqui/base/environments/mqi_tps_env.hpp:1340-1426); here it is checked
against the same closed form evaluated with numpy on the scorer arrays the
run produced:  mu = S1/N, sd = sqrt((S2/N - mu^2)/(N-1)) (float), stat = mean
of sd/mu over voxels with mu > threshold * max mu.  The stopping loop
(run_by_beam_stat, mqi_tps_env.hpp:1243-1336) re-runs the whole plan per
round and finally scales the non-statistical scorers by N_target/N_tracked
(calculate_average_results, :1430-1476): checked against rounds run by hand.
"""
import numpy as np
import pytest

import phantoms
import refbind as rb

pytestmark = pytest.mark.gpu


def _numpy_stat(s1, s2, n, thr):
    n = float(n)
    touched = ~((s1 == 0.0) & (s2 == 0.0))
    mu = (s1 / n).astype(np.float32)
    sd = np.sqrt((((s2 / n) - (s1 / n) * (s1 / n)) / (n - 1)).astype(np.float32))
    lim = float(mu[touched].max()) * float(np.float32(thr))
    keep = touched & (mu.astype(np.float64) > lim)
    return float(np.sum((sd[keep] / mu[keep]).astype(np.float32).astype(np.float64)) / keep.sum())


def _setup(ctx, hist=400):
    g = phantoms.slab_box(24, 24, 140)
    sp = phantoms.spots([110.0, 130.0], hist, sigma=3.0, sigma_e_frac=0.01, positions=[(0, 0, 0.5), (3, 2, 0.5)])
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE | rb.SC_STAT)
    return g, sp


@pytest.mark.parametrize("thr", [0.5, 0.1])
def test_stat_matches_closed_form(ctx, thr):
    _setup(ctx)
    res = ctx.run(ctx.config(seed=31, batched=True))
    s1, s2 = ctx.scorer(1), ctx.scorer(2)
    want = _numpy_stat(s1, s2, res.histories, thr)
    got = ctx.stat(res.histories, thr)
    assert 0 < got < 1
    assert abs(got - want) <= 1e-9 * want


def test_stopping_rounds_rescale_like_hand_rounds(ctx):
    g, sp = _setup(ctx, hist=300)
    crit = 1e-6   # unreachable: the loop stops at max_rounds
    res = ctx.run(ctx.config(seed=32, batched=True, exact=True, stat_criterion=crit, stat_threshold=0.5,
                             max_rounds=3))
    assert res.rounds == 3 and res.histories == 3 * 600
    dose_stat_run = ctx.scorer(0)
    s1_run = ctx.scorer(1)
    # the same three rounds by hand: history offsets 0, 300, 600 (the longest item)
    ctx.zero_scorers()
    for k in range(3):
        ctx.run(ctx.config(seed=32, batched=True, exact=True, hist_offset=k * 300))
    dose_sum = ctx.scorer(0)
    w = np.float64(np.float32(600.0 / 1800.0))
    assert np.allclose(dose_stat_run, dose_sum * w, rtol=1e-9, atol=0)
    # statistics scorers are not rescaled
    assert np.allclose(s1_run, ctx.scorer(1), rtol=1e-9, atol=0)


def test_stopping_reaches_criterion(ctx):
    _setup(ctx, hist=2000)
    res = ctx.run(ctx.config(seed=33, batched=True, stat_criterion=0.2, stat_threshold=0.5, max_rounds=20))
    assert res.stat <= 0.2
    assert 1 <= res.rounds < 20


def test_stat_over_a_statistics_roi(ctx):
    """With a statistics ROI the reference keeps the pair compactly over the
    ROI's voxels and calculate_stat takes the maximum and the mean over those
    (mqi_tps_env.hpp:797-815,1400-1426); here the pair is dense with zeros off
    the ROI: the statistic equals the closed form on the ROI's voxels alone,
    and the stopping loop runs on it."""
    g, sp = _setup(ctx)
    k, j, i = np.meshgrid(np.arange(140), np.arange(24), np.arange(24), indexing="ij")
    roi = (((i - 12) ** 2 + (j - 12) ** 2 < 9) & (k >= 60) & (k < 130)).astype(np.uint8).ravel()
    ctx.set_roi(1, roi)
    try:
        res = ctx.run(ctx.config(seed=34, batched=True))
        s1, s2 = ctx.scorer(1), ctx.scorer(2)
        assert not np.any(s1[roi == 0]) and not np.any(s2[roi == 0])
        inside = roi == 1
        for thr in (0.5, 0.0):
            want = _numpy_stat(s1[inside], s2[inside], res.histories, thr)
            got = ctx.stat(res.histories, thr)
            assert abs(got - want) <= 1e-9 * want, thr
        ctx.zero_scorers()
        res = ctx.run(ctx.config(seed=35, batched=True, stat_criterion=0.2, stat_threshold=0.5, max_rounds=8))
        assert res.stat <= 0.2 or res.rounds == 8
    finally:
        ctx.set_roi(1)
