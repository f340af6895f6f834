"""Launch parameters: results do not depend on them (per-history random
streams), and the autotuner returns a valid, working configuration."""
import numpy as np
import pytest

import phantoms
import refbind as rb

pytestmark = pytest.mark.gpu


def _setup(ctx):
    g = phantoms.slab_box(30, 30, 160)
    sp = phantoms.spots([120.0, 150.0, 90.0], 500, sigma=4.0, sigma_e_frac=0.01,
                        positions=[(0, 0, 0.5), (5, -3, 0.5), (-6, 4, 0.5)])
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE | rb.SC_LETD)
    return g, sp


def test_launch_shape_does_not_change_results(ctx):
    _setup(ctx)
    outs = []
    # refill_lanes 64: a wave takes new histories only when all its lanes are done
    for block, chunk, blocks, refill in [(256, 1, 0, 0), (128, 4, 0, 64), (64, 2, 7, 8), (192, 1, 3, 1),
                                         (128, 2, 0, 64)]:
        ctx.zero_scorers()
        res = ctx.run(ctx.config(seed=41, batched=True, exact=True, block_size=block, fetch_chunk=chunk,
                                 blocks=blocks, refill_lanes=refill))
        assert res.histories == 1500
        outs.append([ctx.scorer(k) for k in range(3)])
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            fin = np.isfinite(a)
            assert np.array_equal(fin, np.isfinite(b))
            assert np.allclose(a[fin], b[fin], rtol=1e-9, atol=0)


def test_production_whole_wave_refills_match_default(ctx):
    """The production kernels with whole-wave refills (the autotuner's pick on
    the CT plan) score the same histories as with the default refills: equal
    up to the FP64 summation order of the atomics."""
    _setup(ctx)
    outs = []
    for refill, chunk in [(0, 0), (64, 2), (32, 1)]:
        ctx.zero_scorers()
        res = ctx.run(ctx.config(seed=47, batched=True, waves_per_simd=3, block_size=128, fetch_chunk=chunk,
                                 refill_lanes=refill))
        assert res.histories == 1500 and res.status == 0
        outs.append([ctx.scorer(k) for k in range(3)])
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            fin = np.isfinite(a)
            assert np.array_equal(fin, np.isfinite(b))
            assert np.allclose(a[fin], b[fin], rtol=1e-9, atol=0)


def test_paired_scoring_matches_the_plain_launch(ctx):
    """Batched production passes add scorers 0 and 1 into one interleaved
    array from pairs of lanes (folded after the pass); the plain per-spot
    launches score each array directly.  Both transport the same histories,
    so every scorer agrees up to the FP64 summation order: dose + LETd (slot
    2 outside the pair) and dose + the statistics pair (Dose derived from
    Dose_stat)."""
    _setup(ctx)
    for mask in (rb.SC_DOSE | rb.SC_LETD, rb.SC_DOSE | rb.SC_STAT):
        ctx.set_scorers(mask)
        outs = []
        for batched, refill in ((True, 8), (True, 64), (False, 0)):
            ctx.zero_scorers()
            res = ctx.run(ctx.config(seed=48, batched=batched, waves_per_simd=3, refill_lanes=refill))
            assert res.histories == 1500 and res.status == 0
            outs.append([ctx.scorer(k) for k in range(ctx.scorer_count())])
        assert len(outs[0]) == 3
        for o in outs[1:]:
            for a, b in zip(outs[0], o):
                fin = np.isfinite(a)
                assert np.array_equal(fin, np.isfinite(b))
                assert np.any(a[fin] != 0)
                assert np.allclose(a[fin], b[fin], rtol=1e-9, atol=0)


def test_paired_scoring_with_per_scenario_arrays(ctx):
    """Paired scoring over scenarios with their own arrays (element offsets of
    one scenario per voxel grid): each scenario's batched scorers equal the
    plain launch's."""
    _setup(ctx)
    ctx.set_scenarios(shifts=[[0, 0, 0], [3, 0, 0], [0, -2, 1]], scales=[1.0, 1.02, 0.97])
    ctx.set_scorers(rb.SC_DOSE | rb.SC_LETD, per_scenario=True)
    outs = []
    for batched in (True, False):
        ctx.zero_scorers()
        res = ctx.run(ctx.config(seed=49, batched=batched, waves_per_simd=3))
        assert res.histories == 3 * 1500 and res.status == 0
        outs.append([[ctx.scorer(k, scenario=sc) for k in range(3)] for sc in range(3)])
    for sc in range(3):
        for a, b in zip(outs[0][sc], outs[1][sc]):
            fin = np.isfinite(a)
            assert np.array_equal(fin, np.isfinite(b))
            assert np.any(a[fin] != 0)
            assert np.allclose(a[fin], b[fin], rtol=1e-9, atol=0)
    # the scenarios differ (shifted and scaled copies of the plan)
    assert not np.allclose(outs[0][0][0], outs[0][1][0])
    ctx.set_scenarios()


def test_production_occupancy_variants_run(ctx):
    _setup(ctx)
    ref = None
    for waves in (2, 3, 4, 5, 6, 8):
        ctx.zero_scorers()
        res = ctx.run(ctx.config(seed=42, batched=True, waves_per_simd=waves))
        assert res.histories == 1500 and res.status == 0
        d = ctx.scorer(0)
        if ref is None:
            ref = d
        # the same kernel source at another register budget: identical arithmetic
        assert np.allclose(d, ref, rtol=1e-9, atol=0)


def test_autotune_picks_a_working_configuration(ctx):
    _setup(ctx)
    cfg = ctx.config(seed=43, batched=True, stack_depth=-1)
    cfg, ms = ctx.autotune(cfg, sample=1.0)
    assert ms > 0
    assert cfg.waves_per_simd in (3, 4, 5)
    assert cfg.block_size in (128, 256)
    assert cfg.fetch_chunk in (1, 2, 4)
    assert cfg.refill_lanes in (1, 2, 4, 8, 16, 32, 60, 62, 64)
    assert cfg.stack_depth in (16, 32, 64, 128, 256, 512)
    # the trial passes left nothing in the scorers
    assert not np.any(ctx.scorer(0))
    res = ctx.run(cfg)
    assert res.histories == 1500 and res.dropped == 0
    with pytest.raises(Exception):
        ctx.autotune(ctx.config(exact=True), sample=0.5)


def test_graph_replays_match_eager(ctx):
    """use_graph: the captured batched pass, replayed several times with eager
    work in between, gives the eager result every time; a new plan is
    captured anew."""
    _setup(ctx)
    eager = []
    for off in (0, 5000):
        ctx.zero_scorers()
        ctx.run(ctx.config(seed=44, batched=True, exact=True, hist_offset=off))
        eager.append(ctx.scorer(0))
    for rep in range(3):
        for k, off in enumerate((0, 5000)):
            ctx.zero_scorers()
            res = ctx.run(ctx.config(seed=44, batched=True, exact=True, use_graph=True, hist_offset=off))
            assert res.histories == 1500
            assert np.allclose(ctx.scorer(0), eager[k], rtol=1e-9, atol=0), (rep, k)
        # eager work between replays
        ctx.zero_scorers()
        ctx.run(ctx.config(seed=45, batched=True))
    # a different plan: recaptured
    g = phantoms.water_box(20, 20, 100)
    sp = phantoms.spots([90.0], 700, sigma=2.0)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE)
    ctx.run(ctx.config(seed=46, batched=True, exact=True))
    want = ctx.scorer(0)
    ctx.zero_scorers()
    res = ctx.run(ctx.config(seed=46, batched=True, exact=True, use_graph=True))
    assert res.histories == 700
    assert np.allclose(ctx.scorer(0), want, rtol=1e-9, atol=0)


def test_graph_with_stat_rounds(ctx):
    """Statistical rounds replay the captured pass with a new history offset
    each round: same scorers as the eager rounds."""
    g = phantoms.slab_box(24, 24, 140)
    sp = phantoms.spots([110.0, 130.0], 300, sigma=3.0, sigma_e_frac=0.01, positions=[(0, 0, 0.5), (3, 2, 0.5)])
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE | rb.SC_STAT)
    out = []
    for graph in (False, True):
        ctx.zero_scorers()
        res = ctx.run(ctx.config(seed=47, batched=True, exact=True, use_graph=graph, stat_criterion=1e-6,
                                 max_rounds=3))
        assert res.rounds == 3
        out.append([ctx.scorer(k) for k in range(3)])
    for a, b in zip(*out):
        assert np.allclose(a, b, rtol=1e-9, atol=0)


def test_atomic_probe_times_and_rejects_bad_sizes(ctx):
    """mqi_atomic_probe (the ceiling bench.py's atomic roofline is priced
    against) runs and reports a positive device time; slot counts beyond
    2^32 and empty sizes are argument errors, not launches."""
    ms = ctx.atomic_probe(1 << 20, 1 << 22)
    assert ms > 0.0
    assert (1 << 22) / (ms * 1e-3) > 1e8    # at least 1e8 atomics/s
    assert ctx.atomic_probe(1 << 20, 1 << 22, window=64) > 0.0
    for n_slots, n_updates, window in [((1 << 32) + 1, 1 << 10, 0), (0, 1 << 10, 0), (1 << 10, 0, 0),
                                       (1 << 10, 1 << 10, 48), (1 << 10, 1 << 10, 1 << 10), (1 << 10, 1 << 10, -1)]:
        with pytest.raises(RuntimeError):
            ctx.atomic_probe(n_slots, n_updates, window)
