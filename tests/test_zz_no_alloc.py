"""mqi_run_vertices_noalloc (device-resident vertices, work list, counters
and secondary stacks in the library's static per-GPU scratch array) against
mqi_run_vertices on the same vertices: the same histories, steps and dropped
secondaries, and the same scorer arrays to FP64 summation order.  Two calls
in a row, the second with fewer vertices, so the second runs over what the
first left in the scratch."""
import numpy as np
import pytest

import phantoms
import refbind as rb

pytestmark = pytest.mark.gpu


def _vertices(n, seed):
    """n proton vertices entering a slab phantom from z = 0.5 mm, directed
    -z with a small spread, 60-160 MeV."""
    rng = np.random.default_rng(seed)
    v = np.zeros((n, 7), np.float32)
    v[:, 0] = rng.uniform(60.0, 160.0, n)
    v[:, 1:3] = rng.normal(0.0, 3.0, (n, 2))
    v[:, 3] = 0.5
    d = np.c_[rng.normal(0.0, 0.01, (n, 2)), -np.ones(n)]
    v[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    return v


def _scores(ctx, n_k):
    return [ctx.scorer(k) for k in range(n_k)]


def _same(a, b):
    for x, y in zip(a, b):
        fin = np.isfinite(x)
        assert np.array_equal(fin, np.isfinite(y))
        np.testing.assert_allclose(y[fin], x[fin], rtol=1e-9, atol=1e-300)


@pytest.mark.parametrize("exact,batched", [(True, True), (False, True), (False, False)])
def test_noalloc_equals_run_vertices_twice(ctx, exact, batched):
    import torch
    g = phantoms.slab_box(32, 32, 160)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(phantoms.spots([100.0, 120.0], 10))
    mask = rb.SC_DOSE | rb.SC_LETD
    ctx.set_scorers(mask)
    n_k = ctx.scorer_count()
    for call, n in enumerate((6000, 2500)):
        v = _vertices(n, 40 + call)
        sid = (np.arange(n) % 2).astype(np.uint32)
        cfg = ctx.config(seed=11 + call, exact=exact, batched=batched)
        print("call %d: %d vertices, exact %d, batched %d" % (call, n, exact, batched), flush=True)
        ctx.zero_scorers()
        want = ctx.run_vertices(v, cfg, spot_ids=sid)
        a = _scores(ctx, n_k)
        ctx.zero_scorers()
        dv = torch.from_numpy(v).to("cuda:0")
        ds = torch.from_numpy(sid.view(np.int32)).to("cuda:0")
        torch.cuda.synchronize()
        got = ctx.run_vertices_noalloc(dv.data_ptr(), n, cfg, d_spot_ids=ds.data_ptr())
        b = _scores(ctx, n_k)
        assert got.histories == want.histories == n
        assert got.steps == want.steps and got.dropped == want.dropped
        assert a[0].sum() > 0
        _same(a, b)
