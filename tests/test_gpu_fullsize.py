"""Size-independent properties at the workloads' full sizes (production
kernels):

- the 5000-spot Dij plan (CT 256x256x128, 1e8 histories): the CSR rows sum
  to the dense dose of the same run, and D^T g and D w equal scipy's FP64
  products of the copied matrix;
- the 2000-spot CT plan (512x512x200, 4.27e7 histories): dose is additive
  over spots (every history's random stream is keyed by seed, spot and
  history, so the plan's dose equals the sum of the doses of its even and odd
  spots, up to FP64 summation order), every history of the plan is
  transported, and the plain per-spot launches give the batched launch's
  dose.
"""
import numpy as np
import pytest
import scipy.sparse as sps

pytestmark = pytest.mark.gpu


def test_full_dij_rows_sum_to_dose_and_products_match_fp64(ctx):
    import torch
    from moquimc_amd import workloads
    w = workloads.dij()
    w.load(ctx)
    print("dij workload: %d spots, %d histories" % (len(w.spots), w.histories), flush=True)
    res = ctx.run(ctx.config(seed=21, batched=True, dij=True, dij_capacity=1 << 19))
    assert res.status == 0 and res.histories == w.histories
    ctx.dij_build(0.0)
    indptr, idx, val, shape = ctx.dij_csr()
    D = sps.csr_matrix((val.astype(np.float64), idx, indptr), shape=shape)
    assert D.shape == (5000, 256 * 256 * 128) and D.nnz > 1e8
    dose = ctx.scorer(0)
    col = np.asarray(D.sum(axis=0)).ravel()
    assert np.allclose(col, dose, rtol=2e-6, atol=1e-6 * dose.max())
    rng = np.random.default_rng(5)
    wt = rng.uniform(0.5, 1.5, D.shape[0]).astype(np.float32)
    g = rng.normal(size=D.shape[1]).astype(np.float32)
    dev = torch.device("cuda", 0)
    tw, tg = torch.from_numpy(wt).to(dev), torch.from_numpy(g).to(dev)
    y = torch.zeros(D.shape[0], dtype=torch.float64, device=dev)
    d = torch.zeros(D.shape[1], dtype=torch.float64, device=dev)
    ctx.dij_matvec(0, 1, tg.data_ptr(), y.data_ptr(), y_double=True)
    ctx.dij_matvec(0, 0, tw.data_ptr(), d.data_ptr(), y_double=True)
    torch.cuda.synchronize()
    want_y = D @ g.astype(np.float64)
    want_d = D.T @ wt.astype(np.float64)
    assert np.allclose(y.cpu().numpy(), want_y, rtol=1e-9, atol=1e-12 * np.abs(want_y).max())
    assert np.allclose(d.cpu().numpy(), want_d, rtol=1e-9, atol=1e-12 * np.abs(want_d).max())


def test_full_ct_plan_dose_is_additive_over_spots(ctx):
    from moquimc_amd import workloads
    w = workloads.ct()
    w.load(ctx)
    print("ct plan: %d spots, %d histories" % (len(w.spots), w.histories), flush=True)
    cfg = ctx.config(seed=13, batched=True)
    doses = []
    for part in (None, 0, 1):
        sp = w.spots.copy()
        if part is not None:
            sp["histories"][np.arange(len(sp)) % 2 != part] = 0
        ctx.set_spots(sp)
        ctx.zero_scorers()
        res = ctx.run(cfg)
        assert res.status == 0 and res.histories == int(sp["histories"].sum())
        doses.append(ctx.scorer(0))
    full, even, odd = doses
    assert np.any(even != 0) and np.any(odd != 0)
    assert np.allclose(full, even + odd, rtol=1e-9, atol=1e-12 * full.max())
    # the plain per-spot form (one launch per spot) transports the same
    # histories: the dose the batched/plain ratio is quoted on is the same
    ctx.set_spots(w.spots)
    ctx.zero_scorers()
    res = ctx.run(ctx.config(seed=13, batched=False))
    assert res.status == 0 and res.histories == w.histories
    plain = ctx.scorer(0)
    assert np.allclose(plain, full, rtol=1e-9, atol=1e-12 * full.max())
