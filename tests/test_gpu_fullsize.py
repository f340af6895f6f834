"""Size-independent properties at the workloads' full sizes (production
kernels):

- the 5000-spot Dij plan (CT 256x256x128, 1e8 histories): the CSR rows sum
  to the dense dose of the same run, and D^T g and D w equal scipy's FP64
  products of the copied matrix; the quadratic objective's gradient through
  the autograd layer equals scipy's FP64 gradient and central differences;
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


def test_full_dij_quadratic_gradient_and_finite_differences(ctx):
    """The 5000-spot Dij (threshold 1e-4 of each spot's maximum): the
    quadratic objective against a prescription in a 60 mm sphere through the
    autograd layer; its value and gradient against scipy in FP64 (gradient
    within 1e-5 of the largest component), and central differences on 64
    random spot weights (exact for a quadratic, up to rounding)."""
    import torch
    from moquimc_amd import dij, workloads
    w = workloads.dij()
    w.load(ctx)
    res = ctx.run(ctx.config(seed=23, batched=True, dij=True, dij_capacity=1 << 19))
    assert res.status == 0
    ctx.dij_build(1e-4)
    indptr, idx, val, shape = ctx.dij_csr()
    D = sps.csr_matrix((val.astype(np.float64), idx, indptr), shape=shape).T.tocsr()   # voxels x spots
    mat = dij.DijMatrix(ctx, 0)
    rng = np.random.default_rng(17)
    w_np = rng.uniform(0.5, 1.5, mat.n_spots).astype(np.float32)
    # 2 Gy in the target, expressed in the Dij's units: the unit-weight plan's
    # mean dose in the target maps to 2
    inside = workloads.target_prescription(w.xe, w.ye, w.ze, radius=30.0, dose=1.0) > 0
    level = float((D @ np.ones(mat.n_spots))[inside].mean())
    p_np = np.where(inside, level, 0.0)
    dev = torch.device("cuda", 0)
    wt = torch.tensor(w_np, device=dev, requires_grad=True)
    p = torch.tensor(p_np, device=dev)
    f = dij.quadratic_objective(mat, wt, p)
    f.backward()
    r = D @ w_np.astype(np.float64) - p_np
    f_ref = float(np.mean(r * r))
    g_ref = 2.0 * (D.T @ r) / D.shape[0]
    got = wt.grad.detach().cpu().numpy().astype(np.float64)
    assert abs(f.item() - f_ref) <= 1e-9 * abs(f_ref)
    assert np.abs(got - g_ref).max() <= 1e-5 * np.abs(g_ref).max()
    h = np.float32(2.0 ** -4)
    for j in rng.choice(mat.n_spots, 64, replace=False):
        wp, wm = w_np.copy(), w_np.copy()
        wp[j] += h
        wm[j] -= h
        with torch.no_grad():
            fp = dij.quadratic_objective(mat, torch.tensor(wp, device=dev), p).item()
            fm = dij.quadratic_objective(mat, torch.tensor(wm, device=dev), p).item()
        fd = (fp - fm) / (float(wp[j]) - float(wm[j]))
        assert abs(fd - got[j]) <= 1e-6 * np.abs(g_ref).max() + 1e-5 * abs(got[j]), (j, fd, got[j])


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


def test_full_robust_plan_nine_scenarios():
    """The 2000-spot CT plan under 9 robust scenarios in one launch (nominal,
    +-3 mm shifts in x, y, z, density x 1.035 and 0.965; the nominal scenario
    with twice the histories; the XShift/YShift/ZShift and DensityScaling
    scenarios of tps_env, mqi_tps_env.hpp:475-477,766-769), Dij scorer with a
    per-scenario dose beside it:
    - every history of every scenario is transported (sum of budget x spot
      histories);
    - each scenario's Dij (threshold 0) sums over its spot rows to that
      scenario's dose (float32 entries of FP64 sums);
    - the worst-case quadratic objective's gradient through the autograd
      layer (moquimc_amd.dij) equals scipy's FP64 gradient of the worst
      scenario to 1e-5 relative."""
    import torch
    from moquimc_amd import _abi, dij, workloads
    w = workloads.robust()
    shifts, scales, budgets = w.scenarios
    c = _abi.Context(0)
    try:
        w.load(c, per_scenario=True)
        print("robust plan: %d spots x %d scenarios" % (len(w.spots), len(budgets)), flush=True)
        res = c.run(c.config(seed=13, batched=True, dij=True, dij_slots_sqrt=4096))
        want_h = sum(int(np.rint(float(b) * h)) for b in budgets for h in w.spots["histories"].astype(np.float64))
        assert res.status == 0 and res.histories == want_h
        print("transported %d histories in %.1f s" % (res.histories, res.kernel_ms / 1e3), flush=True)
        c.dij_build(0.0)
        n_vox = int(np.prod(w.shape))
        for s in range(len(budgets)):
            indptr, idx, val, shape = c.dij_csr(s)
            assert shape == (len(w.spots), n_vox)
            col = np.bincount(idx, weights=val.astype(np.float64), minlength=n_vox)
            dose = c.scorer(0, s)
            assert np.any(dose != 0)
            assert np.allclose(col, dose, rtol=2e-6, atol=1e-6 * dose.max()), s
            del indptr, idx, val, col, dose
            print("scenario %d: rows sum to its dose" % s, flush=True)
        mats = [dij.DijMatrix(c, s, scale=1.0 / float(budgets[s])) for s in range(len(budgets))]
        dev = torch.device("cuda", 0)
        target = torch.from_numpy(workloads.target_prescription(w.xe, w.ye, w.ze, radius=30.0, dose=1.0)).to(dev)
        gen = torch.Generator(device="cpu").manual_seed(5)
        wt = (torch.rand(len(w.spots), generator=gen) + 0.5).to(dev)
        with torch.no_grad():
            presc = target.double() * mats[0].matvec(wt)[target > 0].mean()
        wv = wt.clone().requires_grad_(True)
        f, vals = dij.worst_case_objective(mats, wv, presc)
        f.backward()
        k = int(torch.argmax(vals).item())
        indptr, idx, val, shape = c.dij_csr(k)
        D = sps.csr_matrix((val.astype(np.float64) / float(budgets[k]), idx, indptr), shape=shape)
        wn = wt.double().cpu().numpy()
        r = D.T @ wn - presc.cpu().numpy()
        assert abs(float(f) - float(r @ r) / n_vox) <= 1e-6 * abs(float(r @ r) / n_vox)
        g_ref = 2.0 * (D @ r) / n_vox
        g = wv.grad.double().cpu().numpy()
        err = float(np.abs(g - g_ref).max() / np.abs(g_ref).max())
        print("worst scenario %d, gradient vs scipy FP64: %.3g relative" % (k, err), flush=True)
        assert err <= 1e-5
    finally:
        c.close()
