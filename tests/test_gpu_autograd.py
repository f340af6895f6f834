"""Differentiation through the Dij: the autograd Function against an FP64
CPU gradient (scipy) and central finite differences, for the quadratic
objective and the worst case over robust scenarios."""
import numpy as np
import pytest
import scipy.sparse as sps

import phantoms
import refbind as rb

pytestmark = pytest.mark.gpu


def _plan(ctx, scenarios):
    g = phantoms.slab_box(32, 32, 140)
    pos = [(x, y, 0.5) for x in (-6.0, -2.0, 2.0, 6.0) for y in (-4.0, 0.0, 4.0)]
    energies = [100.0 + 5.0 * k for k in range(len(pos))]
    sp = phantoms.spots(energies, 1500, sigma=3.0, sigma_e_frac=0.01, positions=pos)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    if scenarios:
        shifts = np.array([[0, 0, 0], [2.0, 0, 0], [0, -2.0, 0]], np.float32)
        scales = np.array([1.0, 1.0, 1.035], np.float32)
        ctx.set_scenarios(shifts, scales)
    else:
        ctx.set_scenarios()
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE)
    res = ctx.run(ctx.config(seed=12, batched=True, dij=True, dij_capacity=1 << 16))
    assert res.status == 0
    ctx.dij_build(1e-4)
    return g, sp


def _scipy(ctx, s):
    indptr, idx, val, shape = ctx.dij_csr(s)
    # rows = spots; the voxels x spots matrix is its transpose
    return sps.csr_matrix((val.astype(np.float64), idx, indptr), shape=shape).T.tocsr()


def _prescription(g):
    xc = 0.5 * (g.xe[:-1] + g.xe[1:])
    yc = 0.5 * (g.ye[:-1] + g.ye[1:])
    zc = 0.5 * (g.ze[:-1] + g.ze[1:])
    r2 = (zc[:, None, None] + 70.0) ** 2 / 900.0 + yc[None, :, None] ** 2 / 64.0 + xc[None, None, :] ** 2 / 64.0
    return np.where(r2 <= 1.0, 1e-4, 0.0).reshape(-1)


def test_gradient_matches_fp64_and_finite_differences(ctx):
    import torch
    from moquimc_amd import dij
    g, sp = _plan(ctx, scenarios=False)
    mat = dij.DijMatrix(ctx, 0)
    D = _scipy(ctx, 0)
    assert D.shape == (mat.n_voxels, mat.n_spots)
    rng = np.random.default_rng(3)
    w_np = rng.uniform(0.5, 1.5, mat.n_spots).astype(np.float32)
    p_np = _prescription(g)
    dev = torch.device("cuda", 0)
    w = torch.tensor(w_np, device=dev, requires_grad=True)
    p = torch.tensor(p_np, device=dev)
    f = dij.quadratic_objective(mat, w, p)
    f.backward()
    # FP64 CPU reference
    r = D @ w_np.astype(np.float64) - p_np
    f_ref = np.mean(r * r)
    g_ref = 2.0 * (D.T @ r) / D.shape[0]
    assert abs(f.item() - f_ref) <= 1e-9 * abs(f_ref)
    got = w.grad.detach().cpu().numpy().astype(np.float64)
    assert np.abs(got - g_ref).max() <= 1e-5 * np.abs(g_ref).max()
    # central differences (h a power of two: w +- h is exact in float32)
    h = 2.0 ** -10
    for j in range(mat.n_spots):
        wp, wm = w_np.copy(), w_np.copy()
        wp[j] += h
        wm[j] -= h
        with torch.no_grad():
            fp = dij.quadratic_objective(mat, torch.tensor(wp, device=dev), p).item()
            fm = dij.quadratic_objective(mat, torch.tensor(wm, device=dev), p).item()
        fd = (fp - fm) / (2 * h)
        assert abs(fd - got[j]) <= 1e-6 * np.abs(g_ref).max() + 1e-4 * abs(got[j]), j


def test_worst_case_over_scenarios(ctx):
    import torch
    from moquimc_amd import dij
    g, sp = _plan(ctx, scenarios=True)
    # per-scenario scales (budget normalisation): products use scale * D
    sc = [0.5, 1.0, 1.25]
    mats = [dij.DijMatrix(ctx, s, scale=sc[s]) for s in range(3)]
    Ds = [_scipy(ctx, s) * sc[s] for s in range(3)]
    # the scenarios see different geometry: their matrices differ
    assert (Ds[0] - Ds[1]).count_nonzero() > 0
    rng = np.random.default_rng(4)
    w_np = rng.uniform(0.5, 1.5, mats[0].n_spots).astype(np.float32)
    p_np = _prescription(g)
    dev = torch.device("cuda", 0)
    w = torch.tensor(w_np, device=dev, requires_grad=True)
    fmax, vals = dij.worst_case_objective(mats, w, torch.tensor(p_np, device=dev))
    fmax.backward()
    ref = []
    for D in Ds:
        r = D @ w_np.astype(np.float64) - p_np
        ref.append((np.mean(r * r), 2.0 * (D.T @ r) / D.shape[0]))
    k = int(np.argmax([v for v, _ in ref]))
    assert np.allclose(vals.detach().cpu().numpy(), [v for v, _ in ref], rtol=1e-9, atol=0)
    g_ref = ref[k][1]
    assert np.abs(w.grad.detach().cpu().numpy() - g_ref).max() <= 1e-5 * np.abs(g_ref).max()


def test_fused_quadratic_with_voxel_weights_and_dose_path(ctx):
    """The fused objective (one pass over D w, no dose vector) with voxel
    weights against scipy in FP64 and against the objective formed from the
    differentiable dose; without a graph it forms no gradient buffer."""
    import torch
    from moquimc_amd import dij
    g, sp = _plan(ctx, scenarios=False)
    mat = dij.DijMatrix(ctx, 0, scale=0.75)
    D = _scipy(ctx, 0) * 0.75
    rng = np.random.default_rng(8)
    w_np = rng.uniform(0.5, 1.5, mat.n_spots).astype(np.float32)
    p_np = _prescription(g)
    c_np = rng.uniform(0.0, 2.0, D.shape[0]).astype(np.float32)
    dev = torch.device("cuda", 0)
    w = torch.tensor(w_np, device=dev, requires_grad=True)
    p, c = torch.tensor(p_np, device=dev), torch.tensor(c_np, device=dev)
    f = dij.quadratic_objective(mat, w, p, voxel_weights=c)
    f.backward()
    r = D @ w_np.astype(np.float64) - p_np
    f_ref = np.mean(c_np * r * r)
    g_ref = 2.0 * (D.T @ (c_np * r)) / D.shape[0]
    assert abs(f.item() - f_ref) <= 1e-9 * abs(f_ref)
    got = w.grad.detach().cpu().numpy().astype(np.float64)
    assert np.abs(got - g_ref).max() <= 1e-5 * np.abs(g_ref).max()
    # the same objective through the differentiable dose
    w2 = torch.tensor(w_np, device=dev, requires_grad=True)
    d = dij.dose(mat, w2) - p.double()
    f2 = (c.double() * d * d).mean()
    f2.backward()
    assert abs(f.item() - f2.item()) <= 1e-12 * abs(f2.item())
    assert torch.allclose(w.grad.double(), w2.grad.double(), rtol=1e-6, atol=1e-9 * float(np.abs(g_ref).max()))
    with torch.no_grad():
        f3, g3 = mat.quadratic(w, p, c, want_grad=False)
    assert g3 is None and f3.item() == f.item()
