"""Dose-influence (Dij) matrices and their sparse products on the GPU.

Pins: the CSR rows of a run sum to the dense dose of the same run; with the
verification kernels each row equals the oracle's dose of that spot alone;
the row threshold drops exactly the entries below rel * row max; D w and
D^T g equal an FP64 CPU product (scipy) to 1e-6 relative.
"""
import numpy as np
import pytest
import scipy.sparse as sps

import phantoms
import refbind as rb

pytestmark = pytest.mark.gpu


def _setup(ctx):
    g = phantoms.slab_box(24, 24, 120)
    energies = [90.0, 110.0, 70.0, 100.0, 85.0, 95.0]
    pos = [(-4, -4, 0.5), (4, -4, 0.5), (0, 0, 0.5), (-4, 4, 0.5), (4, 4, 0.5), (0, 6, 0.5)]
    sp = phantoms.spots(energies, 150, sigma=3.0, sigma_e_frac=0.01, positions=pos)
    sp[2]["histories"] = 0   # an empty spot keeps its (empty) row
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_scenarios()
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE)
    return g, sp


def _csr(ctx, scen=0):
    indptr, idx, val, shape = ctx.dij_csr(scen)
    return sps.csr_matrix((val.astype(np.float64), idx, indptr), shape=shape)


def test_dij_rows_sum_to_dose(ctx):
    g, sp = _setup(ctx)
    res = ctx.run(ctx.config(seed=3, batched=True, dij=True, dij_capacity=1 << 15))
    assert res.status == 0
    ctx.dij_build(0.0)
    D = _csr(ctx)
    assert D.shape == (len(sp), g.nvox)
    assert D[2].nnz == 0
    dose = ctx.scorer(0)
    col = np.asarray(D.sum(axis=0)).ravel()
    assert np.allclose(col, dose, rtol=2e-6, atol=1e-30)
    # columns sorted within every row
    for r in range(D.shape[0]):
        c = D.indices[D.indptr[r]:D.indptr[r + 1]]
        assert np.all(np.diff(c) > 0)


def test_dij_rows_match_oracle_per_spot(ctx):
    g, sp = _setup(ctx)
    ctx.run(ctx.config(seed=9, batched=True, exact=True, dij=True, dij_capacity=1 << 15))
    ctx.dij_build(0.0)
    D = _csr(ctx)
    for j in range(len(sp)):
        one = sp.copy()
        for k in range(len(sp)):
            if k != j:
                one[k]["histories"] = 0
        ref, _ = rb.oracle_run_spots(g, one, mask=rb.SC_DOSE, kernel=2, prec=32, rng=1, math=1, debug=0,
                                     master_seed=9, threads=4)
        row = np.asarray(D[j].todense()).ravel()
        want = ref[0].astype(np.float32).astype(np.float64)
        assert np.allclose(row, want, rtol=1e-6, atol=0), j


def test_dij_threshold(ctx):
    g, sp = _setup(ctx)
    ctx.run(ctx.config(seed=4, batched=True, dij=True, dij_capacity=1 << 15))
    ctx.dij_build(0.0)
    full = _csr(ctx)
    ctx.dij_build(1e-2)
    cut = _csr(ctx)
    for r in range(full.shape[0]):
        v = full.data[full.indptr[r]:full.indptr[r + 1]]
        if v.size == 0:
            assert cut.indptr[r + 1] == cut.indptr[r]
            continue
        keep = v >= v.max() * 1e-2 * (1 - 1e-7)
        assert cut.indptr[r + 1] - cut.indptr[r] in (keep.sum(), keep.sum() - 1, keep.sum() + 1)


@pytest.mark.parametrize("histories", [150, 2])
def test_dij_products_fp64(ctx, histories):
    """150 histories per spot: rows of thousands of entries (D^T g over
    16-bit tile-relative columns); 2: rows of a few hundred (the packed
    CSR-Adaptive segments with 32-bit columns)."""
    import torch
    g, sp = _setup(ctx)
    sp["histories"] = np.where(sp["histories"] > 0, histories, 0)
    ctx.set_spots(sp)
    ctx.run(ctx.config(seed=5, batched=True, dij=True, dij_capacity=1 << 15))
    ctx.dij_build(1e-4)
    D = _csr(ctx)   # spots x voxels (rows = spots)
    assert (D.nnz >= 1024 * D.shape[0]) == (histories == 150)
    rng = np.random.default_rng(1)
    w = rng.uniform(0.5, 1.5, D.shape[0]).astype(np.float32)
    gv = rng.normal(size=D.shape[1]).astype(np.float32)
    dev = torch.device("cuda", 0)
    tw = torch.from_numpy(w).to(dev)
    tg = torch.from_numpy(gv).to(dev)
    dose = torch.zeros(D.shape[1], dtype=torch.float64, device=dev)
    grad = torch.zeros(D.shape[0], dtype=torch.float64, device=dev)
    ctx.dij_matvec(0, 0, tw.data_ptr(), dose.data_ptr(), y_double=True)
    ctx.dij_matvec(0, 1, tg.data_ptr(), grad.data_ptr(), y_double=True)
    torch.cuda.synchronize()
    want_dose = D.T @ w.astype(np.float64)
    want_grad = D @ gv.astype(np.float64)
    assert np.allclose(dose.cpu().numpy(), want_dose, rtol=1e-6, atol=1e-6 * np.abs(want_dose).max())
    assert np.allclose(grad.cpu().numpy(), want_grad, rtol=1e-6, atol=1e-6 * np.abs(want_grad).max())


def test_dij_products_long_rows_and_accumulate(ctx):
    """Rows long enough to be cut into several pieces (> 16384 entries), an
    empty row, float32 output and accumulate / alpha against scipy in FP64."""
    import torch
    g = phantoms.slab_box(60, 60, 150)
    sp = phantoms.spots([150.0, 120.0, 130.0], 20000, sigma=5.0, sigma_e_frac=0.01,
                        positions=[(0, 0, 0.5), (8, 5, 0.5), (-6, -4, 0.5)])
    sp[1]["histories"] = 0
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_scenarios()
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE)
    res = ctx.run(ctx.config(seed=8, batched=True, dij=True, dij_capacity=1 << 18))
    assert res.status == 0
    ctx.dij_build(0.0)
    D = _csr(ctx)
    assert np.diff(D.indptr).max() > 3 * 16384
    rng = np.random.default_rng(2)
    w = rng.uniform(0.5, 1.5, D.shape[0]).astype(np.float32)
    gv = rng.normal(size=D.shape[1]).astype(np.float32)
    dev = torch.device("cuda", 0)
    tw, tg = torch.from_numpy(w).to(dev), torch.from_numpy(gv).to(dev)
    want_dose = D.T @ w.astype(np.float64)
    want_grad = D @ gv.astype(np.float64)
    y0 = rng.normal(size=D.shape[0])
    grad = torch.from_numpy(y0.copy()).to(dev)
    ctx.dij_matvec(0, 1, tg.data_ptr(), grad.data_ptr(), y_double=True, alpha=-2.0, accumulate=True)
    dose32 = torch.zeros(D.shape[1], dtype=torch.float32, device=dev)
    ctx.dij_matvec(0, 0, tw.data_ptr(), dose32.data_ptr(), y_double=False)
    torch.cuda.synchronize()
    assert ctx.last_matvec_ms() > 0
    got = grad.cpu().numpy()
    assert np.allclose(got, y0 - 2.0 * want_grad, rtol=1e-6, atol=1e-9 * np.abs(want_grad).max())
    assert np.allclose(dose32.cpu().numpy(), want_dose, rtol=2e-6, atol=1e-6 * np.abs(want_dose).max())
    # the empty spot row
    assert D[1].nnz == 0


def test_scenario_budgets_and_per_scenario_dij(ctx):
    """Per-scenario history budgets: the nominal scenario gets twice the
    histories; each scenario's Dij rows sum to that scenario's dose."""
    g = phantoms.slab_box(24, 24, 120)
    sp = phantoms.spots([90.0, 110.0], 400, sigma=3.0, sigma_e_frac=0.01, positions=[(0, 0, 0.5), (4, -2, 0.5)])
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_scenarios(np.array([[0, 0, 0], [1.5, 0, 0], [0, 0, 0]], np.float32),
                      np.array([1.0, 1.0, 1.05], np.float32), np.array([2.0, 1.0, 0.5], np.float32))
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE, per_scenario=True)
    res = ctx.run(ctx.config(seed=14, batched=True, dij=True, dij_capacity=1 << 15))
    assert res.histories == 2 * (800 + 400 + 200)
    ctx.dij_build(0.0)
    for s in range(3):
        D = _csr(ctx, s)
        col = np.asarray(D.sum(axis=0)).ravel()
        dose = ctx.scorer(0, scenario=s)
        assert np.allclose(col, dose, rtol=2e-6, atol=1e-30), s
    with pytest.raises(Exception):
        ctx.set_scenarios(np.zeros((2, 3)), np.ones(2), np.array([1.0, -1.0]))


def test_sqrt_sized_tables_match_fixed_tables(ctx):
    """Tables sized dij_slots_sqrt * sqrt(histories) per (scenario, spot) give
    the same matrix as one fixed capacity (verification kernels, same
    streams); a table too small for its spot is reported as an overflow."""
    g, sp = _setup(ctx)
    sp["histories"] = [40, 150, 0, 600, 150, 90]
    ctx.set_spots(sp)
    mats = []
    for kw in (dict(dij_capacity=1 << 15), dict(dij_slots_sqrt=1024)):
        ctx.zero_scorers()
        res = ctx.run(ctx.config(seed=12, batched=True, exact=True, dij=True, **kw))
        assert res.status == 0
        ctx.dij_build(0.0)
        mats.append(_csr(ctx))
    a, b = mats
    assert np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
    assert np.allclose(a.data, b.data, rtol=1e-6, atol=0)
    with pytest.raises(RuntimeError, match="overflowed"):
        ctx.run(ctx.config(seed=12, batched=True, dij=True, dij_slots_sqrt=1, dij_capacity=64))


def test_save_npz_and_dense_formats(ctx, tmp_path):
    """mqi_dij_save_npz / mqi_save_scorer: the reference's .npz (rows =
    spots, float64 value * scale, uint32 indices) read by scipy equals the
    device CSR; raw / mhd / mha hold the scorer * scale."""
    g, sp = _setup(ctx)
    ctx.run(ctx.config(seed=6, batched=True, dij=True, dij_capacity=1 << 15))
    ctx.dij_build(1e-4)
    D = _csr(ctx)
    path = tmp_path / "beam_0_Dij.npz"
    ctx.dij_save_npz(str(path), scale=1.5)
    got = sps.load_npz(path)
    assert got.shape == D.shape and np.array_equal(got.indptr, D.indptr) and np.array_equal(got.indices, D.indices)
    assert np.array_equal(got.data, D.data * 1.5)
    dose = ctx.scorer(0)
    for fmt in ("raw", "mhd", "mha"):
        ctx.save_scorer(0, str(tmp_path), "Dose_" + fmt, fmt=fmt, scale=2.0)
    assert np.array_equal(np.fromfile(tmp_path / "Dose_raw.raw", np.float64), dose * 2.0)
    assert np.array_equal(np.fromfile(tmp_path / "Dose_mhd.raw", np.float64), dose * 2.0)
    blob = open(tmp_path / "Dose_mha.mha", "rb").read()
    start = blob.index(b"ElementDataFile = LOCAL\n") + len(b"ElementDataFile = LOCAL\n")
    assert np.array_equal(np.frombuffer(blob[start:], np.float64), dose * 2.0)
    with pytest.raises(Exception):
        ctx.save_scorer(0, str(tmp_path / "no_such_dir"), "x", fmt="raw")


def test_forward_product_is_reproducible_across_builds(ctx):
    """D w runs on a voxel-major copy whose voxel rows list their spots in
    ascending order (a stable sort, not atomic cursors), so two
    mqi_dij_build calls on the same tables give bit-identical D w, and
    repeated products give the same bits."""
    import torch
    g, sp = _setup(ctx)
    sp = np.concatenate([sp, sp])   # more spots per voxel row
    sp["histories"] = 400
    ctx.set_spots(sp)
    ctx.run(ctx.config(seed=4, batched=True, dij=True, dij_capacity=1 << 15))
    dev = torch.device("cuda", 0)
    w = torch.from_numpy(np.random.default_rng(1).uniform(0.5, 1.5, len(sp)).astype(np.float32)).to(dev)
    outs = []
    for _ in range(2):
        ctx.dij_build(0.0)
        for _ in range(2):
            d = torch.zeros(g.nvox, dtype=torch.float64, device=dev)
            ctx.dij_matvec(0, 0, w.data_ptr(), d.data_ptr(), y_double=True)
            torch.cuda.synchronize()
            outs.append(d.cpu().numpy())
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint64), outs[0].view(np.uint64))
    D = _csr(ctx)
    assert np.allclose(outs[0], D.T @ w.cpu().numpy().astype(np.float64), rtol=1e-9, atol=0)


def _adversarial_long(rng):
    """Long rows (the pieces form of D^T g): empty and single-entry rows,
    runs of consecutive columns, deltas at and around the one-byte escape
    (254 / 255 / 256), rows whose every delta is an escape (several escape
    reloads per batch), deltas of exactly 65535 and steps above it (new
    pieces), lengths around the 256-nonzero batch and 4096-nonzero piece."""
    cols = 2_000_000
    rows = []
    rows.append(np.zeros(0, np.int64))
    rows.append(np.array([cols - 1]))
    rows.append(np.arange(4096) + 17)
    rows.append(np.arange(5000) * 300 + 3)
    rows.append(np.cumsum(np.tile([254, 255, 256, 1], 750)))
    rows.append(np.concatenate([np.arange(500) + k * 100_000 for k in range(10)]))
    rows.append(np.sort(rng.choice(cols, 257, replace=False)))
    rows.append(np.sort(rng.choice(cols, 12289, replace=False)))
    rows.append(np.arange(30) * 65535 + 5)
    rows.append(np.arange(20000) + 700_000)
    d = rng.geometric(0.3, 9000)
    far = rng.random(d.size) < 0.03
    d[far] = rng.integers(255, 70_000, int(far.sum()))
    rows.append(np.cumsum(d))
    rows.append(np.cumsum(np.r_[np.ones(2048, np.int64) * 300, np.ones(2049, np.int64)]))
    rows = [r[r < cols] for r in rows]
    return rows, cols


def _to_csr(rows, cols, rng):
    indptr = np.r_[0, np.cumsum([len(r) for r in rows])].astype(np.int64)
    indices = np.concatenate(rows).astype(np.int32) if indptr[-1] else np.zeros(0, np.int32)
    values = rng.uniform(1e-6, 1.0, indices.size).astype(np.float32)
    return sps.csr_matrix((values, indices, indptr), shape=(len(rows), cols))


def _check_products(ctx, D, scen, rng):
    import torch
    dev = torch.device("cuda", 0)
    w = rng.uniform(0.5, 1.5, D.shape[0]).astype(np.float32)
    g = rng.normal(size=D.shape[1]).astype(np.float32)
    tw, tg = torch.from_numpy(w).to(dev), torch.from_numpy(g).to(dev)
    y0 = rng.normal(size=D.shape[0])
    grad = torch.from_numpy(y0.copy()).to(dev)
    ctx.dij_matvec(scen, 1, tg.data_ptr(), grad.data_ptr(), y_double=True, alpha=1.5, accumulate=True)
    dose = torch.zeros(D.shape[1], dtype=torch.float32, device=dev)
    ctx.dij_matvec(scen, 0, tw.data_ptr(), dose.data_ptr(), y_double=False)
    torch.cuda.synchronize()
    D64 = D.astype(np.float64)
    want_grad = y0 + 1.5 * (D64 @ g.astype(np.float64))
    want_dose = D64.T @ w.astype(np.float64)
    got_grad, got_dose = grad.cpu().numpy(), dose.cpu().numpy()
    assert np.allclose(got_grad, want_grad, rtol=1e-6, atol=1e-9 * max(1.0, np.abs(want_grad).max()))
    assert np.allclose(got_dose, want_dose, rtol=2e-6, atol=1e-6 * max(1e-30, np.abs(want_dose).max()))
    return got_grad, got_dose


@pytest.mark.parametrize("kind", ["long", "short", "many_rows", "dense_voxels"])
def test_set_csr_products_match_fp64(ctx, kind, tmp_path):
    """A caller's CSR through mqi_dij_set_csr: both products and the fused
    quadratic objective against scipy in FP64 on matrices built to reach
    every path of the kernels (long rows: the byte-delta pieces with
    escapes, escape reloads and 65535-column steps; short rows: the
    compacted CSR-Adaptive form; more than 65536 rows: 32-bit spot columns
    in D w; voxels with more than 1024 spots: the split rows of D w), the
    matrix read back unchanged, and the same matrix through the .npz writer
    and loader giving the same results bit for bit."""
    rng = np.random.default_rng({"long": 1, "short": 2, "many_rows": 3, "dense_voxels": 6}[kind])
    if kind == "long":
        rows, cols = _adversarial_long(rng)
        D = _to_csr(rows, cols, rng)
        assert D.nnz >= 1024 * D.shape[0]
    elif kind == "short":
        D = sps.random(3000, 100_000, density=2e-4, format="csr", dtype=np.float32, random_state=4)
        D.data = rng.uniform(1e-6, 1.0, D.nnz).astype(np.float32)
    elif kind == "many_rows":
        D = sps.random(70_000, 5000, density=4e-4, format="csr", dtype=np.float32, random_state=5)
        D.data = rng.uniform(1e-6, 1.0, D.nnz).astype(np.float32)
    else:
        D = sps.random(3000, 60, density=0.8, format="csr", dtype=np.float32, random_state=7)
        D.data = rng.uniform(1e-6, 1.0, D.nnz).astype(np.float32)
        assert np.diff(D.tocsc().indptr).max() > 1024
    D.sort_indices()
    ctx.dij_set_csr(D.indptr, D.indices, D.data, D.shape, scenario=0)
    assert ctx.dij_info(0) == (D.nnz, D.shape[0], D.shape[1])
    ip, ix, v, shape = ctx.dij_csr(0)
    assert shape == D.shape and np.array_equal(ip, D.indptr) and np.array_equal(ix, D.indices)
    assert np.array_equal(v.view(np.uint32), D.data.view(np.uint32))
    a = _check_products(ctx, D, 0, np.random.default_rng(9))
    path = str(tmp_path / "d.npz")
    ctx.dij_save_npz(path, 0)
    ctx.dij_load_npz(path, scenario=1)
    b = _check_products(ctx, D, 1, np.random.default_rng(9))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # the fused objective and its dose gradient
    import torch
    from moquimc_amd import dij
    dev = torch.device("cuda", 0)
    w_np = rng.uniform(0.5, 1.5, D.shape[0]).astype(np.float32)
    p_np = rng.uniform(0.0, 2.0, D.shape[1]) * float(np.abs(D.T @ w_np.astype(np.float64)).mean())
    c_np = rng.uniform(0.0, 2.0, D.shape[1]).astype(np.float32)
    mat = dij.DijMatrix(ctx, 0, scale=1.5)
    f, g = mat.quadratic(torch.from_numpy(w_np).to(dev), torch.from_numpy(p_np).to(dev),
                         torch.from_numpy(c_np).to(dev), want_grad=True)
    r = 1.5 * (D.T.astype(np.float64) @ w_np.astype(np.float64)) - p_np
    f_ref = np.mean(c_np * r * r)
    g_ref = 2.0 * c_np * r / D.shape[1]
    assert abs(f.item() - f_ref) <= 1e-9 * abs(f_ref)
    assert np.allclose(g.cpu().numpy(), g_ref, rtol=2e-6, atol=1e-6 * np.abs(g_ref).max())


def test_quadratic_rejects_bad_input(ctx):
    """The fused objective: a scenario without a Dij is a state error and a
    prescription of the wrong length an argument error."""
    import torch
    from moquimc_amd import dij
    from moquimc_amd._abi import MqiError
    D = sps.random(20, 300, density=0.3, format="csr", dtype=np.float32, random_state=1)
    ctx.dij_set_csr(D.indptr, D.indices, D.data, D.shape, scenario=0)
    dev = torch.device("cuda", 0)
    w = torch.ones(20, device=dev)
    f = torch.empty((), dtype=torch.float64, device=dev)
    p = torch.zeros(300, dtype=torch.float64, device=dev)
    with pytest.raises(MqiError, match="no Dij"):
        ctx.dij_quadratic(5, w.data_ptr(), p.data_ptr(), 0, 1.0, 0, f.data_ptr())
    with pytest.raises(ValueError):
        dij.DijMatrix(ctx, 0).quadratic(w, torch.zeros(299, device=dev))


def test_set_csr_rejects_malformed(ctx):
    """Unsorted or repeated columns, columns out of range and row pointers
    that do not start at 0 or decrease are argument errors naming the entry."""
    from moquimc_amd._abi import MqiError
    ip = np.array([0, 2, 4], np.int64)
    v = np.ones(4, np.float32)
    for idx, ipp, msg in [([1, 0, 2, 3], ip, "ascend"), ([0, 1, 2, 2], ip, "ascend"), ([0, 1, 2, 10], ip, "cols"),
                          ([0, 1, 2, -1], ip, "cols"), ([0, 1, 2, 3], np.array([1, 2, 4], np.int64), "indptr"),
                          ([0, 1, 2, 3], np.array([0, 3, 2], np.int64), "indptr")]:
        with pytest.raises(MqiError, match=msg):
            ctx.dij_set_csr(ipp, np.array(idx, np.int32), v, (2, 10))


def test_load_npz_in_reference_row_order(ctx, tmp_path):
    """A Dij .npz laid out as the reference writes it (rows in hash-table
    order, uint32 indices and pointers, float64 values) loads as the
    canonical matrix: both products equal scipy's FP64 products of it."""
    rng = np.random.default_rng(12)
    rows, cols = 30, 20000
    lens = rng.integers(0, 3000, rows)
    indptr = np.r_[0, np.cumsum(lens)].astype(np.uint32)
    indices = np.concatenate([rng.permutation(rng.choice(cols, n, replace=False)) for n in lens]).astype(np.uint32)
    data = rng.uniform(1e-6, 1.0, indices.size)
    path = str(tmp_path / "ref.npz")
    np.savez(path, indices=indices, indptr=indptr, shape=np.array([rows, cols], np.uint32), data=data,
             format=np.array("csr"))
    ctx.dij_load_npz(path, scenario=0)
    D = sps.csr_matrix((data.astype(np.float32), indices.astype(np.int64), indptr.astype(np.int64)),
                       shape=(rows, cols))
    D.sort_indices()
    assert ctx.dij_info(0) == (D.nnz, rows, cols)
    _check_products(ctx, D, 0, np.random.default_rng(13))
