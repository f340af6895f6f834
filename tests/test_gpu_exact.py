"""GPU verification build against the CPU restatement and the reference.

The verification kernels (exact=1) use the same Philox stream and the same
portable math as the oracle with rng=1, math=1, so every history follows the
same trajectory; doses agree to FP64 summation order.
"""
import numpy as np
import pytest

import phantoms
import refbind as rb

pytestmark = pytest.mark.gpu


def test_math_probe_bitwise(ctx):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(1e-6, 10, 4000), rng.uniform(1e-30, 1e-3, 500), [1.0, 0.5, 2.0, 1e-300, 5e-324]])
    xs = np.concatenate([rng.uniform(-8, 8, 4000), rng.uniform(-700, 700, 500)])
    ang = np.concatenate([rng.uniform(0, 2 * np.pi, 4000), rng.uniform(-50, 50, 500)])
    cosv = rng.uniform(-1, 1, 4000)
    y = rng.uniform(-2, 2, len(x))
    lib = rb.oracle()
    cases = [(0, x, None), (1, xs, None), (2, x, y), (3, ang, None), (4, ang, None), (5, cosv, None), (6, x, None)]
    for fn, a, b in cases:
        a = np.ascontiguousarray(a, np.float64)
        bb = a if b is None else np.ascontiguousarray(b, np.float64)
        host = np.zeros_like(a)
        lib.oracle_math_probe(fn, rb.ptr(a, rb.c_d), rb.ptr(bb, rb.c_d), rb.ctypes.c_int64(len(a)), rb.ptr(host, rb.c_d))
        dev = ctx.math_probe(fn, a, bb)
        assert np.array_equal(dev.view(np.uint64), host.view(np.uint64)), fn
    # IEEE division and square root on the device
    a = rng.uniform(-1e3, 1e3, 5000).astype(np.float32).astype(np.float64)
    b = rng.uniform(1e-3, 1e3, 5000).astype(np.float32).astype(np.float64)
    dev = ctx.math_probe(7, a, b)
    assert np.array_equal(dev.astype(np.float32), (a.astype(np.float32) / b.astype(np.float32)))
    dev = ctx.math_probe(8, a, b)
    assert np.array_equal(dev, a / b)
    dev = ctx.math_probe(9, np.abs(a))
    assert np.array_equal(dev.astype(np.float32), np.sqrt(np.abs(a).astype(np.float32)))


def _vacuum_rays(g, n, seed):
    rng = np.random.default_rng(seed)
    v = np.zeros((n, 7), np.float32)
    v[:, 0] = 100.0
    lo = np.array([g.xe[0], g.ye[0], g.ze[0]])
    hi = np.array([g.xe[-1], g.ye[-1], g.ze[-1]])
    k = n // 2
    # rays starting inside the grid, some on edges
    p = rng.uniform(lo, hi, (k, 3))
    snap = rng.random((k, 3)) < 0.3
    p[snap[:, 0], 0] = g.xe[rng.integers(0, len(g.xe), snap[:, 0].sum())]
    p[snap[:, 2], 2] = g.ze[rng.integers(0, len(g.ze), snap[:, 2].sum())]
    # rays starting outside, aimed through the box
    q = rng.uniform(lo - 30, hi + 30, (n - k, 3))
    v[:k, 1:4] = p
    v[k:, 1:4] = q
    d = rng.normal(size=(n, 3))
    axis = rng.random(n) < 0.2
    d[axis] = np.eye(3)[rng.integers(0, 3, axis.sum())] * rng.choice([-1, 1], (axis.sum(), 1))
    d[k:] = (0.5 * (lo + hi) - q) + rng.normal(scale=5, size=(n - k, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    v[:, 4:7] = d
    return v


def _split_ref(ref_cnb, ref_rec, n):
    hist = ref_rec[:, 12].astype(np.int64)
    return [(ref_cnb[hist == h], ref_rec[hist == h][:, :12]) for h in range(n)]


def _assert_trace_equal(cnb, rec, nst, ref_split):
    for h, (rc, rr) in enumerate(ref_split):
        n = nst[h]
        assert n == len(rc), (h, n, len(rc))
        assert np.array_equal(cnb[h, :n], rc), h
        assert np.array_equal(rec[h, :n].view(np.uint32), rr.view(np.uint32)), h


@pytest.mark.parametrize("exact", [True, False])
def test_vacuum_traversal_bit_exact_vs_reference(ctx, exact):
    """Deterministic ray traversal: voxel sequence, exit distances and the
    in-place direction edits of grid3d::intersect, bit-exact against the
    compiled reference transport on a vacuum grid (production and
    verification kernels)."""
    g = phantoms.water_box(20, 24, 30, dx=2.5, dz=3.0, rho=0.0)
    rays = _vacuum_rays(g, 400, 11)
    ref = _split_ref(*rb.ref_trace(1, g, rays), len(rays))
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_scorers(rb.SC_DOSE)
    cfg = ctx.config(seed=5, exact=exact, batched=False)
    cnb, rec, nst = ctx.trace(cfg, len(rays), 200, vertices=rays)
    _assert_trace_equal(cnb, rec, nst, ref)


def _oracle_trace_philox(g, vtx, debug):
    cnb, rec = rb.oracle_trace(g, vtx, cap=400000, debug=debug, prec=32, rng=1, math=1)
    return _split_ref(cnb, rec, len(vtx))


@pytest.mark.parametrize("debug", [0, 1])
def test_exact_physics_trace_matches_oracle(ctx, debug):
    """Full physics, step by step: the verification kernels reproduce the CPU
    restatement (Philox stream, portable math) bit for bit on every step."""
    g = phantoms.slab_box(30, 30, 160)
    rng = np.random.default_rng(7)
    n = 48
    v = np.zeros((n, 7), np.float32)
    v[:, 0] = rng.uniform(60, 200, n)
    v[:, 1:3] = rng.normal(scale=4, size=(n, 2))
    v[:, 3] = 0.5
    v[:, 6] = -1.0
    ref = _oracle_trace_philox(g, v, debug)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_scorers(rb.SC_DOSE)
    cfg = ctx.config(seed=0, exact=True, batched=False, physics_debug=bool(debug))
    cnb, rec, nst = ctx.trace(cfg, n, 6000, vertices=v)
    _assert_trace_equal(cnb, rec, nst, ref)


@pytest.mark.parametrize("debug,batched", [(0, True), (1, True), (0, False)])
def test_exact_dose_letd_matches_oracle(ctx, debug, batched):
    """Spot sampling + transport + scoring on the slab phantom: doses and
    LETd equal the oracle's up to FP64 summation order."""
    g = phantoms.slab_box(30, 30, 200)
    sp = phantoms.spots([120.0, 180.0, 90.0], 300, sigma=4.0, sigma_e_frac=0.01,
                        positions=[(0, 0, 0.5), (5, -3, 0.5), (-6, 4, 0.5)])
    mask = rb.SC_DOSE | rb.SC_LETD
    ref, ntr = rb.oracle_run_spots(g, sp, mask=mask, kernel=2, prec=32, rng=1, math=1, debug=debug, master_seed=42,
                                   threads=4)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(sp)
    ctx.set_scorers(mask)
    res = ctx.run(ctx.config(seed=42, exact=True, batched=batched, physics_debug=bool(debug)))
    assert res.histories == ntr == 900
    for k in range(3):
        got = ctx.scorer(k)
        want = ref[k]
        fin = np.isfinite(want)
        assert np.array_equal(np.isfinite(got), fin)
        assert np.allclose(got[fin], want[fin], rtol=1e-9, atol=0), k


@pytest.mark.parametrize("batched", [True, False])
def test_exact_run_vertices_matches_oracle(ctx, batched):
    """Caller-made vertices with per-vertex spot ids (upload_vertices +
    scorer_offset_vector, mqi_upload_data.hpp:415-420, mqi_transport.hpp:150-154):
    every history keyed by (seed, vertex index, spot id) in both; doses and
    LETd equal the oracle's up to FP64 summation order."""
    g = phantoms.slab_box(30, 30, 200)
    rng = np.random.default_rng(17)
    n = 1200
    v = np.zeros((n, 7), np.float32)
    v[:, 0] = rng.uniform(70.0, 170.0, n)
    v[:, 1:3] = rng.normal(0.0, 3.0, (n, 2))
    v[:, 3] = 0.5
    d = np.c_[rng.normal(0.0, 0.02, (n, 2)), -np.ones(n)]
    v[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    sid = rng.integers(0, 3, n).astype(np.uint32)
    mask = rb.SC_DOSE | rb.SC_LETD
    ref, ntr = rb.oracle_transport(g, v, np.zeros(n, np.int32), spot_ids=sid, mask=mask, kernel=2, threads=4,
                                   prec=32, rng=1, math=1, debug=0, master_seed=23)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(phantoms.spots([100.0, 110.0, 120.0], 1))
    ctx.set_scorers(mask)
    ctx.zero_scorers()
    res = ctx.run_vertices(v, ctx.config(seed=23, exact=True, batched=batched), spot_ids=sid)
    assert res.histories == ntr == n
    for k in range(3):
        got, want = ctx.scorer(k), ref[k]
        fin = np.isfinite(want)
        assert np.array_equal(np.isfinite(got), fin)
        assert np.allclose(got[fin], want[fin], rtol=1e-9, atol=0), k


def test_fresh_context_first_run_binds_work_list():
    """Regression: the work-list arrays are (re)allocated by the run itself;
    the launch must bind them after that, also on a context's first run."""
    from moquimc_amd import _abi
    c = _abi.Context(0)
    try:
        g = phantoms.slab_box(30, 30, 200)
        sp = phantoms.spots([120.0, 180.0, 90.0], 300, sigma=4.0, sigma_e_frac=0.01,
                            positions=[(0, 0, 0.5), (5, -3, 0.5), (-6, 4, 0.5)])
        ref, _ = rb.oracle_run_spots(g, sp, mask=rb.SC_DOSE, kernel=2, prec=32, rng=1, math=1, debug=0,
                                     master_seed=42, threads=4)
        c.set_grid(g.xe, g.ye, g.ze, g.density)
        c.set_spots(sp)
        c.set_scorers(rb.SC_DOSE)
        for batched in (False, True):
            c.zero_scorers()
            res = c.run(c.config(seed=42, exact=True, batched=batched))
            assert res.histories == 900
            assert np.allclose(c.scorer(0), ref[0], rtol=1e-9, atol=0)
    finally:
        c.close()


def test_exact_uniform_phase_space_matches_oracle(ctx):
    """phsp_6d_uniform spots (the phantom driver's beam) through the same
    verification comparison."""
    g = phantoms.water_box(30, 30, 160)
    sp = np.array([rb.make_spot(130.0, 3.0, 400, pos=(0, 0, 0.5), uniform=True),
                   rb.make_spot(90.0, 2.0, 300, pos=(4, -3, 0.5), energy_const=False, sigma_e=0.9, uniform=True)],
                  dtype=rb.SPOT_DTYPE)
    ref, ntr = rb.oracle_run_spots(g, sp, mask=rb.SC_DOSE, kernel=2, prec=32, rng=1, math=1, debug=1, master_seed=7,
                                   threads=4)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE)
    res = ctx.run(ctx.config(seed=7, exact=True, batched=True, physics_debug=True))
    assert res.histories == ntr == 700
    assert np.allclose(ctx.scorer(0), ref[0], rtol=1e-9, atol=0)


def test_cpp_phantom_driver_matches_abi(ctx, tmp_path):
    """The C++ environment layer (mqi_hip::phantom_env through the phantom
    command line) and the ctypes path give the same dose for the same plan."""
    import os
    import subprocess
    exe = os.path.join(rb.ROOT, "moquimc_amd", "host", "phantom_env")
    prefix = str(tmp_path) + "/"
    r = subprocess.run([exe, "--output_prefix", prefix, "--lxyz", "30", "30", "120", "--pxyz", "0", "0", "-60",
                        "--nxyz", "30", "30", "120", "--spot_position", "0", "0", "0.5", "--spot_size", "3", "2",
                        "--spot_energy", "110", "0", "--histories", "600", "--random_seed", "5", "--exact", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    dose_cpp = np.fromfile(prefix + "dose.f64", np.float64)
    g = phantoms.water_box(30, 30, 120, rho=0.0)
    hu = np.zeros(30 * 30 * 120, np.int16)
    ctx.set_grid_hu(g.xe, g.ye, g.ze, hu)
    sp = np.array([rb.make_spot(110.0, 0.0, 600, pos=(0, 0, 0.5), uniform=True)], dtype=rb.SPOT_DTYPE)
    sp[0]["sigma"][:2] = (3.0, 2.0)
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE)
    res = ctx.run(ctx.config(seed=5, exact=True, batched=True, physics_debug=True))
    assert res.histories == 600
    assert dose_cpp.shape == ctx.scorer(0).shape
    assert np.allclose(dose_cpp, ctx.scorer(0), rtol=1e-9, atol=0)


@pytest.mark.parametrize("debug", [1, 0])
def test_exact_beamline_children_match_oracle(ctx, debug):
    """A range shifter and a rotated, translated voxelised aperture before the
    slab phantom: the verification kernels give the oracle's dose (which
    equals the reference's, tests/test_oracle_vs_reference.py), with the
    secondaries made in the range shifter placed by its transform."""
    g = phantoms.slab_box(24, 24, 160)
    sp = phantoms.spots([150.0, 130.0], 300, sigma=6.0, sigma_e_frac=0.01,
                        positions=[(0, 0, 50.0), (3, -2, 50.0)])
    bl = [phantoms.range_shifter(), phantoms.aperture()]
    rb.set_beamline(bl)
    try:
        ref, ntr = rb.oracle_run_spots(g, sp, mask=rb.SC_DOSE | rb.SC_LETD, kernel=2, prec=32, rng=1, math=1,
                                       debug=debug, master_seed=17, threads=4)
    finally:
        rb.set_beamline([])
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_beamline(bl)
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE | rb.SC_LETD)
    res = ctx.run(ctx.config(seed=17, exact=True, batched=True, physics_debug=bool(debug)))
    assert res.histories == ntr == 600
    for k in range(3):
        got, want = ctx.scorer(k), ref[k]
        fin = np.isfinite(want)
        assert np.array_equal(np.isfinite(got), fin)
        assert np.allclose(got[fin], want[fin], rtol=1e-9, atol=0), k
    # the production kernels through the same beamline: same physics, other
    # arithmetic; the aperture still cuts the off-axis dose
    ctx.zero_scorers()
    ctx.run(ctx.config(seed=17, batched=True, physics_debug=bool(debug)))
    d = ctx.scorer(0).reshape(160, 24, 24).sum(0)
    xc = np.arange(24) - 11.5
    r2 = xc[None, :] ** 2 + xc[:, None] ** 2
    assert d[r2 > 100].sum() < 0.2 * d.sum()
    ctx.set_beamline()


@pytest.mark.parametrize("debug", [0, 1])
def test_exact_lett_edep_dm_match_oracle(ctx, debug):
    """LETt numerator/denominator, energy deposit and dose to medium (primaries
    only) from the verification kernels equal the oracle's (which equal the
    reference's, tests/test_oracle_vs_reference.py) up to FP64 summation order."""
    g = phantoms.slab_box(30, 30, 200)
    sp = phantoms.spots([120.0, 170.0], 300, sigma=4.0, sigma_e_frac=0.01, positions=[(0, 0, 0.5), (4, -3, 0.5)])
    mask = rb.SC_DOSE | rb.SC_LETT | rb.SC_EDEP | rb.SC_DM
    ref, ntr = rb.oracle_run_spots(g, sp, mask=mask, kernel=2, prec=32, rng=1, math=1, debug=debug, master_seed=23,
                                   threads=4)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(sp)
    ctx.set_scorers(mask)
    assert ctx.scorer_count() == 5
    res = ctx.run(ctx.config(seed=23, exact=True, batched=True, physics_debug=bool(debug)))
    assert res.histories == ntr == 600
    for k in range(5):
        got, want = ctx.scorer(k), ref[k]
        fin = np.isfinite(want)
        assert np.array_equal(np.isfinite(got), fin)
        assert np.allclose(got[fin], want[fin], rtol=1e-9, atol=0), k
        assert np.any(want[fin] != 0), k
    # more than the kernel's six scorer slots is refused
    with pytest.raises(Exception):
        ctx.set_scorers(rb.SC_DOSE | rb.SC_LETD | rb.SC_LETT | rb.SC_STAT)
    ctx.set_scorers(rb.SC_DOSE)


@pytest.mark.gpu
@pytest.mark.parametrize("debug", [0, 1])
def test_exact_roi_masks_match_oracle(ctx, debug):
    """CONTOUR ROIs (mqi_set_roi) for the scorers and the statistics pair:
    the verification kernels equal the oracle (whose gates equal the
    reference's roi_t, tests/test_oracle_vs_reference.py) up to FP64
    summation order; the production kernels score the same voxels, every
    array is zero off its ROI and the Dij rows hold only ROI voxels."""
    g = phantoms.slab_box(24, 24, 200)
    nvox = g.nx * g.ny * g.nz
    sp = phantoms.spots([150.0, 175.0], 300, sigma=4.0, sigma_e_frac=0.01, positions=[(0, 0, 0.5), (3, -2, 0.5)])
    m0, m1 = phantoms.roi_masks(g.nx, g.ny, g.nz)
    in0, in1 = phantoms.roi_inside(m0), phantoms.roi_inside(m1)
    mask = rb.SC_DOSE | rb.SC_LETD | rb.SC_STAT
    rb.set_roi(0, nvox, m0)
    rb.set_roi(1, nvox, m1)
    try:
        ref, ntr = rb.oracle_run_spots(g, sp, mask=mask, kernel=2, prec=32, rng=1, math=1, debug=debug,
                                       master_seed=31, threads=4)
    finally:
        rb.set_roi(0, nvox)
        rb.set_roi(1, nvox)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(sp)
    ctx.set_scorers(mask)
    with pytest.raises(ValueError):
        ctx.set_roi(0, m0[:-1])
    with pytest.raises(Exception):
        ctx.set_roi(2, m0)
    ctx.set_roi(0, m0)
    ctx.set_roi(1, m1)
    res = ctx.run(ctx.config(seed=31, exact=True, batched=True, physics_debug=bool(debug)))
    assert res.histories == ntr == 600
    for k in range(5):
        got, want = ctx.scorer(k), ref[k]
        fin = np.isfinite(want)
        assert np.array_equal(np.isfinite(got), fin)
        assert np.allclose(got[fin], want[fin], rtol=1e-9, atol=0), k
        inside = in1 if k >= 3 else in0
        assert not np.any(got[~inside]) and np.any(got[inside] != 0), k
    if debug:
        return
    # production kernels: gated equals ungated on the ROI (same histories,
    # scoring does not steer transport), zero elsewhere
    ctx.zero_scorers()
    ctx.run(ctx.config(seed=5, batched=True))
    gated = [ctx.scorer(k) for k in range(5)]
    ctx.set_roi(0)
    ctx.set_roi(1)
    ctx.zero_scorers()
    ctx.run(ctx.config(seed=5, batched=True))
    for k in range(5):
        inside = (in1 if k >= 3 else in0).copy()
        if k < 3:
            inside[0] = False
        free = ctx.scorer(k)
        assert np.allclose(gated[k][inside], free[inside], rtol=1e-9, atol=0), k
        assert not np.any(gated[k][~(in1 if k >= 3 else in0)]), k
    # Dij goes through the scorers' ROI
    ctx.set_roi(0, m0)
    ctx.set_scorers(rb.SC_DOSE)
    ctx.zero_scorers()
    ctx.run(ctx.config(seed=6, batched=True, dij=True, dij_capacity=1 << 15))
    ctx.dij_build(0.0)
    indptr, idx, _, _ = ctx.dij_csr()
    assert indptr[-1] > 0 and np.all(in0[idx[:indptr[-1]]])
    ctx.set_roi(0)


@pytest.mark.gpu
def test_exact_robust_scenarios_match_oracle_on_shifted_and_scaled_grids(ctx):
    """Robust scenarios in one launch (mqi_set_scenarios): the verification
    kernels' dose of a scenario with an edge shift equals the oracle's on the
    grid whose edges are shifted by it (float32 e + s), and of a scenario with
    a density scale the oracle's on the density times it (float32), each with
    the scenario's own random streams; up to FP64 summation order."""
    g = phantoms.slab_box(30, 30, 200)
    sp = phantoms.spots([130.0, 160.0], 250, sigma=4.0, sigma_e_frac=0.01, positions=[(0, 0, 0.5), (3, -2, 0.5)])
    shifts = np.array([[0, 0, 0], [1.3, -0.7, 2.1], [0, 0, 0]], np.float32)
    scales = np.array([1.0, 1.0, 1.035], np.float32)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(sp)
    ctx.set_scenarios(shifts, scales)
    ctx.set_scorers(rb.SC_DOSE | rb.SC_LETD, per_scenario=True)
    res = ctx.run(ctx.config(seed=37, exact=True, batched=True))
    assert res.histories == 3 * 500
    grids = [g,
             rb.Grid(g.xe + shifts[1, 0], g.ye + shifts[1, 1], g.ze + shifts[1, 2], g.density),
             rb.Grid(g.xe, g.ye, g.ze, g.density * scales[2])]
    for s, gs in enumerate(grids):
        ref, ntr = rb.oracle_run_spots(gs, sp, mask=rb.SC_DOSE | rb.SC_LETD, kernel=2, prec=32, rng=1, math=1,
                                       debug=0, master_seed=37, threads=4, scenario=s)
        assert ntr == 500
        for k in range(3):
            got, want = ctx.scorer(k, s), ref[k]
            assert np.allclose(got, want, rtol=1e-9, atol=0), (s, k)
            assert np.any(want != 0)
    # the scenarios differ from each other
    assert not np.allclose(ctx.scorer(0, 1), ctx.scorer(0, 0))
    assert not np.allclose(ctx.scorer(0, 2), ctx.scorer(0, 0))
    ctx.set_scenarios()


@pytest.mark.parametrize("loops,kernel,mask", [
    (0, 0, rb.SC_DOSE | rb.SC_LETD),                  # transport_particles_patient, 3 scorers
    (0, 0, rb.SC_DOSE | rb.SC_LETD | rb.SC_STAT),     # the same with the statistics pair (5 scorers)
    (1, 1, rb.SC_DOSE | rb.SC_LETD | rb.SC_STAT),     # transport_particles_patient_stat
])
def test_exact_reference_scoring_loops_match_oracle(ctx, loops, kernel, mask):
    """The reference's two scoring loops (MQI_LOOPS_REFERENCE_PLAIN: scorers
    0..n-3 scored in both of the plain kernel's loops when n >= 3,
    mqi_transport.hpp:204-225; MQI_LOOPS_REFERENCE_STAT: the last two
    through get_mask_idx, :343-364) on the device equal the oracle's plain
    and _stat kernels; and the plain loop's scorers 0..n-3 are twice the
    once-per-step values (the double count itself)."""
    g = phantoms.slab_box(30, 30, 200)
    sp = phantoms.spots([120.0, 160.0], 300, sigma=4.0, sigma_e_frac=0.01, positions=[(0, 0, 0.5), (5, -3, 0.5)])
    ref, ntr = rb.oracle_run_spots(g, sp, mask=mask, kernel=kernel, prec=32, rng=1, math=1, debug=0, master_seed=44,
                                   threads=4)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(sp)
    ctx.set_scorers(mask)
    n_sc = ctx.scorer_count()
    assert n_sc == len(ref)
    res = ctx.run(ctx.config(seed=44, exact=True, batched=True, loops=loops))
    assert res.histories == ntr == 600
    got = [ctx.scorer(k) for k in range(n_sc)]
    for k in range(n_sc):
        fin = np.isfinite(ref[k])
        assert np.array_equal(np.isfinite(got[k]), fin), k
        assert np.allclose(got[k][fin], ref[k][fin], rtol=1e-9, atol=0), k
    if loops == 0:
        ctx.zero_scorers()
        ctx.run(ctx.config(seed=44, exact=True, batched=True, loops=2))
        for k in range(n_sc - 2):
            once = ctx.scorer(k)
            once[0] = 0.0   # voxel 0 is never scored by the reference's scorers
            assert np.allclose(got[k], 2.0 * once, rtol=1e-9, atol=0), k
