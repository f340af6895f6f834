"""The oracle (CPU restatement) against the reference compiled from its own
sources (oracle/_ref): with the reference's RNG (std::default_random_engine
and libstdc++ distributions) and libm, whole transported histories agree bit
for bit, for both physics builds, every scorer and both kernels.
"""
import numpy as np
import pytest

import phantoms
import refbind as rb

needs_ref = pytest.mark.skipif(not rb.have_reference(), reason="oracle/_ref not built (make -C oracle)")


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _spot_vertices(debug, energy, n, master_seed=7, spot_index=3, sigma=3.0, sigma_e=1.0, angles=(0, 0, 0, 0)):
    sp = rb.make_spot(energy, sigma, n, pos=(0, 0, 0.5), energy_const=(sigma_e == 0), sigma_e=sigma_e)
    return rb.ref_sample_spot(sp, list(angles), [0, 0, 0], master_seed, spot_index, n, debug=debug)


@needs_ref
@pytest.mark.parametrize("debug", [1, 0])
@pytest.mark.parametrize("rho", [1.0e-3, 0.98822e-3, 0.26e-3, 1.85e-3])
def test_transport_bit_exact(debug, rho):
    g = phantoms.water_box(30, 30, 100, rho=rho)
    vtx, seeds = _spot_vertices(debug, 100.0, 200)
    for mask, kernel in [(rb.SC_DOSE | rb.SC_LETD, 0), (rb.SC_DOSE | rb.SC_STAT, 1), (rb.SC_DOSE, 0)]:
        ra, nr = rb.ref_transport(debug, g, vtx, seeds, mask=mask, kernel=kernel)
        oa, no = rb.oracle_transport(g, vtx, seeds, mask=mask, kernel=kernel, debug=debug)
        assert nr == no == 200
        for a, b in zip(ra, oa):
            assert np.array_equal(_bits(a), _bits(b))


@needs_ref
@pytest.mark.parametrize("debug", [1, 0])
def test_slab_transport_bit_exact(debug):
    g = phantoms.slab_box(24, 24, 200)
    vtx, seeds = _spot_vertices(debug, 180.0, 150, sigma_e=1.8)
    ra, _ = rb.ref_transport(debug, g, vtx, seeds, mask=rb.SC_DOSE | rb.SC_LETD, kernel=0)
    oa, _ = rb.oracle_transport(g, vtx, seeds, mask=rb.SC_DOSE | rb.SC_LETD, kernel=0, debug=debug)
    for a, b in zip(ra, oa):
        assert np.array_equal(_bits(a), _bits(b))


@needs_ref
@pytest.mark.parametrize("debug", [1, 0])
def test_lett_edep_dm_scorers_bit_exact(debug):
    """LETt numerator and denominator, energy deposit and dose to medium
    (primaries only), laid out after Dose and LETd and before the statistics
    pair, through both kernels."""
    g = phantoms.slab_box(24, 24, 200)
    vtx, seeds = _spot_vertices(debug, 160.0, 150, sigma_e=1.6)
    for mask, kernel in [(rb.SC_DOSE | rb.SC_LETT | rb.SC_EDEP | rb.SC_DM, 0),
                         (rb.SC_LETT | rb.SC_EDEP | rb.SC_STAT, 1)]:
        ra, nr = rb.ref_transport(debug, g, vtx, seeds, mask=mask, kernel=kernel)
        oa, no = rb.oracle_transport(g, vtx, seeds, mask=mask, kernel=kernel, debug=debug)
        assert nr == no == 150 and len(ra) == len(oa) == rb.n_scorers(mask)
        for a, b in zip(ra, oa):
            assert np.array_equal(_bits(a), _bits(b))
            assert np.any(a != 0)


@needs_ref
@pytest.mark.parametrize("debug", [1, 0])
def test_beamline_children_bit_exact(debug):
    """A range shifter and a rotated, translated voxelised aperture before a
    slab phantom (the world of tps_env::setup_world with beamline objects):
    the child loop, the frame round trips between children, the aperture's
    stop density and the secondaries made in the range shifter agree bit for
    bit."""
    g = phantoms.slab_box(24, 24, 160)
    sp = rb.make_spot(150.0, 6.0, 200, pos=(0, 0, 50.0), energy_const=False, sigma_e=1.5)
    vtx, seeds = rb.ref_sample_spot(sp, [0, 0, 0, 0], [0, 0, 0], 11, 2, 200, debug=debug)
    rb.set_beamline([phantoms.range_shifter(), phantoms.aperture()])
    try:
        for mask, kernel in [(rb.SC_DOSE | rb.SC_LETD, 0), (rb.SC_DOSE | rb.SC_STAT, 1)]:
            ra, nr = rb.ref_transport(debug, g, vtx, seeds, mask=mask, kernel=kernel)
            oa, no = rb.oracle_transport(g, vtx, seeds, mask=mask, kernel=kernel, debug=debug)
            assert nr == no == 200
            for a, b in zip(ra, oa):
                assert np.array_equal(_bits(a), _bits(b))
        # the aperture blocks: less dose than without the beamline, none far off axis
        rb.set_beamline([])
        free, _ = rb.oracle_transport(g, vtx, seeds, mask=rb.SC_DOSE, kernel=0, debug=debug)
        assert oa[0].sum() < free[0].sum()
    finally:
        rb.set_beamline([])


@needs_ref
def test_rotated_translated_grid_bit_exact():
    """Non-identity grid rotation and translation (daughter placement quirk included)."""
    g = phantoms.water_box(20, 20, 80)
    c, s = np.cos(0.3), np.sin(0.3)
    g.rot9 = np.array([c, -s, 0, s, c, 0, 0, 0, 1], np.float32)
    g.t3 = np.array([1.5, -2.0, 3.0], np.float32)
    vtx, seeds = _spot_vertices(1, 120.0, 100)
    ra, _ = rb.ref_transport(1, g, vtx, seeds, mask=rb.SC_DOSE, kernel=0)
    oa, _ = rb.oracle_transport(g, vtx, seeds, mask=rb.SC_DOSE, kernel=0, debug=1)
    assert np.array_equal(_bits(ra[0]), _bits(oa[0]))


@needs_ref
@pytest.mark.parametrize("uniform", [False, True])
@pytest.mark.parametrize("angles", [(0, 0, 0, 0), (0, 0, 0, 90), (15, 30, 10, 90)])
def test_spot_sampling_bit_exact(angles, uniform):
    """phsp_6d and phsp_6d_uniform beamlets with the coordinate transform."""
    sp = rb.make_spot(150.0, 4.0, 500, pos=(2.0, -1.0, 0.5), energy_const=False, sigma_e=1.5,
                      sigma_dir=(0.002, 0.003), corr=(0.3, -0.2), uniform=uniform)
    rot = rb.coordinate_rotation(angles)
    sp2 = sp.copy()
    sp2["rotation"] = rot
    v1, s1 = rb.ref_sample_spot(sp, list(angles), [0, 0, 0], 11, 5, 500)
    v2, s2 = rb.oracle_sample_spot(sp2, 11, 5, 500, rng=0)
    assert np.array_equal(v1.view(np.uint32), v2.view(np.uint32))
    assert np.array_equal(s1, s2)


@needs_ref
def test_vacuum_trace_bit_exact():
    g = phantoms.water_box(16, 18, 20, dx=2.0, dz=2.5, rho=0.0)
    rng = np.random.default_rng(5)
    n = 300
    v = np.zeros((n, 7), np.float32)
    v[:, 0] = 50
    lo = np.array([g.xe[0], g.ye[0], g.ze[0]])
    hi = np.array([g.xe[-1], g.ye[-1], g.ze[-1]])
    v[:, 1:4] = rng.uniform(lo - 10, hi + 10, (n, 3))
    d = (0.5 * (lo + hi) - v[:, 1:4]) + rng.normal(scale=4, size=(n, 3))
    v[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    a = rb.ref_trace(1, g, v)
    b = rb.oracle_trace(g, v, debug=1)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


@needs_ref
def test_deterministic_functions_bit_exact():
    lib_r = rb.reference(1)
    lib_o = rb.oracle()
    hu = np.arange(-1200, 3200, dtype=np.int16)
    a = np.zeros(hu.size, np.float32)
    b = np.zeros(hu.size, np.float32)
    lib_r.ref_hu_to_density(rb.ptr(hu, rb.ctypes.c_void_p), rb.ctypes.c_int64(hu.size), rb.ptr(a, rb.c_f))
    lib_o.oracle_hu_to_density(rb.ptr(hu, rb.ctypes.c_void_p), rb.ctypes.c_int64(hu.size), rb.ptr(b, rb.c_f))
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    rng = np.random.default_rng(2)
    rho = np.concatenate([rng.uniform(1e-7, 3e-3, 4000), [1e-3, 0.9995e-3, 1.0009e-3, 0.26e-3, 0.9e-3]]).astype(np.float32)
    ek = rng.uniform(0.0, 300.0, rho.size).astype(np.float32)
    for dbg in (0, 1):
        lr, lo = rb.reference(dbg), rb.oracle()
        x = np.zeros_like(rho)
        y = np.zeros_like(rho)
        lr.ref_spr(rb.ptr(rho, rb.c_f), rb.ptr(ek, rb.c_f), rb.ctypes.c_int64(rho.size), rb.ptr(x, rb.c_f))
        lo.oracle_spr(0, dbg, rb.ptr(rho, rb.c_f), rb.ptr(ek, rb.c_f), rb.ctypes.c_int64(rho.size), rb.ptr(y, rb.c_f))
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
        lr.ref_radiation_length(rb.ptr(rho, rb.c_f), rb.ctypes.c_int64(rho.size), rb.ptr(x, rb.c_f))
        lo.oracle_radiation_length(0, dbg, rb.ptr(rho, rb.c_f), rb.ctypes.c_int64(rho.size), rb.ptr(y, rb.c_f))
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    e = np.concatenate([np.linspace(0.0, 299.5, 3000), rng.uniform(0.05, 299, 2000)]).astype(np.float32)
    for r in (1e-3, 0.26e-3, 1.85e-3):
        x = np.zeros((e.size, 10), np.float32)
        y = np.zeros((e.size, 10), np.float32)
        rb.reference(0).ref_physics_probe(rb.ptr(e, rb.c_f), rb.ctypes.c_int64(e.size), rb.ctypes.c_float(r),
                                          rb.ptr(x, rb.c_f))
        rb.oracle().oracle_physics_probe(0, rb.ptr(e, rb.c_f), rb.ctypes.c_int64(e.size), rb.ctypes.c_float(r),
                                         rb.ptr(y, rb.c_f))
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    dirs = rng.normal(size=(3000, 3)).astype(np.float32)
    dirs[:200] = [0, 0, 1]
    dirs[200:400] = [0, 0, -1]
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True).astype(np.float32)
    th = rng.uniform(0, 0.5, 3000).astype(np.float32)
    th[::7] = 0
    ph = rng.uniform(0, 2 * np.pi, 3000).astype(np.float32)
    ph[::11] = 0
    x = np.zeros_like(dirs)
    y = np.zeros_like(dirs)
    rb.reference(0).ref_rotate_direction(rb.ptr(dirs, rb.c_f), rb.ptr(th, rb.c_f), rb.ptr(ph, rb.c_f),
                                         rb.ctypes.c_int64(3000), rb.ptr(x, rb.c_f))
    rb.oracle().oracle_rotate_direction(0, rb.ptr(dirs, rb.c_f), rb.ptr(th, rb.c_f), rb.ptr(ph, rb.c_f),
                                        rb.ctypes.c_int64(3000), rb.ptr(y, rb.c_f))
    assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@needs_ref
def test_grid_probe_bit_exact():
    xe = rb.uniform_edges(-10, 10, 20)
    ye = np.cumsum(np.r_[-7.0, np.full(12, 1.2)]).astype(np.float32)
    ze = rb.uniform_edges(-30, 0, 15)
    rng = np.random.default_rng(9)
    n = 5000
    p = rng.uniform([-12, -8, -32], [12, 8, 2], (n, 3)).astype(np.float32)
    snap = rng.random(n) < 0.3
    p[snap, 0] = xe[rng.integers(0, len(xe), snap.sum())] + rng.choice([0, 5e-4, -5e-4, 2e-3], snap.sum())
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[rng.random(n) < 0.1, 1] = 1e-5
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(np.float32)
    cells = np.stack([rng.integers(0, 20, n), rng.integers(0, 12, n), rng.integers(0, 15, n)], 1).astype(np.int32)
    lo = np.stack([xe[cells[:, 0]], ye[cells[:, 1]], ze[cells[:, 2]]], 1)
    hi = np.stack([xe[cells[:, 0] + 1], ye[cells[:, 1] + 1], ze[cells[:, 2] + 1]], 1)
    inside = (lo + rng.uniform(0, 1, (n, 3)) * (hi - lo)).astype(np.float32)   # intersect from inside its cell
    on = rng.random((n, 3)) < 0.2
    inside[on] = hi[on]
    for op in range(4):
        outs = []
        pp = inside if op == 1 else p
        for lib, fn in ((rb.reference(0), "ref_grid_probe"), (rb.oracle(), "oracle_grid_probe")):
            dist = np.zeros(n, np.float32)
            dd = np.zeros((n, 3), np.float32)
            cc = np.zeros((n, 3), np.int32)
            getattr(lib, fn)(op, 20, 12, 15, rb.ptr(xe, rb.c_f), rb.ptr(ye, rb.c_f), rb.ptr(ze, rb.c_f),
                             rb.ptr(np.ascontiguousarray(pp), rb.c_f), rb.ptr(d, rb.c_f), rb.ptr(cells, rb.c_i32),
                             rb.ctypes.c_int64(n), rb.ptr(dist, rb.c_f), rb.ptr(dd, rb.c_f), rb.ptr(cc, rb.c_i32))
            outs.append((dist, dd, cc))
        (a1, a2, a3), (b1, b2, b3) = outs
        assert np.array_equal(a1.view(np.uint32), b1.view(np.uint32)), op
        assert np.array_equal(a2.view(np.uint32), b2.view(np.uint32)), op
        assert np.array_equal(a3, b3), op


@needs_ref
@pytest.mark.parametrize("debug", [1, 0])
def test_roi_masks_bit_exact(debug):
    """CONTOUR ROIs from mask_to_roi's runs for the scorers and the statistics
    pair: the reference's roi_t gates (idx > 0 for the scorers, get_mask_idx
    >= 0 for the statistics pair, whose compact index the harness maps back
    with get_org_idx) and the oracle's agree bit for bit, and the ROI covers
    exactly the voxels a walk over the mask finds."""
    g = phantoms.slab_box(24, 24, 200)
    nvox = g.nx * g.ny * g.nz
    vtx, seeds = _spot_vertices(debug, 160.0, 150, sigma_e=1.6)
    m0, m1 = phantoms.roi_masks(g.nx, g.ny, g.nz)
    in0, in1 = phantoms.roi_inside(m0), phantoms.roi_inside(m1)
    cases = [(rb.SC_DOSE | rb.SC_LETD, 0), (rb.SC_DOSE | rb.SC_LETT | rb.SC_STAT, 1)]
    free = {c: rb.ref_transport(debug, g, vtx, seeds, mask=c[0], kernel=c[1])[0] for c in cases}
    lone = (m0 == 2) & ~in0   # 2s that no 1 opens: outside, yet in the beam
    assert lone.sum() == 5 and np.all(free[cases[0]][0][lone] > 0)
    rb.set_roi(0, nvox, m0)
    rb.set_roi(1, nvox, m1)
    try:
        for mask, kernel in cases:
            ra, nr = rb.ref_transport(debug, g, vtx, seeds, mask=mask, kernel=kernel)
            oa, no = rb.oracle_transport(g, vtx, seeds, mask=mask, kernel=kernel, debug=debug)
            assert nr == no == 150
            for k, (a, b) in enumerate(zip(ra, oa)):
                assert np.array_equal(_bits(a), _bits(b)), k
                stat = mask & rb.SC_STAT and k >= len(ra) - 2
                inside = (in1 if stat else in0).copy()
                if not stat:
                    inside[0] = False   # DIRECT leaves voxel 0 out of the scorers
                f = free[(mask, kernel)][k]
                assert np.array_equal(_bits(a[inside]), _bits(f[inside])), k
                assert not np.any(a[~(in1 if stat else in0)]), k
                assert np.any(a[inside] != 0), k
    finally:
        rb.set_roi(0, nvox)
        rb.set_roi(1, nvox)
