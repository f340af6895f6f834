"""The calibration and material functions on the device against the
reference compiled from its sources (oracle/_ref), value by value.

- HU -> density (patient_material_t::hu_to_density with its 3996-entry
  correction table, mqi_patient_materials.hpp:11-412,515-542) times the
  density scale of a robust scenario (tps_env::setup_world,
  mqi_tps_env.hpp:766-769), through mqi_set_grid_hu and through the
  MetaImage CT loader: every int16 HU in [-1200, 3200), scales 1, 1.035 and
  0.965, bit for bit.
- spr_default and radiation_length_default (mqi_patient_materials.hpp:
  415-473) over densities in every branch (below 0.0012 g/cm^3, up to
  0.26, the 0.26-0.9 blend, from 0.9 on, the phantom build's water
  shortcut), both physics builds: the verification kernels bit for bit;
  the production kernels (hardware log2 / exp2 pow and log) within the
  bounds stated in the test.
- Cross sections, dEdx and relativistic quantities (mqi_p_ionization.hpp:
  254-286, mqi_pp_elastic.hpp:221-235, mqi_relativistic_quantities.hpp):
  the verification kernels bit for bit, the production kernels within the
  stated bounds.
"""
import ctypes
import os

import numpy as np
import pytest

import refbind as rb

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_functions.npz")
needs_ref = pytest.mark.skipif(not rb.have_reference(), reason="oracle/_ref reference libraries not built")


def _ulp(a, b):
    """Distance in units in the last place between float32 arrays of one sign."""
    ia = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def _ref_hu(hu):
    h = np.ascontiguousarray(hu, np.int16)
    out = np.zeros(h.size, np.float32)
    rb.reference(0).ref_hu_to_density(rb.ptr(h, ctypes.c_void_p), ctypes.c_int64(h.size), rb.ptr(out, rb.c_f))
    return out


@needs_ref
@pytest.mark.parametrize("scale", [1.0, 1.035, 0.965])
def test_hu_to_density_on_device_equals_reference(ctx, scale):
    hu = np.arange(-1200, 3200, dtype=np.int16)
    n = hu.size   # 4400 = 40 x 11 x 10
    xe = np.arange(41, dtype=np.float32)
    ye = np.arange(12, dtype=np.float32)
    ze = np.arange(11, dtype=np.float32)
    ctx.set_grid_hu(xe, ye, ze, hu.reshape(10, 11, 40), density_scale=scale)
    got = ctx.density()
    want = _ref_hu(hu) * np.float32(scale)   # float product, as the reference's loop
    assert got.dtype == np.float32 and got.size == n
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, "HU %s: %s vs %s" % (hu[bad[:5]], got[bad[:5]], want[bad[:5]])
    # the committed fixture of the same function (reference, scale 1) agrees
    g = np.load(GOLD)
    assert np.array_equal(g["hu"], hu)
    assert np.array_equal((g["hu_density"] * np.float32(scale)).view(np.uint32), want.view(np.uint32))


@needs_ref
def test_ct_mha_density_equals_reference(ctx, tmp_path):
    """The MetaImage CT path: float32 HU truncated to int16 (read_ct_image,
    mqi_tps_env.hpp:615-701), density = reference hu_to_density x scale."""
    from test_io_formats import _write_ct
    rng = np.random.default_rng(11)
    hu = rng.uniform(-1100, 3100, (7, 9, 11)).astype(np.float32)
    hu[0, 0, :5] = [-1000.9, -0.5, 0.7, 2995.2, -1.0]
    p = tmp_path / "ct.mha"
    _write_ct(p, hu, (-5.0, -4.0, -7.0), (1.0, 1.0, 2.0))
    for scale in (1.0, 1.035):
        ctx.set_grid_ct_mha(str(p), density_scale=scale)
        got = ctx.density()
        want = _ref_hu(np.trunc(hu).astype(np.int16).ravel()) * np.float32(scale)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _material_inputs():
    g = np.load(GOLD)
    rng = np.random.default_rng(12)
    # g/mm^3: below the 0.0012 g/cm^3 cut, the linear part, the blend, the
    # Fippel fit, and the water shortcut of the phantom build
    extra = np.concatenate([rng.uniform(1e-8, 1.2e-6, 200), [0.0, 1.2e-6, 0.26e-3, 0.9e-3, 1e-3, 0.9991e-3,
                                                            1.0009e-3, 1.001e-3, 1.85e-3, 2.2e-3]])
    rho = np.concatenate([g["rho"], extra]).astype(np.float32)
    ek = np.concatenate([g["ek"], rng.uniform(0.0, 300.0, extra.size)]).astype(np.float32)
    return rho, ek


def _ref_spr_x0(dbg, rho, ek):
    L = rb.reference(dbg)
    s = np.zeros_like(rho)
    x = np.zeros_like(rho)
    L.ref_spr(rb.ptr(rho, rb.c_f), rb.ptr(ek, rb.c_f), ctypes.c_int64(rho.size), rb.ptr(s, rb.c_f))
    L.ref_radiation_length(rb.ptr(rho, rb.c_f), ctypes.c_int64(rho.size), rb.ptr(x, rb.c_f))
    return s, x


@needs_ref
@pytest.mark.parametrize("dbg", [0, 1])
def test_spr_and_radiation_length_on_device(ctx, dbg):
    rho, ek = _material_inputs()
    want_s, want_x = _ref_spr_x0(dbg, rho, ek)
    g = np.load(GOLD)
    n0 = g["rho"].size
    assert np.array_equal(want_s[:n0].view(np.uint32), g["spr_%d" % dbg].view(np.uint32))
    assert np.array_equal(want_x[:n0].view(np.uint32), g["x0_%d" % dbg].view(np.uint32))
    # verification kernels: the reference's values bit for bit
    s = ctx.physics_probe(0, rho, ek, exact=True, physics_debug=dbg)
    x = ctx.physics_probe(1, rho, exact=True, physics_debug=dbg)
    assert np.array_equal(s.view(np.uint32), want_s.view(np.uint32))
    assert np.array_equal(x.view(np.uint32), want_x.view(np.uint32))
    # production kernels: stated bounds (pow as exp2(y log2 x) and log on
    # the hardware; spr's (rho^-0.7 - 1) cancels near water, so its bound is
    # absolute, in units of 1e-6 of the ratio)
    s = ctx.physics_probe(0, rho, ek, exact=False, physics_debug=dbg)
    x = ctx.physics_probe(1, rho, exact=False, physics_debug=dbg)
    ds = np.abs(s.astype(np.float64) - want_s)
    ux = _ulp(x, want_x)
    print("production vs reference (debug %d): spr max |d| %.3g, radiation length max %d ulp" % (dbg, ds.max(),
                                                                                               ux.max()))
    assert np.array_equal(s == 0, want_s == 0)
    assert ds.max() <= 2e-6
    assert ux.max() <= 8


@needs_ref
@pytest.mark.parametrize("rho", [1e-3, 1.85e-3])
def test_cross_sections_dedx_relativistic_on_device(ctx, rho):
    g = np.load(GOLD)
    e = np.ascontiguousarray(g["probe_e"])
    want = np.zeros((e.size, 10), np.float32)
    rb.reference(0).ref_physics_probe(rb.ptr(e, rb.c_f), ctypes.c_int64(e.size), ctypes.c_float(rho),
                                      rb.ptr(want, rb.c_f))
    assert np.array_equal(want.view(np.uint32), g["probe_%g" % rho].view(np.uint32))
    r = np.full(e.size, rho, np.float32)
    got = ctx.physics_probe(2, r, e, exact=True)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # production: cross sections and dEdx are table interpolations with the
    # hardware divide (relative 1e-6); Et and gamma are the same float
    # expressions (2 ulp); beta^2 and Te_max use beta^2 = Ek (Ek + 2 Mp) /
    # Et^2 and the momentum sqrt(Ek (Ek + 2 Mp)), which avoid the reference's
    # cancellations in 1 - 1 / gamma^2 and Et^2 - Mp^2 at low energy:
    # compared with the closed form in double (relative 1e-5)
    got = ctx.physics_probe(2, r, e, exact=False)
    t = want.astype(np.float64)
    rel = np.abs(got[:, :5] - t[:, :5]) / np.maximum(np.abs(t[:, :5]), 1e-30)
    print("production: cross sections / dEdx max rel %.3g" % rel.max())
    assert rel.max() <= 1e-6
    assert _ulp(got[:, 6:8], want[:, 6:8]).max() <= 2
    mp, me = 938.272046, 0.510998928
    ek = e.astype(np.float64)
    et = ek + mp
    b2 = ek * (ek + 2 * mp) / et ** 2
    gam = et / mp
    tmax = 2 * me * b2 * gam ** 2 / (1 + 2 * gam * me / mp + (me / mp) ** 2)
    mom = np.sqrt(et ** 2 - mp ** 2)
    live = ek > 0
    for k, exact_v in ((5, b2), (8, tmax), (9, mom)):
        d = np.abs(got[live, k] - exact_v[live]) / exact_v[live]
        print("production column %d: max rel %.3g vs the closed form in double" % (k, d.max()))
        assert d.max() <= 1e-5
