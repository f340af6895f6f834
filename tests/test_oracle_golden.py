"""The oracle against the committed golden fixtures of tests/golden (written
by tests/golden/make_golden.py from the reference compiled from its own
sources).  These run without the reference: they pin the oracle (reference
RNG and libm mode) bit for bit on every machine, and the committed tables of
moquimc_amd/csrc/mqi_tables.inc against the tables the reference compiles.
"""
import os
import re

import numpy as np
import pytest

import phantoms
import refbind as rb

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TABLES_INC = os.path.join(rb.ROOT, "moquimc_amd", "csrc", "mqi_tables.inc")


def _load(name):
    return np.load(os.path.join(GOLD, name))


def _u32(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _u64(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def _parse_tables():
    text = open(TABLES_INC).read()
    out = {}
    for name, body in re.findall(r"const float (\w+)\[\d+\] = \{(.*?)\};", text, re.S):
        vals = [float.fromhex(t.strip().rstrip("f")) for t in body.split(",") if t.strip()]
        out[name] = np.array(vals, np.float32)
    return out


def test_tables_equal_reference():
    g = _load("reference_tables.npz")
    t = _parse_tables()
    pairs = {"cs_p_ion": "kCsPIon", "dedx_restricted": "kDedxRestricted", "range_steps": "kRangeSteps",
             "cs_pp_e": "kCsPPElastic", "cs_po_e": "kCsPOElastic", "cs_po_i": "kCsPOInelastic",
             "density_correction": "kDensityCorrection"}
    for a, b in pairs.items():
        assert np.array_equal(_u32(g[a]), _u32(t[b])), a


def test_deterministic_functions():
    g = _load("reference_functions.npz")
    lib = rb.oracle()
    hu = np.ascontiguousarray(g["hu"])
    y = np.zeros(hu.size, np.float32)
    lib.oracle_hu_to_density(rb.ptr(hu, rb.ctypes.c_void_p), rb.ctypes.c_int64(hu.size), rb.ptr(y, rb.c_f))
    assert np.array_equal(_u32(y), _u32(g["hu_density"]))
    rho, ek = np.ascontiguousarray(g["rho"]), np.ascontiguousarray(g["ek"])
    for dbg in (0, 1):
        y = np.zeros_like(rho)
        lib.oracle_spr(0, dbg, rb.ptr(rho, rb.c_f), rb.ptr(ek, rb.c_f), rb.ctypes.c_int64(rho.size), rb.ptr(y, rb.c_f))
        assert np.array_equal(_u32(y), _u32(g["spr_%d" % dbg])), dbg
        lib.oracle_radiation_length(0, dbg, rb.ptr(rho, rb.c_f), rb.ctypes.c_int64(rho.size), rb.ptr(y, rb.c_f))
        assert np.array_equal(_u32(y), _u32(g["x0_%d" % dbg])), dbg
    e = np.ascontiguousarray(g["probe_e"])
    for r in (1e-3, 1.85e-3):
        y = np.zeros((e.size, 10), np.float32)
        lib.oracle_physics_probe(0, rb.ptr(e, rb.c_f), rb.ctypes.c_int64(e.size), rb.ctypes.c_float(r),
                                 rb.ptr(y, rb.c_f))
        assert np.array_equal(_u32(y), _u32(g["probe_%g" % r])), r
    d, th, ph = (np.ascontiguousarray(g[k]) for k in ("rot_dir", "rot_theta", "rot_phi"))
    y = np.zeros_like(d)
    lib.oracle_rotate_direction(0, rb.ptr(d, rb.c_f), rb.ptr(th, rb.c_f), rb.ptr(ph, rb.c_f),
                                rb.ctypes.c_int64(len(th)), rb.ptr(y, rb.c_f))
    assert np.array_equal(_u32(y), _u32(g["rot_out"]))


def test_grid_probes():
    g = _load("reference_grid.npz")
    xe, ye, ze = (np.ascontiguousarray(g[k]) for k in ("xe", "ye", "ze"))
    d = np.ascontiguousarray(g["d"])
    cells = np.ascontiguousarray(g["cells"])
    n = len(d)
    for op in range(4):
        pp = np.ascontiguousarray(g["p_inside"] if op == 1 else g["p"])
        dist = np.zeros(n, np.float32)
        dd = np.zeros((n, 3), np.float32)
        cc = np.zeros((n, 3), np.int32)
        rb.oracle().oracle_grid_probe(op, len(xe) - 1, len(ye) - 1, len(ze) - 1, rb.ptr(xe, rb.c_f),
                                      rb.ptr(ye, rb.c_f), rb.ptr(ze, rb.c_f), rb.ptr(pp, rb.c_f), rb.ptr(d, rb.c_f),
                                      rb.ptr(cells, rb.c_i32), rb.ctypes.c_int64(n), rb.ptr(dist, rb.c_f),
                                      rb.ptr(dd, rb.c_f), rb.ptr(cc, rb.c_i32))
        assert np.array_equal(_u32(dist), _u32(g["op%d_dist" % op])), op
        assert np.array_equal(_u32(dd), _u32(g["op%d_dir" % op])), op
        assert np.array_equal(cc, g["op%d_cell" % op]), op


def test_vacuum_trace():
    g = _load("reference_trace.npz")
    grid = rb.Grid(g["xe"], g["ye"], g["ze"], np.zeros((len(g["xe"]) - 1) * (len(g["ye"]) - 1) * (len(g["ze"]) - 1),
                                                       np.float32))
    cnb, rec = rb.oracle_trace(grid, g["rays"], debug=1)
    assert np.array_equal(cnb, g["cnb"])
    assert np.array_equal(_u32(rec), _u32(g["rec"]))


@pytest.mark.parametrize("debug", [1, 0])
def test_transport(debug):
    g = _load("reference_transport.npz")
    grid = rb.Grid(g["xe"], g["ye"], g["ze"], g["density"])
    vtx, seeds = g["vtx_%d" % debug], g["seeds_%d" % debug]
    arrs, n = rb.oracle_transport(grid, vtx, seeds, mask=rb.SC_DOSE | rb.SC_LETD, kernel=0, debug=debug)
    assert n == len(seeds)
    for k, a in enumerate(arrs):
        assert np.array_equal(_u64(a), _u64(g["plain_%d_%d" % (debug, k)])), k
    arrs, _ = rb.oracle_transport(grid, vtx, seeds, mask=rb.SC_DOSE | rb.SC_STAT, kernel=1, debug=debug)
    for k, a in enumerate(arrs):
        assert np.array_equal(_u64(a), _u64(g["stat_%d_%d" % (debug, k)])), k


def test_golden_fixture_is_not_trivial():
    """A fixture of zeros would pin nothing: the dose deposits, LETd is
    finite where dose is, and the debug and release builds differ."""
    g = _load("reference_transport.npz")
    d1, d0 = g["plain_1_0"], g["plain_0_0"]
    assert (d1 > 0).sum() > 500 and (d0 > 0).sum() > 500
    assert not np.array_equal(d1, d0)
    num, den = g["plain_0_1"], g["plain_0_2"]
    assert np.all(np.isfinite(num[den > 0]))
