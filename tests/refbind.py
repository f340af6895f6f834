"""ctypes bindings for the test-only checkers built by oracle/Makefile:

  oracle/_ref/libmoqui_ref_{debug,release}.so  reference moqui CPU transport (harness)
  oracle/_ref/liboracle.so                     the CPU restatement (oracle)

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

c_f = ctypes.POINTER(ctypes.c_float)
c_d = ctypes.POINTER(ctypes.c_double)
c_i32 = ctypes.POINTER(ctypes.c_int32)
c_u32 = ctypes.POINTER(ctypes.c_uint32)
c_i64 = ctypes.POINTER(ctypes.c_int64)
c_u64 = ctypes.POINTER(ctypes.c_uint64)

SPOT_DTYPE = np.dtype([
    ("energy", np.float32), ("sigma_energy", np.float32), ("energy_const", np.int32),
    ("phsp", np.uint32), ("mean", np.float32, 6), ("sigma", np.float32, 6),
    ("corr", np.float32, 2), ("rotation", np.float32, 9), ("translation", np.float32, 3),
    ("histories", np.uint64)])
assert SPOT_DTYPE.itemsize == 128

SC_DOSE, SC_LETD, SC_STAT, SC_LOG, SC_LETT, SC_EDEP, SC_DM = 1, 2, 4, 8, 16, 32, 64


def ptr(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def n_scorers(mask):
    return ((1 if mask & SC_DOSE else 0) + (2 if mask & SC_LETD else 0) + (2 if mask & SC_STAT else 0) +
            (2 if mask & SC_LETT else 0) + (1 if mask & SC_EDEP else 0) + (1 if mask & SC_DM else 0))


_libs = {}


def _load(name):
    if name not in _libs:
        path = os.path.join(REF_DIR, name)
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (build with: make -C oracle)")
        _libs[name] = ctypes.CDLL(path)
    return _libs[name]


def have_reference():
    return all(os.path.exists(os.path.join(REF_DIR, n))
               for n in ("libmoqui_ref_debug.so", "libmoqui_ref_release.so"))


def reference(debug):
    return _load("libmoqui_ref_debug.so" if debug else "libmoqui_ref_release.so")


def oracle():
    return _load("liboracle.so")


class Grid:
    """Edges (n+1 per axis, float32) and float32 density (g/mm^3, x fastest)."""

    def __init__(self, xe, ye, ze, density, rot9=None, t3=None):
        self.xe = np.ascontiguousarray(xe, np.float32)
        self.ye = np.ascontiguousarray(ye, np.float32)
        self.ze = np.ascontiguousarray(ze, np.float32)
        self.nx, self.ny, self.nz = len(xe) - 1, len(ye) - 1, len(ze) - 1
        self.density = np.ascontiguousarray(density, np.float32).reshape(-1)
        assert self.density.size == self.nx * self.ny * self.nz
        self.rot9 = None if rot9 is None else np.ascontiguousarray(rot9, np.float32)
        self.t3 = None if t3 is None else np.ascontiguousarray(t3, np.float32)

    @property
    def nvox(self):
        return self.nx * self.ny * self.nz


def uniform_edges(lo, hi, n):
    """Edges as grid3d's (min, max, n+1) constructor computes them in float
    (mqi_grid3d.hpp:134-163): dx = (max - min) / n, e_i = min + i * dx."""
    lo = np.float32(lo)
    hi = np.float32(hi)
    dx = np.float32(np.float32(hi - lo) / np.float32(n))
    i = np.arange(n + 1, dtype=np.float32)
    return (lo + i * dx).astype(np.float32)


def _outs(mask, nvox):
    arrs = [np.zeros(nvox, np.float64) for _ in range(n_scorers(mask))]
    table = (c_d * max(1, len(arrs)))(*[a.ctypes.data_as(c_d) for a in arrs])
    return arrs, table


def ref_transport(debug, grid, vtx, seeds, spot_ids=None, mask=SC_DOSE, kernel=0, threads=1, timing=None):
    """Reference CPU transport; timing (a list) receives the seconds of the transport phase."""
    lib = reference(debug)
    lib.ref_transport.restype = ctypes.c_int
    vtx = np.ascontiguousarray(vtx, np.float32)
    seeds = np.ascontiguousarray(seeds, np.int32)
    n = len(seeds)
    arrs, table = _outs(mask, grid.nvox)
    tracked = np.zeros(1, np.uint64)
    secs = np.zeros(1, np.float64)
    lib.ref_transport(grid.nx, grid.ny, grid.nz, ptr(grid.xe, c_f), ptr(grid.ye, c_f), ptr(grid.ze, c_f),
                      ptr(grid.density, c_f), ptr(grid.rot9, c_f), ptr(grid.t3, c_f), ctypes.c_int64(n),
                      ptr(vtx, c_f), ptr(seeds, c_i32),
                      ptr(None if spot_ids is None else np.ascontiguousarray(spot_ids, np.uint32), c_u32),
                      mask, kernel, threads, table, ptr(tracked, c_u64), None, None, ctypes.c_int64(0), None,
                      ptr(secs, c_d))
    if timing is not None:
        timing.append(float(secs[0]))
    return arrs, int(tracked[0])


def ref_trace(debug, grid, vtx, cap=200000):
    """Per-step log of the reference transport (vacuum or not): cnb and 13 floats/step
    (its.dist, vtx0.pos, vtx0.dir, vtx1.pos, ke0, dE, history)."""
    lib = reference(debug)
    vtx = np.ascontiguousarray(vtx, np.float32)
    n = vtx.shape[0]
    seeds = np.arange(1, n + 1, dtype=np.int32)
    cnb = np.zeros(cap, np.int64)
    rec = np.zeros((cap, 13), np.float32)
    nlog = np.zeros(1, np.int64)
    lib.ref_transport(grid.nx, grid.ny, grid.nz, ptr(grid.xe, c_f), ptr(grid.ye, c_f), ptr(grid.ze, c_f),
                      ptr(grid.density, c_f), ptr(grid.rot9, c_f), ptr(grid.t3, c_f), ctypes.c_int64(n),
                      ptr(vtx, c_f), ptr(seeds, c_i32), None, SC_LOG, 0, 1, None, None,
                      ptr(cnb, c_i64), ptr(rec, c_f), ctypes.c_int64(cap), ptr(nlog, c_i64), None)
    k = min(int(nlog[0]), cap)
    return cnb[:k], rec[:k]


def oracle_transport(grid, vtx, seeds, spot_ids=None, mask=SC_DOSE, kernel=0, threads=1, prec=32,
                     rng=0, math=0, debug=1, master_seed=0):
    lib = oracle()
    vtx = np.ascontiguousarray(vtx, np.float32)
    seeds = np.ascontiguousarray(seeds, np.int32)
    n = len(seeds)
    arrs, table = _outs(mask, grid.nvox)
    tracked = np.zeros(1, np.uint64)
    lib.oracle_transport(prec, rng, math, debug, ctypes.c_uint64(master_seed), grid.nx, grid.ny, grid.nz,
                         ptr(grid.xe, c_f), ptr(grid.ye, c_f), ptr(grid.ze, c_f), ptr(grid.density, c_f),
                         ptr(grid.rot9, c_f), ptr(grid.t3, c_f), ctypes.c_int64(n), ptr(vtx, c_f),
                         ptr(seeds, c_i32),
                         ptr(None if spot_ids is None else np.ascontiguousarray(spot_ids, np.uint32), c_u32),
                         mask, kernel, threads, table, ptr(tracked, c_u64), None, None, ctypes.c_int64(0),
                         None)
    return arrs, int(tracked[0])


def oracle_trace(grid, vtx, cap=200000, debug=1, prec=32, rng=0, math=0):
    lib = oracle()
    vtx = np.ascontiguousarray(vtx, np.float32)
    n = vtx.shape[0]
    seeds = np.arange(1, n + 1, dtype=np.int32)
    cnb = np.zeros(cap, np.int64)
    rec = np.zeros((cap, 13), np.float32)
    nlog = np.zeros(1, np.int64)
    lib.oracle_transport(prec, rng, math, debug, ctypes.c_uint64(0), grid.nx, grid.ny, grid.nz,
                         ptr(grid.xe, c_f), ptr(grid.ye, c_f), ptr(grid.ze, c_f), ptr(grid.density, c_f),
                         ptr(grid.rot9, c_f), ptr(grid.t3, c_f), ctypes.c_int64(n), ptr(vtx, c_f),
                         ptr(seeds, c_i32), None, SC_LOG, 0, 1, None, None, ptr(cnb, c_i64), ptr(rec, c_f),
                         ctypes.c_int64(cap), ptr(nlog, c_i64))
    k = min(int(nlog[0]), cap)
    return cnb[:k], rec[:k]


def oracle_run_spots(grid, spots, mask=SC_DOSE, kernel=0, threads=1, prec=32, rng=1, math=1, debug=0,
                     master_seed=1, scenario=0):
    lib = oracle()
    spots = np.ascontiguousarray(spots, SPOT_DTYPE)
    arrs, table = _outs(mask, grid.nvox)
    tracked = np.zeros(1, np.uint64)
    lib.oracle_run_spots(prec, rng, math, debug, ctypes.c_uint64(master_seed), ctypes.c_uint32(scenario),
                         grid.nx, grid.ny, grid.nz, ptr(grid.xe, c_f), ptr(grid.ye, c_f), ptr(grid.ze, c_f),
                         ptr(grid.density, c_f), spots.ctypes.data_as(ctypes.c_void_p), len(spots), mask,
                         kernel, threads, table, ptr(tracked, c_u64))
    return arrs, int(tracked[0])


def make_spot(energy, sigma_xy, n, pos=(0.0, 0.0, 0.0), sigma_e=0.0, energy_const=True, rotation=None,
              translation=(0.0, 0.0, 0.0), sigma_dir=(0.0, 0.0), corr=(0.0, 0.0), z_mean=None, uniform=False):
    s = np.zeros(1, SPOT_DTYPE)[0]
    s["phsp"] = 1 if uniform else 0
    s["energy"] = energy
    s["sigma_energy"] = sigma_e
    s["energy_const"] = 1 if energy_const else 0
    s["mean"] = [pos[0], pos[1], pos[2] if z_mean is None else z_mean, 0.0, 0.0, -1.0]
    s["sigma"] = [sigma_xy, sigma_xy, 0.0, sigma_dir[0], sigma_dir[1], 0.0]
    s["corr"] = corr
    s["rotation"] = np.eye(3, dtype=np.float32).reshape(-1) if rotation is None else rotation
    s["translation"] = translation
    s["histories"] = n
    return s


def ref_sample_spot(spot, angles, iso, master_seed, spot_index, n, debug=0):
    """Reference beamlet sampling (phsp_6d, or phsp_6d_uniform when the
    spot's phsp is MQI_PHSP_UNIFORM) with run_by_spot's seeding."""
    lib = reference(debug)
    vtx = np.zeros((n, 7), np.float32)
    seeds = np.zeros(n, np.int32)
    mean = np.ascontiguousarray(spot["mean"], np.float32)
    sig = np.ascontiguousarray(spot["sigma"], np.float32)
    corr = np.ascontiguousarray(spot["corr"], np.float32)
    ang = np.ascontiguousarray(angles, np.float32)
    isoa = np.ascontiguousarray(iso, np.float32)
    lib.ref_sample_spot(ctypes.c_float(spot["energy"]), ctypes.c_float(spot["sigma_energy"]),
                        int(spot["energy_const"]), int(spot["phsp"]), ptr(mean, c_f), ptr(sig, c_f), ptr(corr, c_f),
                        ptr(ang, c_f), ptr(isoa, c_f), ctypes.c_uint32(master_seed),
                        ctypes.c_uint32(spot_index), ctypes.c_int64(n), ptr(vtx, c_f), ptr(seeds, c_i32))
    return vtx, seeds


def oracle_sample_spot(spot, master_seed, spot_index, n, rng=0, scenario=0):
    lib = oracle()
    vtx = np.zeros((n, 7), np.float32)
    seeds = np.zeros(n, np.int32)
    sp = np.ascontiguousarray(np.array([spot], SPOT_DTYPE))
    lib.oracle_sample_spot(rng, 0, ctypes.c_uint64(master_seed), ctypes.c_uint32(scenario),
                           ctypes.c_uint32(spot_index), sp.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(n),
                           ptr(vtx, c_f), ptr(seeds, c_i32))
    return vtx, seeds


def coordinate_rotation(angles, math=0):
    out = np.zeros(9, np.float32)
    a = np.ascontiguousarray(angles, np.float32)
    oracle().oracle_coordinate_rotation(math, ptr(a, c_f), ptr(out, c_f))
    return out


def _beamline_args(children):
    """(n, dims int32[3n], edges float32, densities float32, xf float32[12n])
    from Grid objects (rot9/t3 default to identity / zero)."""
    dims, edges, dens, xf = [], [], [], []
    for g in children:
        dims += [g.nx, g.ny, g.nz]
        edges += [g.xe, g.ye, g.ze]
        dens.append(g.density)
        r = g.rot9 if g.rot9 is not None else np.eye(3, dtype=np.float32).ravel()
        t = g.t3 if g.t3 is not None else np.zeros(3, np.float32)
        xf += [np.asarray(r, np.float32).ravel(), np.asarray(t, np.float32).ravel()]
    cat = lambda a, dt: np.ascontiguousarray(np.concatenate(a) if a else np.zeros(1, dt), dt)  # noqa: E731
    return (len(children), np.ascontiguousarray(dims or [0], np.int32), cat(edges, np.float32),
            cat(dens, np.float32), cat(xf, np.float32))


def set_roi(target, nvox, mask=None):
    """ROI (target 0: scorers, 1: statistics pair) of the oracle's and, where
    it is built, the reference harness's next transports; None: DIRECT.  The
    harness refuses a mask whose last run is open at the last voxel or that
    has no run (the reference leaves those undefined)."""
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8).reshape(-1)
    mp = ptr(m, ctypes.POINTER(ctypes.c_uint8))
    o = oracle()
    o.oracle_set_roi.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint8)]
    o.oracle_set_roi(target, nvox, mp)
    if have_reference():
        for debug in (0, 1):
            lib = reference(debug)
            lib.ref_set_roi.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint8)]
            lib.ref_set_roi.restype = ctypes.c_int
            if lib.ref_set_roi(target, nvox, mp) != 0:
                raise ValueError("the reference harness leaves this mask undefined")


def set_beamline(children):
    """Beamline children (before the phantom) of the oracle's and, where it
    is built, the reference harness's next transports; [] clears them."""
    n, dims, edges, dens, xf = _beamline_args(children)
    args = (n, ptr(dims, ctypes.POINTER(ctypes.c_int)), ptr(edges, c_f), ptr(dens, c_f), ptr(xf, c_f))
    oracle().oracle_set_beamline(*args)
    if have_reference():
        for debug in (0, 1):
            reference(debug).ref_set_beamline(*args)
