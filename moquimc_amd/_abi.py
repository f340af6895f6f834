"""ctypes binding of libmqi_hip.so (include/mqi_hip.h).

The library is loaded from this package directory; there is no fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmqi_hip.so")

SPOT_DTYPE = np.dtype([
    ("energy", np.float32), ("sigma_energy", np.float32), ("energy_const", np.int32),
    ("phsp", np.uint32), ("mean", np.float32, 6), ("sigma", np.float32, 6),
    ("corr", np.float32, 2), ("rotation", np.float32, 9), ("translation", np.float32, 3),
    ("histories", np.uint64)])
assert SPOT_DTYPE.itemsize == 128

SCORE_DOSE, SCORE_LETD, SCORE_STAT, SCORE_LETT, SCORE_EDEP, SCORE_DM = 1, 2, 4, 16, 32, 64
LOOPS_REFERENCE_PLAIN, LOOPS_REFERENCE_STAT, LOOPS_ONCE = 0, 1, 2


class RunConfig(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64), ("physics_debug", ctypes.c_int32), ("exact", ctypes.c_int32),
        ("batched", ctypes.c_int32), ("scoring_loops", ctypes.c_int32), ("stack_depth", ctypes.c_int32),
        ("block_size", ctypes.c_int32), ("blocks", ctypes.c_int32), ("waves_per_simd", ctypes.c_int32),
        ("fetch_chunk", ctypes.c_int32), ("refill_lanes", ctypes.c_int32), ("dij_slots_sqrt", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("stat_criterion", ctypes.c_float), ("stat_threshold", ctypes.c_float), ("max_rounds", ctypes.c_int32),
        ("dij", ctypes.c_int32), ("dij_capacity", ctypes.c_uint64), ("hist_offset", ctypes.c_uint64)]


class RunResult(ctypes.Structure):
    _fields_ = [
        ("histories", ctypes.c_uint64), ("dropped", ctypes.c_uint64), ("steps", ctypes.c_uint64),
        ("status", ctypes.c_uint32),
        ("rounds", ctypes.c_int32), ("kernel_ms", ctypes.c_double), ("stat", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        # One HIP runtime per process: PyTorch bundles its own libamdhip64
        # (soname libamdhip64.so.7, loaded by path from torch/lib).  Loading
        # it first makes this library's libamdhip64.so.7 dependency resolve to
        # that same runtime; loading ours first would leave two runtimes in
        # the process and torch would see no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmqi_hip.so is not built (python -m moquimc_amd.build): " + LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, i, u64, f, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_float, ctypes.c_double
        sig = {
            "mqi_create": (i, [i, ctypes.POINTER(vp)]),
            "mqi_destroy": (i, [vp]),
            "mqi_last_error": (ctypes.c_char_p, [vp]),
            "mqi_device_info": (i, [vp, ctypes.POINTER(i), ctypes.POINTER(ctypes.c_size_t)]),
            "mqi_set_grid": (i, [vp, i, i, i, vp, vp, vp, vp, vp, vp]),
            "mqi_set_grid_hu": (i, [vp, i, i, i, vp, vp, vp, vp, f, vp, vp]),
            "mqi_get_density": (i, [vp, vp]),
            "mqi_set_spots": (i, [vp, vp, i]),
            "mqi_set_scenarios": (i, [vp, i, vp, vp, vp]),
            "mqi_set_scorers": (i, [vp, i, i]),
            "mqi_zero_scorers": (i, [vp]),
            "mqi_scorer_count": (i, [vp]),
            "mqi_run": (i, [vp, ctypes.POINTER(RunConfig), ctypes.POINTER(RunResult)]),
            "mqi_run_vertices": (i, [vp, vp, u64, vp, ctypes.POINTER(RunConfig), ctypes.POINTER(RunResult)]),
            "mqi_run_vertices_noalloc": (i, [vp, vp, u64, vp, ctypes.POINTER(RunConfig), ctypes.POINTER(RunResult)]),
            "mqi_atomic_probe": (i, [vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(d)]),
            "mqi_trace": (i, [vp, ctypes.POINTER(RunConfig), vp, u64, i, vp, vp, vp]),
            "mqi_get_scorer": (i, [vp, i, i, vp]),
            "mqi_scorer_device_ptr": (i, [vp, i, i, ctypes.POINTER(vp)]),
            "mqi_stat": (i, [vp, i, u64, f, ctypes.POINTER(d)]),
            "mqi_dij_build": (i, [vp, d]),
            "mqi_dij_info": (i, [vp, i, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(ctypes.c_int64)]),
            "mqi_dij_copy": (i, [vp, i, vp, vp, vp]),
            "mqi_dij_set_csr": (i, [vp, i, ctypes.c_int64, ctypes.c_int64, vp, vp, vp]),
            "mqi_dij_quadratic": (i, [vp, i, vp, vp, vp, d, vp, vp, vp]),
            "mqi_dij_device_ptrs": (i, [vp, i, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]),
            "mqi_dij_matvec": (i, [vp, i, i, vp, vp, i, d, i, vp]),
            "mqi_last_matvec_ms": (i, [vp, ctypes.POINTER(d)]),
            "mqi_autotune": (i, [vp, ctypes.POINTER(RunConfig), d, ctypes.POINTER(d)]),
            "mqi_math_probe": (i, [vp, i, vp, vp, ctypes.c_int64, vp]),
            "mqi_physics_probe": (i, [vp, i, i, i, vp, vp, ctypes.c_int64, vp]),
            "mqi_dij_save_npz": (i, [vp, i, ctypes.c_char_p, d]),
            "mqi_set_grid_ct_mha": (i, [vp, ctypes.c_char_p, f, vp]),
            "mqi_set_beamline": (i, [vp, i, vp]),
            "mqi_set_roi": (i, [vp, i, vp]),
            "mqi_set_roi_mha": (i, [vp, i, i, ctypes.POINTER(ctypes.c_char_p)]),
            "mqi_save_scorer": (i, [vp, i, i, ctypes.c_char_p, ctypes.c_char_p, i, d]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


EXPORTED = ["mqi_create", "mqi_destroy", "mqi_last_error", "mqi_device_info", "mqi_set_grid", "mqi_set_grid_hu",
            "mqi_get_density", "mqi_set_spots", "mqi_set_scenarios", "mqi_set_scorers", "mqi_zero_scorers",
            "mqi_scorer_count", "mqi_run", "mqi_run_vertices", "mqi_trace", "mqi_get_scorer",
            "mqi_scorer_device_ptr", "mqi_stat", "mqi_dij_build", "mqi_dij_info", "mqi_dij_copy", "mqi_dij_set_csr", "mqi_dij_quadratic",
            "mqi_dij_device_ptrs", "mqi_dij_matvec", "mqi_last_matvec_ms", "mqi_math_probe", "mqi_autotune",
            "mqi_dij_save_npz", "mqi_save_scorer", "mqi_set_grid_ct_mha", "mqi_set_beamline", "mqi_set_roi", "mqi_set_roi_mha",
            "mqi_physics_probe", "mqi_run_vertices_noalloc", "mqi_atomic_probe"]


class GridDesc(ctypes.Structure):
    """mqi_grid_desc: a beamline child grid (include/mqi_hip_types.h)."""
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("xe", ctypes.c_void_p), ("ye", ctypes.c_void_p), ("ze", ctypes.c_void_p),
                ("density", ctypes.c_void_p), ("rotation", ctypes.c_float * 9), ("translation", ctypes.c_float * 3)]

FORMATS = {"raw": 0, "mhd": 1, "mha": 2}


class MqiError(RuntimeError):
    pass


def csr_canonical(indptr, indices, data):
    """The CSR with each row's columns ascending and repeated columns summed
    (in float64, then float32): the reference's io::save_to_npz writes each
    row in hash-table order (mqi_io.hpp:249-320), and mqi_dij_set_csr takes
    rows with strictly ascending columns."""
    indptr = np.asarray(indptr, np.int64)
    indices = np.asarray(indices, np.int64)[:indptr[-1]]
    data = np.asarray(data, np.float64)[:indptr[-1]]
    rows = np.repeat(np.arange(indptr.size - 1, dtype=np.int64), np.diff(indptr))
    order = np.lexsort((indices, rows))
    rows, indices, data = rows[order], indices[order], data[order]
    if indices.size:
        first = np.r_[True, (rows[1:] != rows[:-1]) | (indices[1:] != indices[:-1])]
        starts = np.flatnonzero(first)
        data = np.add.reduceat(data, starts)
        rows, indices = rows[starts], indices[starts]
    out_ptr = np.zeros(indptr.size, np.int64)
    np.add.at(out_ptr, rows + 1, 1)
    return np.cumsum(out_ptr), indices.astype(np.int32), data.astype(np.float32)


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class Context:
    """One device context of libmqi_hip.so (mqi_create / mqi_destroy)."""

    def __init__(self, device=0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.mqi_create(device, ctypes.byref(h))
        if rc != 0:
            raise MqiError("mqi_create(%d) failed with code %d" % (device, rc))
        self.h = h
        self.nvox = 0
        self.n_scen = 1
        self.shape = None

    def close(self):
        if self.h:
            self._L.mqi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self._L.mqi_last_error(self.h)
            raise MqiError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def device_info(self):
        cu = ctypes.c_int()
        mem = ctypes.c_size_t()
        self._chk(self._L.mqi_device_info(self.h, ctypes.byref(cu), ctypes.byref(mem)), "mqi_device_info")
        return cu.value, mem.value

    def set_grid(self, xe, ye, ze, density, rot=None, trans=None):
        xe, ye, ze = (np.ascontiguousarray(a, np.float32) for a in (xe, ye, ze))
        density = np.ascontiguousarray(density, np.float32).reshape(-1)
        nx, ny, nz = len(xe) - 1, len(ye) - 1, len(ze) - 1
        assert density.size == nx * ny * nz
        rot = None if rot is None else np.ascontiguousarray(rot, np.float32)
        trans = None if trans is None else np.ascontiguousarray(trans, np.float32)
        self._chk(self._L.mqi_set_grid(self.h, nx, ny, nz, _p(xe), _p(ye), _p(ze), _p(density), _p(rot), _p(trans)),
                  "mqi_set_grid")
        self.nvox = nx * ny * nz
        self.shape = (nz, ny, nx)

    def set_grid_hu(self, xe, ye, ze, hu, density_scale=1.0, rot=None, trans=None):
        xe, ye, ze = (np.ascontiguousarray(a, np.float32) for a in (xe, ye, ze))
        hu = np.ascontiguousarray(hu, np.int16).reshape(-1)
        nx, ny, nz = len(xe) - 1, len(ye) - 1, len(ze) - 1
        assert hu.size == nx * ny * nz
        rot = None if rot is None else np.ascontiguousarray(rot, np.float32)
        trans = None if trans is None else np.ascontiguousarray(trans, np.float32)
        self._chk(self._L.mqi_set_grid_hu(self.h, nx, ny, nz, _p(xe), _p(ye), _p(ze), _p(hu),
                                          ctypes.c_float(density_scale), _p(rot), _p(trans)), "mqi_set_grid_hu")
        self.nvox = nx * ny * nz
        self.shape = (nz, ny, nx)

    def set_beamline(self, children=()):
        """Beamline children before the patient, in order; each has xe, ye,
        ze (edges in its frame), density (g/mm^3, x fastest) and optional
        rot9 (frame -> world) and t3 (translation).  No arguments: none."""
        keep, descs = [], (GridDesc * max(1, len(children)))()
        for k, g in enumerate(children):
            xe, ye, ze = (np.ascontiguousarray(a, np.float32) for a in (g.xe, g.ye, g.ze))
            dens = np.ascontiguousarray(g.density, np.float32).reshape(-1)
            keep += [xe, ye, ze, dens]
            d = descs[k]
            d.nx, d.ny, d.nz = len(xe) - 1, len(ye) - 1, len(ze) - 1
            d.xe, d.ye, d.ze, d.density = (a.ctypes.data for a in (xe, ye, ze, dens))
            rot = getattr(g, "rot9", None)
            tr = getattr(g, "t3", None)
            d.rotation[:] = list(np.asarray(rot if rot is not None else np.eye(3), np.float32).ravel())
            d.translation[:] = list(np.asarray(tr if tr is not None else np.zeros(3), np.float32).ravel())
        self._chk(self._L.mqi_set_beamline(self.h, len(children), ctypes.cast(descs, ctypes.c_void_p)),
                  "mqi_set_beamline")

    def set_roi(self, target, mask=None):
        """ROI of the scorers (target 0) or of the statistics pair (target 1)
        from a uint8 voxel mask of the grid's size (x fastest; 1 opens a run,
        0 closes it, other values continue it); None restores DIRECT."""
        m = None
        if mask is not None:
            m = np.ascontiguousarray(mask, np.uint8).reshape(-1)
            if m.size != self.nvox:
                raise ValueError("mask has %d voxels, the grid %d" % (m.size, self.nvox))
        self._chk(self._L.mqi_set_roi(self.h, int(target), _p(m)), "mqi_set_roi")

    def set_roi_mha(self, target, paths=()):
        """ROI from structure mask files (MetaImage, uint8 after the header),
        summed as the reference's mask reader does; () restores DIRECT."""
        arr = (ctypes.c_char_p * max(1, len(paths)))(*[os.fsencode(p) for p in paths])
        self._chk(self._L.mqi_set_roi_mha(self.h, int(target), len(paths), arr), "mqi_set_roi_mha")

    def set_grid_ct_mha(self, path, density_scale=1.0, shift=None):
        """CT from a MetaImage file (float32 HU after the header)."""
        sh = None if shift is None else np.ascontiguousarray(shift, np.float32).reshape(3)
        self._chk(self._L.mqi_set_grid_ct_mha(self.h, os.fsencode(path), ctypes.c_float(density_scale), _p(sh)),
                  "mqi_set_grid_ct_mha")
        nx = ny = nz = 0
        with open(path, "rb") as f:   # the grid's size, for the download helpers
            for raw in f:
                key, _, val = raw.decode("latin-1").partition("=")
                if key.strip().lower() == "dimsize":
                    nx, ny, nz = (int(v) for v in val.split()[:3])
                if key.strip().lower() == "elementdatafile":
                    break
        self.nvox = nx * ny * nz
        self.shape = (nz, ny, nx)

    def density(self):
        out = np.zeros(self.nvox, np.float32)
        self._chk(self._L.mqi_get_density(self.h, _p(out)), "mqi_get_density")
        return out

    def set_spots(self, spots):
        spots = np.ascontiguousarray(spots, SPOT_DTYPE)
        self._chk(self._L.mqi_set_spots(self.h, _p(spots), len(spots)), "mqi_set_spots")
        self.n_spots = len(spots)

    def set_scenarios(self, shifts=None, scales=None, budgets=None):
        """Robust scenarios: shifts (n, 3) mm, density scales (n), optional
        history budget factors (n); no arguments: the nominal scenario."""
        if shifts is None:
            self._chk(self._L.mqi_set_scenarios(self.h, 0, None, None, None), "mqi_set_scenarios")
            self.n_scen = 1
            return
        shifts = np.ascontiguousarray(shifts, np.float32).reshape(-1, 3)
        scales = np.ascontiguousarray(scales, np.float32).reshape(-1)
        b = None if budgets is None else np.ascontiguousarray(budgets, np.float32).reshape(-1)
        if len(shifts) != len(scales) or (b is not None and len(b) != len(scales)):
            raise ValueError("scenario arrays differ in length")
        self._chk(self._L.mqi_set_scenarios(self.h, len(scales), _p(shifts), _p(scales), _p(b)), "mqi_set_scenarios")
        self.n_scen = len(scales)

    def set_scorers(self, mask, per_scenario=False):
        self._chk(self._L.mqi_set_scorers(self.h, mask, 1 if per_scenario else 0), "mqi_set_scorers")

    def zero_scorers(self):
        self._chk(self._L.mqi_zero_scorers(self.h), "mqi_zero_scorers")

    def scorer_count(self):
        return self._L.mqi_scorer_count(self.h)

    @staticmethod
    def config(seed=1, physics_debug=False, exact=False, batched=True, loops=LOOPS_ONCE, stack_depth=0,
               block_size=0, blocks=0, waves_per_simd=0, fetch_chunk=0, refill_lanes=0, use_graph=False,
               stat_criterion=0.0, stat_threshold=0.5, max_rounds=0, dij=False, dij_capacity=0, hist_offset=0,
               dij_slots_sqrt=0):
        return RunConfig(seed, int(physics_debug), int(exact), int(batched), loops, stack_depth, block_size, blocks,
                         waves_per_simd, fetch_chunk, refill_lanes, dij_slots_sqrt, int(use_graph), stat_criterion,
                         stat_threshold, max_rounds, int(dij), dij_capacity, hist_offset)

    def run(self, cfg):
        res = RunResult()
        self._chk(self._L.mqi_run(self.h, ctypes.byref(cfg), ctypes.byref(res)), "mqi_run")
        return res

    def autotune(self, cfg, sample=0.1):
        """Fill cfg's launch parameters with the fastest measured on a sample
        of the plan; returns (cfg, device ms of the best trial)."""
        ms = ctypes.c_double()
        self._chk(self._L.mqi_autotune(self.h, ctypes.byref(cfg), ctypes.c_double(sample), ctypes.byref(ms)),
                  "mqi_autotune")
        return cfg, ms.value

    def run_vertices(self, vertices, cfg, spot_ids=None):
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 7)
        sid = None if spot_ids is None else np.ascontiguousarray(spot_ids, np.uint32)
        res = RunResult()
        self._chk(self._L.mqi_run_vertices(self.h, _p(v), len(v), _p(sid), ctypes.byref(cfg), ctypes.byref(res)),
                  "mqi_run_vertices")
        return res

    def run_vertices_noalloc(self, d_vertices, n, cfg, d_spot_ids=None):
        """mqi_run_vertices_noalloc: d_vertices (n x 7 float32) and d_spot_ids
        (n uint32, or None) are device addresses (e.g. a torch tensor's
        data_ptr()); nothing is allocated by the library."""
        res = RunResult()
        self._chk(self._L.mqi_run_vertices_noalloc(self.h, ctypes.c_void_p(d_vertices), n,
                                                   None if d_spot_ids is None else ctypes.c_void_p(d_spot_ids),
                                                   ctypes.byref(cfg), ctypes.byref(res)), "mqi_run_vertices_noalloc")
        return res

    def trace(self, cfg, n_histories, max_steps, vertices=None):
        v = None if vertices is None else np.ascontiguousarray(vertices, np.float32).reshape(-1, 7)
        if v is not None:
            n_histories = len(v)
        cnb = np.zeros((n_histories, max_steps), np.int64)
        rec = np.zeros((n_histories, max_steps, 12), np.float32)
        nst = np.zeros(n_histories, np.int32)
        self._chk(self._L.mqi_trace(self.h, ctypes.byref(cfg), _p(v), n_histories, max_steps, _p(cnb), _p(rec),
                                    _p(nst)), "mqi_trace")
        return cnb, rec, nst

    def scorer(self, k, scenario=0):
        out = np.zeros(self.nvox, np.float64)
        self._chk(self._L.mqi_get_scorer(self.h, k, scenario, _p(out)), "mqi_get_scorer")
        return out

    def scorer_device_ptr(self, k, scenario=0):
        p = ctypes.c_void_p()
        self._chk(self._L.mqi_scorer_device_ptr(self.h, k, scenario, ctypes.byref(p)), "mqi_scorer_device_ptr")
        return p.value

    def stat(self, n_histories, threshold=0.5, scenario=0):
        out = ctypes.c_double()
        self._chk(self._L.mqi_stat(self.h, scenario, n_histories, ctypes.c_float(threshold), ctypes.byref(out)),
                  "mqi_stat")
        return out.value

    def dij_build(self, rel_threshold=1e-4):
        self._chk(self._L.mqi_dij_build(self.h, ctypes.c_double(rel_threshold)), "mqi_dij_build")

    def dij_info(self, scenario=0):
        nnz, rows, cols = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._chk(self._L.mqi_dij_info(self.h, scenario, ctypes.byref(nnz), ctypes.byref(rows), ctypes.byref(cols)),
                  "mqi_dij_info")
        return nnz.value, rows.value, cols.value

    def dij_set_csr(self, indptr, indices, values, shape, scenario=0):
        """A CSR (rows x cols; columns ascending within rows) as the Dij of
        scenario (mqi_dij_set_csr); values are taken as float32."""
        rows, cols = int(shape[0]), int(shape[1])
        indptr = np.ascontiguousarray(indptr, np.int64)
        indices = np.ascontiguousarray(indices, np.int32)
        values = np.ascontiguousarray(values, np.float32)
        if indptr.size != rows + 1 or indices.size < indptr[-1] or values.size < indptr[-1]:
            raise ValueError("indptr needs rows + 1 entries and indices/values indptr[-1]")
        self._chk(self._L.mqi_dij_set_csr(self.h, scenario, rows, cols, _p(indptr), _p(indices), _p(values)),
                  "mqi_dij_set_csr")

    def dij_load_npz(self, path, scenario=0):
        """A scipy-style CSR .npz (the layout io::save_to_npz and
        mqi_dij_save_npz write: rows = spots, columns = voxels) as the Dij of
        scenario: rows put in column order (csr_canonical), values as
        float32."""
        with np.load(path, allow_pickle=False) as z:
            fmt = z["format"][()]
            fmt = fmt.decode() if isinstance(fmt, bytes) else str(fmt)
            if fmt != "csr":
                raise ValueError("%s is not a CSR matrix (format %r)" % (path, fmt))
            shape = [int(v) for v in z["shape"]]
            ip, ix, v = csr_canonical(z["indptr"], z["indices"], z["data"])
            self.dij_set_csr(ip, ix, v, shape, scenario)

    def dij_csr(self, scenario=0):
        nnz, rows, cols = self.dij_info(scenario)
        indptr = np.zeros(rows + 1, np.int64)
        idx = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1), np.float32)
        self._chk(self._L.mqi_dij_copy(self.h, scenario, _p(indptr), _p(idx), _p(val)), "mqi_dij_copy")
        return indptr, idx[:nnz], val[:nnz], (rows, cols)

    def dij_save_npz(self, path, scenario=0, scale=1.0):
        """The reference's scipy-style CSR .npz of a scenario's Dij."""
        self._chk(self._L.mqi_dij_save_npz(self.h, scenario, os.fsencode(path), ctypes.c_double(scale)),
                  "mqi_dij_save_npz")

    def save_scorer(self, k, directory, name, fmt="raw", scenario=0, scale=1.0):
        """Scorer k as <directory>/<name>.raw, .mhd + .raw or .mha (float64)."""
        self._chk(self._L.mqi_save_scorer(self.h, k, scenario, os.fsencode(directory), os.fsencode(name),
                                          FORMATS[fmt], ctypes.c_double(scale)), "mqi_save_scorer")

    def dij_matvec(self, scenario, transpose, x_ptr, y_ptr, y_double=True, alpha=1.0, accumulate=False, stream=None):
        self._chk(self._L.mqi_dij_matvec(self.h, scenario, int(transpose), ctypes.c_void_p(x_ptr),
                                         ctypes.c_void_p(y_ptr), int(y_double), ctypes.c_double(alpha),
                                         int(accumulate), None if stream is None else ctypes.c_void_p(stream)),
                  "mqi_dij_matvec")

    def dij_quadratic(self, scenario, w_ptr, p_ptr, c_ptr, alpha, g_ptr, f_ptr, stream=None):
        """f = mean c (alpha D w - p)^2 and g = 2 c (alpha D w - p) / n on the
        device (mqi_dij_quadratic; c_ptr / g_ptr may be 0)."""
        self._chk(self._L.mqi_dij_quadratic(self.h, scenario, ctypes.c_void_p(w_ptr), ctypes.c_void_p(p_ptr),
                                            ctypes.c_void_p(c_ptr or None), ctypes.c_double(alpha),
                                            ctypes.c_void_p(g_ptr or None), ctypes.c_void_p(f_ptr),
                                            None if stream is None else ctypes.c_void_p(stream)),
                  "mqi_dij_quadratic")

    def last_matvec_ms(self):
        ms = ctypes.c_double()
        self._chk(self._L.mqi_last_matvec_ms(self.h, ctypes.byref(ms)), "mqi_last_matvec_ms")
        return ms.value

    def math_probe(self, fn, x, y=None):
        x = np.ascontiguousarray(x, np.float64)
        y = x if y is None else np.ascontiguousarray(y, np.float64)
        out = np.zeros_like(x)
        self._chk(self._L.mqi_math_probe(self.h, fn, _p(x), _p(y), len(x), _p(out)), "mqi_math_probe")
        return out

    def atomic_probe(self, n_slots, n_updates, window=0):
        """Device ms of n_updates FP64 atomic adds to hashed slots of an
        n_slots array, each wave's adds inside one window of that many slots
        when window > 0 (mqi_atomic_probe)."""
        ms = ctypes.c_double()
        self._chk(self._L.mqi_atomic_probe(self.h, n_slots, n_updates, window, ctypes.byref(ms)), "mqi_atomic_probe")
        return ms.value

    def physics_probe(self, fn, rho, ek=None, exact=True, physics_debug=False):
        """fn 0: spr(rho, ek); 1: radiation length(rho); 2: (n, 10) cross
        sections, dEdx and relativistic quantities (mqi_physics_probe)."""
        rho = np.ascontiguousarray(rho, np.float32)
        ek = rho if ek is None else np.ascontiguousarray(ek, np.float32)
        out = np.zeros((rho.size, 10) if fn == 2 else rho.size, np.float32)
        self._chk(self._L.mqi_physics_probe(self.h, int(exact), int(physics_debug), fn, _p(rho), _p(ek), rho.size,
                                            _p(out)), "mqi_physics_probe")
        return out
