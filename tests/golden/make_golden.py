"""Write the golden fixtures of tests/golden from the reference built from its
own sources (oracle/_ref/libmoqui_ref_{debug,release}.so, see oracle/Makefile).

    make -C oracle && python tests/golden/make_golden.py [--stats]

Fixtures (inputs and the reference's outputs, nothing else):
  reference_tables.npz      the seven tables as compiled
  reference_functions.npz   hu_to_density, spr_default, radiation_length_default,
                            cross sections / dEdx / relativistic quantities,
                            update_post_vertex_direction on seeded inputs
  reference_grid.npz        grid3d index / intersect / index-update probes
  reference_trace.npz       per-step vacuum ray traversal of the transport kernel
  reference_transport.npz   dose / LETd / stat scorers of short seeded runs
                            (vertices and seeds included), debug and release
  reference_profiles.npz    (--stats) high-statistics dose of the water and
                            slab workloads for the statistical GPU tests
  reference_ct.npz          (--ct) a 100-spot sub-plan of the CT workload on
                            its cropped CT: inputs and the reference's dose
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import dosecmp  # noqa: E402
import phantoms  # noqa: E402
import refbind as rb  # noqa: E402


def grid_arrays():
    xe = rb.uniform_edges(-10, 10, 20)
    ye = np.cumsum(np.r_[-7.0, np.full(12, 1.2)]).astype(np.float32)
    ze = rb.uniform_edges(-30, 0, 15)
    return xe, ye, ze


def functions():
    rng = np.random.default_rng(20)
    out = {}
    hu = np.arange(-1200, 3200, dtype=np.int16)
    a = np.zeros(hu.size, np.float32)
    rb.reference(0).ref_hu_to_density(rb.ptr(hu, rb.ctypes.c_void_p), rb.ctypes.c_int64(hu.size), rb.ptr(a, rb.c_f))
    out["hu"], out["hu_density"] = hu, a
    rho = np.concatenate([rng.uniform(1e-7, 3e-3, 2000), [1e-3, 0.9995e-3, 1.0009e-3, 0.26e-3, 0.9e-3]])
    rho = rho.astype(np.float32)
    ek = rng.uniform(0.0, 300.0, rho.size).astype(np.float32)
    out["rho"], out["ek"] = rho, ek
    for dbg in (0, 1):
        x = np.zeros_like(rho)
        rb.reference(dbg).ref_spr(rb.ptr(rho, rb.c_f), rb.ptr(ek, rb.c_f), rb.ctypes.c_int64(rho.size),
                                  rb.ptr(x, rb.c_f))
        out["spr_%d" % dbg] = x
        y = np.zeros_like(rho)
        rb.reference(dbg).ref_radiation_length(rb.ptr(rho, rb.c_f), rb.ctypes.c_int64(rho.size), rb.ptr(y, rb.c_f))
        out["x0_%d" % dbg] = y
    e = np.concatenate([np.linspace(0.0, 299.5, 1500), rng.uniform(0.05, 299, 1000)]).astype(np.float32)
    out["probe_e"] = e
    for r in (1e-3, 1.85e-3):
        x = np.zeros((e.size, 10), np.float32)
        rb.reference(0).ref_physics_probe(rb.ptr(e, rb.c_f), rb.ctypes.c_int64(e.size), rb.ctypes.c_float(r),
                                          rb.ptr(x, rb.c_f))
        out["probe_%g" % r] = x
    dirs = rng.normal(size=(2000, 3)).astype(np.float32)
    dirs[:100] = [0, 0, 1]
    dirs[100:200] = [0, 0, -1]
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True).astype(np.float32)
    th = rng.uniform(0, 0.5, 2000).astype(np.float32)
    th[::7] = 0
    ph = rng.uniform(0, 2 * np.pi, 2000).astype(np.float32)
    ph[::11] = 0
    x = np.zeros_like(dirs)
    rb.reference(0).ref_rotate_direction(rb.ptr(dirs, rb.c_f), rb.ptr(th, rb.c_f), rb.ptr(ph, rb.c_f),
                                         rb.ctypes.c_int64(2000), rb.ptr(x, rb.c_f))
    out["rot_dir"], out["rot_theta"], out["rot_phi"], out["rot_out"] = dirs, th, ph, x
    return out


def grid_probes():
    xe, ye, ze = grid_arrays()
    rng = np.random.default_rng(21)
    n = 3000
    p = rng.uniform([-12, -8, -32], [12, 8, 2], (n, 3)).astype(np.float32)
    snap = rng.random(n) < 0.3
    p[snap, 0] = xe[rng.integers(0, len(xe), snap.sum())] + rng.choice([0, 5e-4, -5e-4, 2e-3], snap.sum())
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[rng.random(n) < 0.1, 1] = 1e-5
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(np.float32)
    cells = np.stack([rng.integers(0, 20, n), rng.integers(0, 12, n), rng.integers(0, 15, n)], 1).astype(np.int32)
    lo = np.stack([xe[cells[:, 0]], ye[cells[:, 1]], ze[cells[:, 2]]], 1)
    hi = np.stack([xe[cells[:, 0] + 1], ye[cells[:, 1] + 1], ze[cells[:, 2] + 1]], 1)
    inside = (lo + rng.uniform(0, 1, (n, 3)) * (hi - lo)).astype(np.float32)
    on = rng.random((n, 3)) < 0.2
    inside[on] = hi[on]
    out = {"xe": xe, "ye": ye, "ze": ze, "p": p, "p_inside": inside, "d": d, "cells": cells}
    for op in range(4):
        pp = inside if op == 1 else p
        dist = np.zeros(n, np.float32)
        dd = np.zeros((n, 3), np.float32)
        cc = np.zeros((n, 3), np.int32)
        rb.reference(0).ref_grid_probe(op, 20, 12, 15, rb.ptr(xe, rb.c_f), rb.ptr(ye, rb.c_f), rb.ptr(ze, rb.c_f),
                                       rb.ptr(np.ascontiguousarray(pp), rb.c_f), rb.ptr(d, rb.c_f),
                                       rb.ptr(cells, rb.c_i32), rb.ctypes.c_int64(n), rb.ptr(dist, rb.c_f),
                                       rb.ptr(dd, rb.c_f), rb.ptr(cc, rb.c_i32))
        out["op%d_dist" % op], out["op%d_dir" % op], out["op%d_cell" % op] = dist, dd, cc
    return out


def trace_fixture():
    g = phantoms.water_box(16, 18, 20, dx=2.0, dz=2.5, rho=0.0)
    rng = np.random.default_rng(22)
    n = 300
    v = np.zeros((n, 7), np.float32)
    v[:, 0] = 50
    lo = np.array([g.xe[0], g.ye[0], g.ze[0]])
    hi = np.array([g.xe[-1], g.ye[-1], g.ze[-1]])
    v[:, 1:4] = rng.uniform(lo - 10, hi + 10, (n, 3))
    d = (0.5 * (lo + hi) - v[:, 1:4]) + rng.normal(scale=4, size=(n, 3))
    v[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    cnb, rec = rb.ref_trace(1, g, v)
    return {"xe": g.xe, "ye": g.ye, "ze": g.ze, "rays": v, "cnb": cnb, "rec": rec}


def transport_fixture():
    out = {}
    g = phantoms.slab_box(10, 10, 160)
    out["xe"], out["ye"], out["ze"], out["density"] = g.xe, g.ye, g.ze, g.density
    for dbg in (1, 0):
        sp = rb.make_spot(130.0, 2.0, 100, pos=(0, 0, 0.5), energy_const=False, sigma_e=1.3)
        vtx, seeds = rb.ref_sample_spot(sp, [0, 0, 0, 0], [0, 0, 0], 5, 2, 100, debug=dbg)
        out["vtx_%d" % dbg], out["seeds_%d" % dbg] = vtx, seeds
        ra, _ = rb.ref_transport(dbg, g, vtx, seeds, mask=rb.SC_DOSE | rb.SC_LETD, kernel=0)
        for k, a in enumerate(ra):
            out["plain_%d_%d" % (dbg, k)] = a
        ra, _ = rb.ref_transport(dbg, g, vtx, seeds, mask=rb.SC_DOSE | rb.SC_STAT, kernel=1)
        for k, a in enumerate(ra):
            out["stat_%d_%d" % (dbg, k)] = a
    return out


def _batches(g, spots, dbg, per_spot, n_batches, mask, threads, seed0):
    """n_batches independent reference runs of the spot list (per_spot
    histories per spot each); returns the per-batch scorer arrays divided by
    the histories of the batch.

    Vertices come from the reference's beamlet sampling; the per-history
    transport seeds do not: run_by_spot draws them as consecutive outputs of
    a minstd engine (mqi_tps_env.hpp:1537-1550) and the CPU transport seeds
    each history's minstd engine with one, so history i+1 replays history
    i's random numbers shifted by one draw and the histories of a batch are
    strongly correlated (batch means scatter ~25x more than independent
    histories allow).  Independent seeds from a 64-bit generator keep the
    reference's per-history algorithm and make the statistics nominal."""
    out = []
    for b in range(n_batches):
        vtx, seeds = [], []
        for k, sp in enumerate(spots):
            v, _ = rb.ref_sample_spot(sp, [0, 0, 0, 0], [0, 0, 0], seed0 + b, k, per_spot, debug=dbg)
            sd = np.random.default_rng([seed0 + b, k]).integers(1, 2 ** 31 - 1, size=per_spot).astype(np.int32)
            vtx.append(v)
            seeds.append(sd)
        vtx = np.concatenate(vtx)
        seeds = np.concatenate(seeds)
        arrs, tracked = rb.ref_transport(dbg, g, vtx, seeds, mask=mask, kernel=0, threads=threads)
        out.append([a / tracked for a in arrs])
        print("  batch %d: %d histories" % (b, tracked), flush=True)
    return out


def _mean_se(xs):
    a = np.array(xs)
    return a.mean(0), a.std(0, ddof=1) / np.sqrt(len(xs))


def profiles_fixture(threads, only=None):
    """High-statistics reference runs for the statistical GPU tests.

    c1  water 200x200x350 @ 0.5x0.5x1 mm, one 200 MeV spot (sigma 3 mm, no
        energy spread), debug physics, 8 batches x 1.25e6 histories: dose in
        the box above 10% of the maximum (plus 3 voxels), the integral depth
        dose and the integral dose with their standard errors.
    c2  the slab phantom with the 25-spot plan (70-230 MeV, 1% spread,
        sigma 4 mm), release physics: dose from 8 batches x 6.25e5 histories
        per spot (central xz / yz planes, integral depth dose); LETd
        (numerator and denominator scored without the dose scorer) from 4
        batches x 1.25e5 per spot (per-batch sums per depth over the field,
        central planes).
    only="c2" recomputes c2 and keeps the c1 arrays of the existing file.
    Dose is per history (Gy)."""
    from moquimc_amd import workloads
    out = {}
    if only == "c2":
        prev = np.load(os.path.join(HERE, "reference_profiles.npz"))
        out.update({k: prev[k] for k in prev.files if k.startswith("c1_")})
    if only != "c2":
        t0 = time.time()
        g = phantoms.water_box(200, 200, 350, dx=0.5, dz=1.0)
        sp = rb.make_spot(200.0, 3.0, 0, pos=(0, 0, 0.5))
        runs = _batches(g, [sp], 1, 1250000, 8, rb.SC_DOSE, threads, 1000)
        m, se = _mean_se([r[0] for r in runs])
        d3, s3 = m.reshape(350, 200, 200), se.reshape(350, 200, 200)
        # the region above 10 % of the maximum plus 3 voxels (the 1 mm search of
        # the gamma index), as float16 relative to the maximum (error <= 2.5e-4)
        hot = np.argwhere(d3 >= 0.1 * d3.max())
        lo = np.maximum(hot.min(0) - 3, 0)
        hi = np.minimum(hot.max(0) + 4, d3.shape)
        out["c1_box"] = np.concatenate([lo, hi]).astype(np.int32)   # z0 y0 x0 z1 y1 x1
        out["c1_dose_scale"] = np.array(d3.max())
        out["c1_dose_rel"] = (d3[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] / d3.max()).astype(np.float16)
        out["c1_dose_se_rel"] = (s3[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] / d3.max()).astype(np.float16)
        idd = np.array([r[0].reshape(350, -1).sum(1) for r in runs])
        out["c1_idd"], out["c1_idd_se"] = _mean_se(idd)
        tot = np.array([r[0].sum() for r in runs])
        out["c1_total"], out["c1_total_se"] = _mean_se(tot)
        out["c1_histories"] = np.array(8 * 1250000)
        print("c1 done in %.0f s" % (time.time() - t0), flush=True)

    t0 = time.time()
    w = workloads.slabs()
    g = rb.Grid(w.xe, w.ye, w.ze, w.density)
    spots = [s.copy() for s in w.spots]
    runs = _batches(g, spots, 0, 625000, 8, rb.SC_DOSE, threads, 2000)
    m, se = _mean_se([r[0] for r in runs])
    d3, s3 = m.reshape(350, 200, 200), se.reshape(350, 200, 200)
    out["c2_xz"], out["c2_xz_se"] = d3[:, 100, :].astype(np.float32), s3[:, 100, :].astype(np.float32)
    out["c2_yz"], out["c2_yz_se"] = d3[:, :, 100].astype(np.float32), s3[:, :, 100].astype(np.float32)
    idd = np.array([r[0].reshape(350, -1).sum(1) for r in runs])
    out["c2_idd"], out["c2_idd_se"] = _mean_se(idd)
    tot = np.array([r[0].sum() for r in runs])
    out["c2_total"], out["c2_total_se"] = _mean_se(tot)
    out["c2_histories"] = np.array(8 * 625000 * len(spots))
    del runs
    # LETd: per-batch dose-weighted sums per depth over the whole field (numerator
    # and denominator), and the central planes of the summed batches
    runs = _batches(g, spots, 0, 125000, 4, rb.SC_LETD, threads, 3000)
    out["c2_let_num_z"] = np.array([r[0].reshape(350, -1).sum(1) for r in runs])
    out["c2_let_den_z"] = np.array([r[1].reshape(350, -1).sum(1) for r in runs])
    num = np.sum([r[0] for r in runs], 0).reshape(350, 200, 200)
    den = np.sum([r[1] for r in runs], 0).reshape(350, 200, 200)
    out["c2_let_num_xz"], out["c2_let_den_xz"] = num[:, 100, :].astype(np.float32), den[:, 100, :].astype(np.float32)
    out["c2_let_num_yz"], out["c2_let_den_yz"] = num[:, :, 100].astype(np.float32), den[:, :, 100].astype(np.float32)
    out["c2_let_histories"] = np.array(4 * 125000 * len(spots))
    print("c2 done in %.0f s" % (time.time() - t0), flush=True)
    return out


def ct_fixture(threads, per_spot=10000, n_batches=8, let_per_spot=0, let_batches=0):
    """The 100-spot sub-plan on the cropped CT of the 2000-spot plan
    (phantoms.ct_crop_plan), release physics, n_batches x per_spot histories
    per spot.  Vertices from the beamlet sampling of the spot records
    (oracle_sample_spot with the reference's generator, equal to the
    reference's sampling bit for bit, tests/test_oracle_vs_reference.py),
    independent per-history transport seeds as in _batches; density from
    the reference's hu_to_density.  Stored: the inputs (HU crop, edges,
    spots) and the dose summed over z (x-y map), its integral depth dose
    along y and the integral dose, with batch standard errors (Gy per
    history); with let_batches, the per-batch LETd numerator and denominator
    summed per depth."""
    t0 = time.time()
    xe, ye, ze, hu, spots = phantoms.ct_crop_plan(per_spot)
    nz, ny, nx = hu.shape
    rho = np.zeros(hu.size, np.float32)
    h = np.ascontiguousarray(hu.ravel())
    rb.reference(0).ref_hu_to_density(rb.ptr(h, rb.ctypes.c_void_p), rb.ctypes.c_int64(h.size), rb.ptr(rho, rb.c_f))
    g = rb.Grid(xe, ye, ze, rho)
    runs = []
    for b in range(n_batches):
        vtx, seeds = [], []
        for k, sp in enumerate(spots):
            v, _ = rb.oracle_sample_spot(sp, 4000 + b, k, per_spot, rng=0)
            vtx.append(v)
            seeds.append(np.random.default_rng([4000 + b, k]).integers(1, 2 ** 31 - 1, size=per_spot).astype(np.int32))
        arrs, tracked = rb.ref_transport(0, g, np.concatenate(vtx), np.concatenate(seeds), mask=rb.SC_DOSE, kernel=0,
                                         threads=threads)
        runs.append(arrs[0].reshape(nz, ny, nx) / tracked)
        print("  batch %d: %d histories, %.0f s" % (b, tracked, time.time() - t0), flush=True)
    out = {"c3_xe": xe, "c3_ye": ye, "c3_ze": ze, "c3_hu": hu, "c3_spots": spots,
           "c3_histories": np.array(n_batches * per_spot * len(spots))}
    if let_batches:
        # LETd numerator and denominator (scored without the dose scorer): per-batch
        # sums per depth along the beam (y)
        lets = []
        for b in range(let_batches):
            vtx, seeds = [], []
            for k, sp in enumerate(spots):
                v, _ = rb.oracle_sample_spot(sp, 5000 + b, k, let_per_spot, rng=0)
                vtx.append(v)
                seeds.append(np.random.default_rng([5000 + b, k]).integers(1, 2 ** 31 - 1,
                                                                            size=let_per_spot).astype(np.int32))
            arrs, tracked = rb.ref_transport(0, g, np.concatenate(vtx), np.concatenate(seeds), mask=rb.SC_LETD,
                                             kernel=0, threads=threads)
            lets.append([a.reshape(nz, ny, nx).sum((0, 2)) for a in arrs])
            print("  LETd batch %d: %d histories, %.0f s" % (b, tracked, time.time() - t0), flush=True)
        out["c3_let_num_y"] = np.array([x[0] for x in lets])
        out["c3_let_den_y"] = np.array([x[1] for x in lets])
        out["c3_let_histories"] = np.array(let_batches * let_per_spot * len(spots))
        if not n_batches:
            return out
    out["c3_xy"], out["c3_xy_se"] = _mean_se([r.sum(0) for r in runs])
    out["c3_idd"], out["c3_idd_se"] = _mean_se([r.sum((0, 2)) for r in runs])
    out["c3_total"], out["c3_total_se"] = _mean_se([r.sum() for r in runs])
    print("c3 done in %.0f s" % (time.time() - t0), flush=True)
    return out


DOSE_Q = 16384.0   # volume fixtures: dose / max stored as round(value * DOSE_Q) in uint16


def _quantise(d, s):
    """Dose relative to its maximum as uint16 steps of 1 / DOSE_Q (error
    <= 3.1e-5 of the maximum), the standard error relative to the same
    maximum as float16 with 4 significand bits (relative error <= 3 %,
    enough for a 2-sigma bar; the dropped bits are noise that would not
    compress)."""
    q = np.rint(np.clip(d, 0.0, 1.0) * DOSE_Q).astype(np.uint16)
    sb = s.astype(np.float16).view(np.uint16)
    sb = ((sb.astype(np.uint32) + 0x20) & 0xffc0).astype(np.uint16)   # round to 4 significand bits
    return q, sb.view(np.float16)


def _box_rel(d3, s3, margin):
    """The bounding box of the voxels at or above 10 % of the maximum, grown
    by `margin` voxels per axis (the gamma search reach), and the mean and
    standard error inside it relative to the maximum (_quantise)."""
    margin = np.broadcast_to(np.asarray(margin), (3,))
    hot = np.argwhere(d3 >= 0.1 * d3.max())
    lo = np.maximum(hot.min(0) - margin, 0)
    hi = np.minimum(hot.max(0) + margin + 1, d3.shape)
    sl = tuple(slice(a, b) for a, b in zip(lo, hi))
    q, s = _quantise(d3[sl] / d3.max(), s3[sl] / d3.max())
    return np.concatenate([lo, hi]).astype(np.int32), np.array(d3.max()), q, s


def volumes_fixture(threads, c2_per_spot=312500, c3_per_spot=40000, n_batches=8, only=None):
    """Volumes for 3-D gamma tests.

    c2  the slab workload (25 spots 70-230 MeV, release physics), n_batches x
        c2_per_spot histories per spot, dose; the 0.5 x 0.5 x 1 mm voxels
        rebinned laterally 2 x 2 into 1 x 1 x 1 mm (dosecmp.rebin_xy), kept in the box
        above 10 % of the maximum plus 2 voxels.
    c3  the 100-spot sub-plan on the cropped CT (phantoms.ct_crop_plan), n_batches x
        c3_per_spot histories per spot, dose on the 0.98 x 0.98 x 2 mm voxels,
        kept in the box above 10 % of the maximum plus 2 voxels (y, x) and
        1 voxel (z).
    Mean and batch standard error per voxel relative to the maximum
    (_quantise; Gy per history = maximum x dose_rel / DOSE_Q)."""
    from moquimc_amd import workloads
    path = os.path.join(HERE, "reference_volumes.npz")
    out = dict(np.load(path)) if (only and os.path.exists(path)) else {}
    if only in (None, "c2"):
        t0 = time.time()
        w = workloads.slabs()
        g = rb.Grid(w.xe, w.ye, w.ze, w.density)
        runs = _batches(g, [s.copy() for s in w.spots], 0, c2_per_spot, n_batches, rb.SC_DOSE, threads, 6000)
        vols = [dosecmp.rebin_xy(r[0].reshape(350, 200, 200), 2) for r in runs]
        m, se = _mean_se(vols)
        box, scale, d, s = _box_rel(m, se, 2)
        out.update({"c2_box": box, "c2_dose_scale": scale, "c2_dose_rel": d, "c2_dose_se_rel": s,
                    "c2_histories": np.array(n_batches * c2_per_spot * len(w.spots)),
                    "c2_batches": np.array(n_batches)})
        print("c2 volume done in %.0f s" % (time.time() - t0), flush=True)
    if only in (None, "c3"):
        t0 = time.time()
        xe, ye, ze, hu, spots = phantoms.ct_crop_plan(c3_per_spot)
        nz, ny, nx = hu.shape
        rho = np.zeros(hu.size, np.float32)
        h = np.ascontiguousarray(hu.ravel())
        rb.reference(0).ref_hu_to_density(rb.ptr(h, rb.ctypes.c_void_p), rb.ctypes.c_int64(h.size),
                                          rb.ptr(rho, rb.c_f))
        g = rb.Grid(xe, ye, ze, rho)
        vols = []
        for b in range(n_batches):
            vtx, seeds = [], []
            for k, sp in enumerate(spots):
                v, _ = rb.oracle_sample_spot(sp, 7000 + b, k, c3_per_spot, rng=0)
                vtx.append(v)
                seeds.append(np.random.default_rng([7000 + b, k]).integers(1, 2 ** 31 - 1,
                                                                            size=c3_per_spot).astype(np.int32))
            arrs, tracked = rb.ref_transport(0, g, np.concatenate(vtx), np.concatenate(seeds), mask=rb.SC_DOSE,
                                             kernel=0, threads=threads)
            vols.append(arrs[0].reshape(nz, ny, nx) / tracked)
            print("  c3 batch %d: %d histories, %.0f s" % (b, tracked, time.time() - t0), flush=True)
        m, se = _mean_se(vols)
        box, scale, d, s = _box_rel(m, se, [1, 2, 2])
        out.update({"c3_box": box, "c3_dose_scale": scale, "c3_dose_rel": d, "c3_dose_se_rel": s,
                    "c3_histories": np.array(n_batches * c3_per_spot * len(spots)),
                    "c3_batches": np.array(n_batches)})
        print("c3 volume done in %.0f s" % (time.time() - t0), flush=True)
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--volumes", choices=["all", "c2", "c3"], default=None,
                    help="write reference_volumes.npz (3-D dose of the slab workload and the CT sub-plan)")
    ap.add_argument("--stats", action="store_true")
    ap.add_argument("--only", choices=["c2"], default=None, help="with --stats: recompute the slab workload only")
    ap.add_argument("--ct", action="store_true", help="write reference_ct.npz (the cropped-CT sub-plan)")
    ap.add_argument("--ct-per-spot", type=int, default=10000)
    ap.add_argument("--ct-let-only", action="store_true", help="with --ct: add the LETd arrays to reference_ct.npz")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    if not rb.have_reference():
        raise SystemExit("build the reference harness first: make -C oracle")
    if args.volumes:
        volumes_fixture(args.threads, only=None if args.volumes == "all" else args.volumes)
        return
    if args.ct:
        path = os.path.join(HERE, "reference_ct.npz")
        if args.ct_let_only:
            out = dict(np.load(path))
            new = ct_fixture(args.threads, per_spot=args.ct_per_spot, n_batches=0, let_per_spot=args.ct_per_spot,
                             let_batches=4)
            out.update({k: v for k, v in new.items() if k.startswith("c3_let")})
        else:
            out = ct_fixture(args.threads, per_spot=args.ct_per_spot)
        np.savez_compressed(path, **out)
        return
    if args.stats:
        out = profiles_fixture(args.threads, only=args.only)
        np.savez_compressed(os.path.join(HERE, "reference_profiles.npz"), **out)
        return
    tables = [np.zeros(600, np.float32) for _ in range(6)] + [np.zeros(3996, np.float32)]
    rb.reference(0).ref_tables(*[rb.ptr(t, rb.c_f) for t in tables])
    names = ["cs_p_ion", "dedx_restricted", "range_steps", "cs_pp_e", "cs_po_e", "cs_po_i", "density_correction"]
    np.savez_compressed(os.path.join(HERE, "reference_tables.npz"), **dict(zip(names, tables)))
    np.savez_compressed(os.path.join(HERE, "reference_functions.npz"), **functions())
    np.savez_compressed(os.path.join(HERE, "reference_grid.npz"), **grid_probes())
    np.savez_compressed(os.path.join(HERE, "reference_trace.npz"), **trace_fixture())
    np.savez_compressed(os.path.join(HERE, "reference_transport.npz"), **transport_fixture())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
