"""Benchmark: proton transport histories/s on one MI355X, batched persistent
launch vs plain per-spot launches, on the 2000-spot synthetic CT plan; plus
the Dij adjoint (D^T g) bandwidth on the 5000-spot Dij workload, the
9-scenario robust Dij leg, and the water and slab phantom workloads.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one transport pass over every history of the plan (4.27e7
histories).  Prints one JSON line.  The CPU baseline is the reference's own
CPU transport (oracle/_ref, built from the reference sources) timed on this
host's cores on a bounded sample of the same plan.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "histories/s, 1 MI355X, gamma≥99% 1%/1mm, batched vs plain; Dij adjoint GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")
TRAFFIC_SOURCE = "profiles/hbm_counters_ct2000_dij5000.md (rocprofv3 FETCH_SIZE and WRITE_SIZE, copy-calibrated)"
COUNTER_SOURCE = "profiles/transport_counters_ct2000.md (rocprofv3 SQ_* and TCC_ATOMIC of the same pass)"


def measured_traffic(key):
    """HBM bytes per launch of a kernel from the committed counter runs, or
    None when they are absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)[key]
    except (OSError, KeyError, ValueError):
        return None


def transport_roofline(steps, kernel_ms, launches, n_scorers, atomic_ceiling):
    """Algorithmic bytes of the transport: per voxel step one density load
    (4 B) and one FP64 read-modify-write per scorer (8 B each), over the
    average duration of one persistent launch (HIP events on its stream).
    Beside it the two ceilings the counters point at: the VALU issue share
    and lane activity of the pass (SQ counters), and its FP64 atomic rate
    (TCC_ATOMIC per voxel step from the counter run, scaled to this launch)
    against the rate mqi_atomic_probe measures for random FP64 atomics on
    this device (atomic_ceiling, atomics/s, None when not measured)."""
    bytes_per_step = 4.0 + 8.0 * n_scorers
    per_launch = steps / launches * bytes_per_step
    ms_per_launch = kernel_ms / launches
    achieved = per_launch / (ms_per_launch * 1e-3) / 1e9
    m = measured_traffic("k_transport")
    # counters per voxel step, scaled to this launch's steps
    traffic = None if m is None else round(m["traffic_bytes"] / m["voxel_steps"] * steps / launches)
    out = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
           "traffic_source": TRAFFIC_SOURCE if m else None,
           "kernel": "k_transport<release, dynamic> (persistent)", "launch_ms": round(ms_per_launch, 3),
           "voxel_steps_per_launch": int(steps / launches), "bytes_per_step": bytes_per_step,
           "scorer_deposits_per_s": round(steps * n_scorers / (kernel_ms * 1e-3), 1)}
    sq = measured_traffic("k_transport_sq")
    if sq:
        out["valu_issue_frac"] = sq["valu_issue_frac"]
        out["valu_lanes_active"] = sq["valu_lanes_active"]
    if m and "atomics" in m:
        rate = m["atomics"] / m["voxel_steps"] * steps / launches / (ms_per_launch * 1e-3)
        out["atomic_roofline"] = {"achieved": float("%.4g" % rate), "peak": atomic_ceiling, "unit": "atomics/s",
                                  "frac": round(rate / atomic_ceiling, 4) if atomic_ceiling else None,
                                  "atomics_per_voxel_step": round(m["atomics"] / m["voxel_steps"], 4),
                                  "peak_source": ("mqi_atomic_probe, uniformly random FP64 atomic adds on this "
                                                  "device, this run (clustered streams run faster: "
                                                  "profiles/atomic_probes.md)"),
                                  "counter_source": COUNTER_SOURCE}
    out["limiter"] = ("VALU issue (valu_issue_frac of the issue slots busy); the FP64 atomic requests of the "
                      "paired scorers run below the device's random-atomic rate (atomic_roofline) and HBM "
                      "bandwidth is far from its ceiling (" + COUNTER_SOURCE + ")")
    return out


def atomic_ceiling(ctx, n_slots=2 * 512 * 512 * 200, n_updates=1 << 30):
    """Random FP64 atomic adds per second on this device (mqi_atomic_probe
    over two CT-dose-sized arrays)."""
    ms = ctx.atomic_probe(n_slots, n_updates)
    return float("%.4g" % (n_updates / (ms * 1e-3)))


def dose_agreement(batched, plain, w, dd=0.01, dta=1.0, threshold=0.1):
    """Agreement of the batched and the plain form's dose on the same
    histories: the largest difference relative to the maximum, and the
    global 3-D gamma pass rate (dd of the maximum, dta mm, voxels above
    threshold of the maximum; the evaluated dose searched over the voxel
    offsets within dta, no interpolation), computed on the GPU."""
    import itertools

    import torch
    dev = torch.device("cuda", 0)
    nz, ny, nx = w.shape
    spacing = (float(w.ze[1] - w.ze[0]), float(w.ye[1] - w.ye[0]), float(w.xe[1] - w.xe[0]))
    r = torch.from_numpy(batched.reshape(nz, ny, nx)).to(dev)
    e = torch.from_numpy(plain.reshape(nz, ny, nx)).to(dev)
    dmax = r.max()
    reach = [int(dta // sp) for sp in spacing]
    ep = torch.nn.functional.pad(e[None, None], (reach[2], reach[2], reach[1], reach[1], reach[0], reach[0]),
                                 value=float("inf"))[0, 0]
    best = torch.full_like(r, float("inf"))
    for off in itertools.product(*[range(-k, k + 1) for k in reach]):
        d2 = sum((o * sp) ** 2 for o, sp in zip(off, spacing)) / (dta * dta)
        if d2 > 1.0:
            continue
        sh = ep[reach[0] + off[0]:reach[0] + off[0] + nz, reach[1] + off[1]:reach[1] + off[1] + ny,
                reach[2] + off[2]:reach[2] + off[2] + nx]
        best = torch.minimum(best, d2 + ((sh - r) / (dd * dmax)) ** 2)
    mask = r >= threshold * dmax
    out = {"gamma_1pct_1mm": round(float((best[mask] <= 1.0).double().mean()), 6), "voxels": int(mask.sum()),
           "max_abs_diff_of_max": float("%.3g" % float((r - e).abs().max() / dmax))}
    del r, e, ep, best, mask
    torch.cuda.empty_cache()
    return out


def _stdout_to_stderr(fn, *a):
    """Run fn with file descriptor 1 on stderr (C stdio buffers flushed
    before it is restored): the reference library prints to stdout, and the
    one JSON line is the only thing bench.py writes there."""
    import ctypes
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        return fn(*a)
    finally:
        ctypes.CDLL(None).fflush(None)
        os.dup2(saved, 1)
        os.close(saved)


def cpu_model():
    """The host CPU's model name (first /proc/cpuinfo entry), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(w, cores, target_s=20.0):
    """Reference CPU transport (oracle/_ref, release physics) on a bounded
    sample of the plan: every k-th spot with its full history count, on all
    the cores this process may run on, plus a one-core rate."""
    return _stdout_to_stderr(_cpu_baseline, w, cores, target_s)


def _cpu_baseline(w, cores, target_s):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refbind as rb
    if not os.path.exists(os.path.join(rb.REF_DIR, "libmoqui_ref_release.so")):
        return None
    import ctypes
    lib = rb.reference(0)
    hu = np.ascontiguousarray(w.hu.reshape(-1))
    rho = np.zeros(hu.size, np.float32)
    lib.ref_hu_to_density(hu.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(hu.size),
                          rho.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    grid = rb.Grid(w.xe, w.ye, w.ze, rho)
    # pilot: one spot on all cores to size the sample
    stride = 10
    picks = list(range(0, len(w.spots), stride))
    vtx, seeds = [], []
    for j in picks:
        s = w.spots[j]
        # coordinate_transform(angles = iec2dicom 90 deg, isocentre = the spot's
        # translation) is the plan's beam-to-world map
        v, sd = rb.ref_sample_spot(s, [0, 0, 0, 90], s["translation"], 3, j, int(s["histories"]), debug=0)
        vtx.append(v)
        seeds.append(sd)
    vtx = np.concatenate(vtx)
    seeds = np.concatenate(seeds)
    pilot = min(len(seeds), 4000 * cores)
    t = []
    mask = rb.SC_DOSE | rb.SC_STAT   # transport_particles_patient_stat with the bench's scorers
    rb.ref_transport(0, grid, vtx[:pilot], seeds[:pilot], mask=mask, kernel=1, threads=cores, timing=t)
    rate = pilot / max(t[0], 1e-6)
    n = int(min(len(seeds), max(pilot, rate * target_s)))
    t = []
    _, tracked = rb.ref_transport(0, grid, vtx[:n], seeds[:n], mask=mask, kernel=1, threads=cores, timing=t)
    # one core, on the first histories of the same sample
    n1 = int(min(n, max(2000, rate / cores * 5.0)))
    t1 = []
    _, tracked1 = rb.ref_transport(0, grid, vtx[:n1], seeds[:n1], mask=mask, kernel=1, threads=1, timing=t1)
    return {"value": round(tracked / t[0], 1), "unit": "histories/s", "cores": cores, "kind": "reference",
            "cpu_model": cpu_model(), "per_core_histories_per_s": round(tracked1 / t1[0], 1),
            "sample": "%d histories of %d spots (every %dth spot of the plan, full history counts), reference "
                      "transport_particles_patient_stat on the CPU (dose + stat scorers), release physics, "
                      "%d threads with a world and scorer tables each (independent, as separate processes), "
                      "%.1f s; one-core rate from %d of those histories in %.1f s"
                      % (tracked, len(picks), stride, cores, t[0], tracked1, t1[0])}


def adjoint_bench(ctx_cls, reps=20):
    """Dij workload: transport with the Dij scorer, CSR build (1e-4 of spot
    max), then D^T g and D w timed with HIP events: the median of reps
    products after reps untimed ones."""
    import ctypes

    from moquimc_amd import _abi, workloads
    import torch
    w = workloads.dij()
    ctx = ctx_cls(0)
    w.load(ctx)
    t0 = time.time()
    # per-spot hash of 2^19 entries (12 B each, 31 GB for 5000 spots): a 180 MeV
    # spot of 2e4 histories touches ~1.5e5 voxels of this grid
    res = ctx.run(ctx.config(seed=7, batched=True, dij=True, dij_capacity=1 << 19))
    t_tr = time.time() - t0
    ctx.dij_build(1e-4)
    nnz, rows, cols = ctx.dij_info(0)
    dev = torch.device("cuda", 0)
    g = torch.randn(cols, device=dev, dtype=torch.float32)
    y = torch.zeros(rows, device=dev, dtype=torch.float64)
    for _ in range(reps):   # warm-up: clocks settle over the first tens of products
        ctx.dij_matvec(0, 1, g.data_ptr(), y.data_ptr(), y_double=True)
    ms = []
    for _ in range(reps):
        ctx.dij_matvec(0, 1, g.data_ptr(), y.data_ptr(), y_double=True)
        ms.append(ctx.last_matvec_ms())
    ms_med = float(np.median(ms))
    alg_bytes = 8.0 * nnz + 8.0 * (rows + 1) + 8.0 * rows + 4.0 * cols
    gbs = alg_bytes / (ms_med * 1e-3) / 1e9
    # forward D w on the voxel-major copy (float32 dose out)
    w8 = torch.rand(rows, device=dev, dtype=torch.float32) + 0.5
    d = torch.zeros(cols, device=dev, dtype=torch.float32)
    for _ in range(reps):
        ctx.dij_matvec(0, 0, w8.data_ptr(), d.data_ptr(), y_double=False)
    fms = []
    for _ in range(reps):
        ctx.dij_matvec(0, 0, w8.data_ptr(), d.data_ptr(), y_double=False)
        fms.append(ctx.last_matvec_ms())
    f_med = float(np.median(fms))
    f_bytes = 8.0 * nnz + 8.0 * (cols + 1) + 4.0 * cols + 4.0 * rows
    ctx.close()
    return {"dij_workload": w.name, "dij_histories": int(res.histories), "dij_transport_s": round(t_tr, 2),
            "dij_nnz": int(nnz), "dij_rows": int(rows), "dij_cols": int(cols),
            "adjoint_ms": round(ms_med, 4), "adjoint_GBs": round(gbs, 1),
            "forward_ms": round(f_med, 4), "forward_GBs": round(f_bytes / (f_med * 1e-3) / 1e9, 1),
            "adjoint_roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(gbs / HBM_PEAK_GBS, 4),
                                 "traffic": (round(measured_traffic("k_pieces_adjoint")["traffic_bytes"])
                                             if measured_traffic("k_pieces_adjoint") else None),
                                 "traffic_source": TRAFFIC_SOURCE,
                                 "kernel": "k_gather_x + k_pieces_staged + k_piece_rows (D^T g)",
                                 "algorithmic_bytes": int(alg_bytes)}}


def phantom_bench(ctx_cls, reps=3):
    """The two phantom workloads: water (200x200x350, one 200 MeV spot, 1e5
    histories, dose, phantom-build physics) and slabs (25 spots 70-230 MeV,
    1e6 histories each, dose + LETd).  One batched launch per run, launch
    shape from the autotuner; histories/s over the wall time of the best of
    reps runs (each run synchronised)."""
    from moquimc_amd import workloads
    out = {}
    for key, w, sample in (("water", workloads.water(), 1.0), ("slabs", workloads.slabs(), 0.1)):
        ctx = ctx_cls(0)
        w.load(ctx)
        cfg = ctx.config(seed=5, batched=True, physics_debug=w.debug)
        cfg, _ = ctx.autotune(cfg, sample=sample)
        ctx.zero_scorers()
        ctx.run(cfg)
        best, kms = None, None
        for i in range(reps):
            ctx.zero_scorers()
            cfg.hist_offset = (i + 1) * (1 << 24)
            t0 = time.perf_counter()
            r = ctx.run(cfg)
            dt = time.perf_counter() - t0
            if best is None or dt < best:
                best, kms = dt, r.kernel_ms
        ctx.close()
        out[key] = {"workload": w.name, "physics": "phantom build" if w.debug else "release",
                    "histories": int(w.histories), "wall_s": round(best, 4),
                    "histories_per_s": round(w.histories / best, 1),
                    "kernel_histories_per_s": round(w.histories / (kms / 1e3), 1)}
        log("%s: %.4f s wall, %.0f histories/s" % (key, best, w.histories / best))
    return {"phantoms": out}


def transport_pass(waves=3, block=128, fetch=1, refill=4):
    """One production batched pass of the 2000-spot CT plan (dose + the
    statistics pair) with a fixed launch shape: a single k_transport dispatch
    for the counter runs."""
    from moquimc_amd import _abi, workloads
    w = workloads.ct()
    w.mask = _abi.SCORE_DOSE | _abi.SCORE_STAT
    ctx = _abi.Context(0)
    w.load(ctx)
    cfg = ctx.config(seed=11, batched=True)
    cfg.waves_per_simd, cfg.block_size, cfg.fetch_chunk, cfg.refill_lanes = waves, block, fetch, refill
    r = ctx.run(cfg)
    ctx.close()
    print(json.dumps({"histories": int(r.histories), "voxel_steps": int(r.steps), "kernel_ms": round(r.kernel_ms, 2),
                      "launch": {"waves_per_simd": waves, "block_size": block, "fetch_chunk": fetch,
                                 "refill_lanes": refill}}), flush=True)


def atomic_probes(n_slots=2 * 512 * 512 * 200, n_updates=1 << 30):
    """FP64 atomic adds per second into two CT-dose-sized arrays, uniform and
    with each wave's 64 adds inside windows of 64 .. 262144 slots: how
    address clustering raises the rate the transport's scoring runs at."""
    from moquimc_amd import _abi
    ctx = _abi.Context(0)
    out = {}
    for window in (0, 64, 512, 4096, 32768, 262144):
        ms = ctx.atomic_probe(n_slots, n_updates, window)
        out["uniform" if window == 0 else "window_%d" % window] = float("%.4g" % (n_updates / (ms * 1e-3)))
    ctx.close()
    print(json.dumps({"atomics_per_s": out, "n_slots": n_slots, "n_updates": n_updates}), flush=True)


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain device copy of known size: FETCH_SIZE / WRITE_SIZE of its
    dispatches against nbytes give the scale of the HBM counters."""
    import torch
    dev = torch.device("cuda", 0)
    x = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}), flush=True)


def robust_bench(ctx_cls, reps=5):
    """The 2000-spot CT plan x 9 robust scenarios (nominal 2x histories) in
    one launch with the Dij scorer (tables of 4096 sqrt(histories) slots per
    scenario and spot); 9 CSR matrices; worst-case quadratic objective and
    its gradient through the autograd layer, checked against an FP64 CPU
    gradient (scipy) of the worst scenario."""
    import scipy.sparse as sps
    import torch

    from moquimc_amd import dij, workloads
    w = workloads.robust()
    ctx = ctx_cls(0)
    w.load(ctx)
    log("robust: %s, %d histories per scenario unit" % (w.name, w.histories))
    t0 = time.time()
    res = ctx.run(ctx.config(seed=13, batched=True, dij=True, dij_slots_sqrt=4096))
    t_tr = time.time() - t0
    log("robust: transport %.1f s, %d histories, status %d" % (t_tr, res.histories, res.status))
    t0 = time.time()
    ctx.dij_build(1e-4)
    t_build = time.time() - t0
    log("robust: Dij build %.1f s" % t_build)
    budgets = w.scenarios[2]
    # per unit history budget: every scenario on the dose scale of budget 1
    mats = [dij.DijMatrix(ctx, s, scale=1.0 / float(budgets[s])) for s in range(len(budgets))]
    dev = torch.device("cuda", 0)
    n_vox = mats[0].n_voxels
    target = torch.from_numpy(workloads.target_prescription(w.xe, w.ye, w.ze, radius=30.0, centre=(0.0, 0.0, 0.0),
                                                            dose=1.0)).to(dev)
    gen = torch.Generator(device="cpu").manual_seed(5)
    wt = (torch.rand(mats[0].n_spots, generator=gen) + 0.5).to(dev)
    with torch.no_grad():
        d0 = mats[0].matvec(wt)
        scale = d0[target > 0].mean()
    presc = target.double() * scale
    ms = []
    for _ in range(reps + 1):
        wv = wt.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f, vals = dij.worst_case_objective(mats, wv, presc)
        f.backward()
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    k = int(torch.argmax(vals).item())
    log("robust: objective+gradient %.2f ms, worst scenario %d; FP64 CPU check" % (float(np.median(ms[1:])), k))
    indptr, idx, val, shape = ctx.dij_csr(k)
    D = sps.csr_matrix((val.astype(np.float64) / float(budgets[k]), idx, indptr), shape=shape)   # spots x voxels
    wn = wt.double().cpu().numpy()
    r = D.T @ wn - presc.cpu().numpy()
    g_ref = 2.0 * (D @ r) / n_vox
    g = wv.grad.double().cpu().numpy()
    err = float(np.abs(g - g_ref).max() / np.abs(g_ref).max())
    out = {"robust_workload": w.name, "robust_histories": int(res.histories),
           "robust_transport_s": round(t_tr, 2), "robust_kernel_s": round(res.kernel_ms / 1e3, 2),
           "robust_histories_per_s": round(res.histories / (res.kernel_ms / 1e3), 1),
           "robust_dij_build_s": round(t_build, 2), "robust_nnz": [int(m.nnz) for m in mats],
           "robust_objective_grad_ms": round(float(np.median(ms[1:])), 2), "robust_worst_scenario": k,
           "robust_objective_per_scenario": [float("%.4g" % v) for v in vals.detach().tolist()],
           "robust_grad_rel_err_fp64": err}
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-passes", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-robust", action="store_true", help="skip the 9-scenario robust Dij leg")
    ap.add_argument("--robust-only", action="store_true", help="only the 9-scenario robust Dij leg")
    ap.add_argument("--exact", action="store_true", help="time the verification kernels instead")
    ap.add_argument("--no-tune", action="store_true", help="default launch parameters instead of the autotuner's")
    ap.add_argument("--no-graph", action="store_true", help="launch the batched rounds directly, not from a hipGraph")
    ap.add_argument("--adjoint-only", action="store_true", help="only the Dij workload and its products")
    ap.add_argument("--no-phantoms", action="store_true", help="skip the water and slab phantom leg")
    ap.add_argument("--phantoms-only", action="store_true", help="only the water and slab phantom leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (the reference kernel of the HBM counter runs)")
    ap.add_argument("--atomic-probes", action="store_true",
                    help="only the FP64 atomic rate, uniform and clustered (mqi_atomic_probe)")
    ap.add_argument("--transport-pass", action="store_true",
                    help="only one production pass of the CT plan with a fixed launch shape (the counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would take its own slice "
              "of the (scenario, spot) work list, and the dose grids would be summed once at the end).", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return
    if args.transport_pass:
        transport_pass()
        return
    if args.atomic_probes:
        atomic_probes()
        return
    if args.robust_only:
        from moquimc_amd import _abi
        print(json.dumps(robust_bench(_abi.Context)), flush=True)
        return
    if args.adjoint_only:
        from moquimc_amd import _abi
        print(json.dumps(adjoint_bench(_abi.Context)), flush=True)
        return
    if args.phantoms_only:
        from moquimc_amd import _abi
        print(json.dumps(phantom_bench(_abi.Context)), flush=True)
        return

    from moquimc_amd import _abi, workloads
    t0 = time.time()
    w = workloads.ct()
    w.mask = _abi.SCORE_DOSE | _abi.SCORE_STAT
    log("workload generated in %.1f s: %d spots, %d histories" % (time.time() - t0, len(w.spots), w.histories))
    ctx = _abi.Context(0)
    w.load(ctx)
    n_scorers = ctx.scorer_count()
    # statistical stopping: mean sd/mean <= 1 % over voxels above 50 % of the maximum
    stat = dict(stat_criterion=0.01, stat_threshold=0.5, max_rounds=16)
    cfg = ctx.config(seed=11, batched=True, exact=args.exact, **stat)
    tuned = None
    dropped = 0
    if not args.exact and not args.no_tune:
        t0 = time.time()
        cfg, best = ctx.autotune(cfg, sample=0.1)
        tuned = {"waves_per_simd": cfg.waves_per_simd, "block_size": cfg.block_size, "fetch_chunk": cfg.fetch_chunk,
                 "refill_lanes": cfg.refill_lanes, "autotune_s": round(time.time() - t0, 2)}
        log("autotuned: %s" % tuned)
    # the batched round is captured in a hipGraph once and replayed per round
    cfg.use_graph = 0 if args.no_graph else 1
    for i in range(args.warmup):
        ctx.zero_scorers()
        r = ctx.run(cfg)
        log("warmup %d: %.3f s kernel, %d histories, %d rounds, stat %.4f" % (i, r.kernel_ms / 1e3, r.histories,
                                                                             r.rounds, r.stat))
    kernel_ms, hist, launches, steps, last = 0.0, 0, 0, 0, None
    tw0 = time.perf_counter()
    for i in range(args.steps):
        ctx.zero_scorers()
        cfg.hist_offset = (i + 1) * (1 << 24)
        last = ctx.run(cfg)
        kernel_ms += last.kernel_ms
        hist += last.histories
        launches += last.rounds
        steps += last.steps
        dropped += last.dropped
        log("step %d: %.3f s kernel, %d histories, %d rounds, stat %.4f, %d voxel steps" % (
            i, last.kernel_ms / 1e3, last.histories, last.rounds, last.stat, last.steps))
    wall = time.perf_counter() - tw0
    ms_per_step = wall / args.steps * 1e3
    value = hist / wall

    # the dose of one batched pass (untimed) on the histories of the first plain pass
    ctx.zero_scorers()
    cfg.hist_offset = 1 << 24
    ctx.run(cfg)
    dose_batched = ctx.scorer(0)

    # plain per-spot launches (one launch and host sync per spot, as run_by_spot), same criterion
    plain_ms, plain_hist, plain_wall, agree = 0.0, 0, 0.0, None
    for i in range(args.plain_passes):
        ctx.zero_scorers()
        tp0 = time.perf_counter()
        r = ctx.run(ctx.config(seed=11, batched=False, exact=args.exact, hist_offset=(i + 1) * (1 << 24), **stat))
        plain_wall += time.perf_counter() - tp0
        plain_ms += r.kernel_ms
        plain_hist += r.histories
        log("plain pass %d: %.3f s kernel, %d rounds" % (i, r.kernel_ms / 1e3, r.rounds))
        if i == 0:
            agree = dose_agreement(dose_batched, ctx.scorer(0), w)
            log("batched vs plain dose: %s" % agree)
    del dose_batched
    plain_rate = plain_hist / plain_wall if plain_hist else None

    out = {"metric": METRIC, "value": round(value, 1), "unit": "histories/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2), "higher_is_better": True,
           "scaling": "strong", "vs_baseline": None, "dtype": "f32 (FP64 dose accumulation)",
           "data": "synthetic (seeded CT and plan generators in moquimc_amd/workloads.py)",
           "config": {"workload": w.name + ", one persistent launch per round, stat stop 1% above 50% max, "
                                           "release physics",
                      "histories_per_round": int(w.histories), "spots": int(len(w.spots)),
                      "grid": "512x512x200", "scorers": "dose, dose_stat, dose_square_stat", "tuned": tuned,
                      "hip_graph": bool(cfg.use_graph), "stack_depth": "10 (the reference's release depth)"},
           "rounds_per_step": round(launches / args.steps, 2),
           "kernel_histories_per_s": round(hist / (kernel_ms / 1e3), 1),
           "plain_histories_per_s": round(plain_rate, 1) if plain_rate else None,
           "plain_kernel_histories_per_s": round(plain_hist / (plain_ms / 1e3), 1) if plain_ms else None,
           "batched_over_plain": round(value / plain_rate, 3) if plain_rate else None,
           "batched_vs_plain_dose": agree,
           "secondaries_dropped_per_step": round(dropped / args.steps, 1),
           "steps_per_history": round(steps / max(1, hist), 1),
           "roofline": transport_roofline(steps, kernel_ms, launches, n_scorers, atomic_ceiling(ctx))}
    ctx.close()
    if not args.no_adjoint:
        try:
            out.update(adjoint_bench(_abi.Context))
        except Exception as e:  # reported, not hidden
            out["adjoint_error"] = repr(e)
    if not args.no_robust:
        try:
            out.update(robust_bench(_abi.Context))
        except Exception as e:  # reported, not hidden
            out["robust_error"] = repr(e)
    if not args.no_phantoms:
        try:
            out.update(phantom_bench(_abi.Context))
        except Exception as e:  # reported, not hidden
            out["phantoms_error"] = repr(e)
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        try:
            out["cpu_baseline"] = cpu_baseline(w, cores)
        except Exception as e:
            out["cpu_baseline"] = {"error": repr(e)}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
