"""Benchmark: proton transport histories/s on one MI355X, batched persistent
launch vs plain per-spot launches, on the 2000-spot synthetic CT plan; plus
the Dij adjoint (D^T g) bandwidth on the 5000-spot Dij workload.

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


def transport_roofline(res, ms_per_pass, n_scorers):
    # algorithmic bytes of the transport: per voxel step one density load
    # (4 B) and one FP64 read-modify-write per scorer (8 B each)
    bytes_per_pass = res.steps * (4.0 + 8.0 * n_scorers)
    achieved = bytes_per_pass / (ms_per_pass * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
            "kernel": "k_transport (batched)", "steps_per_pass": int(res.steps),
            "bytes_per_step": 4 + 8 * n_scorers}


def cpu_baseline(w, cores, target_s=20.0):
    """Reference CPU transport (oracle/_ref, release physics) on a bounded
    sample of the plan: every k-th spot with its full history count."""
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
    rb.ref_transport(0, grid, vtx[:pilot], seeds[:pilot], mask=rb.SC_DOSE, kernel=0, threads=cores, timing=t)
    rate = pilot / max(t[0], 1e-6)
    n = int(min(len(seeds), max(pilot, rate * target_s)))
    t = []
    _, tracked = rb.ref_transport(0, grid, vtx[:n], seeds[:n], mask=rb.SC_DOSE, kernel=0, threads=cores, timing=t)
    return {"value": round(tracked / t[0], 1), "unit": "histories/s", "cores": cores, "kind": "reference",
            "sample": "%d histories of %d spots (every %dth spot of the plan, full history counts), reference "
                      "transport_particles_patient on the CPU, release physics, %d threads, %.1f s"
                      % (tracked, len(picks), stride, cores, t[0])}


def adjoint_bench(ctx_cls, reps=20):
    """Dij workload: transport with the Dij scorer, CSR build (1e-4 of spot
    max), then time D^T g with HIP events."""
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
    for _ in range(3):
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
    for _ in range(3):
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
                                 "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                                 "kernel": "k_csr_adaptive + k_csr_split_rows (D^T g)", "algorithmic_bytes": int(alg_bytes)}}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-passes", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--exact", action="store_true", help="time the verification kernels instead")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (the reference kernel of the HBM counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would take its own slice "
              "of the (scenario, spot) work list; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return

    from moquimc_amd import _abi, workloads
    t0 = time.time()
    w = workloads.ct()
    log("workload generated in %.1f s: %d spots, %d histories" % (time.time() - t0, len(w.spots), w.histories))
    ctx = _abi.Context(0)
    w.load(ctx)
    cfg = ctx.config(seed=11, batched=True, exact=args.exact)
    for i in range(args.warmup):
        ctx.zero_scorers()
        r = ctx.run(cfg)
        log("warmup %d: %.3f s kernel, %d histories" % (i, r.kernel_ms / 1e3, r.histories))
    kernel_ms, hist, last = 0.0, 0, None
    tw0 = time.perf_counter()
    for i in range(args.steps):
        ctx.zero_scorers()
        cfg.hist_offset = (i + 1) * 200000
        last = ctx.run(cfg)
        kernel_ms += last.kernel_ms
        hist += last.histories
        log("step %d: %.3f s kernel, %d histories, %d steps" % (i, last.kernel_ms / 1e3, last.histories, last.steps))
    wall = time.perf_counter() - tw0
    ms_per_step = wall / args.steps * 1e3
    value = hist / wall

    # plain per-spot launches
    plain_ms, plain_hist = 0.0, 0
    tp0 = time.perf_counter()
    for i in range(args.plain_passes):
        ctx.zero_scorers()
        r = ctx.run(ctx.config(seed=11, batched=False, exact=args.exact, hist_offset=(i + 1) * 200000))
        plain_ms += r.kernel_ms
        plain_hist += r.histories
        log("plain pass %d: %.3f s kernel" % (i, r.kernel_ms / 1e3))
    plain_wall = time.perf_counter() - tp0
    plain_rate = plain_hist / plain_wall if plain_hist else None

    out = {"metric": METRIC, "value": round(value, 1), "unit": "histories/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2), "higher_is_better": True,
           "scaling": "strong", "vs_baseline": None, "dtype": "f32 (FP64 dose accumulation)",
           "data": "synthetic (seeded CT and plan generators in moquimc_amd/workloads.py)",
           "config": {"workload": w.name + ", one persistent launch per pass, release physics",
                      "histories_per_step": int(w.histories), "spots": int(len(w.spots)),
                      "grid": "512x512x200", "scorers": "dose"},
           "kernel_histories_per_s": round(hist / (kernel_ms / 1e3), 1),
           "plain_histories_per_s": round(plain_rate, 1) if plain_rate else None,
           "plain_kernel_histories_per_s": round(plain_hist / (plain_ms / 1e3), 1) if plain_ms else None,
           "batched_over_plain": round(value / plain_rate, 3) if plain_rate else None,
           "steps_per_history": round(last.steps / max(1, last.histories), 1),
           "roofline": transport_roofline(last, kernel_ms / args.steps, 1)}
    ctx.close()
    if not args.no_adjoint:
        try:
            out.update(adjoint_bench(_abi.Context))
        except Exception as e:  # reported, not hidden
            out["adjoint_error"] = repr(e)
    if not args.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        try:
            out["cpu_baseline"] = cpu_baseline(w, cores)
        except Exception as e:
            out["cpu_baseline"] = {"error": repr(e)}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
