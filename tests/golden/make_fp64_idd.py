"""FP32 stepping against FP64: the oracle runs the water workload (200 MeV,
sigma 3 mm, debug physics, 200x200x350 at 0.5x0.5x1 mm) twice with the
same Philox streams and portable math, once in float (the kernels'
arithmetic, with the reference's float/double promotions) and once in
double, and writes the integral depth doses to reference_fp64_idd.npz.
With --batches B it instead adds B independent double runs of
--batch-histories each (distinct master seeds) to the same file, the FP64
depth dose with batch errors that the production kernels are compared with.

    make -C oracle && python tests/golden/make_fp64_idd.py [--histories N] [--threads T]
    python tests/golden/make_fp64_idd.py --batches 8 --batch-histories 100000
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--histories", type=int, default=200000)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--seed", type=int, default=77)
    ap.add_argument("--batches", type=int, default=0)
    ap.add_argument("--batch-histories", type=int, default=100000)
    args = ap.parse_args()
    g = phantoms.water_box(200, 200, 350, dx=0.5, dz=1.0)
    path = os.path.join(HERE, "reference_fp64_idd.npz")
    if args.batches:
        out = dict(np.load(path))
        sp = np.array([rb.make_spot(200.0, 3.0, args.batch_histories, pos=(0, 0, 0.5))], dtype=rb.SPOT_DTYPE)
        idd = []
        for b in range(args.batches):
            t0 = time.time()
            arrs, n = rb.oracle_run_spots(g, sp, mask=rb.SC_DOSE, kernel=2, prec=64, rng=1, math=1, debug=1,
                                          master_seed=1000 + b, threads=args.threads)
            idd.append(arrs[0].reshape(350, -1).sum(1) / n)
            print("batch %d: %d histories in %.0f s" % (b, n, time.time() - t0), flush=True)
        out["idd64_batches"] = np.array(idd)
        out["batch_histories"] = np.array(args.batch_histories)
        out["batch_seeds"] = 1000 + np.arange(args.batches)
        np.savez_compressed(path, **out)
        return
    sp = np.array([rb.make_spot(200.0, 3.0, args.histories, pos=(0, 0, 0.5))], dtype=rb.SPOT_DTYPE)
    out = {"histories": np.array(args.histories), "seed": np.array(args.seed)}
    depth = -0.5 * (g.ze[:-1] + g.ze[1:]).astype(np.float64)[::-1]
    for prec in (32, 64):
        t0 = time.time()
        arrs, n = rb.oracle_run_spots(g, sp, mask=rb.SC_DOSE, kernel=2, prec=prec, rng=1, math=1, debug=1,
                                      master_seed=args.seed, threads=args.threads)
        idd = arrs[0].reshape(350, -1).sum(1) / n
        out["idd%d" % prec] = idd
        out["r80_%d" % prec] = np.array(dosecmp.r80(depth, idd[::-1]))
        out["total%d" % prec] = np.array(arrs[0].sum() / n)
        print("prec %d: %d histories in %.0f s, R80 %.4f mm" % (prec, n, time.time() - t0, out["r80_%d" % prec]),
              flush=True)
    d = np.abs(out["idd32"] - out["idd64"]) / out["idd64"].max()
    print("FP32 vs FP64: max |dIDD| %.2e of max, dR80 %.4f mm, integral ratio %.6f" % (
        d.max(), out["r80_32"] - out["r80_64"], out["total32"] / out["total64"]))
    np.savez_compressed(path, **out)


if __name__ == "__main__":
    main()
