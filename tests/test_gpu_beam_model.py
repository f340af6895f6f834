"""The PBS beam model on the device path: a machine file and a spot map go
through the host pbs_model (moquimc_amd/host/mqi_hip_pbs.hpp, after
treatment_machine_pbs::characterize_beamlet / characterize_history and
treatment_machine_ion::create_beamsource, mqi_treatment_machine_pbs.hpp:
135-143,199-235, mqi_treatment_machine_ion.hpp:83-134,324-333) into
beamlets; the beamlets equal the numpy restatement (oracle/beam_model.py)
bit for bit, and the verification kernels transporting them give the
oracle's dose and LETd for the same beamlets (FP64 summation order)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import phantoms
import refbind as rb
from test_beam_model import _oracle, _same, _table, _write_machine

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import beam_model as bm  # noqa: E402

f32 = np.float32


def test_pbs_beamlets_transport_like_the_oracle(ctx, tmp_path):
    exe = str(tmp_path / "pbs_driver")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "moquimc_amd", "host"), os.path.join(ROOT, "tests", "pbs_driver.cpp"),
                    "-o", exe], check=True)
    table = _table(7)
    sad, z, pph = (2000.0, 1800.0), 30.0, 2.0e5
    mf = str(tmp_path / "machine.txt")
    _write_machine(mf, table, sad)
    rng = np.random.default_rng(8)
    spots = [(float(f32(e)), float(f32(x)), float(f32(y)), float(f32(mu)))
             for e, x, y, mu in zip(rng.uniform(95.0, 150.0, 12), rng.uniform(-5, 5, 12), rng.uniform(-5, 5, 12),
                                    rng.uniform(0.1, 0.4, 12))]
    sf = tmp_path / "spots.txt"
    sf.write_text("".join(" ".join(float(f32(v)).hex() for v in s) + "\n" for s in spots))
    out = tmp_path / "beamsource.bin"
    r = subprocess.run([exe, mf, float(f32(z)).hex(), float(f32(pph)).hex(), str(sf), str(out)], capture_output=True,
                       text=True, check=True)
    bs = np.frombuffer(out.read_bytes(), dtype=rb.SPOT_DTYPE).copy()
    assert len(bs) == len(spots) and r.stdout.splitlines()[-1] == "beamsource %d %d" % (len(bs),
                                                                                        int(bs["histories"].sum()))
    model = bm.BeamModel.from_file(mf, z)
    for s, b in zip(spots, bs):
        want, h = _oracle(model, s, pph)
        got = [b["energy"], b["sigma_energy"]] + list(b["mean"]) + list(b["sigma"])
        assert _same(got, want) and int(b["histories"]) == h
    assert 50 <= bs["histories"].min() and bs["histories"].sum() <= 20000
    g = phantoms.slab_box(24, 24, 160)
    mask = rb.SC_DOSE | rb.SC_LETD
    ref, n_ref = rb.oracle_run_spots(g, bs, mask=mask, kernel=2, prec=32, rng=1, math=1, debug=0, master_seed=9,
                                     threads=4)
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(bs)
    ctx.set_scorers(mask)
    res = ctx.run(ctx.config(seed=9, exact=True, batched=True))
    assert res.histories == n_ref == int(bs["histories"].sum())
    for k in range(3):
        got, want = ctx.scorer(k), ref[k]
        fin = np.isfinite(want)
        assert np.array_equal(np.isfinite(got), fin)
        assert np.allclose(got[fin], want[fin], rtol=1e-9, atol=0), k
    assert ctx.scorer(0).sum() > 0
