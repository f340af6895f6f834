"""Loading a MetaImage CT through the C ABI (mqi_set_grid_ct_mha, after
tps_env::read_ct_image, moqui/base/environments/mqi_tps_env.hpp:615-701):
the device density equals mqi_set_grid_hu of the same HU and edges.
Structure mask files (mqi_set_roi_mha, after mask_reader::read_mask_files,
moqui/base/mqi_file_handler.hpp:92-113) score as mqi_set_roi of their sum."""
import numpy as np
import pytest

import phantoms
import refbind as rb
from test_io_formats import _edges, _write_ct, write_mask

pytestmark = pytest.mark.gpu


def test_ct_mha_loads_like_set_grid_hu(ctx, tmp_path):
    rng = np.random.default_rng(7)
    hu = rng.uniform(-1000, 2500, (6, 7, 8)).astype(np.float32)
    origin, spacing = (-3.92, -4.5, -6.0), (0.98, 1.25, 2.0)
    p = tmp_path / "ct.mha"
    _write_ct(p, hu, origin, spacing)
    ctx.set_grid_ct_mha(str(p), density_scale=1.035)
    got = ctx.density()
    xe = _edges(origin[0], spacing[0], 8)
    ye = _edges(origin[1], spacing[1], 7)
    z = [np.float32(np.float32(origin[2]) - np.float32(spacing[2]) / 2.0)]
    for _ in range(6):
        z.append(np.float32(z[-1] + np.float32(spacing[2])))
    ctx.set_grid_hu(xe, ye, np.float32(z), np.trunc(hu).astype(np.int16), density_scale=1.035)
    want = ctx.density()
    assert np.array_equal(got, want)
    with pytest.raises(Exception):
        ctx.set_grid_ct_mha(str(tmp_path / "missing.mha"))


def test_roi_mask_files_score_like_their_sum(ctx, tmp_path):
    g = phantoms.slab_box(24, 24, 120)
    k, j, i = np.meshgrid(np.arange(120), np.arange(24), np.arange(24), indexing="ij")
    r2 = (i - 12) ** 2 + (j - 12) ** 2
    outer = ((r2 < 25) & (k >= 20) & (k < 100)).astype(np.uint8)
    inner = ((r2 < 4) & (k >= 40) & (k < 60)).astype(np.uint8)   # inside outer: sums to 2
    pa, pb = tmp_path / "outer.mha", tmp_path / "inner.mha"
    write_mask(pa, outer)
    write_mask(pb, inner)
    sp = phantoms.spots([110.0], 400, sigma=3.0, positions=[(0, 0, 0.5)])
    ctx.set_grid(g.xe, g.ye, g.ze, g.density)
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE | rb.SC_STAT)
    got = []
    for use_files in (True, False):
        if use_files:
            ctx.set_roi_mha(0, [str(pa), str(pb)])
            ctx.set_roi_mha(1, [str(pb)])
        else:
            ctx.set_roi(0, (outer.astype(int) + inner).ravel())
            ctx.set_roi(1, inner.ravel())
        ctx.zero_scorers()
        ctx.run(ctx.config(seed=17, exact=True, batched=True))
        got.append([ctx.scorer(s) for s in range(3)])
    for a, b in zip(*got):
        assert np.array_equal(a == 0, b == 0)
        assert np.allclose(a, b, rtol=1e-12, atol=0)
    assert not np.any(got[0][0][outer.ravel() == 0]) and not np.any(got[0][1][inner.ravel() == 0])
    assert np.any(got[0][1] != 0)
    # a mask of another size is refused; n = 0 restores DIRECT
    write_mask(tmp_path / "small.mha", outer[:-1])
    with pytest.raises(Exception):
        ctx.set_roi_mha(0, [str(tmp_path / "small.mha")])
    ctx.set_roi_mha(0)
    ctx.set_roi_mha(1)


def test_ct_env_writes_what_the_abi_scores(ctx, tmp_path):
    """mqi_hip::ct_env (tps_env without DICOM, moquimc_amd/host/mqi_hip_ct_env.hpp)
    through a small driver: MetaImage CT, a scoring mask file, Dose + LETd
    with the statistics pair (stopping rounds, saved), mhd output named
    <beam>_<child>_<scorer> and scaled by particles per history x RBE x
    fractions (1 x 1.1f x 1), equal to the ctypes path with the same seed;
    a Dij run writes <beam>_0_Dij.npz equal to the ctypes CSR x 1.1f."""
    import os
    import subprocess
    import scipy.sparse as sps
    from moquimc_amd import _abi
    root = rb.ROOT
    exe = str(tmp_path / "ct_env_driver")
    lib = os.path.join(root, "moquimc_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(root, "include"), "-I", os.path.join(lib, "host"),
                    os.path.join(root, "tests", "ct_env_driver.cpp"), "-L", lib, "-lmqi_hip", "-Wl,-rpath," + lib,
                    "-o", exe], check=True)
    rng = np.random.default_rng(9)
    nz, ny, nx = 100, 24, 24
    hu = rng.uniform(-150, 150, (nz, ny, nx)).astype(np.float32)
    ct = tmp_path / "ct.mha"
    _write_ct(ct, hu, (-23.0, -23.0, -199.0), (2.0, 2.0, 2.0))
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    mask = (((i - 11.5) ** 2 + (j - 11.5) ** 2 < 9.0) & (k >= 1)).astype(np.uint8)
    mp = tmp_path / "mask.mha"
    write_mask(mp, mask)
    sp = np.array([rb.make_spot(110.0, 3.0, 400, pos=(-4.0, 0.0, 0.5)),
                   rb.make_spot(130.0, 3.0, 300, pos=(4.0, 0.0, 0.5))], dtype=rb.SPOT_DTYPE)
    spf = tmp_path / "spots.bin"
    spf.write_bytes(sp.tobytes())
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([exe, str(ct), str(spf), str(out), "mhd", "7", "Dose,LETd", "1", str(mp)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert int(r.stdout.split()[-1]) >= 700
    ctx.set_grid_ct_mha(str(ct))
    ctx.set_spots(sp)
    ctx.set_scorers(rb.SC_DOSE | rb.SC_LETD | rb.SC_STAT)
    ctx.set_roi_mha(0, [str(mp)])
    # tps_env's kernel with the statistics pair: transport_particles_patient_stat
    ctx.run(ctx.config(seed=7, exact=True, batched=True, stat_criterion=0.5, stat_threshold=0.5,
                       loops=_abi.LOOPS_REFERENCE_STAT))
    names = ["Dose", "LETd_numer", "LETd_denom", "Dose_stat", "DoseSquare_stat"]
    scale = float(np.float32(1.1))   # RBE is a float parameter, as the reference's
    for kk, n in enumerate(names):
        got = np.fromfile(out / ("beam_0_%s.raw" % n), np.float64)
        want = ctx.scorer(kk) * scale
        assert (out / ("beam_0_%s.mhd" % n)).exists()
        assert np.allclose(got, want, rtol=1e-9, atol=0), n
        assert np.any(got != 0), n
    assert not np.any(np.fromfile(out / "beam_0_Dose.raw", np.float64)[mask.ravel() == 0])
    ctx.set_roi_mha(0)
    # Dij alone, npz
    dout = tmp_path / "dij"
    dout.mkdir()
    r = subprocess.run([exe, str(ct), str(spf), str(dout), "raw", "7", "Dij", "0", "-"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ctx.set_scorers(rb.SC_DOSE)
    ctx.run(ctx.config(seed=7, exact=True, batched=True, dij=True, dij_slots_sqrt=4096,
                       loops=_abi.LOOPS_REFERENCE_PLAIN))
    ctx.dij_build(0.0)
    indptr, idx, val, shape = ctx.dij_csr()
    D = sps.load_npz(dout / "beam_0_Dij.npz")
    assert D.shape == shape
    want = sps.csr_matrix((val.astype(np.float64) * scale, idx, indptr), shape=shape)
    assert np.array_equal(D.indptr, want.indptr) and np.array_equal(D.indices, want.indices)
    assert np.allclose(D.data, want.data, rtol=1e-6, atol=0)
