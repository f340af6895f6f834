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
