"""Loading a MetaImage CT through the C ABI (mqi_set_grid_ct_mha, after
tps_env::read_ct_image, moqui/base/environments/mqi_tps_env.hpp:615-701):
the device density equals mqi_set_grid_hu of the same HU and edges."""
import numpy as np
import pytest

from test_io_formats import _edges, _write_ct

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
