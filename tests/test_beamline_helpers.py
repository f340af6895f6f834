"""Beamline helpers (moquimc_amd/host/mqi_hip_beamline.hpp) against the
geometry they restate (tps_env::create_rangeshifter and
create_voxelized_aperture, moqui/base/environments/mqi_tps_env.hpp:1606-1772):
the range shifter is the one-voxel box pos +- volume / 2 with density * 1e-3;
the aperture has ceil(volume) 1 mm voxels and a column is open when its
centre is inside the opening by the crossing-number rule, the last opening
deciding when there are several (as there)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
f32 = np.float32


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bl") / "beamline_driver")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "moquimc_amd", "host"), os.path.join(ROOT, "tests", "beamline_driver.cpp"),
                    "-o", exe], check=True)
    return exe


def test_range_shifter(driver):
    out = subprocess.run([driver, "rs"], capture_output=True, text=True, check=True).stdout.split()
    v = [float.fromhex(x) for x in out]
    for (lo, hi), (pos, vol) in zip([(v[0], v[1]), (v[2], v[3]), (v[4], v[5])],
                                    [(0.5, 60.0), (-1.25, 50.0), (30.0, 20.3)]):
        assert lo == f32(f32(pos) - f32(vol) / f32(2))
        assert abs(hi - (pos + vol / 2)) <= 4e-6 * max(1.0, abs(hi))
    # float density times 1e-3 in double, rounded once
    assert v[6] == f32(np.float64(f32(1.19)) * 1e-3)


def _pnpoly(x, y, poly):
    c = False
    n = len(poly)
    j = n - 1
    for i in range(n):
        (x0, y0), (x1, y1) = poly[i], poly[j]
        if ((y0 <= y < y1) or (y1 <= y < y0)) and x < (x1 - x0) * (y - y0) / (y1 - y0) + x0:
            c = not c
        j = i
    return c


def test_aperture_openings(driver, tmp_path):
    square = [(-5, -5), (5, -5), (5, 5), (-5, 5)]
    tri = [(8, 0), (14, 0), (8, 9)]
    for openings, last in (([square], square), ([square, tri], tri)):
        f = tmp_path / "op.txt"
        f.write_text("\n".join(" ".join("%g %g" % p for p in o) for o in openings) + "\n")
        out = subprocess.run([driver, "ap", str(f)], capture_output=True, text=True, check=True).stdout.split("\n")
        nx, ny, nz = (int(v) for v in out[0].split())
        assert (nx, ny, nz) == (41, 30, 6)      # ceil(40.5), 30, 6
        grid = np.array([[int(ch) for ch in row] for row in out[1:1 + ny]])
        pos = (2.0, -3.0)
        xe0, ye0 = 2.0 - 40.5 / 2, -3.0 - 30.0 / 2
        want = np.ones((ny, nx), int)
        for j in range(ny):
            for i in range(nx):
                x, y = xe0 + i + 0.5, ye0 + j + 0.5
                poly = [(px + pos[0], py + pos[1]) for px, py in last]
                if _pnpoly(x, y, poly):
                    want[j, i] = 0
        assert np.array_equal(grid, want)
        # the square opening is 10 x 10 voxels when it is the deciding one
        if last is square:
            assert (grid == 0).sum() == 100
