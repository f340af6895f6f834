"""The host beam model (moquimc_amd/host/mqi_hip_pbs.hpp, after
moqui/base/mqi_treatment_machine_pbs.hpp:95-235,401-516 and
mqi_treatment_machine_ion.hpp:83-134,324-333) against the numpy restatement
in oracle/beam_model.py, bit for bit, and both against closed forms: table
values at the nodes, linear interpolation (energy and ratio over the nominal
energies, spreads over the E column), the start point on the ray from the
virtual source, unit directions, the truncated history count.

The reference's beam model cannot be compiled here (it includes GDCM
through mqi_dataset.hpp), so these closed forms are its pin.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import beam_model as bm  # noqa: E402

f32 = np.float32


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("pbs") / "pbs_driver")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "moquimc_amd", "host"), os.path.join(ROOT, "tests", "pbs_driver.cpp"),
                    "-o", exe], check=True)
    return exe


def _table(seed=0, offset=-0.3):
    """Nominal energies 70..230 MeV every 10 MeV; the E column sits `offset`
    MeV from the key; smooth spreads and ratios."""
    rng = np.random.default_rng(seed)
    rows = []
    for e in np.arange(70.0, 231.0, 10.0):
        rows.append((float(e), [e + offset, 0.4 + 0.004 * e + 0.05 * rng.random(), 3.0 + 0.02 * e,
                                3.2 + 0.018 * e, 0.002 + 1e-5 * e, 0.0021 + 1.1e-5 * e,
                                1.0e8 * (1.0 + 0.003 * e + 0.1 * rng.random())]))
    # exactly what the file will hold: 9 significant digits round-trip float32
    return [(float(f32(e)), [float(f32(v)) for v in r]) for e, r in rows]


def _write_machine(path, table, sad):
    with open(path, "w") as f:
        f.write("# generic PBS machine\n[geometry]\nSAD(mm) %r %r\n" % sad)
        f.write("rangeshifter(mm) 300 300 # ignored\n[rangeshifter_thickness]\n\"RS1\" 40\n[spot]\n")
        for e, r in table:
            f.write(" ".join("%.9g" % v for v in [e] + r) + "\n")
        f.write("[time]\nI(nA) 1.0\n")


def _run(driver, tmp_path, table, sad, z, pph, spots):
    mf = str(tmp_path / "machine.txt")
    _write_machine(mf, table, sad)
    sf = str(tmp_path / "spots.txt")
    with open(sf, "w") as f:
        for s in spots:
            f.write(" ".join(float(f32(v)).hex() for v in s) + "\n")
    out = subprocess.run([driver, mf, float(f32(z)).hex(), float(f32(pph)).hex(), sf], capture_output=True,
                         text=True, check=True).stdout.splitlines()
    rows = []
    for ln in out[:-1]:
        if ln == "error":
            rows.append(None)
        else:
            t = ln.split()
            rows.append(([float.fromhex(v) for v in t[:14]], int(t[14])))
    n, total = out[-1].split()[1:]
    return rows, int(n), int(total), bm.BeamModel.from_file(mf, z)


def _oracle(model, s, pph):
    e, x, y, mu = s
    me, de, mean, sigma = model.characterize_beamlet(e, x, y)
    return [float(me), float(de)] + [float(v) for v in mean] + [float(v) for v in sigma], \
        model.characterize_history(e, mu, pph)


def _same(a, b):
    a, b = np.array(a, np.float32), np.array(b, np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("sad,z", [((2000.0, 1800.0), 390.0), ((2300.5, 2300.5), 420.25)])
def test_beamlets_match_oracle_bitwise(driver, tmp_path, sad, z):
    table = _table(1)
    rng = np.random.default_rng(2)
    spots = [(rng.uniform(70.001, 230.0), rng.uniform(-100, 100), rng.uniform(-100, 100), rng.uniform(0.01, 5.0))
             for _ in range(300)]
    rows, n, total, model = _run(driver, tmp_path, table, sad, z, 1000.0, spots)
    assert n == len(spots)
    for s, r in zip(spots, rows):
        want, h = _oracle(model, s, 1000.0)
        assert r is not None
        assert _same(r[0], want), (s, r[0], want)
        assert r[1] == h
    assert total == sum(r[1] for r in rows)


def test_beam_model_closed_forms(driver, tmp_path):
    table = _table(3, offset=-0.3)
    keys = np.array([e for e, _ in table])
    T = np.array([r for _, r in table])
    S = (2000.0, 1800.0)
    z = 390.0
    spots = []
    for k in range(1, len(keys)):
        spots.append((keys[k], 10.0 * k, -7.0 * k, 1.0))                        # at a node
        spots.append((0.5 * (keys[k - 1] + keys[k]), -5.0, 12.0, 2.0))          # midway between nodes
        spots.append((T[k, 0], 0.0, 0.0, 1.0))                                  # at row k's E column value
    rows, n, _, model = _run(driver, tmp_path, table, S, z, 1.0, spots)
    for i, k in enumerate(range(1, len(keys))):
        at_node, mid, at_E = rows[3 * i], rows[3 * i + 1], rows[3 * i + 2]
        # energy at a node is the node's E (the interpolation lands on y1)
        assert abs(at_node[0][0] - T[k, 0]) <= 2 * np.spacing(f32(T[k, 0]))
        # energy midway is the mean of the two rows' E; ratio likewise
        assert abs(mid[0][0] - 0.5 * (T[k - 1, 0] + T[k, 0])) <= 4 * np.spacing(f32(T[k, 0]))
        r_mid = 0.5 * (T[k - 1, 6] + T[k, 6])
        assert abs(mid[1] - 2.0 * r_mid) <= 4e-7 * 2.0 * r_mid + 1
        # spreads interpolate over the E column: at E_k they are row k's
        # (interpolating over the nominal energies instead would miss by
        # 0.3 MeV / 10 MeV of the step)
        for j, col in ((1, 1), (8, 2), (9, 3), (11, 4), (12, 5)):
            assert abs(at_E[0][j] - T[k, col]) <= 4 * np.spacing(f32(T[k, col])), (k, j)
        # history count: meterset * ratio / particles_per_history, truncated
        assert abs(at_node[1] - T[k, 6]) <= 2e-7 * T[k, 6] + 1
    # geometry of every beamlet: the start lies on the ray from the virtual
    # source (0, 0, SAD) to the isocenter-plane point, at z; the direction is
    # the unit vector along (x / SAD_x, y / SAD_y, -1); zero spreads in z
    for s, r in zip(spots, rows):
        v = np.array(r[0])
        x, y = float(f32(s[1])), float(f32(s[2]))
        assert np.allclose(v[2:5], [x * (S[0] - z) / S[0], y * (S[1] - z) / S[1], z], rtol=2e-7, atol=1e-5)
        d = np.array([x / S[0], y / S[1], -1.0])
        d /= np.linalg.norm(d)
        assert np.allclose(v[5:8], d, rtol=0, atol=3e-7)
        assert abs(np.linalg.norm(v[5:8]) - 1.0) <= 3e-7
        assert v[10] == 0.0 and v[13] == 0.0


def test_out_of_range_energies_and_infinite_sad(driver, tmp_path):
    table = _table(4)
    keys = [e for e, _ in table]
    spots = [(keys[0], 0, 0, 1), (keys[0] - 5, 0, 0, 1), (keys[-1] + 0.5, 0, 0, 1), (keys[-1], 1, 2, 1),
             (keys[0] + 1e-3, 3, 4, 1)]
    rows, n, _, model = _run(driver, tmp_path, table, (2000.0, 2000.0), 390.0, 1.0, spots)
    # lower_bound semantics: the lowest nominal energy and anything outside
    # the table have no bracketing rows
    assert rows[0] is None and rows[1] is None and rows[2] is None
    assert rows[3] is not None and rows[4] is not None and n == 2
    for s in spots[:3]:
        with pytest.raises(ValueError):
            model.characterize_beamlet(*s[:3])
    # SAD 0 in the file means infinity; the start position is then
    # x * (inf - z) / inf = NaN, as in the reference
    rows, _, _, model = _run(driver, tmp_path, table, (0.0, 0.0), 390.0, 1.0, [(150.0, 10.0, 0.0, 1.0)])
    v = rows[0][0]
    assert np.isnan(v[2]) and np.isnan(v[3]) and v[4] == 390.0
    assert _same(v[:2], _oracle(model, (150.0, 10.0, 0.0, 1.0), 1.0)[0][:2])


def test_unit_histories_without_particles_per_history(driver, tmp_path):
    table = _table(5)
    spots = [(100.0 + k, 0, 0, 3.0) for k in range(7)]
    rows, n, total, _ = _run(driver, tmp_path, table, (2000.0, 2000.0), 390.0, -1.0, spots)
    # create_beamsource: particles_per_history == -1 gives one history per beamlet
    assert n == 7 and total == 7
