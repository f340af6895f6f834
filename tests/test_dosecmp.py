"""The dose-comparison helpers against cases with known answers."""
import numpy as np

import dosecmp


def _gauss_field(shape, spacing, centre, sigma):
    grids = np.meshgrid(*[(np.arange(n) + 0.5) * s for n, s in zip(shape, spacing)], indexing="ij")
    r2 = sum((g - c) ** 2 for g, c in zip(grids, centre))
    return np.exp(-0.5 * r2 / sigma ** 2)


def test_gamma_identical_and_scaled():
    a = _gauss_field((40, 40, 40), (0.5, 0.5, 1.0), (10, 10, 20), 3.0)
    assert dosecmp.gamma_pass_rate(a, a, (0.5, 0.5, 1.0))[0] == 1.0
    # 0.5 % scaling passes everywhere, 3 % fails near the maximum (flat top:
    # no neighbour within 1 mm is 2 % higher)
    assert dosecmp.gamma_pass_rate(a, 1.005 * a, (0.5, 0.5, 1.0))[0] == 1.0
    rate, n = dosecmp.gamma_pass_rate(a, 1.03 * a, (0.5, 0.5, 1.0))
    assert rate < 0.9 and n > 100


def test_gamma_shift_within_and_beyond_dta():
    sp = (0.5, 0.5, 0.5)
    a = _gauss_field((60, 60, 60), sp, (15, 15, 15), 2.0)
    b = _gauss_field((60, 60, 60), sp, (15.5, 15, 15), 2.0)   # 0.5 mm: passes
    c = _gauss_field((60, 60, 60), sp, (17.0, 15, 15), 2.0)   # 2 mm on a steep field: fails
    assert dosecmp.gamma_pass_rate(a, b, sp)[0] == 1.0
    assert dosecmp.gamma_pass_rate(a, c, sp)[0] < 0.8


def test_gamma_refinement_recovers_sub_voxel_shift():
    # a 0.5 mm shift on a 1 mm grid: without interpolation the nearest
    # evaluated points are 0.5 mm off the matching position (dose differs by
    # up to 7 % of max on this gradient) or 1 mm away; refined, they pass
    sp = (1.0, 1.0)
    a = _gauss_field((80, 80), sp, (40, 40), 4.0)
    b = _gauss_field((80, 80), sp, (40.5, 40), 4.0)
    coarse = dosecmp.gamma_pass_rate(a, b, sp, refine=1)[0]
    fine = dosecmp.gamma_pass_rate(a, b, sp, refine=4)[0]
    assert fine == 1.0 and coarse < 0.9


def test_r80_linear_falloff():
    z = np.arange(100) + 0.5
    idd = np.where(z < 60, 1.0, np.clip(1.0 - (z - 60) / 10.0, 0, 1))
    # 80 % of 1.0 at z = 62 exactly on the linear falloff
    assert abs(dosecmp.r80(z, idd) - 62.0) < 1e-12


def test_within_2sigma():
    rng = np.random.default_rng(1)
    a = rng.normal(size=200000)
    b = rng.normal(size=200000)
    frac = dosecmp.within_2sigma(a, np.ones_like(a) / np.sqrt(2) * np.sqrt(2), b, np.ones_like(b), np.ones(a.size, bool))
    # |a - b| ~ N(0, 2): within 2 sqrt(2) with probability 95.45 %
    assert abs(frac - 0.9545) < 0.005


def test_two_sigma_expectation_matches_student_t():
    """Batch means with estimated errors: the fraction within 2 estimated
    sigma follows the t distribution, which the expectation reproduces."""
    rng = np.random.default_rng(3)
    nvox, nb = 20000, 8
    a = rng.normal(size=(nb, nvox))
    b = rng.normal(size=(nb, nvox)) * 0.3
    ma, mb = a.mean(0), b.mean(0)
    sa, sb = a.std(0, ddof=1) / np.sqrt(nb), b.std(0, ddof=1) / np.sqrt(nb)
    mask = np.ones(nvox, bool)
    frac = dosecmp.within_2sigma(ma, sa, mb, sb, mask)
    want = dosecmp.two_sigma_expectation(sa, sb, nb - 1, nb - 1, mask)
    assert 0.90 < want < 0.93
    assert abs(frac - want) < 0.01


def test_r80_takes_the_distal_falloff_of_a_multi_peak_curve():
    # a dip to 75 % of the maximum between two peaks, then a linear distal
    # falloff from 100 % at z = 60 to 0 at z = 70: R80 = 62, not the dip
    z = np.arange(0.0, 80.0, 1.0)
    idd = np.where(z < 20, 1.0, np.where(z < 40, 0.75, np.where(z <= 60, 1.0, np.clip(1.0 - (z - 60) / 10.0, 0, 1))))
    assert abs(dosecmp.r80(z, idd) - 62.0) < 1e-12
