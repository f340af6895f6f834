"""Pins of what the reference cannot pin bit for bit: the counter-based RNG and
the portable math library that the verification kernels share with the
oracle (moquimc_amd/csrc/mqi_philox.hpp, mqi_pmath.hpp).

  Philox4x32-10   the published Random123 known-answer vectors.
  portable math   against the host libm over wide ranges (ulps).
  distributions   spot sampling with the Philox stream against the sampling
                  with the reference's own RNG (std::default_random_engine):
                  the same moments within their standard errors.
"""
import ctypes

import numpy as np
import pytest

import refbind as rb


def _philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, np.uint32)
    key = np.ascontiguousarray(key, np.uint32)
    out = np.zeros_like(ctr)
    rb.oracle().oracle_philox(ctr.ctypes.data_as(ctypes.c_void_p), key.ctypes.data_as(ctypes.c_void_p),
                              ctypes.c_int64(len(ctr)), out.ctypes.data_as(ctypes.c_void_p))
    return out


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    ctr = [[0, 0, 0, 0], [0xFFFFFFFF] * 4, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]]
    key = [[0, 0], [0xFFFFFFFF] * 2, [0xA4093822, 0x299F31D0]]
    want = [[0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8],
            [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD],
            [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]]
    assert np.array_equal(_philox(ctr, key), np.array(want, np.uint32))


def test_philox_bits_are_uniform():
    n = 1 << 16
    ctr = np.zeros((n, 4), np.uint32)
    ctr[:, 0] = np.arange(n)
    key = np.tile(np.array([[7, 11]], np.uint32), (n, 1))
    out = _philox(ctr, key).reshape(-1)
    bits = np.unpackbits(out.view(np.uint8)).reshape(-1, 32).mean(0)
    # each bit is set with probability 1/2: 4 n draws, sd 1 / (2 sqrt(4n))
    assert np.all(np.abs(bits - 0.5) < 5 * 0.5 / np.sqrt(4 * n))


def _probe(fn, x, y=None):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(x if y is None else y, np.float64)
    out = np.zeros_like(x)
    rb.oracle().oracle_math_probe(fn, x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.c_int64(x.size), out.ctypes.data_as(ctypes.c_void_p))
    return out


def _ulps(a, b):
    return np.abs(a - b) / np.spacing(np.abs(b))


@pytest.mark.parametrize("fn,gen,ref,tol", [
    (0, lambda r: np.exp(r.uniform(-700, 700, 100000)), np.log, 2),
    (1, lambda r: r.uniform(-700, 700, 100000), np.exp, 2),
    (3, lambda r: r.uniform(-10, 10, 100000), np.sin, 3),
    (4, lambda r: r.uniform(-10, 10, 100000), np.cos, 3),
    (5, lambda r: r.uniform(-1, 1, 100000), np.arccos, 2),
    (6, lambda r: r.uniform(0, 1e6, 100000), np.sqrt, 0),
])
def test_portable_math_against_libm(fn, gen, ref, tol):
    x = gen(np.random.default_rng(fn))
    assert _ulps(_probe(fn, x), ref(x)).max() <= tol


def test_portable_pow_and_special_cases():
    r = np.random.default_rng(9)
    x = r.uniform(0.01, 300, 100000)
    y = r.uniform(-3, 3, 100000)
    assert (np.abs(_probe(2, x, y) / np.power(x, y) - 1)).max() < 1e-14
    assert np.array_equal(_probe(0, [1.0]), [0.0])
    assert np.array_equal(_probe(1, [0.0]), [1.0])
    assert np.array_equal(_probe(5, [1.0]), [0.0])
    assert _probe(5, [-1.0])[0] == np.pi
    assert np.array_equal(_probe(2, [2.0, 9.0], [10.0, 0.5]), [1024.0, 3.0])


def test_philox_spot_sampling_matches_reference_distribution():
    """Same spot, reference RNG versus Philox: energy, position and direction
    moments agree within 5 standard errors."""
    n = 200000
    sp = rb.make_spot(150.0, 4.0, n, pos=(2.0, -1.0, 0.5), energy_const=False, sigma_e=1.5,
                      sigma_dir=(0.002, 0.003), corr=(0.3, -0.2))
    a, _ = rb.oracle_sample_spot(sp, 11, 5, n, rng=0)
    b, _ = rb.oracle_sample_spot(sp, 11, 5, n, rng=1)
    assert not np.array_equal(a, b)
    for k in range(7):
        sa, sb = a[:, k].astype(np.float64), b[:, k].astype(np.float64)
        se = np.sqrt(sa.var() / n + sb.var() / n)
        assert abs(sa.mean() - sb.mean()) <= 5 * se + 1e-12, k
        if sa.std() > 0:
            assert abs(sa.std() / sb.std() - 1) < 0.02, k
    # closed forms: energy N(150, 1.5), x N(2, 4), y N(-1, 4)
    for arr in (a, b):
        assert abs(arr[:, 0].mean() - 150.0) < 5 * 1.5 / np.sqrt(n)
        assert abs(arr[:, 0].std() / 1.5 - 1) < 0.01
        assert abs(arr[:, 1].mean() - 2.0) < 5 * 4.0 / np.sqrt(n)
        assert abs(arr[:, 1].std() / 4.0 - 1) < 0.01
        assert abs(arr[:, 2].mean() + 1.0) < 5 * 4.0 / np.sqrt(n)


def test_philox_uniform_phase_space_matches_reference_distribution():
    """phsp_6d_uniform: x = mean + sigma U(-1, 1): the Philox stream and the
    reference RNG give the same moments; the support is [mean - sigma, mean + sigma]."""
    n = 200000
    sp = rb.make_spot(100.0, 3.0, n, pos=(1.0, -2.0, 0.5), sigma_dir=(0.001, 0.002), corr=(0.4, 0.1), uniform=True)
    a, _ = rb.oracle_sample_spot(sp, 5, 1, n, rng=0)
    b, _ = rb.oracle_sample_spot(sp, 5, 1, n, rng=1)
    for arr in (a, b):
        x, y = arr[:, 1].astype(np.float64), arr[:, 2].astype(np.float64)
        assert x.min() >= 1.0 - 3.0 - 1e-5 and x.max() <= 1.0 + 3.0 + 1e-5
        assert abs(x.mean() - 1.0) < 5 * 3.0 / np.sqrt(3 * n)
        assert abs(x.std() / (3.0 / np.sqrt(3.0)) - 1) < 0.01
        assert abs(y.mean() + 2.0) < 5 * 3.0 / np.sqrt(3 * n)
        assert np.all(arr[:, 0] == np.float32(100.0))
    for k in range(1, 7):
        sa, sb = a[:, k].astype(np.float64), b[:, k].astype(np.float64)
        se = np.sqrt(sa.var() / n + sb.var() / n)
        assert abs(sa.mean() - sb.mean()) <= 5 * se + 1e-12, k
