"""Dose comparison used by the statistical tests: global gamma index, R80,
integral dose and the per-voxel 2-sigma test.

gamma (Low, Harms, Mutic and Purdy, Med. Phys. 25 (1998) 656): for every
reference voxel above `threshold` x max(ref),
    G = min over evaluated voxels e of sqrt(|x_e - x_r|^2 / dta^2 + (D_e - D_r)^2 / (dd * max(ref))^2),
searched over the evaluated grid, optionally refined by linear interpolation
(`refine` sub-steps per voxel along each axis); the voxel passes when G <= 1.
"""
import itertools

import numpy as np


def _upsample(a, f, axes):
    """Linear interpolation by an integer factor f along the given axes (the
    original samples are kept; f - 1 points are inserted between neighbours)."""
    for ax in axes:
        n = a.shape[ax]
        if n < 2 or f == 1:
            continue
        t = np.arange((n - 1) * f + 1) / f
        i0 = np.minimum(np.floor(t).astype(int), n - 2)
        w = (t - i0).reshape([-1 if k == ax else 1 for k in range(a.ndim)])
        a = np.take(a, i0, axis=ax) * (1 - w) + np.take(a, i0 + 1, axis=ax) * w
    return a


def gamma_pass_rate(ref, ev, spacing, dd=0.01, dta=1.0, threshold=0.1, refine=1):
    """Fraction of reference voxels above threshold * max(ref) with G <= 1.
    ref, ev: arrays of the same shape on the same grid; spacing: voxel size
    per axis (mm)."""
    ref = np.asarray(ref, np.float64)
    ev = np.asarray(ev, np.float64)
    assert ref.shape == ev.shape
    dmax = ref.max()
    mask = ref >= threshold * dmax
    f = max(1, int(refine))
    evf = _upsample(ev, f, range(ev.ndim))
    step = np.asarray(spacing, np.float64) / f
    reach = [int(np.floor(dta / s)) for s in step]
    best = np.full(ref.shape, np.inf)
    sl_ref = tuple(slice(0, n) for n in ref.shape)
    for off in itertools.product(*[range(-r, r + 1) for r in reach]):
        dist2 = float(np.sum((np.asarray(off) * step) ** 2)) / (dta * dta)
        if dist2 > 1.0:
            continue
        # evaluated value at reference voxel index i plus the fine offset
        idx = []
        valid = np.ones(ref.shape, bool)
        for ax, o in enumerate(off):
            pos = np.arange(ref.shape[ax]) * f + o
            ok = (pos >= 0) & (pos < evf.shape[ax])
            shape = [1] * ref.ndim
            shape[ax] = -1
            valid &= ok.reshape(shape)
            idx.append(np.clip(pos, 0, evf.shape[ax] - 1))
        e = evf[np.ix_(*idx)]
        g2 = dist2 + ((e - ref) / (dd * dmax)) ** 2
        g2 = np.where(valid, g2, np.inf)
        np.minimum(best, g2, out=best)
    del sl_ref
    return float(np.mean(best[mask] <= 1.0)), int(mask.sum())


def rebin_xy(a3, f):
    """Sum of f x f lateral (y, x) voxel blocks of a (z, y, x) array divided
    by f * f: the dose of the coarser voxel."""
    nz, ny, nx = a3.shape
    return a3.reshape(nz, ny // f, f, nx // f, f).sum((2, 4)) / (f * f)


def r80(depth_centres, idd):
    """Distal depth where the integral depth dose falls to 80 % of its
    maximum for the last time (linear interpolation between voxel centres).
    For a single Bragg peak this is the one crossing beyond the maximum; a
    field of several energies through heterogeneous slabs can touch the 80 %
    level between its peaks, where the first crossing would flip on noise,
    while the distal falloff is steep and well defined."""
    idd = np.asarray(idd, np.float64)
    k = int(np.argmax(idd))
    level = 0.8 * idd[k]
    for i in range(len(idd) - 2, k - 1, -1):
        if idd[i] >= level > idd[i + 1]:
            t = (idd[i] - level) / (idd[i] - idd[i + 1])
            return float(depth_centres[i] + t * (depth_centres[i + 1] - depth_centres[i]))
    raise ValueError("no distal 80% crossing")


def within_2sigma(a, sa, b, sb, mask):
    """Fraction of masked voxels where |a - b| <= 2 sqrt(sa^2 + sb^2)."""
    s = np.sqrt(np.asarray(sa, np.float64) ** 2 + np.asarray(sb, np.float64) ** 2)
    d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))
    return float(np.mean(d[mask] <= 2.0 * s[mask]))


def two_sigma_expectation(sa, sb, dof_a, dof_b, mask):
    """Expected fraction of |a - b| <= 2 sqrt(sa^2 + sb^2) when a and b have
    the same mean and sa, sb are batch standard errors with dof_a, dof_b
    degrees of freedom: the standardised difference then follows Student's
    t with the Welch-Satterthwaite degrees of freedom (for 8 batches a side
    about 8-14, i.e. 0.92-0.94 rather than the normal 0.954)."""
    from scipy import stats
    va = np.asarray(sa, np.float64)[mask] ** 2
    vb = np.asarray(sb, np.float64)[mask] ** 2
    den = va * va / dof_a + vb * vb / dof_b
    nu = np.where(den > 0, (va + vb) ** 2 / np.where(den > 0, den, 1.0), 1e6)
    return float(np.mean(2.0 * stats.t.cdf(2.0, nu) - 1.0))

