"""Seeded synthetic inputs of the benchmark and parity workloads.

  water   200x200x350 water phantom (0.5 x 0.5 x 1 mm), one 200 MeV spot,
          Gaussian sigma 3 mm, 1e5 histories, dose (debug physics, as the
          reference's phantom build).
  slabs   same grid with water / lung (0.26 g/cm3) / bone (1.85 g/cm3) / water
          slabs along depth, 25 spots on a 5x5 grid at 10 mm, energies
          U(70, 230) MeV with 1 % spread, sigma 4 mm, 1e6 histories per spot,
          dose + LETd.
  ct      512x512x200 int16 CT at 0.98 x 0.98 x 2 mm: an elliptic body
          (180 x 120 mm semi-axes) of Gaussian-smoothed noise (sigma 8 mm)
          mapped to [-900, 1500] HU in air (-1000); 2000 spots in 20 layers at
          100-180 MeV on a 5 mm hexagonal grid, histories log-uniform in
          [1e3, 1e5]; the beam travels +y and enters at y = -130 mm.
  dij     the same CT generator on 256x256x128 at 1.96 x 1.96 x 4 mm (air
          beyond |z| = 200 mm), 5000 spots (25 layers x 200) of 2e4 histories.
  robust  the ct plan under 9 scenarios: nominal, +-3 mm shifts in x, y, z,
          density x (1 +- 0.035); the nominal scenario gets twice the
          histories of the others.

Positions are in mm, densities in g/mm^3 (the transport's unit).
"""
import numpy as np

from . import _abi


def uniform_edges(lo, hi, n):
    """Edges as the reference's grid3d(min, max, n+1) constructor makes them in
    float: dx = (max - min) / n, e_i = min + i * dx (mqi_grid3d.hpp:134-163)."""
    lo = np.float32(lo)
    hi = np.float32(hi)
    dx = np.float32(np.float32(hi - lo) / np.float32(n))
    i = np.arange(n + 1, dtype=np.float32)
    return (lo + i * dx).astype(np.float32)


# beam frame -z -> world +y (a 90 degree rotation about x, the iec2dicom angle
# tps_env sets, mqi_tps_env.hpp:974-976)
ROT_PLUS_Y = np.array([1, 0, 0, 0, 0, -1, 0, 1, 0], np.float32)
IDENTITY = np.eye(3, dtype=np.float32).reshape(-1)


def spot(energy, sigma_xy, histories, pos=(0.0, 0.0), z=0.0, sigma_e=0.0, rotation=IDENTITY,
         translation=(0.0, 0.0, 0.0)):
    s = np.zeros(1, _abi.SPOT_DTYPE)[0]
    s["energy"] = energy
    s["sigma_energy"] = sigma_e
    s["energy_const"] = 1 if sigma_e == 0.0 else 0
    s["mean"] = [pos[0], pos[1], z, 0.0, 0.0, -1.0]
    s["sigma"] = [sigma_xy, sigma_xy, 0.0, 0.0, 0.0, 0.0]
    s["rotation"] = rotation
    s["translation"] = translation
    s["histories"] = int(histories)
    return s


class Workload:
    def __init__(self, name, xe, ye, ze, spots, density=None, hu=None, mask=_abi.SCORE_DOSE, debug=False,
                 scenarios=None):
        self.name = name
        self.xe, self.ye, self.ze = xe, ye, ze
        self.density = density
        self.hu = hu
        self.spots = np.asarray(spots, _abi.SPOT_DTYPE)
        self.mask = mask
        self.debug = debug
        self.scenarios = scenarios   # (shifts[n,3], scales[n], history budgets[n]) or None

    @property
    def shape(self):
        return (len(self.ze) - 1, len(self.ye) - 1, len(self.xe) - 1)

    @property
    def histories(self):
        return int(self.spots["histories"].sum())

    def load(self, ctx, per_scenario=False):
        if self.hu is not None:
            ctx.set_grid_hu(self.xe, self.ye, self.ze, self.hu)
        else:
            ctx.set_grid(self.xe, self.ye, self.ze, self.density)
        if self.scenarios is not None:
            ctx.set_scenarios(*self.scenarios)
        ctx.set_spots(self.spots)
        ctx.set_scorers(self.mask, per_scenario)


def water(histories=100000):
    xe = uniform_edges(-50, 50, 200)
    ye = uniform_edges(-50, 50, 200)
    ze = uniform_edges(-350, 0, 350)
    rho = np.full(200 * 200 * 350, 1.0e-3, np.float32)
    sp = [spot(200.0, 3.0, histories, z=0.5)]
    return Workload("water 200x200x350 @0.5x0.5x1 mm, 1 spot 200 MeV sigma 3 mm, %d histories, dose" % histories,
                    xe, ye, ze, sp, density=rho, mask=_abi.SCORE_DOSE, debug=True)


def slab_density(ze, nxy):
    zc = 0.5 * (ze[:-1] + ze[1:])
    depth = -zc
    rho_z = np.where(depth < 50, 1.0e-3, np.where(depth < 120, 0.26e-3, np.where(depth < 140, 1.85e-3, 1.0e-3)))
    return np.repeat(rho_z.astype(np.float32), nxy)


def slabs(histories_per_spot=1000000, seed=2):
    xe = uniform_edges(-50, 50, 200)
    ye = uniform_edges(-50, 50, 200)
    ze = uniform_edges(-350, 0, 350)
    rho = slab_density(ze, 200 * 200)
    rng = np.random.default_rng(seed)
    energies = rng.uniform(70.0, 230.0, 25)
    sp = []
    k = 0
    for iy in range(5):
        for ix in range(5):
            e = float(np.float32(energies[k]))
            sp.append(spot(e, 4.0, histories_per_spot, pos=(-20.0 + 10 * ix, -20.0 + 10 * iy), z=0.5,
                           sigma_e=0.01 * e))
            k += 1
    return Workload("slabs water/lung/bone/water 200x200x350, 25 spots 70-230 MeV, %d histories/spot, dose+LETd"
                    % histories_per_spot, xe, ye, ze, sp, density=rho,
                    mask=_abi.SCORE_DOSE | _abi.SCORE_LETD, debug=False)


def _smooth_noise(shape, spacing, sigma_mm, rng):
    """White noise smoothed with a separable Gaussian of sigma_mm (FFT per axis)."""
    import os

    import scipy.fft as sfft
    workers = min(16, os.cpu_count() or 1)
    a = rng.standard_normal(shape, dtype=np.float32)
    for axis, (n, d) in enumerate(zip(shape, spacing)):
        k = np.fft.rfftfreq(n, d=d)
        g = np.exp(-2.0 * (np.pi * k * sigma_mm) ** 2).astype(np.float32)
        sh = [1] * len(shape)
        sh[axis] = len(k)
        f = sfft.rfft(a, axis=axis, workers=workers)
        f *= g.reshape(sh)
        a = sfft.irfft(f, n=n, axis=axis, workers=workers).astype(np.float32)
    return a


def synthetic_ct(nx, ny, nz, dx, dy, dz, seed=3, z_half=None):
    """int16 HU volume [nz, ny, nx]: body ellipse (180 x 120 mm) of smoothed
    noise mapped to [-900, 1500] HU, air (-1000) outside."""
    rng = np.random.default_rng(seed)
    noise = _smooth_noise((nz, ny, nx), (dz, dy, dx), 8.0, rng)
    lo, hi = np.percentile(noise, [0.5, 99.5])
    hu = -900.0 + (np.clip(noise, lo, hi) - lo) / (hi - lo) * 2400.0
    x = (np.arange(nx) + 0.5) * dx - nx * dx / 2
    y = (np.arange(ny) + 0.5) * dy - ny * dy / 2
    z = (np.arange(nz) + 0.5) * dz - nz * dz / 2
    body = (x[None, :] / 180.0) ** 2 + (y[:, None] / 120.0) ** 2 <= 1.0
    mask = np.broadcast_to(body[None, :, :], (nz, ny, nx))
    if z_half is not None:
        mask = mask & (np.abs(z)[:, None, None] <= z_half)
    out = np.where(mask, hu, -1000.0)
    return np.round(out).astype(np.int16)


def hex_positions(n, pitch, rng_order=None):
    """n positions of a hexagonal grid of the given pitch, closest to the centre."""
    m = int(np.ceil(np.sqrt(n))) + 4
    pts = []
    for j in range(-m, m + 1):
        for i in range(-m, m + 1):
            pts.append(((i + 0.5 * (j & 1)) * pitch, j * pitch * np.sqrt(3.0) / 2.0))
    pts = np.array(pts)
    order = np.lexsort((pts[:, 1], pts[:, 0], np.hypot(pts[:, 0], pts[:, 1]).round(6)))
    return pts[order[:n]]


def ct_plan(n_layers, per_layer, e_lo, e_hi, hist_fn, seed):
    rng = np.random.default_rng(seed)
    pos = hex_positions(per_layer, 5.0)
    sp = []
    for e in np.linspace(e_lo, e_hi, n_layers):
        e = float(np.float32(e))
        for p in pos:
            sp.append(spot(e, 4.0, hist_fn(rng), pos=(float(p[0]), float(p[1])), sigma_e=0.01 * e,
                           rotation=ROT_PLUS_Y, translation=(0.0, -130.0, 0.0)))
    return sp


def ct(seed=3, hist_scale=1.0):
    nx, ny, nz = 512, 512, 200
    dx, dy, dz = 0.98, 0.98, 2.0
    xe = uniform_edges(-nx * dx / 2, nx * dx / 2, nx)
    ye = uniform_edges(-ny * dy / 2, ny * dy / 2, ny)
    ze = uniform_edges(-nz * dz / 2, nz * dz / 2, nz)
    hu = synthetic_ct(nx, ny, nz, dx, dy, dz, seed)
    lo, hi = np.log(1e3), np.log(1e5)
    sp = ct_plan(20, 100, 100.0, 180.0, lambda r: max(1, int(round(np.exp(r.uniform(lo, hi)) * hist_scale))), seed)
    return Workload("CT 512x512x200 @0.98x0.98x2 mm, 2000 spots 100-180 MeV, histories log-U[1e3,1e5]",
                    xe, ye, ze, sp, hu=hu, mask=_abi.SCORE_DOSE, debug=False)


def dij(seed=3, histories=20000):
    nx, ny, nz = 256, 256, 128
    dx, dy, dz = 1.96, 1.96, 4.0
    xe = uniform_edges(-nx * dx / 2, nx * dx / 2, nx)
    ye = uniform_edges(-ny * dy / 2, ny * dy / 2, ny)
    ze = uniform_edges(-nz * dz / 2, nz * dz / 2, nz)
    hu = synthetic_ct(nx, ny, nz, dx, dy, dz, seed, z_half=200.0)
    sp = ct_plan(25, 200, 100.0, 180.0, lambda r: histories, seed)
    return Workload("CT 256x256x128 @1.96x1.96x4 mm, 5000 spots x %d histories, Dij" % histories,
                    xe, ye, ze, sp, hu=hu, mask=_abi.SCORE_DOSE, debug=False)


def robust_scenarios():
    shifts = np.zeros((9, 3), np.float32)
    scales = np.ones(9, np.float32)
    for k, (axis, sgn) in enumerate([(0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)]):
        shifts[1 + k, axis] = 3.0 * sgn
    scales[7] = 1.035
    scales[8] = 0.965
    budgets = np.array([2.0] + [1.0] * 8, np.float32)
    return shifts, scales, budgets


def robust(seed=3):
    """The ct plan (2000 spots, log-uniform histories, 512x512x200 CT) under
    the 9 robust scenarios in one launch, with the Dij scorer."""
    w = ct(seed)
    w.name = ("CT 512x512x200 @0.98x0.98x2 mm, 2000-spot plan (histories log-U[1e3,1e5]) x 9 robust scenarios "
              "(nominal 2x histories), Dij")
    w.scenarios = robust_scenarios()
    return w


def target_prescription(xe, ye, ze, radius=30.0, centre=(0.0, 0.0, 0.0), dose=2.0):
    """2 Gy inside a sphere of the given radius (mm), 0 elsewhere, x fastest."""
    x = 0.5 * (xe[:-1] + xe[1:]) - centre[0]
    y = 0.5 * (ye[:-1] + ye[1:]) - centre[1]
    z = 0.5 * (ze[:-1] + ze[1:]) - centre[2]
    r2 = z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2
    return np.where(r2 <= radius * radius, dose, 0.0).astype(np.float32).reshape(-1)
