"""Small seeded inputs shared by the tests."""
import numpy as np

import refbind as rb


def water_box(nx=40, ny=40, nz=120, dx=1.0, dz=1.0, rho=1.0e-3):
    xe = rb.uniform_edges(-nx * dx / 2, nx * dx / 2, nx)
    ye = rb.uniform_edges(-ny * dx / 2, ny * dx / 2, ny)
    ze = rb.uniform_edges(-nz * dz, 0.0, nz)
    return rb.Grid(xe, ye, ze, np.full(nx * ny * nz, rho, np.float32))


def slab_box(nx=40, ny=40, nz=200, dx=1.0, dz=1.0):
    """Water / lung / bone / water along depth (depth = -z), as the slab workload."""
    g = water_box(nx, ny, nz, dx, dz)
    zc = 0.5 * (g.ze[:-1] + g.ze[1:])
    depth = -zc
    rho_z = np.where(depth < 50, 1.0e-3, np.where(depth < 120, 0.26e-3, np.where(depth < 140, 1.85e-3, 1.0e-3)))
    g.density = np.ascontiguousarray(np.repeat(rho_z.astype(np.float32), nx * ny), np.float32)
    return g


def spots(energies, n, sigma=3.0, sigma_e_frac=0.0, positions=None):
    out = []
    for k, e in enumerate(energies):
        pos = (0.0, 0.0, 0.5) if positions is None else positions[k]
        out.append(rb.make_spot(e, sigma, n, pos=pos, sigma_e=e * sigma_e_frac, energy_const=(sigma_e_frac == 0.0)))
    return np.array(out, dtype=rb.SPOT_DTYPE)
