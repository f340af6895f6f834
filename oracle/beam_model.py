"""Test infrastructure, not product code: a numpy restatement of moqui's
generic pencil-beam-scanning beam model, the checker for the host beam model
(moquimc_amd/host/mqi_hip_pbs.hpp).  Only tests import it.

Follows moqui/base/mqi_treatment_machine_pbs.hpp and
moqui/base/mqi_treatment_machine_ion.hpp step by step in float32 (numpy
scalar float32 arithmetic rounds every operation to float, as the
reference's float code does on the host):

  intpl                   mqi_treatment_machine_pbs.hpp:95-97
  characterize_history    mqi_treatment_machine_pbs.hpp:135-143
  characterize_beamlet    mqi_treatment_machine_pbs.hpp:199-235
  beam_starting_position  mqi_treatment_machine_ion.hpp:324-333
  [spot] / [geometry] rows of load_beamdata  mqi_treatment_machine_pbs.hpp:401-516

The reference itself cannot be compiled here (mqi_treatment_machine_ion.hpp
pulls in mqi_dataset.hpp, which needs GDCM), so this restatement is pinned by
closed forms in tests/test_beam_model.py: table values at the nodes, linear
interpolation between them, the start point on the ray from the virtual
source, unit directions and the truncated history count.
"""
import bisect

import numpy as np

f32 = np.float32


def intpl(x, x0, x1, y0, y1):
    x, x0, x1, y0, y1 = f32(x), f32(x0), f32(x1), f32(y0), f32(y1)
    if x1 == x0:
        return y0
    return f32(y0 + f32(f32(f32(x - x0) * f32(y1 - y0)) / f32(x1 - x0)))


class BeamModel:
    """keys: sorted nominal energies; rows[k] = (E, dE, x, y, xp, yp, ratio)."""

    def __init__(self, table, sad=(np.inf, np.inf), source_to_isocenter=390.0):
        table = sorted((f32(e), tuple(f32(v) for v in row)) for e, row in table)
        self.keys = [e for e, _ in table]
        self.rows = [r for _, r in table]
        self.sad = (f32(sad[0]), f32(sad[1]))
        self.z = f32(source_to_isocenter)

    @classmethod
    def from_file(cls, path, source_to_isocenter=390.0):
        section, table, sad = None, [], (0.0, 0.0)
        for line in open(path):
            line = line.split("#", 1)[0]
            if "[" in line and "]" in line:
                section = line[line.index("[") + 1:line.index("]")]
                continue
            tok = line.split()
            if not tok:
                continue
            if section == "geometry" and tok[0] == "SAD(mm)":
                sad = (float(tok[1]), float(tok[2]))
            elif section == "spot":
                table.append((float(tok[0]), [float(t) for t in tok[1:8]]))
        sad = tuple(np.inf if s == 0.0 else s for s in sad)
        return cls(table, sad, source_to_isocenter)

    def _bracket(self, e):
        k = bisect.bisect_left(self.keys, f32(e))   # lower_bound
        if k == 0 or k == len(self.keys):
            raise ValueError("energy %r outside the table" % e)
        return k - 1, k

    def beam_starting_position(self, iso, z):
        with np.errstate(invalid="ignore"):
            bx = f32(f32(f32(iso[0]) * f32(self.sad[0] - f32(z))) / self.sad[0])
            by = f32(f32(f32(iso[1]) * f32(self.sad[1] - f32(z))) / self.sad[1])
        return bx, by, f32(z)

    def characterize_history(self, e, meterset, scale):
        d, u = self._bracket(e)
        ratio = intpl(e, self.keys[d], self.keys[u], self.rows[d][6], self.rows[u][6])
        h = f32(f32(f32(meterset) * ratio) / f32(scale))
        return int(np.floor(h))

    def characterize_beamlet(self, e, x, y):
        """(energy, sigma_energy, mean[6], sigma[6]) of the beamlet."""
        d, u = self._bracket(e)
        D, U = self.rows[d], self.rows[u]
        me = intpl(e, self.keys[d], self.keys[u], D[0], U[0])
        mid = [intpl(e, D[0], U[0], D[i], U[i]) for i in range(1, 6)]   # dE, x, y, xp, yp
        bx, by, bz = self.beam_starting_position((f32(x), f32(y)), self.z)
        with np.errstate(invalid="ignore"):
            dx, dy, dz = f32(f32(x) - bx), f32(f32(y) - by), f32(f32(0.0) - bz)
            n = f32(np.sqrt(f32(f32(f32(dx * dx) + f32(dy * dy)) + f32(dz * dz))))
            dx, dy, dz = f32(dx / n), f32(dy / n), f32(dz / n)
        mean = [bx, by, bz, dx, dy, dz]
        sigma = [mid[1], mid[2], f32(0.0), mid[3], mid[4], f32(0.0)]
        return me, mid[0], mean, sigma
