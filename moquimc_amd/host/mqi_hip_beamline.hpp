// Beamline children as tps_env builds them, for mqi_set_beamline:
//
//   range_shifter     tps_env::create_rangeshifter (mqi_tps_env.hpp:1606-1650):
//                     one voxel, the box pos +- volume / 2 (grid3d(min, max, 2)
//                     per axis), density = rangeshifter density (g/cm^3) * 1e-3
//   voxel_aperture    tps_env::create_voxelized_aperture (mqi_tps_env.hpp:1652-1736):
//                     1 mm voxels over pos +- volume / 2, ceil(volume) per
//                     axis; a voxel column is open (1e-8 g/mm^3) when its
//                     centre lies in the opening polygons (offset by pos),
//                     blocked (100 g/mm^3, where stepping stops a track)
//                     otherwise
//   in_opening        tps_env::is_inside / sol1_1 (mqi_tps_env.hpp:1738-1772):
//                     crossing-number test per polygon; as there, each
//                     opening's result replaces the previous one, so the
//                     last opening decides
//
// The rotation and translation are the beam's coordinate transform with the
// iec2dicom angle set to 90 degrees (p_final there); coordinate_rotation()
// in mqi_hip_env.hpp gives the rotation from the angles.
#ifndef MQI_HIP_BEAMLINE_HPP
#define MQI_HIP_BEAMLINE_HPP

#include <array>
#include <cmath>
#include <vector>

#include "mqi_hip_env.hpp"

namespace mqi_hip
{

// a child grid with its arrays; desc() points into them
struct beamline_grid {
    int                  nx = 0, ny = 0, nz = 0;
    std::vector<float>   xe, ye, ze, density;
    std::array<float, 9> rotation{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
    std::array<float, 3> translation{{0, 0, 0}};

    mqi_grid_desc
    desc() const {
        mqi_grid_desc d{};
        d.nx      = nx;
        d.ny      = ny;
        d.nz      = nz;
        d.xe      = xe.data();
        d.ye      = ye.data();
        d.ze      = ze.data();
        d.density = density.data();
        for (int k = 0; k < 9; ++k) d.rotation[k] = rotation[k];
        for (int k = 0; k < 3; ++k) d.translation[k] = translation[k];
        return d;
    }
};

inline beamline_grid
range_shifter(const std::array<float, 3>& pos, const std::array<float, 3>& volume, float density_g_cm3,
              const std::array<float, 9>& rotation, const std::array<float, 3>& translation) {
    beamline_grid g;
    g.nx = g.ny = g.nz = 1;
    g.xe               = uniform_edges(pos[0] - volume[0] / 2, pos[0] + volume[0] / 2, 1);
    g.ye               = uniform_edges(pos[1] - volume[1] / 2, pos[1] + volume[1] / 2, 1);
    g.ze               = uniform_edges(pos[2] - volume[2] / 2, pos[2] + volume[2] / 2, 1);
    g.density.assign(1, (float) (density_g_cm3 * 1e-3));   // the product in double, as there
    g.rotation    = rotation;
    g.translation = translation;
    return g;
}

// crossing-number point-in-polygon of one opening (vertices in order)
inline bool
in_polygon(float x, float y, const std::vector<std::array<float, 2>>& v) {
    bool      c = false;
    const int n = (int) v.size();
    for (int i = 0, j = n - 1; i < n; j = i++) {
        const std::array<float, 2>& p0 = v[i];
        const std::array<float, 2>& p1 = v[j];
        if ((((p0[1] <= y) && (y < p1[1])) || ((p1[1] <= y) && (y < p0[1]))) &&
            (x < (p1[0] - p0[0]) * (y - p0[1]) / (p1[1] - p0[1]) + p0[0]))
            c = !c;
    }
    return c;
}

inline bool
in_opening(float x, float y, const std::vector<std::vector<std::array<float, 2>>>& openings) {
    bool inside = false;
    for (const auto& o : openings) inside = in_polygon(x, y, o);   // the last opening decides
    return inside;
}

// openings: polygons in the aperture plane (mm, relative to pos)
inline beamline_grid
voxel_aperture(const std::array<float, 3>& pos, const std::array<float, 3>& volume,
               const std::vector<std::vector<std::array<float, 2>>>& openings, const std::array<float, 9>& rotation,
               const std::array<float, 3>& translation) {
    const float   vox = 1.0f;
    beamline_grid g;
    g.nx = (int) std::ceil(volume[0] / vox);
    g.ny = (int) std::ceil(volume[1] / vox);
    g.nz = (int) std::ceil(volume[2] / vox);
    g.xe.resize(g.nx + 1);
    g.ye.resize(g.ny + 1);
    g.ze.resize(g.nz + 1);
    for (int i = 0; i <= g.nx; ++i) g.xe[i] = (pos[0] - volume[0] / 2) + i * vox;
    for (int i = 0; i <= g.ny; ++i) g.ye[i] = (pos[1] - volume[1] / 2) + i * vox;
    for (int i = 0; i <= g.nz; ++i) g.ze[i] = (pos[2] - volume[2] / 2) + i * vox;
    std::vector<std::vector<std::array<float, 2>>> seg = openings;
    for (auto& o : seg)
        for (auto& p : o) {
            p[0] = p[0] + pos[0];
            p[1] = p[1] + pos[1];
        }
    g.density.assign((size_t) g.nx * g.ny * g.nz, 0.0f);
    for (int i = 0; i < g.nx; ++i) {
        const float x = g.xe[i] + vox * 0.5f;
        for (int j = 0; j < g.ny; ++j) {
            const float y   = g.ye[j] + vox * 0.5f;
            const float rho = in_opening(x, y, seg) ? 1e-8f : 100.0f;
            for (int k = 0; k < g.nz; ++k) g.density[(size_t) k * g.nx * g.ny + (size_t) j * g.nx + i] = rho;
        }
    }
    g.rotation    = rotation;
    g.translation = translation;
    return g;
}

}   // namespace mqi_hip

#endif
