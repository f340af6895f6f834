// Generic pencil-beam-scanning beam model on the host: a plan spot (nominal
// energy, position at the isocenter plane, meterset) becomes a beamlet (the
// spot table entry the transport samples histories from) and a history
// count.  Mirrors mqi::pbs<T> (moqui/base/mqi_treatment_machine_pbs.hpp) and
// the parts of mqi::treatment_machine_ion<T> it relies on:
//
//   pbs::characterize_beamlet(spot)      mqi_treatment_machine_pbs.hpp:199-235
//   pbs::characterize_history(spot, s)   mqi_treatment_machine_pbs.hpp:135-143
//   pbs::intpl                           mqi_treatment_machine_pbs.hpp:95-97
//   pbs::load_beamdata ([geometry] SAD and [spot] only)   :401-516
//   treatment_machine_ion::beam_starting_position       mqi_treatment_machine_ion.hpp:324-333
//   treatment_machine_ion::create_beamsource (MODULATED) mqi_treatment_machine_ion.hpp:83-134
//
// Arithmetic is float, in the reference's order, so the beamlets are the
// reference's to the bit.  Kept on purpose: energy and ratio interpolate
// against the table's nominal energies (the map keys) while dE, sigma and
// divergence interpolate against its E column; an infinite SAD (0 in the
// file) gives NaN start positions, as in the reference.  Different on
// purpose: an energy at or below the lowest nominal energy, or above the
// highest, throws (the reference prints "out-of-bound" and reads outside
// the table); sections of the machine file may come in any order.
#ifndef MQI_HIP_PBS_HPP
#define MQI_HIP_PBS_HPP

#include <array>
#include <cmath>
#include <cstddef>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "mqi_hip_env.hpp"

namespace mqi_hip
{

// beam_module_ion::spot (mqi_beam_module_ion.hpp:31-38)
struct ion_spot {
    float e;          // nominal energy (MeV)
    float x;          // position at the isocenter plane (mm)
    float y;
    float fwhm_x;     // not used by the MODULATED beam model
    float fwhm_y;
    float meterset;   // MU (or number of particles)
};

// one row of the machine file's [spot] table (pbs::spot_specification,
// mqi_treatment_machine_pbs.hpp:58-66)
struct spot_spec {
    float E;       // mean energy (MeV)
    float dE;      // energy spread, 1 sigma (MeV)
    float x, y;    // spot sigma (mm)
    float xp, yp;  // divergence sigma (rad)
    float ratio;   // particles per MU
};

class pbs_model
{
public:
    std::map<float, spot_spec> beamdata;   // nominal energy -> row
    std::array<float, 2>       sad = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity()};
    float                      source_to_isocenter_mm = 390.0f;   // where histories start (create_beamsource)

    pbs_model() = default;

    explicit pbs_model(const std::string& machine_file) {
        load(machine_file);
    }

    static float
    intpl(float x, float x0, float x1, float y0, float y1) {
        return (x1 == x0) ? y0 : y0 + (x - x0) * (y1 - y0) / (x1 - x0);
    }

    // [geometry] "SAD(mm) sx sy" (0: infinity) and [spot] rows
    // "nominal_E E dE x y xp yp ratio"; '#' starts a comment; other sections
    // (rangeshifter, aperture, time) are skipped.
    void
    load(const std::string& path) {
        std::ifstream f(path);
        if (!f) throw error(MQI_E_ARG, "cannot open beam model " + path);
        std::string line, section;
        std::array<float, 2> s = {0.0f, 0.0f};
        bool                 have_sad = false;
        while (std::getline(f, line)) {
            line = line.substr(0, line.find_first_of('#'));
            const size_t a = line.find('['), b = line.find(']');
            if (a != std::string::npos && b != std::string::npos && b > a) {
                section = line.substr(a + 1, b - a - 1);
                continue;
            }
            if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
            std::stringstream d(line);
            if (section == "geometry") {
                std::string key;
                float       v1 = 0.0f, v2 = 0.0f;
                d >> key >> v1 >> v2;
                if (key == "SAD(mm)") {
                    s        = {v1, v2};
                    have_sad = true;
                }
            } else if (section == "spot") {
                float     e;
                spot_spec r;
                if (!(d >> e >> r.E >> r.dE >> r.x >> r.y >> r.xp >> r.yp >> r.ratio))
                    throw error(MQI_E_ARG, "bad [spot] row in " + path + ": " + line);
                beamdata.insert(std::make_pair(e, r));
            }
        }
        if (have_sad)
            for (int k = 0; k < 2; ++k) sad[k] = s[k] == 0.0f ? std::numeric_limits<float>::infinity() : s[k];
        if (beamdata.size() < 2) throw error(MQI_E_ARG, "beam model " + path + " has fewer than two [spot] rows");
    }

    // start of the beam at height z on the ray from the virtual source
    // (0, 0, SAD) through the isocenter-plane point iso
    std::array<float, 3>
    beam_starting_position(const std::array<float, 3>& iso, float z) const {
        std::array<float, 3> beam = {0.0f, 0.0f, z};
        beam[0]                   = iso[0] * (sad[0] - beam[2]) / sad[0];
        beam[1]                   = iso[1] * (sad[1] - beam[2]) / sad[1];
        return beam;
    }

    // histories of a spot: meterset * ratio / particles_per_history, truncated
    size_t
    characterize_history(const ion_spot& s, float scale) const {
        const auto  r     = bracket(s.e);
        const float ratio = intpl(s.e, r.first->first, r.second->first, r.first->second.ratio, r.second->second.ratio);
        const float h     = s.meterset * ratio / scale;
        if (!(h >= 0.0f) || h >= 1.8e19f) throw error(MQI_E_ARG, "spot history count out of range");
        return (size_t) h;
    }

    // the beamlet of a spot: Gaussian energy (mean, sigma) and Gaussian 6-D
    // phase space centred on the start position and the unit direction
    // towards the isocenter-plane point; histories left 0
    mqi_spot
    characterize_beamlet(const ion_spot& s) const {
        const auto       r    = bracket(s.e);
        const spot_spec& down = r.first->second;
        const spot_spec& up   = r.second->second;
        const float      e    = intpl(s.e, r.first->first, r.second->first, down.E, up.E);
        const float      de   = intpl(s.e, down.E, up.E, down.dE, up.dE);
        const float      sx   = intpl(s.e, down.E, up.E, down.x, up.x);
        const float      sy   = intpl(s.e, down.E, up.E, down.y, up.y);
        const float      sxp  = intpl(s.e, down.E, up.E, down.xp, up.xp);
        const float      syp  = intpl(s.e, down.E, up.E, down.yp, up.yp);

        const std::array<float, 3> iso  = {s.x, s.y, 0.0f};
        const std::array<float, 3> beam = beam_starting_position(iso, source_to_isocenter_mm);
        float                      dx = iso[0] - beam[0], dy = iso[1] - beam[1], dz = iso[2] - beam[2];
        const float                n  = std::sqrt(dx * dx + dy * dy + dz * dz);
        dx /= n;
        dy /= n;
        dz /= n;

        mqi_spot b{};
        b.energy       = e;
        b.sigma_energy = de;
        b.energy_const = 0;
        b.phsp         = MQI_PHSP_GAUSS;
        const float mean[6]  = {beam[0], beam[1], beam[2], dx, dy, dz};
        const float sigma[6] = {sx, sy, 0.0f, sxp, syp, 0.0f};
        for (int k = 0; k < 6; ++k) {
            b.mean[k]  = mean[k];
            b.sigma[k] = sigma[k];
        }
        b.corr[0] = b.corr[1] = 0.0f;
        for (int k = 0; k < 9; ++k) b.rotation[k] = (k % 4 == 0) ? 1.0f : 0.0f;
        return b;
    }

    // the spot table of a MODULATED beam: one beamlet per spot with its
    // history count (1 each when particles_per_history == -1) and the beam's
    // coordinate transform (rotation = coordinate_rotation(angles),
    // translation = isocenter)
    std::vector<mqi_spot>
    create_beamsource(const std::vector<ion_spot>& spots, const std::array<float, 9>& rotation,
                      const std::array<float, 3>& translation, float particles_per_history = -1.0f) const {
        std::vector<mqi_spot> out;
        out.reserve(spots.size());
        for (const ion_spot& s : spots) {
            mqi_spot b  = characterize_beamlet(s);
            b.histories = particles_per_history == -1.0f ? 1 : characterize_history(s, particles_per_history);
            for (int k = 0; k < 9; ++k) b.rotation[k] = rotation[k];
            for (int k = 0; k < 3; ++k) b.translation[k] = translation[k];
            out.push_back(b);
        }
        return out;
    }

private:
    // (the row at or above e, the row below it), as lower_bound and prev
    std::pair<std::map<float, spot_spec>::const_iterator, std::map<float, spot_spec>::const_iterator>
    bracket(float e) const {
        auto up = beamdata.lower_bound(e);
        if (up == beamdata.begin() || up == beamdata.end())
            throw error(MQI_E_ARG, "spot energy " + std::to_string(e) + " MeV is outside the beam model's table");
        return std::make_pair(std::prev(up), up);
    }
};

}   // namespace mqi_hip

#endif
