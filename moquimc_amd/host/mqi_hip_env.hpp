// C++ environment classes over the C ABI of libmqi_hip.so, mirroring moqui's
// environment interface so a moqui driver ports by swapping the base class:
//
//   mqi::x_environment<R>  (moqui/base/environments/mqi_xenvironment.hpp:70-146)
//       setup_world / setup_materials / setup_beamsource / initialize / run /
//       finalize                                  -> mqi_hip::x_environment
//   mqi::phantom_env<R>    (moqui/base/environments/mqi_phantom_env.hpp:21-428)
//       same parameters (lxyz, pos, nxyz, spot_position, spot_size,
//       spot_energy, spot_angles, n_histories, random_seed, phantom_path)
//                                                 -> mqi_hip::phantom_env
//
// Differences by design: the world is the patient/phantom grid alone (the
// reference also transports through a 600 mm air box around it); beamlet
// vertices are sampled on the device from the spot table instead of on the
// host; errors throw mqi_hip::error instead of calling exit().
#ifndef MQI_HIP_ENV_HPP
#define MQI_HIP_ENV_HPP

#include <array>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mqi_hip.h"

namespace mqi_hip
{

struct error : std::runtime_error {
    int code;
    error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void
check(const mqi_ctx* ctx, int rc, const char* what) {
    if (rc != MQI_OK) throw error(rc, std::string(what) + ": " + mqi_last_error(ctx));
}

// grid3d(min, max, n + 1) edges in float: dx = (max - min) / n, e_i = min + i dx
// (mqi_grid3d.hpp:134-163)
inline std::vector<float>
uniform_edges(float lo, float hi, int n) {
    std::vector<float> e(n + 1);
    const float        d = (hi - lo) / (float) n;
    for (int i = 0; i <= n; ++i) e[i] = lo + (float) i * d;
    return e;
}

// coordinate_transform rotation: iec2dicom (x) * couch (z) * gantry (y) *
// collimator (z), angles in degrees, float (mqi_coordinate_transform.hpp:42-60)
inline std::array<float, 9>
coordinate_rotation(const std::array<float, 4>& deg) {
    typedef std::array<float, 9> m3;
    const float                  k = (float) (M_PI / 180.0);
    // one axis rotation applied in place to the identity as mat3x3::rotate_x/y/z
    // do it (mqi_matrix.hpp:212-273), so that even the signs of the zero
    // entries (c * 0 with c = cos(90 deg) < 0) come out the same
    auto axis = [](int ax, float a) {
        m3 m = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        if (a == 0) return m;
        const float c = std::cos(a), s = std::sin(a);
        // rows (i, j) rotated into each other: x: (1, 2), y: (2, 0), z: (0, 1)
        const int i = ax == 0 ? 1 : (ax == 1 ? 2 : 0), j = ax == 0 ? 2 : (ax == 1 ? 0 : 1);
        for (int k = 0; k < 3; ++k) {
            const float u = m[3 * i + k], v = m[3 * j + k];
            m[3 * i + k]  = c * u - s * v;
            m[3 * j + k]  = s * u + c * v;
        }
        return m;
    };
    auto mul = [](const m3& a, const m3& b) {
        m3 r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
        return r;
    };
    return mul(mul(mul(axis(0, deg[3] * k), axis(2, deg[2] * k)), axis(1, deg[1] * k)), axis(2, deg[0] * k));
}

// The environment: subclasses describe the world and the beam; initialize()
// uploads them (upload_node / upload_materials), run() transports every
// history of every spot, finalize() downloads the scorers (download_node).
class x_environment
{
public:
    int            gpu_id = 0;
    mqi_run_config config{};   // seed, physics build, batching, statistics, Dij ...
    mqi_run_result result{};

    // world: the voxel grid (edges, and density in g/mm^3 or HU), placement
    std::vector<float>   xe, ye, ze;
    std::vector<float>   density;   // used when hu is empty
    std::vector<int16_t> hu;
    float                density_scale = 1.0f;
    std::array<float, 9> rotation      = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::array<float, 3> translation   = {0, 0, 0};
    int                  scorer_mask   = MQI_SCORE_DOSE;
    // beam source: one spot per beamlet
    std::vector<mqi_spot> spots;
    // downloaded scorers (Dose, LETd numerator / denominator, Dose_stat, DoseSquare_stat)
    std::vector<std::vector<double>> scorers;

    x_environment() { config.batched = 1; }
    x_environment(const x_environment&) = delete;
    x_environment& operator=(const x_environment&) = delete;
    virtual ~x_environment() {
        if (ctx_) mqi_destroy(ctx_);
    }

    virtual void setup_world()      = 0;
    virtual void setup_materials()  {}   // the patient material is the device HU calibration
    virtual void setup_beamsource() = 0;

    virtual void
    initialize() {
        setup_world();
        setup_materials();
        setup_beamsource();
        if (!ctx_) check(nullptr, mqi_create(gpu_id, &ctx_), "mqi_create");
        const int nx = (int) xe.size() - 1, ny = (int) ye.size() - 1, nz = (int) ze.size() - 1;
        if (nx < 1 || ny < 1 || nz < 1) throw error(MQI_E_ARG, "setup_world left an empty grid");
        if (!hu.empty())
            check(ctx_, mqi_set_grid_hu(ctx_, nx, ny, nz, xe.data(), ye.data(), ze.data(), hu.data(), density_scale,
                                        rotation.data(), translation.data()),
                  "mqi_set_grid_hu");
        else
            check(ctx_, mqi_set_grid(ctx_, nx, ny, nz, xe.data(), ye.data(), ze.data(), density.data(),
                                     rotation.data(), translation.data()),
                  "mqi_set_grid");
        check(ctx_, mqi_set_spots(ctx_, spots.data(), (int) spots.size()), "mqi_set_spots");
        check(ctx_, mqi_set_scorers(ctx_, scorer_mask, 0), "mqi_set_scorers");
    }

    virtual void
    run() {
        check(ctx_, mqi_run(ctx_, &config, &result), "mqi_run");
    }

    virtual void
    finalize() {
        const size_t n = (xe.size() - 1) * (ye.size() - 1) * (ze.size() - 1);
        const int    k = mqi_scorer_count(ctx_);
        scorers.assign(k, std::vector<double>(n));
        for (int i = 0; i < k; ++i) check(ctx_, mqi_get_scorer(ctx_, i, 0, scorers[i].data()), "mqi_get_scorer");
    }

    mqi_ctx*
    context() {
        return ctx_;
    }

protected:
    mqi_ctx* ctx_ = nullptr;
};

// The water / CT phantom driver: a box of nxyz voxels of size lxyz centred at
// pos, one beamlet at spot_position travelling -z in its frame, uniform
// phase space of half-width spot_size (phsp_6d_uniform), constant energy
// spot_energy[0], the coordinate transform of spot_angles, __PHYSICS_DEBUG__
// physics as the reference's phantom build (mqi_phantom_env.hpp:205-300).
class phantom_env : public x_environment
{
public:
    std::array<float, 3> lxyz          = {512.0f, 512.0f, 400.0f};
    std::array<float, 3> pos           = {-256.0f, 0.0f, 0.0f};
    std::array<int, 3>   nxyz          = {512, 512, 200};
    std::array<float, 3> spot_position = {1.0f, 0.0f, 0.0f};
    std::array<float, 2> spot_size     = {0.0f, 0.0f};
    std::array<float, 2> spot_energy   = {230.0f, 0.0f};
    std::array<float, 4> spot_angles   = {0.0f, 0.0f, 0.0f, 0.0f};
    size_t               n_histories   = 10000;
    int                  random_seed   = 0;
    std::string          phantom_path;   // int16 HU, x fastest; empty: water (HU 0)

    void
    setup_world() override {
        xe = uniform_edges(pos[0] - 0.5f * lxyz[0], pos[0] + 0.5f * lxyz[0], nxyz[0]);
        ye = uniform_edges(pos[1] - 0.5f * lxyz[1], pos[1] + 0.5f * lxyz[1], nxyz[1]);
        ze = uniform_edges(pos[2] - 0.5f * lxyz[2], pos[2] + 0.5f * lxyz[2], nxyz[2]);
        const size_t n = (size_t) nxyz[0] * nxyz[1] * nxyz[2];
        hu.assign(n, 0);
        if (!phantom_path.empty()) {
            std::ifstream f(phantom_path, std::ios::binary);
            if (!f) throw error(MQI_E_ARG, "cannot open phantom " + phantom_path);
            f.read(reinterpret_cast<char*>(hu.data()), (std::streamsize) (n * sizeof(int16_t)));
            if ((size_t) f.gcount() != n * sizeof(int16_t)) throw error(MQI_E_ARG, "short phantom file " + phantom_path);
        }
    }

    void
    setup_beamsource() override {
        mqi_spot s{};
        s.energy       = spot_energy[0];
        s.sigma_energy = spot_energy[1];
        s.energy_const = 1;   // const_1d: the spread is ignored
        s.phsp         = MQI_PHSP_UNIFORM;
        const float m[6] = {spot_position[0], spot_position[1], spot_position[2], 0.0f, 0.0f, -1.0f};
        const float g[6] = {spot_size[0], spot_size[1], 0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = 0; k < 6; ++k) {
            s.mean[k]  = m[k];
            s.sigma[k] = g[k];
        }
        const std::array<float, 9> r = coordinate_rotation(spot_angles);
        for (int k = 0; k < 9; ++k) s.rotation[k] = r[k];
        s.histories = n_histories;
        spots.assign(1, s);
    }

    void
    run() override {
        config.seed          = (uint64_t) (uint32_t) random_seed;
        config.physics_debug = 1;
        x_environment::run();
    }

    // dose-to-water per voxel (Gy, summed over histories), x fastest
    const std::vector<double>&
    dose() const {
        return scorers.at(0);
    }
};

}   // namespace mqi_hip

#endif
