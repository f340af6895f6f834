// The phantom driver on the GPU library: the command line of moqui's phantom
// executable (tests/mc/phantom/phantom_env.cpp with mqi::phantom_env's
// options, mqi_phantom_env.hpp:36-171) over mqi_hip::phantom_env.
//
//   phantom_env --output_prefix out/ [--phantom_path ct.raw] --lxyz 100 100 350
//               --pxyz 0 0 -175 --nxyz 200 200 350 --spot_position 0 0 0.5
//               --spot_size 3 3 --spot_energy 200 0 --spot_angles 0 0 0 0
//               --histories 100000 --random_seed 1 [--gpu_id 0] [--exact 0|1]
//
// Writes <output_prefix>dose.f64 (float64 dose-to-water per voxel, Gy summed
// over histories, x fastest) and prints the run summary.  Exit code 0 on
// success, 1 with the library's message on failure.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "mqi_hip_env.hpp"

namespace
{
// --name v1 v2 ... as mqi::cli reads it (mqi_cli.hpp)
std::map<std::string, std::vector<std::string>>
parse(int argc, char** argv) {
    std::map<std::string, std::vector<std::string>> opts;
    std::string                                     key;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
            key       = a;
            opts[key] = {};
        } else if (!key.empty()) {
            opts[key].push_back(a);
        }
    }
    return opts;
}

template<size_t N, typename T, typename F>
void
take(const std::map<std::string, std::vector<std::string>>& o, const char* name, std::array<T, N>& dst, F conv) {
    auto it = o.find(name);
    if (it == o.end()) return;
    if (it->second.size() < N) throw mqi_hip::error(MQI_E_ARG, std::string(name) + " needs " + std::to_string(N) + " values");
    for (size_t k = 0; k < N; ++k) dst[k] = conv(it->second[k]);
}
}   // namespace

int
main(int argc, char** argv) {
    try {
        const auto o = parse(argc, argv);
        auto       f = [](const std::string& s) { return std::stof(s); };
        auto       n = [](const std::string& s) { return std::stoi(s); };
        mqi_hip::phantom_env env;
        take<3>(o, "--lxyz", env.lxyz, f);
        take<3>(o, "--pxyz", env.pos, f);
        take<3>(o, "--nxyz", env.nxyz, n);
        take<3>(o, "--spot_position", env.spot_position, f);
        take<2>(o, "--spot_size", env.spot_size, f);
        take<2>(o, "--spot_energy", env.spot_energy, f);
        take<4>(o, "--spot_angles", env.spot_angles, f);
        if (o.count("--histories") && !o.at("--histories").empty()) env.n_histories = std::stoull(o.at("--histories")[0]);
        if (o.count("--random_seed") && !o.at("--random_seed").empty()) env.random_seed = std::stoi(o.at("--random_seed")[0]);
        if (o.count("--gpu_id") && !o.at("--gpu_id").empty()) env.gpu_id = std::stoi(o.at("--gpu_id")[0]);
        if (o.count("--phantom_path") && !o.at("--phantom_path").empty()) env.phantom_path = o.at("--phantom_path")[0];
        if (o.count("--exact") && !o.at("--exact").empty()) env.config.exact = std::stoi(o.at("--exact")[0]);
        if (!o.count("--output_prefix") || o.at("--output_prefix").empty())
            throw mqi_hip::error(MQI_E_ARG, "--output_prefix is required");
        const std::string prefix = o.at("--output_prefix")[0];

        env.initialize();
        env.run();
        env.finalize();
        std::ofstream out(prefix + "dose.f64", std::ios::binary);
        out.write(reinterpret_cast<const char*>(env.dose().data()), (std::streamsize) (env.dose().size() * sizeof(double)));
        if (!out) throw mqi_hip::error(MQI_E_ARG, "cannot write " + prefix + "dose.f64");
        std::printf("histories %llu steps %llu dropped %llu kernel %.3f ms\n", (unsigned long long) env.result.histories,
                    (unsigned long long) env.result.steps, (unsigned long long) env.result.dropped, env.result.kernel_ms);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "phantom_env: %s\n", e.what());
        return 1;
    }
}
