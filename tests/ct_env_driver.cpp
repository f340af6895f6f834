// Test driver for mqi_hip::ct_env (moquimc_amd/host/mqi_hip_ct_env.hpp):
//   ct_env_driver <ct.mha> <spots.bin> <out_dir> <format> <seed> <scorers> <stat> <mask.mha|->
// spots.bin holds mqi_spot records; scorers is a comma list (Dose, LETd,
// LETt or Dij); stat 1 adds the statistics pair (criterion 0.5, threshold
// 0.5, saved); the verification kernels run, batched.  Prints the histories.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <sstream>
#include <string>
#include <vector>

#include "mqi_hip_ct_env.hpp"

int
main(int argc, char** argv) {
    if (argc != 9) return 2;
    try {
        mqi_hip::ct_env env;
        env.ct_path = argv[1];
        std::ifstream     f(argv[2], std::ios::binary);
        std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        env.beamlets.resize(b.size() / sizeof(mqi_spot));
        std::copy(b.begin(), b.begin() + env.beamlets.size() * sizeof(mqi_spot),
                  reinterpret_cast<char*>(env.beamlets.data()));
        env.output_dir    = argv[3];
        env.output_format = argv[4];
        env.random_seed   = std::strtoull(argv[5], nullptr, 10);
        env.scorer_names.clear();
        std::stringstream ss(argv[6]);
        for (std::string t; std::getline(ss, t, ',');) env.scorer_names.push_back(t);
        env.record_statistics = std::atoi(argv[7]) != 0;
        env.stat_criteria     = 0.5f;
        env.stat_threshold    = 0.5f;
        env.save_statistics   = true;
        if (std::string(argv[8]) != "-") env.mask_files = {argv[8]};
        env.config.exact   = 1;
        env.config.batched = 1;
        env.initialize();
        env.run();
        env.finalize();
        std::printf("%llu\n", (unsigned long long) env.result.histories);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
