// Test driver for the host beam model (moquimc_amd/host/mqi_hip_pbs.hpp):
//   pbs_driver <machine file> <source_to_isocenter (hex float)> <particles_per_history (hex float)> <spots file>
//              [beamsource.bin]
// spots file: one "e x y meterset" line per spot, hex floats.  Prints per
// spot either "error" or the beamlet as hex floats:
//   energy sigma_energy mean[6] sigma[6] histories
// and, last, the create_beamsource table of the whole list with identity
// rotation and zero translation as "beamsource <n> <sum of histories>"; with
// a fifth argument the table's mqi_spot records are also written there.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "mqi_hip_pbs.hpp"

int
main(int argc, char** argv) {
    if (argc != 5 && argc != 6) return 2;
    try {
        mqi_hip::pbs_model m(argv[1]);
        m.source_to_isocenter_mm = std::strtof(argv[2], nullptr);
        const float   pph        = std::strtof(argv[3], nullptr);
        std::ifstream f(argv[4]);
        std::string   t[4];
        std::vector<mqi_hip::ion_spot> ok;
        while (f >> t[0] >> t[1] >> t[2] >> t[3]) {
            mqi_hip::ion_spot s{std::strtof(t[0].c_str(), nullptr), std::strtof(t[1].c_str(), nullptr),
                                std::strtof(t[2].c_str(), nullptr), 0.0f, 0.0f, std::strtof(t[3].c_str(), nullptr)};
            try {
                mqi_spot     b = m.characterize_beamlet(s);
                const size_t h = pph == -1.0f ? 1 : m.characterize_history(s, pph);
                std::printf("%a %a", b.energy, b.sigma_energy);
                for (int k = 0; k < 6; ++k) std::printf(" %a", b.mean[k]);
                for (int k = 0; k < 6; ++k) std::printf(" %a", b.sigma[k]);
                std::printf(" %zu\n", h);
                ok.push_back(s);
            } catch (const mqi_hip::error&) {
                std::printf("error\n");
            }
        }
        const std::array<float, 9> rot = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        std::vector<mqi_spot>      bs  = m.create_beamsource(ok, rot, {0.0f, 0.0f, 0.0f}, pph);
        unsigned long long         sum = 0;
        for (const mqi_spot& b : bs) sum += b.histories;
        std::printf("beamsource %zu %llu\n", bs.size(), sum);
        if (argc == 6) {
            std::ofstream o(argv[5], std::ios::binary);
            o.write(reinterpret_cast<const char*>(bs.data()), (std::streamsize) (bs.size() * sizeof(mqi_spot)));
            if (!o) return 1;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
