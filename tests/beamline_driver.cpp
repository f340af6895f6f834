// Test driver for the beamline helpers (moquimc_amd/host/mqi_hip_beamline.hpp):
//   beamline_driver rs                prints the range shifter's edges and density (hex floats)
//   beamline_driver ap <openings.txt> prints nx ny nz and the aperture's first z plane (0 open, 1 blocked)
// openings.txt: one opening per line, "x0 y0 x1 y1 ..." (mm, relative to pos)
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>

#include "mqi_hip_beamline.hpp"

int
main(int argc, char** argv) {
    const std::array<float, 9> I = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const std::string          mode = argc > 1 ? argv[1] : "";
    if (mode == "rs") {
        mqi_hip::beamline_grid g = mqi_hip::range_shifter({0.5f, -1.25f, 30.0f}, {60.0f, 50.0f, 20.3f}, 1.19f, I, {0, 0, 0});
        for (const std::vector<float>* e : {&g.xe, &g.ye, &g.ze}) std::printf("%a %a\n", (*e)[0], (*e)[1]);
        std::printf("%a\n", g.density[0]);
        return 0;
    }
    if (mode == "ap" && argc == 3) {
        std::vector<std::vector<std::array<float, 2>>> op;
        std::ifstream f(argv[2]);
        std::string   line;
        while (std::getline(f, line)) {
            std::istringstream s(line);
            std::vector<std::array<float, 2>> poly;
            float x, y;
            while (s >> x >> y) poly.push_back({x, y});
            if (!poly.empty()) op.push_back(poly);
        }
        mqi_hip::beamline_grid g = mqi_hip::voxel_aperture({2.0f, -3.0f, 10.0f}, {40.5f, 30.0f, 6.0f}, op, I, {0, 0, 5});
        std::printf("%d %d %d\n", g.nx, g.ny, g.nz);
        for (int j = 0; j < g.ny; ++j) {
            for (int i = 0; i < g.nx; ++i) std::printf("%d", g.density[(size_t) j * g.nx + i] > 1.0f ? 1 : 0);
            std::printf("\n");
        }
        return 0;
    }
    return 2;
}
