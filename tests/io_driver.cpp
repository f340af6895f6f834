// Test driver for the output writers (moquimc_amd/host/mqi_hip_io.hpp):
//   io_driver crc <file>                       prints the CRC-32 of the file's bytes (hex)
//   io_driver npz <out.npz> <rows> <cols> <scale> <indptr.i64> <indices.i32> <values.f32>
//   io_driver ct <ct.mha> <hu.out> [sx sy sz]  reads a MetaImage CT, prints dims and edges
//   io_driver dense <dir> <name> <nx> <ny> <nz> <scale> <values.f64> <xe.f32> <ye.f32> <ze.f32>
//       writes <name>_raw.raw, <name>_mhd.mhd + .raw and <name>_mha.mha
//   io_driver mask <out.u8> <nx> <ny> <nz> <mask.mha>...   sum of the masks (mask_total) to <out>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "mqi_hip_io.hpp"

template<typename T>
static std::vector<T>
slurp(const char* path) {
    std::ifstream     f(path, std::ios::binary);
    std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<T>    v(b.size() / sizeof(T));
    if (!v.empty()) std::memcpy(v.data(), b.data(), v.size() * sizeof(T));
    return v;
}

int
main(int argc, char** argv) {
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "crc" && argc == 3) {
            std::vector<char> b = slurp<char>(argv[2]);
            std::printf("%08x\n", mqi_hip::io::crc32_update(0, b.data(), b.size()));
            return 0;
        }
        if (mode == "npz" && argc == 9) {
            auto ip = slurp<int64_t>(argv[6]);
            auto ix = slurp<int32_t>(argv[7]);
            auto v  = slurp<float>(argv[8]);
            mqi_hip::io::save_npz_csr(argv[2], (uint32_t) std::atol(argv[3]), (uint32_t) std::atol(argv[4]), ip.data(),
                                      ix.data(), v.data(), std::atof(argv[5]));
            return 0;
        }
        if (mode == "ct" && (argc == 4 || argc == 7)) {
            // prints nx ny nz, then the edges as hex floats; HU to <out>
            float sh[3];
            if (argc == 7)
                for (int k = 0; k < 3; ++k) sh[k] = (float) std::atof(argv[4 + k]);
            mqi_hip::io::ct_image ct = mqi_hip::io::read_mha_ct(argv[2], argc == 7 ? sh : nullptr);
            std::printf("%d %d %d\n", ct.nx, ct.ny, ct.nz);
            for (const std::vector<float>* e : {&ct.xe, &ct.ye, &ct.ze}) {
                for (float v : *e) std::printf("%a ", v);
                std::printf("\n");
            }
            std::ofstream o(argv[3], std::ios::binary);
            o.write(reinterpret_cast<const char*>(ct.hu.data()), (std::streamsize) (ct.hu.size() * 2));
            return 0;
        }
        if (mode == "mask" && argc >= 7) {
            std::vector<std::string> paths(argv + 6, argv + argc);
            const std::vector<uint8_t> t =
              mqi_hip::io::read_mask_total(paths, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
            std::ofstream o(argv[2], std::ios::binary);
            o.write(reinterpret_cast<const char*>(t.data()), (std::streamsize) t.size());
            return 0;
        }
        if (mode == "dense" && argc == 12) {
            const int nx = std::atoi(argv[4]), ny = std::atoi(argv[5]), nz = std::atoi(argv[6]);
            const double scale = std::atof(argv[7]);
            auto d  = slurp<double>(argv[8]);
            auto xe = slurp<float>(argv[9]);
            auto ye = slurp<float>(argv[10]);
            auto ze = slurp<float>(argv[11]);
            const std::string dir = argv[2], name = argv[3];
            mqi_hip::io::save_to_bin(d.data(), d.size(), scale, dir, name + "_raw");
            mqi_hip::io::save_metaimage(d.data(), nx, ny, nz, xe.data(), ye.data(), ze.data(), scale, dir, name + "_mhd", false);
            mqi_hip::io::save_metaimage(d.data(), nx, ny, nz, xe.data(), ye.data(), ze.data(), scale, dir, name + "_mha", true);
            return 0;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 2;
}
