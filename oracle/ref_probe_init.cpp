// ref_probe_init.cpp -- dumps the initial fields produced by the reference's
// GPU-tree Inamuro constructor (GPU/Inamuro.cpp:491-566, init_mode /
// init_profile keys), compiled from /root/reference by oracle/Makefile.
// TEST INFRASTRUCTURE ONLY.  The GPU-tree Inamuro.cpp has no CUDA includes,
// so g++ builds it.
//
// usage: ref_probe_init PARAMS OUT
//   OUT: int32 {magic, lx, ly, lz}; 6 float64 fields (fei, rho, u, v, w, p),
//        each [z][y][x] x-fastest.
#include "Inamuro.hpp"

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

struct InitProbe : public Inamuro {
    explicit InitProbe(const std::string& file) : Inamuro(file) {}

    void dump(std::FILE* out)
    {
        const std::int32_t hdr[4] = {0x4c424d32, lx, ly, lz};
        std::fwrite(hdr, sizeof(hdr), 1, out);
        std::vector<double> tmp(static_cast<std::size_t>(lx) * ly * lz);
        for (Vector3D* a : {&fei, &rho, &u, &v, &w, &p}) {
            for (int z = 0; z < lz; ++z)
                for (int y = 0; y < ly; ++y)
                    for (int x = 0; x < lx; ++x)
                        tmp[(static_cast<std::size_t>(z) * ly + y) * lx + x] = (*a)[x][y][z + 1];
            std::fwrite(tmp.data(), sizeof(double), tmp.size(), out);
        }
    }
};

} // namespace

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s PARAMS OUT\n", argv[0]);
        return 2;
    }
    std::ofstream devnull("/dev/null");
    std::streambuf* saved = std::cout.rdbuf(devnull.rdbuf());
    std::streambuf* saved_err = std::cerr.rdbuf(devnull.rdbuf());
    InitProbe probe(argv[1]);
    std::cout.rdbuf(saved);
    std::cerr.rdbuf(saved_err);
    std::FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    probe.dump(out);
    std::fclose(out);
    return 0;
}
