// ref_probe.cpp -- drives the reference CPU solver (CPU/Inamuro.cpp, compiled
// from /root/reference by oracle/Makefile) to produce golden vectors for the
// oracle.  TEST INFRASTRUCTURE ONLY.
//
// A derived class reads and overwrites the reference's protected fields; no
// reference file is edited.  One simulation per process, because the
// reference keeps function-static state (CPU/Inamuro.cpp:150, 254, 627-628).
//
// usage: ref_probe PARAMS INIT|- STEPS OUT [DUMP_DISTS [OUTPUT_DIR]]
//   INIT: 6 float64 fields (fei, rho, u, v, w, p), each [z][y][x] x-fastest,
//         written into the reference's macro arrays at z+1 after construction
//         ("-" keeps the reference's own two-droplet initialisation).
//   OUT:  int32 {magic, lx, ly, lz, steps, dump}; uint64 iterations[steps];
//         double step_ms[steps]; 6 fields; optionally f, g, h as [q][z][y][x].
//   OUTPUT_DIR: after the steps, the reference's own writeResults(STEPS)
//         writes its Tecplot file and debug CSV there (CPU/Inamuro.cpp:365-381).
#include "Inamuro.hpp"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

struct Probe : public Inamuro {
    explicit Probe(const std::string& file) : Inamuro(file) {}

    Vector3D* field(int i)
    {
        switch (i) {
        case 0: return &fei;
        case 1: return &rho;
        case 2: return &u;
        case 3: return &v;
        case 4: return &w;
        default: return &p;
        }
    }

    void load(const std::vector<double>& buf)
    {
        const std::size_t n = static_cast<std::size_t>(lx) * ly * lz;
        for (int f = 0; f < 6; ++f) {
            Vector3D& a = *field(f);
            for (int z = 0; z < lz; ++z)
                for (int y = 0; y < ly; ++y)
                    for (int x = 0; x < lx; ++x)
                        a[x][y][z + 1] = buf[f * n + (static_cast<std::size_t>(z) * ly + y) * lx + x];
        }
    }

    void dump(std::FILE* out, bool dists)
    {
        std::vector<double> tmp(static_cast<std::size_t>(lx) * ly * lz);
        for (int f = 0; f < 6; ++f) {
            Vector3D& a = *field(f);
            for (int z = 0; z < lz; ++z)
                for (int y = 0; y < ly; ++y)
                    for (int x = 0; x < lx; ++x)
                        tmp[(static_cast<std::size_t>(z) * ly + y) * lx + x] = a[x][y][z + 1];
            std::fwrite(tmp.data(), sizeof(double), tmp.size(), out);
        }
        if (!dists) return;
        for (Vector4D* d : {&ff, &gg, &hh})
            for (int q = 0; q < D3Q15::Q; ++q) {
                for (int z = 0; z < lz; ++z)
                    for (int y = 0; y < ly; ++y)
                        for (int x = 0; x < lx; ++x)
                            tmp[(static_cast<std::size_t>(z) * ly + y) * lx + x] = (*d)[q][x][y][z];
                std::fwrite(tmp.data(), sizeof(double), tmp.size(), out);
            }
    }

    int nx() const { return lx; }
    int ny() const { return ly; }
    int nz() const { return lz; }
};

} // namespace

int main(int argc, char** argv)
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s PARAMS INIT|- STEPS OUT [DUMP_DISTS]\n", argv[0]);
        return 2;
    }
    const std::string params = argv[1];
    const std::string init = argv[2];
    const int steps = std::atoi(argv[3]);
    const std::string out_path = argv[4];
    const bool dists = argc > 5 && std::atoi(argv[5]) != 0;
    const std::string output_dir = argc > 6 ? argv[6] : "";

    std::ofstream devnull("/dev/null");
    std::streambuf* saved = std::cout.rdbuf(devnull.rdbuf());
    try {
        Probe probe(params);
        const std::size_t n = static_cast<std::size_t>(probe.nx()) * probe.ny() * probe.nz();
        if (init != "-") {
            std::vector<double> buf(6 * n);
            std::FILE* in = std::fopen(init.c_str(), "rb");
            if (!in || std::fread(buf.data(), sizeof(double), buf.size(), in) != buf.size()) {
                std::fprintf(stderr, "cannot read %s\n", init.c_str());
                return 2;
            }
            std::fclose(in);
            probe.load(buf);
        }
        std::vector<std::uint64_t> iters(steps);
        std::vector<double> step_ms(steps);
        for (int s = 0; s < steps; ++s) {
            const auto t0 = std::chrono::steady_clock::now();
            probe.performTimeStep();
            const auto t1 = std::chrono::steady_clock::now();
            step_ms[s] = std::chrono::duration<double, std::milli>(t1 - t0).count();
            iters[s] = probe.getLastPressureSolveDiagnostics().iteration_count;
        }
        if (!output_dir.empty()) {
            Inamuro::OutputConfig oc;
            oc.enable_tecplot = true;
            oc.enable_debug_field_csv = true;
            oc.output_dir = output_dir;
            probe.setOutputConfig(oc);
            probe.writeResults(steps);
        }
        std::cout.rdbuf(saved);
        std::FILE* out = std::fopen(out_path.c_str(), "wb");
        if (!out) {
            std::fprintf(stderr, "cannot write %s\n", out_path.c_str());
            return 2;
        }
        const std::int32_t hdr[6] = {0x4c424d31, probe.nx(), probe.ny(), probe.nz(), steps, dists ? 1 : 0};
        std::fwrite(hdr, sizeof(hdr), 1, out);
        std::fwrite(iters.data(), sizeof(std::uint64_t), iters.size(), out);
        std::fwrite(step_ms.data(), sizeof(double), step_ms.size(), out);
        probe.dump(out, dists);
        std::fclose(out);
    } catch (const std::exception& e) {
        std::cout.rdbuf(saved);
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
