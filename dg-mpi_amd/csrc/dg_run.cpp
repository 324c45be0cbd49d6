// dg_run: counterpart of the reference driver (src/exec/main.cpp:26-119).
// Reads a TOML case (-input, default input.toml), projects the initial
// condition, runs Solver::solve and writes the final coefficients.
//
//   dg_run [-input case.toml] [-output data.h5]
//
// Output: the reference's data.h5 layout (src/io/writer.cpp:10-98) when the
// tools are built with HDF5, and for any output name ending in ".dgs" (or
// without HDF5) a raw file: 8-byte magic "DGAMDSOL", int64 E, int32 nb,
// int32 ns, double time, int32 stored_layout (1 = row major, the
// reference's "Stored Layout" flag), then E*nb*ns doubles.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "dg_amd.h"
#ifdef DG_WITH_HDF5
#include "dg_amd_h5.h"
#endif

static const char* opt(int argc, char** argv, const char* key, const char* dflt) {
    for (int i = 1; i + 1 < argc; ++i)
        if (!std::strcmp(argv[i], key)) return argv[i + 1];
    return dflt;
}

static bool ends_with(const std::string& s, const char* suf) {
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

static int write_raw(dg_solver* s, const dg_info& info, const std::string& out) {
    std::vector<double> U((size_t)info.num_elems * info.num_basis * info.num_state_vars);
    if (dg_get_state(s, U.data(), DG_LAYOUT_RIGHT) != DG_OK) {
        std::fprintf(stderr, "error: %s\n", dg_last_error(s));
        return 1;
    }
    std::ofstream o(out, std::ios::binary);
    const char magic[8] = {'D', 'G', 'A', 'M', 'D', 'S', 'O', 'L'};
    const int64_t E = info.num_elems;
    const int32_t nb = info.num_basis, ns = info.num_state_vars, layout = 1;
    o.write(magic, 8);
    o.write((const char*)&E, 8);
    o.write((const char*)&nb, 4);
    o.write((const char*)&ns, 4);
    o.write((const char*)&info.time, 8);
    o.write((const char*)&layout, 4);
    o.write((const char*)U.data(), U.size() * sizeof(double));
    if (!o) {
        std::fprintf(stderr, "cannot write %s\n", out.c_str());
        return 1;
    }
    return 0;
}

int main(int argc, char** argv) {
    const std::string in = opt(argc, argv, "-input", "input.toml");
#ifdef DG_WITH_HDF5
    const std::string out = opt(argc, argv, "-output", "data.h5");
#else
    const std::string out = opt(argc, argv, "-output", "data.dgs");
#endif
    std::ifstream f(in);
    if (!f) { std::fprintf(stderr, "cannot open %s\n", in.c_str()); return 2; }
    std::stringstream ss;
    ss << f.rdbuf();
    dg_solver* s = nullptr;
#ifdef DG_WITH_HDF5
    // Mesh.file cases read their mesh file; PeriodicBox cases go to dg_create
    if (dg_create_from_mesh_file(ss.str().c_str(), nullptr, &s) != DG_OK) {
        std::fprintf(stderr, "dg_create: %s\n", dg_h5_last_error());
        return 1;
    }
#else
    if (dg_create(ss.str().c_str(), &s) != DG_OK) {
        std::fprintf(stderr, "dg_create: %s\n", dg_last_error(nullptr));
        return 1;
    }
#endif
    int rc = dg_init_state(s);
    dg_info info;
    dg_get_info(s, &info);
    std::printf("Number of elements = %ld, p = %d, dt = %.17g\n", info.num_elems, info.order, info.timestep);
    std::printf("Number of iterations = %d\n", info.num_iterations);
    auto t0 = std::chrono::steady_clock::now();
    if (rc == DG_OK) rc = dg_solve(s);
    auto t1 = std::chrono::steady_clock::now();
    if (rc != DG_OK) { std::fprintf(stderr, "error: %s\n", dg_last_error(s)); dg_destroy(s); return 1; }
    const double sec = std::chrono::duration<double>(t1 - t0).count();
    dg_get_info(s, &info);
    std::printf("Time to Solution: %.6f s (%.3f ms per step), time = %.17g\n", sec,
                info.num_iterations ? 1e3 * sec / info.num_iterations : 0.0, info.time);
    int wrc = 0;
#ifdef DG_WITH_HDF5
    if (!ends_with(out, ".dgs")) {
        if (dg_write_solution(s, out.c_str()) != DG_OK) {
            std::fprintf(stderr, "error: %s\n", dg_h5_last_error());
            wrc = 1;
        }
    } else
#endif
        wrc = write_raw(s, info, out);
    if (wrc == 0) std::printf("Solution written to %s\n", out.c_str());
    dg_destroy(s);
    return wrc;
}
