// dg_post: counterpart of the reference post-processor (src/post/main.cpp:
// 203-253): reads the case and the solution file written by dg_run and
// prints the volume-normalised L2 error of each conserved variable.
//
//   dg_post [-input case.toml] [-data data.h5]
//
// data.h5 files are read as Solver::read_in_coefficients reads them
// (src/solver/base.cpp:233-279); names ending in ".dgs" are dg_run's raw
// format.  The exact solution is evaluated at the file's "Solver Final
// Time"; the reference leaves its solver time at 0 here, which only matters
// for time-dependent exact solutions.
#include <cfloat>
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

static int read_raw(dg_solver* s, const std::string& data) {
    std::ifstream d(data, std::ios::binary);
    char magic[8];
    int64_t E = 0;
    int32_t nb = 0, ns = 0, layout = 1;
    double time = 0.0;
    d.read(magic, 8);
    d.read((char*)&E, 8);
    d.read((char*)&nb, 4);
    d.read((char*)&ns, 4);
    d.read((char*)&time, 8);
    d.read((char*)&layout, 4);
    dg_info info;
    dg_get_info(s, &info);
    if (!d || std::memcmp(magic, "DGAMDSOL", 8) || E != info.num_elems || nb != info.num_basis ||
        ns != info.num_state_vars) {
        std::fprintf(stderr, "%s does not match the case\n", data.c_str());
        return 1;
    }
    std::vector<double> U((size_t)E * nb * ns);
    d.read((char*)U.data(), U.size() * sizeof(double));
    if (!d) {
        std::fprintf(stderr, "%s is truncated\n", data.c_str());
        return 1;
    }
    if (dg_set_state(s, U.data(), layout ? DG_LAYOUT_RIGHT : DG_LAYOUT_LEFT) != DG_OK ||
        dg_set_time(s, time) != DG_OK) {
        std::fprintf(stderr, "error: %s\n", dg_last_error(s));
        return 1;
    }
    return 0;
}

int main(int argc, char** argv) {
    const std::string in = opt(argc, argv, "-input", "input.toml");
#ifdef DG_WITH_HDF5
    const std::string data = opt(argc, argv, "-data", "data.h5");
#else
    const std::string data = opt(argc, argv, "-data", "data.dgs");
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
    int rc = 0;
#ifdef DG_WITH_HDF5
    if (!ends_with(data, ".dgs")) {
        if (dg_read_solution(s, data.c_str(), 1) != DG_OK) {
            std::fprintf(stderr, "error: %s\n", dg_h5_last_error());
            rc = 1;
        }
    } else
#endif
        rc = read_raw(s, data);
    if (rc) {
        dg_destroy(s);
        return rc;
    }
    dg_info info;
    dg_get_info(s, &info);
    const double time = info.time;
    std::printf("time=%g\n", time);
    std::printf("solver time = %f\n", time);
    double err[5] = {0., 0., 0., 0., 0.};
    if (dg_l2_error(s, time, err) != DG_OK) {
        std::fprintf(stderr, "error: %s\n", dg_last_error(s));
        dg_destroy(s);
        return 1;
    }
    for (int i = 0; i < info.num_state_vars; ++i)
        std::printf("norm_total_error(%d)=%.*e\n", i, DBL_DECIMAL_DIG, err[i]);
    dg_destroy(s);
    return 0;
}
