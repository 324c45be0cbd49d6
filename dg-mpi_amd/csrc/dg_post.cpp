// dg_post: counterpart of the reference post-processor (src/post/main.cpp:
// 203-275): reads the case and a solution file written by dg_run and prints
// the volume-normalised L2 error of each conserved variable.
//
//   dg_post [-input case.toml] [-data data.dgs]
#include <cfloat>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "dg_amd.h"

static const char* opt(int argc, char** argv, const char* key, const char* dflt) {
    for (int i = 1; i + 1 < argc; ++i)
        if (!std::strcmp(argv[i], key)) return argv[i + 1];
    return dflt;
}

int main(int argc, char** argv) {
    const std::string in = opt(argc, argv, "-input", "input.toml");
    const std::string data = opt(argc, argv, "-data", "data.dgs");
    std::ifstream f(in);
    if (!f) { std::fprintf(stderr, "cannot open %s\n", in.c_str()); return 2; }
    std::stringstream ss;
    ss << f.rdbuf();
    dg_solver* s = nullptr;
    if (dg_create(ss.str().c_str(), &s) != DG_OK) {
        std::fprintf(stderr, "dg_create: %s\n", dg_last_error(nullptr));
        return 1;
    }
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
    if (!d || std::memcmp(magic, "DGAMDSOL", 8) || E != info.num_elems || nb != info.num_basis || ns != 5) {
        std::fprintf(stderr, "%s does not match the case\n", data.c_str());
        dg_destroy(s);
        return 1;
    }
    std::vector<double> U((size_t)E * nb * ns);
    d.read((char*)U.data(), U.size() * sizeof(double));
    dg_set_state(s, U.data(), layout ? DG_LAYOUT_RIGHT : DG_LAYOUT_LEFT);
    dg_set_time(s, time);
    std::printf("solver time = %f\n", time);
    double err[5];
    if (dg_l2_error(s, time, err) != DG_OK) {
        std::fprintf(stderr, "error: %s\n", dg_last_error(s));
        dg_destroy(s);
        return 1;
    }
    for (int i = 0; i < 5; ++i) std::printf("norm_total_error(%d)=%.*e\n", i, DBL_DECIMAL_DIG, err[i]);
    dg_destroy(s);
    return 0;
}
