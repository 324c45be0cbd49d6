// The solver handle behind the C ABI (shared by the host files of the
// library).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "dg_config.hpp"
#include "dg_kernels.hpp"

struct dg_solver {
    dg::CaseConfig cfg;
    dg::KTables tab;
    dg::KGeom geo;
    dg::KIC ic;
    int P = 0;
    int ns = 5;                        // state variables: dim + 2
    size_t ncoef = 0;                  // E * nb * ns
    double dt = 0.0;                   // case time step (0 until known)
    double time = 0.0;
    int device = 0;
    hipStream_t stream = nullptr;      // stream the handle launches on
    hipStream_t own_stream = nullptr;  // stream created by dg_create
    double* U0 = nullptr;              // solution (Solver::Uc)
    double* Ua = nullptr;              // stage states
    double* Ub = nullptr;
    double* acc = nullptr;             // RK accumulator
    float* cmp = nullptr;              // Stepper.compensated: rounding error of U0 (FP32)
    double* work = nullptr;            // per-element reduction output
    double* X = nullptr;               // mapped box: element node coordinates [E][8][3]
    double* imm = nullptr;             // mapped box: inverse mass matrices [E][nb][nb]
    unsigned* frec = nullptr;          // general path: face records [E][6][2]
    double* nev = nullptr;             // general path: normal evaluation points
    dg::KTables* dtab = nullptr;       // device copy of tab
    // learned-closure stepping (dg_closure_step): feature rows [E][70],
    // corrections [E][40] and the fallback counter, allocated on first use
    float* cl_X = nullptr;
    float* cl_C = nullptr;
    unsigned long long* cl_cnt = nullptr;
    std::vector<double> mesh_xyz;      // given mesh: node coordinates [n][dim]
    std::vector<long> mesh_e2n;        // given mesh: element nodes [E][8] (2D: [E][4])
    std::string err;
};
