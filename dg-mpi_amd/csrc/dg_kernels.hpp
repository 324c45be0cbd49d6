// Host-visible declarations of the HIP kernels of the DG residual + RK path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dg {

constexpr int PMAX = 5;            // highest order with a compiled kernel
constexpr int NVAR = 5;            // rho, rho u, rho v, rho w, rho E
// projection matrix of the bf16 form: (p+1)^2 rows and 3 (p+2)^2 columns,
// padded to 16
constexpr int KPROJ_ROWS = ((PMAX + 1) * (PMAX + 1) + 15) / 16 * 16;
constexpr int KPROJ_COLS = (3 * (PMAX + 2) * (PMAX + 2) + 15) / 16 * 16;

// Reference-element operators of one case (1D factors of the tensor bases).
// Kept in device memory (one copy per solver) and passed to the kernels by
// pointer: uniform reads go through the scalar cache, and lane-dependent
// reads are plain cached loads instead of a per-lane copy of the struct.
struct KTables {
    double B[PMAX + 2][PMAX + 1];   // phi_i(x_a), a: (p+2)-point GL rule
    double D[PMAX + 2][PMAX + 1];   // phi_i'(x_a)
    double phim[PMAX + 1];          // phi_i(-1)
    double phip[PMAX + 1];          // phi_i(+1)
    double w[PMAX + 2];             // GL weights of the (p+2)-point rule
    double minv[PMAX + 1][PMAX + 1];// inverse of the 1D mass matrix
    double BQ[2 * PMAX + 2][PMAX + 1];  // phi_i at the (2p+2)-point rule
    double xq[2 * PMAX + 2];        // (2p+2)-point GL nodes
    double wq[2 * PMAX + 2];        // (2p+2)-point GL weights
    double xv[PMAX + 2];            // (p+2)-point GL nodes (volume and face points)
    double BF[PMAX + 1][PMAX + 1];  // 2D faces: phi_i at the (p+1)-point segment rule
    double wf[PMAX + 1];            // 2D faces: weights of that rule
    int minv_diag;                  // 1: minv is diagonal (Legendre)
    double one[PMAX + 1];           // the constant 1 in the 1D basis (Legendre 1, 0, ..; Lagrange all 1)
    int one_kind;                   // that vector: 0 = (1, 0, ..), 1 = all ones, 2 = otherwise
    // single-precision copies (round to nearest) for the FP32 stage kernel
    float B32[PMAX + 2][PMAX + 1];
    float D32[PMAX + 2][PMAX + 1];
    float phim32[PMAX + 1];
    float phip32[PMAX + 1];
    float w32[PMAX + 2];
    // the 1D factors with the (p+2)-point weight folded in (w_a B[a][i],
    // w_a D[a][i]): the weak-gradient projection and the face lift of the
    // box stage contract with these, so no weight products are formed per
    // quadrature point
    double Bw[PMAX + 2][PMAX + 1];
    double Dw[PMAX + 2][PMAX + 1];
    float Bw32[PMAX + 2][PMAX + 1];
    float Dw32[PMAX + 2][PMAX + 1];
    // The x-y part of the weak-gradient projection as one matrix for the
    // bf16 matrix-core form: Kproj[part][m][kk], row m = j (p+1) + i, column
    // kk = dir (p+2)^2 + b (p+2) + a, the value Dw[a][i] Bw[b][j] (dir 0, the
    // x flux), Bw[a][i] Dw[b][j] (dir 1) or Bw[a][i] Bw[b][j] (dir 2),
    // rounded to bf16 (part 0) with the rounding error in bf16 (part 1);
    // rows and columns padded with zeros to whole 16 x 16 tiles
    uint16_t Kproj[2][KPROJ_ROWS][KPROJ_COLS];
};

// Structured periodic box and physics constants.  dim = 2: nx x ny box of
// quadrilaterals (ne[2] = 1), 4 state variables.
struct KGeom {
    int dim;                        // 2 or 3
    int ne[3];                      // elements per direction
    unsigned E;                     // number of elements
    double lo[3], h[3];             // box corner, element widths
    double period[3];               // box lengths
    double gvol[3];                 // (d xi_d / d x_d) |J|: contravariant flux scale
    double farea[3];                // |N| of faces normal to d: h_t1 h_t2 / 4
    double jdet, jinv;              // |J| = hx hy hz / 8 and its inverse
    double gamma, R;
    int xcd_remap;                  // 1: map workgroups to elements XCD-contiguously
    int tile[3];                    // log2 brick edges of the stage-kernel traversal
    // general path (warped box, meshes from files): trilinear elements with
    // face records instead of box neighbours
    int fp32;                       // 1: the stage kernel computes in FP32 (box only)
    int proj;                       // FP32 form: x-y projection on bf16 matrix cores,
                                    // 0 none, 1 one bf16 term, 3 hi/lo split (3 terms)
    // FP32 form: the reference state subtracted from the coefficients (in
    // FP64) before they are rounded to FP32 -- conserved variables, zero for
    // none -- with its velocity (uref[1 + d] = uref[0] vref[d] exactly) and
    // its pressure (gamma - 1)(uref[4] - uref[1..3].vref / 2)
    double uref[5], vref[3], pref;
    int mapped;                     // 1: the fields below are used
    int exact_normals;              // 1: face normals at the flux points
    const double* X;                // [E][8][3] node coordinates, node 4k+2j+i
    const double* imm;              // [E][nb][nb] inverse mass matrices (symmetric)
    const unsigned* frec;           // [E][6][2] face records, slot d*2 + side
    const double* nev;              // [48][(p+2)^2][2] normal evaluation points
};

// Initial condition / exact solution selector and data.
struct KIC {
    int ic;                         // enum dg::IC
    double data[24];
};

// Epilogue of the fused stage kernel.
enum StageMode {
    MODE_RESIDUAL = 0,   // R = volume - face (Solver::get_residual)
    MODE_RK_FIRST = 1,   // Uout = U0 + alpha dU; acc = dU          (RK4 stage 1)
    MODE_RK_MID = 2,     // Uout = U0 + alpha dU; acc += 2 dU       (RK4 stages 2, 3)
    MODE_RK_LAST = 3,    // U0 += (acc + dU) / 6                    (RK4 stage 4)
    MODE_FE = 4          // Uout = U0 + dU                          (forward Euler)
};
// With a compensation register (StageArgs::cmp) the last RK4 stage and
// forward Euler form Uout = U0 + (inc + cmp) by TwoSum and keep its
// rounding error in cmp, so U0 + cmp carries the solution to about twice
// the FP64 precision over many small steps.  The error is below half an ulp
// of U0, so FP32 holds it to 24 more bits (half the traffic of FP64).

struct StageArgs {
    const double* Ust;   // state the residual is evaluated at (neighbours read it)
    const double* U0;    // base state of the RK update
    double* Uout;        // next stage state
    double* acc;         // RK accumulator
    double* R;           // residual output (MODE_RESIDUAL)
    float* cmp;          // compensated update: rounding error of U0 (null: plain)
    double dt;           // dU = dt M^-1 R
    double alpha;        // stage coefficient
    int mode;
};

// One case of a batched stage launch (device memory): its RK registers as
// byte offsets from the base pointer the launch receives (pointers derived
// from a kernel argument are known to the compiler to be global memory,
// unlike pointers read from memory), its time step, its geometry and the
// first position of the launch that works on it (positions first_block ..
// first_block + E - 1).
struct BatchCase {
    long long off_U0, off_Ua, off_Ub, off_acc;
    long long off_cmp;       // compensation register, -1: plain update
    double dt;
    KGeom g;
    unsigned first_block;
};

// Launchers; P in 1..PMAX, flux 0 LLF / 1 HLLC / 2 Roe.  Return the launch
// status (hipGetLastError after the launch).
// `t` points to the case's KTables in device memory.
hipError_t launch_stage(int P, int flux, const StageArgs& a, const KTables* t, const KGeom& g,
                        hipStream_t st);
// RK stage of an affine FP64 p = 1 box with dU = dt (M^-1 R + C): C the
// learned closure's FP32 correction, laid out like the state.
hipError_t launch_stage_corrected(int P, int flux, const StageArgs& a, const float* C, const KTables* t,
                                  const KGeom& g, hipStream_t st);
// Batched stage over the cases of one (order, basis, flux, precision)
// group (tables t shared): nblocks = total elements of the group,
// blk_case[q] = case of traversal position q, stage 0..3 of RK4; variant
// 0 or 1 picks one of two team sizes per order (batch_team).
hipError_t launch_stage_batch(int P, int flux, int fp32, int proj, int variant, const BatchCase* cases,
                              const unsigned* blk_case, unsigned nblocks, int stage, const KTables* t, hipStream_t st);
// threads per element of the batched launch's variant (0 or 1) at order P
int batch_team(int P, int variant);
// zero `bytes` (a multiple of 4) of device memory with a kernel on st
hipError_t launch_zero(void* p, size_t bytes, hipStream_t st);
hipError_t launch_inv_mass(int P, double dt, const double* R, double* dU, const KTables* t,
                           const KGeom& g, hipStream_t st);
hipError_t launch_init_state(int P, double time, double* U, const KTables* t, const KGeom& g,
                             const KIC& ic, hipStream_t st);
// mode 0: per-element squared L2 error vs exact solution at time t and the
//         element volume (sum of w |J|) -> out[E][6]
// mode 1: per-element integrals (rho, rho u, rho v, rho w, rho E, KE) -> out[E][6]
//         (2D: rho, rho u, rho v, rho E, KE, 0)
hipError_t launch_eval(int P, int mode, double time, const double* U, double* out,
                       const KTables* t, const KGeom& g, const KIC& ic, hipStream_t st);
// mapped box: per-element mass matrices on the (2p+2)^3 rule, inverted in
// place -> G.imm (written through `imm`); jmin[E] = smallest |J| at the
// points (<= 0: invalid element)
hipError_t launch_mass(int P, double* imm, double* jmin, const KTables* t, const KGeom& g, hipStream_t st);
// per-element max(|u| + c) at the volume points -> out[E]
hipError_t launch_wavespeed(int P, const double* U, double* out, const KTables* t, const KGeom& g,
                            hipStream_t st);

}  // namespace dg
