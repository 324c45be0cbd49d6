// HIP kernels (gfx950) of the DG residual + RK path: 3D Euler on periodic
// hexahedral boxes, FP64.
//
// Design (a team of 2-4 wavefronts per element, all exchanges in LDS):
//   * The whole RK stage is one launch: face traces of the element and of
//     its six neighbours, the Riemann fluxes, the lift, the sum-factorised
//     volume integral, the inverse mass matrix and the RK register update.
//     The state is read once from HBM per stage (neighbour blocks come back
//     through L2/MALL); only the RK registers add HBM traffic.
//   * All global loads of an element (its block, the neighbour lines that
//     feed the neighbour traces, the RK registers) are issued together at
//     the start of the workgroup.
//   * No atomics: every element recomputes the trace of each neighbour on
//     the shared face with exactly the arithmetic the neighbour uses for its
//     own trace, and evaluates the Riemann flux with the same (UL, UR, N), so
//     both sides see bit-identical fluxes (exact discrete conservation).
//   * Tensor-product sum factorisation: each contraction step is a set of
//     (line, variable) tasks; a thread keeps its line in registers, basis
//     matrices are read from a device-resident table through the scalar
//     cache, per-thread-indexed table entries from an LDS copy.
//
// Reference semantics restated here (paths relative to the reference root):
//   residual            src/solver/base.cpp:401-561
//   volume flux         src/solver/flux_functors_impl.h:46-115, tools.cpp:6-35
//   face flux / lift    src/solver/flux_functors_impl.h:168-264, base.cpp:287-374,
//                       src/solver/tools.cpp:37-63
//   LLF / HLLC          src/physics/base/functions.cpp:33-89,
//                       src/physics/euler/functions.cpp:298-399
//   inverse mass, RK    src/solver/tools.cpp:124-139,
//                       src/numerics/timestepping/stepper.cpp:23-142
//   IC projection       src/solver/base.cpp:113-223, src/solver/tools.cpp:66-122
//   L2 error            src/post/main.cpp:30-201
#include <hip/hip_runtime.h>

#include "dg_kernels.hpp"

namespace dg {
namespace {

constexpr int V = NVAR;
constexpr int WAVE = 64;
// Threads that cooperate on one element in the fused stage kernel, by
// order (measured, 64^3 Legendre Roe RK4, ms per stage for 64 / 128 / 256 /
// 384 / 512 threads: p=1 0.70 / 0.84 / 1.18, p=2 1.51 / 1.30 / 1.58,
// p=3 3.47 / 2.48 / 2.12 (320: 3.33), p=4 10.0 / 6.21 / 4.54 / 5.41 / 5.31,
// p=5 20.7 / 11.4 / 8.00 / 12.2 / 6.79).  DG_STAGE_TEAM overrides.
template <int P>
constexpr int stage_team() {
#ifdef DG_STAGE_TEAM
    return DG_STAGE_TEAM;
#else
    return P == 1 ? 64 : (P == 2 ? 128 : (P <= 4 ? 256 : 512));
#endif
}

template <int P>
struct Dims {
    static constexpr int N = P + 1, M = P + 2;
    static constexpr int NN = N * N, MM = M * M, NNN = N * N * N, MMM = M * M * M;
    static constexpr int Q1 = 2 * P + 2;
};

// The few table entries that the stage kernel indexes with per-thread
// indices, copied to LDS once per workgroup.  Reading them from global
// memory would issue vector loads, and waiting for those would also wait
// for every prefetch load issued before them.
// R: the compute type of the stage kernel (the M^-1 diagonal stays FP64).
template <class R>
struct LaneTabT {
    R w[PMAX + 2];           // GL weights of the (p+2)-point rule
    R phim[PMAX + 1];        // phi_i(-1)
    R phip[PMAX + 1];        // phi_i(+1)
    double mdiag[PMAX + 1];  // diagonal of the 1D inverse mass matrix
    R xv[PMAX + 2];          // GL nodes of the (p+2)-point rule
    double uref[sizeof(R) == 4 ? V : 1];         // FP32 form: the reference state
};
using LaneTab = LaneTabT<double>;

// Reference-element tables in the compute type R of the stage kernel
// (FP32: the rounded copies the host stores beside the FP64 tables).
template <class R>
__device__ __forceinline__ R tB(const KTables& T, int a, int i) {
    if constexpr (sizeof(R) == 4) return T.B32[a][i];
    else return T.B[a][i];
}
template <class R>
__device__ __forceinline__ R tD(const KTables& T, int a, int i) {
    if constexpr (sizeof(R) == 4) return T.D32[a][i];
    else return T.D[a][i];
}
template <class R>
__device__ __forceinline__ R tw(const KTables& T, int a) {
    if constexpr (sizeof(R) == 4) return T.w32[a];
    else return T.w[a];
}
template <class R>
__device__ __forceinline__ R tphim(const KTables& T, int i) {
    if constexpr (sizeof(R) == 4) return T.phim32[i];
    else return T.phim[i];
}
template <class R>
__device__ __forceinline__ R tphip(const KTables& T, int i) {
    if constexpr (sizeof(R) == 4) return T.phip32[i];
    else return T.phip[i];
}

template <class R>
__device__ __forceinline__ R tBw(const KTables& T, int a, int i) {
    if constexpr (sizeof(R) == 4) return T.Bw32[a][i];
    else return T.Bw[a][i];
}
template <class R>
__device__ __forceinline__ R tDw(const KTables& T, int a, int i) {
    if constexpr (sizeof(R) == 4) return T.Dw32[a][i];
    else return T.Dw[a][i];
}

template <int T, int TS = WAVE, class F>
__device__ __forceinline__ void team_for(int lane, F&& f) {
#pragma unroll
    for (int t0 = 0; t0 < T; t0 += TS) {
        const int t = t0 + lane;
        if (t < T) f(t);
    }
}

__device__ __forceinline__ void sync() { __syncthreads(); }

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// ---------------------------------------------------------------------------
// Pointwise physics
// ---------------------------------------------------------------------------

// src/physics/euler/functions.cpp:10-23
__device__ __forceinline__ double pressure(double gam, const double* U) {
    const double rKE = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
    return (gam - 1.0) * (U[4] - 0.5 * rKE);
}

// Pressure with 1/rho supplied (one reciprocal per state, reused below).
template <class R>
__device__ __forceinline__ R pressure_r(R gam, const R* U, R r1) {
    return (gam - R(1)) * (U[4] - R(0.5) * ((U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * r1));
}

// Reciprocal and square root of the flux computations: hardware estimates
// refined by Newton-Raphson (reciprocal: two steps r <- r + r (1 - x r))
// and by Goldschmidt's iteration with two residual corrections (square
// root, from the reciprocal square-root estimate).  For the positive normal
// arguments they receive here they are within an ulp or two of the
// correctly rounded results, at about half the instructions of the exact
// IEEE expansions.  Both sides of a face evaluate the same code, so the
// flux stays bit-identical between them.
__device__ __forceinline__ double frcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    r = fma(fma(-x, r, 1.0), r, r);
    return r;
}

__device__ __forceinline__ double fsqrt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    // zero and negative arguments (not physical here) as sqrt has them
    return x > 0.0 ? g : (x == 0.0 ? 0.0 : __builtin_nan(""));
}

// Single precision: the hardware reciprocal refined by one Newton-Raphson
// step, and the hardware square root (both within about an ulp).
__device__ __forceinline__ float frcp(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
}

__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Reciprocal square root of a positive normal argument: the hardware
// estimate refined by two Newton-Raphson steps y <- y + y (1 - x y^2) / 2
// (FP32: the hardware estimate, about an ulp).
__device__ __forceinline__ double frsq(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * y;
    y = fma(y, fma(-x * y, h, 0.5), y);
    h = 0.5 * y;
    y = fma(y, fma(-x * y, h, 0.5), y);
    return y;
}

__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// src/physics/euler/functions.cpp:25-40
__device__ __forceinline__ double max_wave_speed(double gam, const double* U) {
    const double rho1 = 1. / U[0];
    const double sqrtrKE = sqrt(U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * rho1;
    return sqrtrKE + sqrt(gam * pressure_r(gam, U, rho1) * rho1);
}

// Euler flux along axis D (one block of src/physics/euler/functions.cpp:64-91)
// with 1/rho and p supplied.
template <int D, class R>
__device__ __forceinline__ void flux_axis(const R* U, R r1, R P, R* F) {
    const R un = U[1 + D];
    F[0] = un;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const R t = un * U[1 + i] * r1;
        F[1 + i] = (i == D) ? t + P : t;
    }
    F[4] = (U[4] + P) * un * r1;
}

// Local Lax-Friedrichs with N = nmag e_D (src/physics/base/functions.cpp:33-89).
template <int D, class R>
__device__ __forceinline__ void riemann_llf(R gam, const R* UL, const R* UR, R nmag, R* F) {
    const R r1L = frcp(UL[0]), r1R = frcp(UR[0]);
    const R PL = pressure_r(gam, UL, r1L), PR = pressure_r(gam, UR, r1R);
    R FL[V], FR[V];
    flux_axis<D>(UL, r1L, PL, FL);
    flux_axis<D>(UR, r1R, PR, FR);
    R a = fsqrt(UL[1] * UL[1] + UL[2] * UL[2] + UL[3] * UL[3]) * r1L + fsqrt(gam * PL * r1L);
    const R aR = fsqrt(UR[1] * UR[1] + UR[2] * UR[2] + UR[3] * UR[3]) * r1R + fsqrt(gam * PR * r1R);
    if (aR > a) a = aR;
#pragma unroll
    for (int s = 0; s < V; ++s) F[s] = nmag * (R(0.5) * (FL[s] + FR[s]) - R(0.5) * a * (UR[s] - UL[s]));
}

// HLLC with PVRS wave speeds, N = nmag e_D (src/physics/euler/functions.cpp:298-399).
template <int D, class R>
__device__ __forceinline__ void riemann_hllc(R gam, const R* UL, const R* UR, R nmag, R* F) {
    const R djac = nmag;
    const R rhoL = UL[0], rhoL1 = frcp(rhoL);
    const R unL = UL[1 + D] * rhoL1;
    const R PL = pressure_r(gam, UL, rhoL1);
    const R aL = fsqrt(gam * PL * rhoL1);
    const R rhoR = UR[0], rhoR1 = frcp(rhoR);
    const R unR = UR[1 + D] * rhoR1;
    const R PR = pressure_r(gam, UR, rhoR1);
    const R aR = fsqrt(gam * PR * rhoR1);
    const R rAvg = R(0.5) * (rhoL + rhoR);
    const R aAvg = R(0.5) * (aL + aR);
    const R pStar = fmax(R(0.), R(0.5) * (PL + PR - (unR - unL) * rAvg * aAvg));
    const R gq = (gam + R(1.)) / (R(2.) * gam);
    R qL = R(1.0);
    const R xL = pStar * frcp(PL);
    if (xL > R(1.)) qL = fsqrt(R(1.) + gq * (xL - R(1.)));
    R qR = R(1.0);
    const R xR = pStar * frcp(PR);
    if (xR > R(1.)) qR = fsqrt(R(1.) + gq * (xR - R(1.)));
    const R sL = unL - aL * qL;
    const R sR = unR + aR * qR;
    const R sStar = (PR - PL + rhoL * unL * (sL - unL) - rhoR * unR * (sR - unR)) *
                         frcp(rhoL * (sL - unL) - rhoR * (sR - unR));
    if (sL >= R(0.)) {
        F[0] = djac * rhoL * unL;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const R t = UL[1 + i] * unL;
            F[1 + i] = djac * ((i == D) ? t + PL : t);
        }
        F[4] = djac * ((UL[4] + PL) * unL);
    } else if (sR <= R(0.)) {
        F[0] = djac * rhoR * unR;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const R t = UR[1 + i] * unR;
            F[1 + i] = djac * ((i == D) ? t + PR : t);
        }
        F[4] = djac * (UR[4] + PR) * unR;
    } else if ((sL < R(0.)) && (sStar >= R(0.))) {
        const R c = (sL - unL) * frcp(sL - sStar);
        const R f1 = sStar - unL;
        const R f2 = sStar + PL * frcp(rhoL * (sL - unL));
        F[0] = djac * (rhoL * unL + rhoL * sL * (c - R(1.)));
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const R t = UL[1 + i] * unL;
            const R u = UL[1 + i] * (c - R(1.));
            F[1 + i] = djac * ((i == D) ? (t + PL + sL * (u + rhoL * c * f1)) : (t + sL * u));
        }
        F[4] = djac * ((UL[4] + PL) * unL + sL * (UL[4] * (c - R(1.)) + rhoL * c * f1 * f2));
    } else if ((sStar < R(0.)) && (sR > R(0.))) {
        const R c = (sR - unR) * frcp(sR - sStar);
        const R f1 = sStar - unR;
        const R f2 = sStar + PR * frcp(rhoR * (sR - unR));
        F[0] = djac * (rhoR * unR + rhoR * sR * (c - R(1.)));
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const R t = UR[1 + i] * unR;
            const R u = UR[1 + i] * (c - R(1.));
            F[1 + i] = djac * ((i == D) ? (t + PR + sR * (u + rhoR * c * f1)) : (t + sR * u));
        }
        F[4] = djac * ((UR[4] + PR) * unR + sR * (UR[4] * (c - R(1.)) + rhoR * c * f1 * f2));
    } else {
        // the reference prints and asserts (functions.cpp:389-398); NaN
        // propagates to the host-side finiteness check instead
#pragma unroll
        for (int s = 0; s < V; ++s) F[s] = __builtin_nan("");
    }
}

// Roe-Pike flux with a Harten-Hyman-type entropy fix on the acoustic waves
// (no reference counterpart; the algorithm of the oracle's flux_roe,
// arranged around four reciprocals and five square roots).
template <class R>
__device__ __forceinline__ R entropy_fix(R lam, R lamL, R lamR) {
    const R delta = fmax(R(0), fmax(lam - lamL, lamR - lam));
    R a = fabs(lam);
    if (a < delta) a = R(0.5) * (lam * lam + delta * delta) * frcp(delta);   // rare: transonic expansion
    return a;
}

template <int D, class R>
__device__ __forceinline__ void riemann_roe(R gam, const R* UL, const R* UR, R nmag, R* F) {
    const R rL = UL[0], rR = UR[0];
    const R r1L = frcp(rL), r1R = frcp(rR);
    R uL[3], uR[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        uL[d] = UL[1 + d] * r1L;
        uR[d] = UR[1 + d] * r1R;
    }
    const R gm1 = gam - R(1.);
    const R pL = gm1 * (UL[4] - R(0.5) * (UL[1] * uL[0] + UL[2] * uL[1] + UL[3] * uL[2]));
    const R pR = gm1 * (UR[4] - R(0.5) * (UR[1] * uR[0] + UR[2] * uR[1] + UR[3] * uR[2]));
    const R HL = (UL[4] + pL) * r1L, HR = (UR[4] + pR) * r1R;
    // Roe weights sqrt(rL) / (sqrt(rL) + sqrt(rR)) and sqrt(rR) / (...) from
    // one square root k = sqrt(rR / rL); sqrt(rL rR) = rL k
    const R k = fsqrt(rR * r1L);
    const R wl = frcp(R(1.) + k), wr = k * wl;
    R ut[3];
#pragma unroll
    for (int d = 0; d < 3; d++) ut[d] = fma(wl, uL[d], wr * uR[d]);
    const R Ht = fma(wl, HL, wr * HR);
    const R q2 = ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2];
    const R c2 = gm1 * (Ht - R(0.5) * q2);
    // c = sqrt(c2) and 1 / c2 from one reciprocal square root
    const R yc = frsq(c2);
    const R ct = c2 * yc;
    const R ic2 = yc * yc;
    const R rt = rL * k;
    const R unt = ut[D], unL = uL[D], unR = uR[D];
    const R drho = rR - rL, dp = pR - pL, dun = unR - unL;
    const R a1 = (dp - rt * ct * dun) * (R(0.5) * ic2);
    const R a5 = (dp + rt * ct * dun) * (R(0.5) * ic2);
    const R a2 = drho - dp * ic2;
    const R cL = fsqrt(gam * pL * r1L), cR = fsqrt(gam * pR * r1R);
    const R l1 = entropy_fix(unt - ct, unL - cL, unR - cR);
    const R l2 = fabs(unt);
    const R l5 = entropy_fix(unt + ct, unL + cL, unR + cR);
    const R w1 = l1 * a1, w5 = l5 * a5, w2 = l2 * a2, ws = l2 * rt;
    // sum over the waves of w_k r_k, grouped by eigenvector component:
    // the acoustic pair enters as (w1 + w5) and (w5 - w1) c
    const R wsum = w1 + w5, wdc = (w5 - w1) * ct;
    R Dv[V];
    Dv[0] = wsum + w2;
    R udu = R(0.);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const R dut = (d == D) ? (uR[d] - uL[d]) - dun : (uR[d] - uL[d]);
        const R v = fma(Dv[0], ut[d], ws * dut);
        Dv[1 + d] = (d == D) ? v + wdc : v;
        udu = fma(ut[d], dut, udu);
    }
    Dv[4] = fma(wsum, Ht, fma(wdc, unt, fma(w2 * R(0.5), q2, ws * udu)));
    R FL[V], FR[V];
    flux_axis<D>(UL, r1L, pL, FL);
    flux_axis<D>(UR, r1R, pR, FR);
    const R hn = R(0.5) * nmag;
#pragma unroll
    for (int s = 0; s < V; ++s) F[s] = hn * ((FL[s] + FR[s]) - Dv[s]);
}

// ---------------------------------------------------------------------------
// FP32 form: the state as a deviation from a reference state
// ---------------------------------------------------------------------------
// The FP32 form subtracts a constant reference state Ur (conserved
// variables, chosen by the host: the domain mean) from the coefficients in
// FP64 before rounding them to FP32; traces and interpolations then carry
// U' = U - Ur, and every flux is evaluated as the difference F(Ur + U') -
// F(Ur) by the identities below, exact in U'.  The constant flux F(Ur) has
// no weak divergence (the volume term and the face terms of a constant
// flux cancel under the (p+2)-point rule), so the residual is unchanged,
// while FP32 rounding scales with the deviation U' instead of the state.
// Ur = 0 gives the plain fluxes.

// uref[s] of the reference state by selects on values (a dynamic index,
// or a select of the members' addresses, would copy the kernel-argument
// struct to scratch memory)
__device__ __forceinline__ double uref_of(const KGeom& G, int s) {
    const double u0 = G.uref[0], u1 = G.uref[1], u2 = G.uref[2], u3 = G.uref[3], u4 = G.uref[4];
    return s == 0 ? u0 : (s == 1 ? u1 : (s == 2 ? u2 : (s == 3 ? u3 : u4)));
}

template <class R>
struct SplitRef {
    R rho, v[3], m[3], E, p, H;    // m = rho v, p = (gamma - 1)(E - m.v / 2), H = E + p
    // momentum components cycled so that direction d comes first (the
    // face-normal frame of the x-normal Riemann solver)
    __device__ __forceinline__ SplitRef cycled(int d) const {
        // selects on values, not on the members' addresses
        const R v0 = v[0], v1 = v[1], v2 = v[2], m0 = m[0], m1 = m[1], m2 = m[2];
        SplitRef r;
        r.rho = rho;
        r.E = E;
        r.p = p;
        r.H = H;
        r.v[0] = d == 0 ? v0 : (d == 1 ? v1 : v2);
        r.v[1] = d == 0 ? v1 : (d == 1 ? v2 : v0);
        r.v[2] = d == 0 ? v2 : (d == 1 ? v0 : v1);
        r.m[0] = d == 0 ? m0 : (d == 1 ? m1 : m2);
        r.m[1] = d == 0 ? m1 : (d == 1 ? m2 : m0);
        r.m[2] = d == 0 ? m2 : (d == 1 ? m0 : m1);
        return r;
    }
};

template <class R>
__device__ __forceinline__ SplitRef<R> split_ref(const KGeom& G) {
    SplitRef<R> r;
    r.rho = R(G.uref[0]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        r.v[d] = R(G.vref[d]);
        r.m[d] = R(G.uref[1 + d]);
    }
    r.E = R(G.uref[4]);
    r.p = R(G.pref);
    r.H = R(G.uref[4] + G.pref);
    return r;
}

// Point quantities of Ur + U' from the deviation U'.
template <class R>
struct SplitPt {
    R rho, r1;        // density, 1 / density
    R v[3], dv[3];    // velocity and its deviation v - vr = (m' - vr rho') / rho
    R dp, p;          // pressure deviation and pressure
};

template <class R>
__device__ __forceinline__ void split_point(R gm1, const SplitRef<R>& rf, const R* U, SplitPt<R>& q) {
    q.rho = rf.rho + U[0];
    q.r1 = frcp(q.rho);
    R ke = R(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        q.dv[d] = fma(-rf.v[d], U[0], U[1 + d]) * q.r1;
        q.v[d] = rf.v[d] + q.dv[d];
        // m.v - mr.vr = m'.v + mr.v'
        ke = fma(U[1 + d], q.v[d], fma(rf.m[d], q.dv[d], ke));
    }
    q.dp = gm1 * fma(R(-0.5), ke, U[4]);
    q.p = rf.p + q.dp;
}

// F_D(Ur + U') - F_D(Ur): m'_D; m'_i v_D + mr_i v'_D + p' delta_iD;
// (E' + p') v_D + Hr v'_D
template <int D, class R>
__device__ __forceinline__ void flux_axis_split(const SplitRef<R>& rf, const R* U, const SplitPt<R>& q, R* F) {
    const R vn = q.v[D], dvn = q.dv[D];
    F[0] = U[1 + D];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const R t = fma(U[1 + i], vn, rf.m[i] * dvn);
        F[1 + i] = (i == D) ? t + q.dp : t;
    }
    F[4] = fma(U[4] + q.dp, vn, rf.H * dvn);
}

// The three numerical fluxes minus F(Ur).N from deviation states: the
// formulas of riemann_llf, riemann_hllc and riemann_roe, with the central
// fluxes, the jumps and the star-state corrections formed from U'.
template <int D, class R>
__device__ __forceinline__ void riemann_llf_split(R gam, const SplitRef<R>& rf, const R* UL, const R* UR, R nmag,
                                                  R* F) {
    SplitPt<R> qL, qR;
    split_point(gam - R(1), rf, UL, qL);
    split_point(gam - R(1), rf, UR, qR);
    R FL[V], FR[V];
    flux_axis_split<D>(rf, UL, qL, FL);
    flux_axis_split<D>(rf, UR, qR, FR);
    R a = fsqrt(qL.v[0] * qL.v[0] + qL.v[1] * qL.v[1] + qL.v[2] * qL.v[2]) + fsqrt(gam * qL.p * qL.r1);
    const R aR = fsqrt(qR.v[0] * qR.v[0] + qR.v[1] * qR.v[1] + qR.v[2] * qR.v[2]) + fsqrt(gam * qR.p * qR.r1);
    if (aR > a) a = aR;
#pragma unroll
    for (int s = 0; s < V; ++s) F[s] = nmag * (R(0.5) * (FL[s] + FR[s]) - R(0.5) * a * (UR[s] - UL[s]));
}

template <int D, class R>
__device__ __forceinline__ void riemann_hllc_split(R gam, const SplitRef<R>& rf, const R* UL, const R* UR, R nmag,
                                                   R* F) {
    SplitPt<R> qL, qR;
    split_point(gam - R(1), rf, UL, qL);
    split_point(gam - R(1), rf, UR, qR);
    const R djac = nmag;
    const R rhoL = qL.rho, rhoL1 = qL.r1, unL = qL.v[D], PL = qL.p;
    const R rhoR = qR.rho, rhoR1 = qR.r1, unR = qR.v[D], PR = qR.p;
    const R aL = fsqrt(gam * PL * rhoL1);
    const R aR = fsqrt(gam * PR * rhoR1);
    const R rAvg = R(0.5) * (rhoL + rhoR);
    const R aAvg = R(0.5) * (aL + aR);
    const R pStar = fmax(R(0.), R(0.5) * (PL + PR - (qR.dv[D] - qL.dv[D]) * rAvg * aAvg));
    const R gq = (gam + R(1.)) / (R(2.) * gam);
    R qsL = R(1.0);
    const R xL = pStar * frcp(PL);
    if (xL > R(1.)) qsL = fsqrt(R(1.) + gq * (xL - R(1.)));
    R qsR = R(1.0);
    const R xR = pStar * frcp(PR);
    if (xR > R(1.)) qsR = fsqrt(R(1.) + gq * (xR - R(1.)));
    const R sL = unL - aL * qsL;
    const R sR = unR + aR * qsR;
    const R sStar = ((qR.dp - qL.dp) + rhoL * unL * (sL - unL) - rhoR * unR * (sR - unR)) *
                    frcp(rhoL * (sL - unL) - rhoR * (sR - unR));
    R FL[V], FR[V];
    if (sL >= R(0.)) {
        flux_axis_split<D>(rf, UL, qL, FL);
#pragma unroll
        for (int s = 0; s < V; ++s) F[s] = djac * FL[s];
    } else if (sR <= R(0.)) {
        flux_axis_split<D>(rf, UR, qR, FR);
#pragma unroll
        for (int s = 0; s < V; ++s) F[s] = djac * FR[s];
    } else if ((sL < R(0.)) && (sStar >= R(0.))) {
        // F_L + sL (U*_L - U_L) with U*_L - U_L from the full state
        flux_axis_split<D>(rf, UL, qL, FL);
        const R c = (sL - unL) * frcp(sL - sStar);
        const R f1 = sStar - unL;
        const R f2 = sStar + PL * frcp(rhoL * (sL - unL));
        F[0] = djac * (FL[0] + rhoL * sL * (c - R(1.)));
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const R u = (rf.m[i] + UL[1 + i]) * (c - R(1.));
            F[1 + i] = djac * ((i == D) ? (FL[1 + i] + sL * (u + rhoL * c * f1)) : (FL[1 + i] + sL * u));
        }
        F[4] = djac * (FL[4] + sL * ((rf.E + UL[4]) * (c - R(1.)) + rhoL * c * f1 * f2));
    } else if ((sStar < R(0.)) && (sR > R(0.))) {
        flux_axis_split<D>(rf, UR, qR, FR);
        const R c = (sR - unR) * frcp(sR - sStar);
        const R f1 = sStar - unR;
        const R f2 = sStar + PR * frcp(rhoR * (sR - unR));
        F[0] = djac * (FR[0] + rhoR * sR * (c - R(1.)));
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const R u = (rf.m[i] + UR[1 + i]) * (c - R(1.));
            F[1 + i] = djac * ((i == D) ? (FR[1 + i] + sR * (u + rhoR * c * f1)) : (FR[1 + i] + sR * u));
        }
        F[4] = djac * (FR[4] + sR * ((rf.E + UR[4]) * (c - R(1.)) + rhoR * c * f1 * f2));
    } else {
#pragma unroll
        for (int s = 0; s < V; ++s) F[s] = __builtin_nan("");
    }
}

template <int D, class R>
__device__ __forceinline__ void riemann_roe_split(R gam, const SplitRef<R>& rf, const R* UL, const R* UR, R nmag,
                                                  R* F) {
    const R gm1 = gam - R(1.);
    SplitPt<R> qL, qR;
    split_point(gm1, rf, UL, qL);
    split_point(gm1, rf, UR, qR);
    const R rL = qL.rho, rR = qR.rho, r1L = qL.r1, r1R = qR.r1;
    const R pL = qL.p, pR = qR.p;
    const R HL = (rf.H + UL[4] + qL.dp) * r1L, HR = (rf.H + UR[4] + qR.dp) * r1R;
    const R k = fsqrt(rR * r1L);
    const R wl = frcp(R(1.) + k), wr = k * wl;
    R ut[3];
#pragma unroll
    for (int d = 0; d < 3; d++) ut[d] = fma(wl, qL.v[d], wr * qR.v[d]);
    const R Ht = fma(wl, HL, wr * HR);
    const R q2 = ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2];
    const R c2 = gm1 * (Ht - R(0.5) * q2);
    const R yc = frsq(c2);
    const R ct = c2 * yc;
    const R ic2 = yc * yc;
    const R rt = rL * k;
    const R unt = ut[D], unL = qL.v[D], unR = qR.v[D];
    const R drho = UR[0] - UL[0], dp = qR.dp - qL.dp, dun = qR.dv[D] - qL.dv[D];
    const R a1 = (dp - rt * ct * dun) * (R(0.5) * ic2);
    const R a5 = (dp + rt * ct * dun) * (R(0.5) * ic2);
    const R a2 = drho - dp * ic2;
    const R cL = fsqrt(gam * pL * r1L), cR = fsqrt(gam * pR * r1R);
    const R l1 = entropy_fix(unt - ct, unL - cL, unR - cR);
    const R l2 = fabs(unt);
    const R l5 = entropy_fix(unt + ct, unL + cL, unR + cR);
    const R w1 = l1 * a1, w5 = l5 * a5, w2 = l2 * a2, ws = l2 * rt;
    const R wsum = w1 + w5, wdc = (w5 - w1) * ct;
    R Dv[V];
    Dv[0] = wsum + w2;
    R udu = R(0.);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const R du = qR.dv[d] - qL.dv[d];
        const R dut = (d == D) ? du - dun : du;
        const R v = fma(Dv[0], ut[d], ws * dut);
        Dv[1 + d] = (d == D) ? v + wdc : v;
        udu = fma(ut[d], dut, udu);
    }
    Dv[4] = fma(wsum, Ht, fma(wdc, unt, fma(w2 * R(0.5), q2, ws * udu)));
    R FL[V], FR[V];
    flux_axis_split<D>(rf, UL, qL, FL);
    flux_axis_split<D>(rf, UR, qR, FR);
    const R hn = R(0.5) * nmag;
#pragma unroll
    for (int s = 0; s < V; ++s) F[s] = hn * ((FL[s] + FR[s]) - Dv[s]);
}

template <int FLUX, int D, class R>
__device__ __forceinline__ void riemann_split(R gam, const SplitRef<R>& rf, const R* UL, const R* UR, R nmag,
                                              R* F) {
    if constexpr (FLUX == 0) riemann_llf_split<D>(gam, rf, UL, UR, nmag, F);
    else if constexpr (FLUX == 1) riemann_hllc_split<D>(gam, rf, UL, UR, nmag, F);
    else riemann_roe_split<D>(gam, rf, UL, UR, nmag, F);
}

template <int FLUX, int D, class R>
__device__ __forceinline__ void riemann(R gam, const R* UL, const R* UR, R nmag, R* F) {
    if constexpr (FLUX == 0) riemann_llf<D>(gam, UL, UR, nmag, F);
    else if constexpr (FLUX == 1) riemann_hllc<D>(gam, UL, UR, nmag, F);
    else riemann_roe<D>(gam, UL, UR, nmag, F);
}

// ---------------------------------------------------------------------------
// Non-affine (trilinear) elements
// ---------------------------------------------------------------------------

// Columns a[c][r] = dx_r/dxi_c of the trilinear map of an element with nodes
// X[n][3] (n = 4k + 2j + i, src/numerics/basis/tools.cpp:215-247) at the
// reference point (u0, u1, u2); optionally the mapped point x.
__device__ __forceinline__ void trilinear(const double* X, double u0, double u1, double u2, double a[3][3],
                                          double* x = nullptr) {
    const double l0[2] = {0.5 * (1. - u0), 0.5 * (1. + u0)};
    const double l1[2] = {0.5 * (1. - u1), 0.5 * (1. + u1)};
    const double l2[2] = {0.5 * (1. - u2), 0.5 * (1. + u2)};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double c0 = 0., c1 = 0., c2 = 0.;
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int o = 0; o < 2; ++o) {
                c0 += 0.5 * (X[(4 * o + 2 * q + 1) * 3 + r] - X[(4 * o + 2 * q) * 3 + r]) * (l1[q] * l2[o]);
                c1 += 0.5 * (X[(4 * o + 2 + q) * 3 + r] - X[(4 * o + q) * 3 + r]) * (l0[q] * l2[o]);
                c2 += 0.5 * (X[(4 + 2 * o + q) * 3 + r] - X[(2 * o + q) * 3 + r]) * (l0[q] * l1[o]);
            }
        a[0][r] = c0;
        a[1][r] = c1;
        a[2][r] = c2;
        if (x) {
            double v = 0.;
#pragma unroll
            for (int n = 0; n < 8; ++n) v += X[n * 3 + r] * (l0[n & 1] * l1[(n >> 1) & 1] * l2[n >> 2]);
            x[r] = v;
        }
    }
}

__device__ __forceinline__ void cross3(const double* u, const double* v, double* w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

// Area-weighted normal of the + face in direction d of an element at the
// face's tangential reference coordinates (u1, u2) (t1 < t2 the other two
// directions): the cofactor row d, a_{d+1} x a_{d+2}, on xi_d = +1 -- the
// reference's dx/ds x dx/dt of the left element's face
// (src/numerics/basis/shape.cpp:482-532), outward from it.
__device__ __forceinline__ void face_normal(const double* X, int d, double u1, double u2, double* N) {
    double a[3][3];
    if (d == 0) trilinear(X, 1., u1, u2, a);
    else if (d == 1) trilinear(X, u1, 1., u2, a);
    else trilinear(X, u1, u2, 1., a);
    cross3(a[(d + 1) % 3], a[(d + 2) % 3], N);
}

// Euler flux dotted with the unit normal n, from 1/rho and p.
__device__ __forceinline__ void flux_dot(const double* U, double r1, double P, const double* n, double* F) {
    const double un = (U[1] * n[0] + U[2] * n[1] + U[3] * n[2]) * r1;
    F[0] = U[0] * un;
#pragma unroll
    for (int i = 0; i < 3; ++i) F[1 + i] = U[1 + i] * un + P * n[i];
    F[4] = (U[4] + P) * un;
}

// The three numerical fluxes for a general area-weighted normal N, |N|
// F(n_hat): the formulas of the axis-aligned versions above
// (src/physics/base/functions.cpp:33-89, src/physics/euler/functions.cpp:
// 298-399, the Roe-Pike flux of riemann_roe).
__device__ __forceinline__ void riemann_llf_n(double gam, const double* UL, const double* UR, const double* N,
                                              double* F) {
    const double nmag = fsqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]), in = frcp(nmag);
    const double n[3] = {N[0] * in, N[1] * in, N[2] * in};
    const double r1L = frcp(UL[0]), r1R = frcp(UR[0]);
    const double PL = pressure_r(gam, UL, r1L), PR = pressure_r(gam, UR, r1R);
    double FL[V], FR[V];
    flux_dot(UL, r1L, PL, n, FL);
    flux_dot(UR, r1R, PR, n, FR);
    double a = fsqrt(UL[1] * UL[1] + UL[2] * UL[2] + UL[3] * UL[3]) * r1L + fsqrt(gam * PL * r1L);
    const double aR = fsqrt(UR[1] * UR[1] + UR[2] * UR[2] + UR[3] * UR[3]) * r1R + fsqrt(gam * PR * r1R);
    if (aR > a) a = aR;
#pragma unroll
    for (int s = 0; s < V; ++s) F[s] = nmag * (0.5 * (FL[s] + FR[s]) - 0.5 * a * (UR[s] - UL[s]));
}

__device__ __forceinline__ void riemann_hllc_n(double gam, const double* UL, const double* UR, const double* N,
                                               double* F) {
    const double djac = fsqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]), in = frcp(djac);
    const double n[3] = {N[0] * in, N[1] * in, N[2] * in};
    const double rhoL = UL[0], rhoL1 = frcp(rhoL);
    const double unL = (UL[1] * n[0] + UL[2] * n[1] + UL[3] * n[2]) * rhoL1;
    const double PL = pressure_r(gam, UL, rhoL1);
    const double aL = fsqrt(gam * PL * rhoL1);
    const double rhoR = UR[0], rhoR1 = frcp(rhoR);
    const double unR = (UR[1] * n[0] + UR[2] * n[1] + UR[3] * n[2]) * rhoR1;
    const double PR = pressure_r(gam, UR, rhoR1);
    const double aR = fsqrt(gam * PR * rhoR1);
    const double rAvg = 0.5 * (rhoL + rhoR);
    const double aAvg = 0.5 * (aL + aR);
    const double pStar = fmax(0., 0.5 * (PL + PR - (unR - unL) * rAvg * aAvg));
    const double gq = (gam + 1.) / (2. * gam);
    double qL = 1.0;
    const double xL = pStar * frcp(PL);
    if (xL > 1.) qL = fsqrt(1. + gq * (xL - 1.));
    double qR = 1.0;
    const double xR = pStar * frcp(PR);
    if (xR > 1.) qR = fsqrt(1. + gq * (xR - 1.));
    const double sL = unL - aL * qL;
    const double sR = unR + aR * qR;
    const double sStar = (PR - PL + rhoL * unL * (sL - unL) - rhoR * unR * (sR - unR)) *
                         frcp(rhoL * (sL - unL) - rhoR * (sR - unR));
    if (sL >= 0.) {
        F[0] = djac * rhoL * unL;
#pragma unroll
        for (int i = 0; i < 3; i++) F[1 + i] = djac * (UL[1 + i] * unL + PL * n[i]);
        F[4] = djac * ((UL[4] + PL) * unL);
    } else if (sR <= 0.) {
        F[0] = djac * rhoR * unR;
#pragma unroll
        for (int i = 0; i < 3; i++) F[1 + i] = djac * (UR[1 + i] * unR + PR * n[i]);
        F[4] = djac * (UR[4] + PR) * unR;
    } else if ((sL < 0.) && (sStar >= 0.)) {
        const double c = (sL - unL) * frcp(sL - sStar);
        const double f1 = sStar - unL;
        const double f2 = sStar + PL * frcp(rhoL * (sL - unL));
        F[0] = djac * (rhoL * unL + rhoL * sL * (c - 1.));
#pragma unroll
        for (int i = 0; i < 3; i++)
            F[1 + i] = djac * (UL[1 + i] * unL + PL * n[i] + sL * (UL[1 + i] * (c - 1.) + rhoL * c * f1 * n[i]));
        F[4] = djac * ((UL[4] + PL) * unL + sL * (UL[4] * (c - 1.) + rhoL * c * f1 * f2));
    } else if ((sStar < 0.) && (sR > 0.)) {
        const double c = (sR - unR) * frcp(sR - sStar);
        const double f1 = sStar - unR;
        const double f2 = sStar + PR * frcp(rhoR * (sR - unR));
        F[0] = djac * (rhoR * unR + rhoR * sR * (c - 1.));
#pragma unroll
        for (int i = 0; i < 3; i++)
            F[1 + i] = djac * (UR[1 + i] * unR + PR * n[i] + sR * (UR[1 + i] * (c - 1.) + rhoR * c * f1 * n[i]));
        F[4] = djac * ((UR[4] + PR) * unR + sR * (UR[4] * (c - 1.) + rhoR * c * f1 * f2));
    } else {
#pragma unroll
        for (int s = 0; s < V; ++s) F[s] = __builtin_nan("");
    }
}

__device__ __forceinline__ void riemann_roe_n(double gam, const double* UL, const double* UR, const double* N,
                                              double* F) {
    const double nmag = fsqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]), in = frcp(nmag);
    const double n[3] = {N[0] * in, N[1] * in, N[2] * in};
    const double rL = UL[0], rR = UR[0];
    const double r1L = frcp(rL), r1R = frcp(rR);
    double uL[3], uR[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        uL[d] = UL[1 + d] * r1L;
        uR[d] = UR[1 + d] * r1R;
    }
    const double gm1 = gam - 1.;
    const double pL = gm1 * (UL[4] - 0.5 * (UL[1] * uL[0] + UL[2] * uL[1] + UL[3] * uL[2]));
    const double pR = gm1 * (UR[4] - 0.5 * (UR[1] * uR[0] + UR[2] * uR[1] + UR[3] * uR[2]));
    const double HL = (UL[4] + pL) * r1L, HR = (UR[4] + pR) * r1R;
    const double sL = fsqrt(rL), sR = fsqrt(rR), isum = frcp(sL + sR);
    double ut[3];
#pragma unroll
    for (int d = 0; d < 3; d++) ut[d] = (sL * uL[d] + sR * uR[d]) * isum;
    const double Ht = (sL * HL + sR * HR) * isum;
    const double q2 = ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2];
    const double c2 = gm1 * (Ht - 0.5 * q2);
    const double ct = fsqrt(c2);
    const double ic2 = frcp(c2);
    const double rt = sL * sR;
    const double unt = ut[0] * n[0] + ut[1] * n[1] + ut[2] * n[2];
    const double unL = uL[0] * n[0] + uL[1] * n[1] + uL[2] * n[2];
    const double unR = uR[0] * n[0] + uR[1] * n[1] + uR[2] * n[2];
    const double drho = rR - rL, dp = pR - pL, dun = unR - unL;
    const double a1 = (dp - rt * ct * dun) * (0.5 * ic2);
    const double a5 = (dp + rt * ct * dun) * (0.5 * ic2);
    const double a2 = drho - dp * ic2;
    const double cL = fsqrt(gam * pL * r1L), cR = fsqrt(gam * pR * r1R);
    const double l1 = entropy_fix(unt - ct, unL - cL, unR - cR);
    const double l2 = fabs(unt);
    const double l5 = entropy_fix(unt + ct, unL + cL, unR + cR);
    const double w1 = l1 * a1, w5 = l5 * a5, w2 = l2 * a2, ws = l2 * rt;
    // grouped by eigenvector component, as in riemann_roe
    const double wsum = w1 + w5, wdc = (w5 - w1) * ct;
    double Dv[V];
    Dv[0] = wsum + w2;
    double udu = 0.;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double dut = (uR[d] - uL[d]) - dun * n[d];
        Dv[1 + d] = fma(Dv[0], ut[d], fma(wdc, n[d], ws * dut));
        udu = fma(ut[d], dut, udu);
    }
    Dv[4] = fma(wsum, Ht, fma(wdc, unt, fma(w2 * 0.5, q2, ws * udu)));
    double FL[V], FR[V];
    flux_dot(UL, r1L, pL, n, FL);
    flux_dot(UR, r1R, pR, n, FR);
#pragma unroll
    for (int s = 0; s < V; ++s) F[s] = nmag * (0.5 * (FL[s] + FR[s]) - 0.5 * Dv[s]);
}

template <int FLUX>
__device__ __forceinline__ void riemann_n(double gam, const double* UL, const double* UR, const double* N,
                                          double* F) {
    if constexpr (FLUX == 0) riemann_llf_n(gam, UL, UR, N, F);
    else if constexpr (FLUX == 1) riemann_hllc_n(gam, UL, UR, N, F);
    else riemann_roe_n(gam, UL, UR, N, F);
}

// ---------------------------------------------------------------------------
// Initial conditions / exact solutions (same formulas as the oracle's
// orc_ic_eval; reference: src/physics/base/base.cpp:150-194,
// src/physics/euler/functions.cpp:94-191)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wrap(double x, double L) { return x - L * round(x / L); }

__device__ __forceinline__ void prim_to_cons(double gam, double rho, const double* u, double p, double* U) {
    U[0] = rho;
    U[1] = rho * u[0];
    U[2] = rho * u[1];
    U[3] = rho * u[2];
    U[4] = p / (gam - 1.0) + 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / rho;
}

__device__ void ic_eval(const KIC& ic, const KGeom& g, const double* x, double t, double* U) {
    const double gam = g.gamma;
    const double* D = ic.data;
    switch (ic.ic) {
    case 0:  // Uniform
        for (int s = 0; s < V; s++) U[s] = D[s];
        break;
    case 1: {  // Sphere
        const double v1 = 1., v0 = 0., R = 0.5, w = 0.5;
        const double magR = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        for (int s = 0; s < V; s++) U[s] = 0.5 * (v1 - v0) * (tanh(M_PI * (R - magR) / w) + 1.0) + v0;
        break;
    }
    case 2: {  // reference 3D isentropic vortex
        const double rho0 = 1., sos0 = 1., L = 1.;
        const double Mach = 1.0, radius = L / 10., sigma = 1;
        const double beta = Mach * 5. * sqrt(2.) / (4. * M_PI) * exp(0.5);
        const double u0 = Mach * sos0;
        const double x1 = x[0] - 0.5, y1 = x[2] - 0.5;
        const double f = -0.5 / (sigma * sigma) * (x1 * x1 + y1 * y1) / (radius * radius);
        const double Om = beta * exp(f);
        const double du = -y1 / radius * Om, dv = x1 / radius * Om;
        const double dT = -(gam - 1.) / 2. * Om * Om;
        const double p = 1. / gam * pow(1 + dT, gam / (gam - 1.));
        U[0] = pow(rho0 + dT, 1. / (gam - 1.));
        U[1] = U[0] * (u0 + du);
        U[2] = 0.;
        U[3] = U[0] * dv;
        const double rKE = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
        U[4] = p / (gam - 1.0) + 0.5 * rKE;
        break;
    }
    case 3: {  // z-extruded 2D vortex, advected with periodic minimum image;
               // D[5], D[6]: initial centre (0 when not given)
        const double rhob = D[0], ub = D[1], vb = D[2], pb = D[3], vs = D[4];
        const double Rg = g.R;
        const double Tb = pb / (rhob * Rg);
        const double s = pb / (pow(rhob, gam));
        const double xr = wrap((x[0] - D[5]) - ub * t, g.period[0]);
        const double yr = wrap((x[1] - D[6]) - vb * t, g.period[1]);
        const double r = sqrt(xr * xr + yr * yr);
        const double dU = vs / (2.0 * M_PI) * exp(0.5 * (1.0 - r * r));
        const double du = dU * -1.0 * yr, dv = dU * xr;
        const double dT = -1.0 * (gam - 1.0) * vs * vs / (8.0 * gam * M_PI * M_PI) * exp(1.0 - r * r);
        const double u = ub + du, v = vb + dv, T = Tb + dT;
        const double rho = pow(T / s, 1.0 / (gam - 1.0));
        const double rhou = rho * u, rhov = rho * v;
        U[0] = rho;
        U[1] = rhou;
        U[2] = rhov;
        U[3] = 0.;
        U[4] = rho * Rg / (gam - 1.0) * T + 0.5 * (rhou * rhou + rhov * rhov) / rho;
        break;
    }
    case 4: {  // Gaussian density bump carried by uniform flow
        const double A = D[0], W = D[1];
        const double Vv[3] = {D[2], D[3], D[4]};
        double r2 = 0.;
        for (int d = 0; d < 3; d++) {
            const double xd = wrap(x[d] - D[6 + d] - Vv[d] * t, g.period[d]);
            r2 += xd * xd;
        }
        prim_to_cons(gam, 1. + A * exp(-r2 / (W * W)), Vv, D[5], U);
        break;
    }
    case 7: {  // up to four z-extruded vortices and a Gaussian pressure pulse
        // data [ub, vb, (xc, yc, beta) x 4, A, px, py, pz, w]: on rho = p = 1
        // (T = 1, p / rho^gamma = 1) flowing at (ub, vb, 0), the Shu-vortex
        // velocity and temperature perturbations of each vortex added (beta
        // 0: absent), rho = T^(1 / (gamma - 1)), p = rho T; then an acoustic
        // pulse p' = A exp(-r^2 / w^2), rho' = p' / gamma.  Periodic minimum
        // images; no exact solution (t is ignored).
        double du = 0., dv = 0., dT = 0.;
        for (int k = 0; k < 4; ++k) {
            const double vs = D[4 + 3 * k];
            const double xr = wrap(x[0] - D[2 + 3 * k], g.period[0]), yr = wrap(x[1] - D[3 + 3 * k], g.period[1]);
            const double r2 = xr * xr + yr * yr;
            const double a = vs / (2.0 * M_PI) * exp(0.5 * (1.0 - r2));
            du -= a * yr;
            dv += a * xr;
            dT -= (gam - 1.0) * vs * vs / (8.0 * gam * M_PI * M_PI) * exp(1.0 - r2);
        }
        const double T = 1.0 + dT;
        double rho = pow(T, 1.0 / (gam - 1.0)), p = rho * T;
        double r2 = 0.;
        for (int d = 0; d < 3; ++d) {
            const double xd = wrap(x[d] - D[15 + d], g.period[d]);
            r2 += xd * xd;
        }
        const double pp = D[14] * exp(-r2 / (D[18] * D[18]));
        rho += pp / gam;
        p += pp;
        const double u[3] = {D[0] + du, D[1] + dv, 0.};
        prim_to_cons(gam, rho, u, p, U);
        break;
    }
    default: {  // Taylor-Green vortex at Mach D[0]
        const double Mn = D[0];
        const double p0 = 1. / (gam * Mn * Mn);
        const double u[3] = {sin(x[0]) * cos(x[1]) * cos(x[2]), -cos(x[0]) * sin(x[1]) * cos(x[2]), 0.};
        const double p = p0 + (cos(2. * x[0]) + cos(2. * x[2])) * (cos(2. * x[1]) + 2.) / 16.;
        prim_to_cons(gam, 1., u, p, U);
        break;
    }
    }
}

// ---------------------------------------------------------------------------
// Element indexing on the periodic box
// ---------------------------------------------------------------------------

// traversal position -> element of a box case
__device__ __forceinline__ unsigned brick_element(const KGeom& g, unsigned p) {
    // bricks of 2^lx x 2^ly x 2^lz elements,
    // x fastest inside a brick and across bricks (all edges 1: e = p)
    const unsigned lx = g.tile[0], ly = g.tile[1], lz = g.tile[2];
    const unsigned lxy = lx + ly;
    const unsigned q = p & ((1u << (lxy + lz)) - 1u), tp = p >> (lxy + lz);
    const unsigned ntx = (unsigned)g.ne[0] >> lx, nty = (unsigned)g.ne[1] >> ly;
    const unsigned tx = tp % ntx, t2 = tp / ntx, ty = t2 % nty, tz = t2 / nty;
    const unsigned x = (tx << lx) | (q & ((1u << lx) - 1u));
    const unsigned y = (ty << ly) | ((q >> lx) & ((1u << ly) - 1u));
    const unsigned z = (tz << lz) | (q >> lxy);
    return x + (unsigned)g.ne[0] * (y + (unsigned)g.ne[1] * z);
}

// the same traversal, also returning the element's box coordinates
__device__ __forceinline__ unsigned brick_element_xyz(const KGeom& g, unsigned p, int* c) {
    const unsigned lx = g.tile[0], ly = g.tile[1], lz = g.tile[2];
    const unsigned lxy = lx + ly;
    const unsigned q = p & ((1u << (lxy + lz)) - 1u), tp = p >> (lxy + lz);
    const unsigned ntx = (unsigned)g.ne[0] >> lx, nty = (unsigned)g.ne[1] >> ly;
    const unsigned tx = tp % ntx, t2 = tp / ntx, ty = t2 % nty, tz = t2 / nty;
    const unsigned x = (tx << lx) | (q & ((1u << lx) - 1u));
    const unsigned y = (ty << ly) | ((q >> lx) & ((1u << ly) - 1u));
    const unsigned z = (tz << lz) | (q >> lxy);
    c[0] = (int)x;
    c[1] = (int)y;
    c[2] = (int)z;
    return x + (unsigned)g.ne[0] * (y + (unsigned)g.ne[1] * z);
}

__device__ __forceinline__ unsigned block_element_xyz(const KGeom& g, int* c) {
    const unsigned w = blockIdx.x;
    const unsigned p = g.xcd_remap ? (w & 7u) * (g.E >> 3) + (w >> 3) : w;
    return brick_element_xyz(g, p, c);
}

__device__ __forceinline__ unsigned block_element(const KGeom& g) {
    const unsigned w = blockIdx.x;
    // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous range of the traversal so neighbour blocks hit its L2.
    const unsigned p = g.xcd_remap ? (w & 7u) * (g.E >> 3) + (w >> 3) : w;
    return brick_element(g, p);
}

__device__ __forceinline__ void elem_ijk(unsigned e, const KGeom& g, int* c) {
    c[0] = (int)(e % (unsigned)g.ne[0]);
    const unsigned r = e / (unsigned)g.ne[0];
    c[1] = (int)(r % (unsigned)g.ne[1]);
    c[2] = (int)(r / (unsigned)g.ne[1]);
}

__device__ __forceinline__ unsigned elem_at(int ix, int iy, int iz, const KGeom& g) {
    ix = ix < 0 ? ix + g.ne[0] : (ix >= g.ne[0] ? ix - g.ne[0] : ix);
    iy = iy < 0 ? iy + g.ne[1] : (iy >= g.ne[1] ? iy - g.ne[1] : iy);
    iz = iz < 0 ? iz + g.ne[2] : (iz >= g.ne[2] ? iz - g.ne[2] : iz);
    return (unsigned)ix + (unsigned)g.ne[0] * ((unsigned)iy + (unsigned)g.ne[1] * (unsigned)iz);
}

// basis index of (normal index r, tangential t1 < t2) for faces normal to D
template <int N, int D>
__device__ __forceinline__ constexpr int bidx(int r, int t1, int t2) {
    return D == 0 ? (t2 * N * N + t1 * N + r) : (D == 1 ? (t2 * N * N + r * N + t1) : (r * N * N + t2 * N + t1));
}

// ---------------------------------------------------------------------------
// LDS layout of the stage kernel
// ---------------------------------------------------------------------------
template <int P, bool MAPPED = false>
struct StageLDS {
    using Dm = Dims<P>;
    static constexpr int N = Dm::N, M = Dm::M;
    // face phase, all three directions at once; "sl" indexes the four
    // traces of a direction: own on the - face, own on the + face,
    // neighbour across the - face, neighbour across the + face
    static constexpr int U = Dm::NNN * V;            // own coefficients
    static constexpr int SLAB = 12 * Dm::NN * V;      // [d][sl][t2][t1][s]
    static constexpr int F1 = 12 * N * M * V;         // [d][sl][t2][a1][s]
    static constexpr int F2 = 12 * Dm::MM * V;        // [d][sl][a2][a1][s]
    static constexpr int FH = 6 * Dm::MM * V;         // [d][side][a2][a1][s]
    static constexpr int L1 = 6 * N * M * V;          // [d][side][t2][a1][s]
    static constexpr int L2 = 6 * Dm::NN * V;         // [d][side][t2][t1][s]
    // volume phase
    static constexpr int X1 = N * N * M * V;
    static constexpr int X2 = N * M * M * V;
    static constexpr int UQ = Dm::MMM * V;
    static constexpr int QW = MAPPED ? 17 : 9;      // point data per volume point
    static constexpr int QD = Dm::MMM * QW;
    static constexpr int C = 3 * N * M * M * V;
    static constexpr int BV = 2 * N * N * M * V;
    // bank A holds U at offset 0 during the face phase
    static constexpr int A = cmax(cmax(U + cmax(cmax(F1, FH), L2), X2), cmax(QD, BV));
    static constexpr int Bk = cmax(cmax(cmax(SLAB, F2), L1), cmax(cmax(X1, UQ), C));
    static constexpr int R = Dm::NNN * V;
    static constexpr int TOTAL = R + A + Bk;
    // Riemann tasks per direction, padded to whole waves so that every
    // wave works on a single direction
    static constexpr int RP = ((2 * Dm::MM + WAVE - 1) / WAVE) * WAVE;
};

// basis index of (normal index r, tangential t1, t2) for faces normal to d:
// d = 0: (i, j, k) = (r, t1, t2); d = 1: (t1, r, t2); d = 2: (t1, t2, r)
template <int N>
__device__ __forceinline__ int bidx_d(int d, int r, int t1, int t2) {
    const int sr = d == 0 ? 1 : (d == 1 ? N : N * N);
    const int s1 = d == 0 ? N : 1;
    const int s2 = d == 2 ? N : N * N;
    return r * sr + t1 * s1 + t2 * s2;
}

// Reference face ids (shape.h:80-86: 0 z-, 1 y-, 2 x+, 3 y+, 4 x-, 5 z+)
// as (normal direction d, side: 1 the + face), from packed bit tables.
__device__ __forceinline__ int face_dir(int f) { return (2118 >> (2 * f)) & 3; }
__device__ __forceinline__ int face_side(int f) { return (44 >> f) & 1; }

// Face record of an element face (slot d*2 + side) on the general path:
// frec[2 slot] = neighbour element, frec[2 slot + 1] = neighbour face id
// (bits 0-2) | map code (bits 3-5) | this element is the face's left one
// (bit 6) | left face id * 8 + left orientation (bits 8-13).  The map
// code takes this face's point (a1, a2) to the neighbour face's (b1, b2):
// swap (bit 0), then reverse b1 (bit 1), reverse b2 (bit 2).
__device__ __forceinline__ void map_point(int code, int m, int a1, int a2, int& b1, int& b2) {
    const int c1 = (code & 1) ? a2 : a1, c2 = (code & 1) ? a1 : a2;
    b1 = (code & 2) ? m - 1 - c1 : c1;
    b2 = (code & 4) ? m - 1 - c2 : c2;
}

// Outward area-weighted normal of face (d, side) of an element at the
// face's tangential reference coordinates (u1, u2): the cofactor row d on
// xi_d = +-1, negated on the - face.
__device__ __forceinline__ void face_normal_gen(const double* X, int d, int side, double u1, double u2,
                                                double* N) {
    const double x = side ? 1. : -1.;
    double a[3][3];
    if (d == 0) trilinear(X, x, u1, u2, a);
    else if (d == 1) trilinear(X, u1, x, u2, a);
    else trilinear(X, u1, u2, x, a);
    cross3(a[(d + 1) % 3], a[(d + 2) % 3], N);
    if (!side) {
        N[0] = -N[0];
        N[1] = -N[1];
        N[2] = -N[2];
    }
}

// ---------------------------------------------------------------------------
// Face phase: traces (own and neighbour) on all six faces, Riemann fluxes,
// lift, accumulated into sR.  The neighbour slab uses exactly the
// arithmetic of the neighbour's own slab, so both sides of a face see the
// same trace and the same flux.
// ---------------------------------------------------------------------------
template <int P, int FLUX, int TS, bool MAPPED = false, class R = double>
__device__ __forceinline__ void face_phase(const KTables& T, const LaneTabT<R>& LT, const KGeom& G, int lane,
                                           const R* sU, R* sR, R* bankA, R* bankB,
                                           const double* XL = nullptr, unsigned eid = 0) {
    using Dm = Dims<P>;
    using L = StageLDS<P>;
    constexpr int N = Dm::N, M = Dm::M, NN = Dm::NN, MM = Dm::MM, NNN = Dm::NNN;
    R* slab = bankB;
    R* f1 = bankA + L::U;
    R* f2 = bankB;
    R* fh = bankA + L::U;
    R* l1 = bankB;
    R* l2 = bankA + L::U;

    // 1. own slabs of both sides (the neighbour slabs are already in `slab`)
    team_for<3 * NN * V, TS>(lane, [&](int t) {
        const int s = t % V, u = t / V, t1 = u % N, u2 = u / N, t2 = u2 % N, d = u2 / N;
        R um = 0., up = 0.;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            const R x = sU[bidx_d<N>(d, r, t1, t2) * V + s];
            um = fma(tphim<R>(T, r), x, um);
            up = fma(tphip<R>(T, r), x, up);
        }
        slab[((d * 4 + 0) * NN + t2 * N + t1) * V + s] = um;
        slab[((d * 4 + 1) * NN + t2 * N + t1) * V + s] = up;
    });
    sync();
    // 2. interpolate along t1 to the face quadrature points
    team_for<12 * N * V, TS>(lane, [&](int t) {
        const int s = t % V, u = t / V, t2 = u % N, ds = u / N;
        R x[N];
#pragma unroll
        for (int t1 = 0; t1 < N; ++t1) x[t1] = slab[(ds * NN + t2 * N + t1) * V + s];
#pragma unroll
        for (int a = 0; a < M; ++a) {
            R acc = 0.;
#pragma unroll
            for (int t1 = 0; t1 < N; ++t1) acc = fma(tB<R>(T, a, t1), x[t1], acc);
            f1[((ds * N + t2) * M + a) * V + s] = acc;
        }
    });
    sync();
    // 3. interpolate along t2
    team_for<12 * M * V, TS>(lane, [&](int t) {
        const int s = t % V, u = t / V, a1 = u % M, ds = u / M;
        R x[N];
#pragma unroll
        for (int t2 = 0; t2 < N; ++t2) x[t2] = f1[((ds * N + t2) * M + a1) * V + s];
#pragma unroll
        for (int a2 = 0; a2 < M; ++a2) {
            R acc = 0.;
#pragma unroll
            for (int t2 = 0; t2 < N; ++t2) acc = fma(tB<R>(T, a2, t2), x[t2], acc);
            f2[((ds * M + a2) * M + a1) * V + s] = acc;
        }
    });
    sync();
    // 4. Riemann flux at the face points; face weight and the sign of the
    //    element's side folded in (Fq = F w; -Fq on the left element, +Fq
    //    on the right).  Each wave handles one direction.
    team_for<3 * L::RP, TS>(lane, [&](int t) {
        const int d = t / L::RP, q = t % L::RP;
        if (q >= 2 * MM) return;
        const int a1 = q % M, u = q / M, a2 = u % M, side = u / M;
        const R* own = f2 + (((d * 4 + side) * M + a2) * M + a1) * V;
        const R* nbv = f2 + (((d * 4 + 2 + side) * M + a2) * M + a1) * V;
        R UL[V], UR[V], F[V];
#pragma unroll
        for (int s = 0; s < V; ++s) {
            UL[s] = side == 0 ? nbv[s] : own[s];
            UR[s] = side == 0 ? own[s] : nbv[s];
        }
        R sgn = side == 0 ? 1. : -1.;   // the box: the + face's element is the left one
        if constexpr (MAPPED) {
            // general connectivity: the neighbour's trace sits at its own
            // face point (b1, b2); the flux is evaluated with the face's
            // left state first and the left element's normal, at the point
            // where the reference evaluates it (table nev: each left face's
            // (s, t) swapped, basis.cpp:332; or the flux point itself)
            const unsigned meta = G.frec[((size_t)eid * 6 + (d * 2 + side)) * 2 + 1];
            const bool isL = (meta >> 6) & 1;
            int b1, b2;
            map_point((meta >> 3) & 7, M, a1, a2, b1, b2);
            const R* nbp = f2 + (((d * 4 + 2 + side) * M + b2) * M + b1) * V;
#pragma unroll
            for (int s = 0; s < V; ++s) {
                UL[s] = isL ? own[s] : nbp[s];
                UR[s] = isL ? nbp[s] : own[s];
            }
            const int lfo = (meta >> 8) & 63, fL = lfo >> 3;
            const int c1 = isL ? a1 : b1, c2 = isL ? a2 : b2;
            const R* ev = G.nev + ((size_t)lfo * MM + c2 * M + c1) * 2;
            R Nv[3];
            face_normal_gen(XL + (isL ? 0 : (1 + d * 2 + side) * 24), face_dir(fL), face_side(fL), ev[0], ev[1],
                            Nv);
            riemann_n<FLUX>(R(G.gamma), UL, UR, Nv, F);
            sgn = isL ? -1. : 1.;
        } else {
            if (d == 0) riemann<FLUX, 0>(R(G.gamma), UL, UR, R(G.farea[0]), F);
            else if (d == 1) riemann<FLUX, 1>(R(G.gamma), UL, UR, R(G.farea[1]), F);
            else riemann<FLUX, 2>(R(G.gamma), UL, UR, R(G.farea[2]), F);
        }
        const R wq = LT.w[a1] * LT.w[a2];
        R* o = fh + (((d * 2 + side) * M + a2) * M + a1) * V;
#pragma unroll
        for (int s = 0; s < V; ++s) o[s] = sgn * (F[s] * wq);
    });
    sync();
    // 5. lift: transpose of the interpolation, along t2 then t1
    team_for<6 * M * V, TS>(lane, [&](int t) {
        const int s = t % V, u = t / V, a1 = u % M, ds = u / M;
        R f[M];
#pragma unroll
        for (int a2 = 0; a2 < M; ++a2) f[a2] = fh[((ds * M + a2) * M + a1) * V + s];
#pragma unroll
        for (int t2 = 0; t2 < N; ++t2) {
            R acc = 0.;
#pragma unroll
            for (int a2 = 0; a2 < M; ++a2) acc = fma(tB<R>(T, a2, t2), f[a2], acc);
            l1[((ds * N + t2) * M + a1) * V + s] = acc;
        }
    });
    sync();
    team_for<6 * N * V, TS>(lane, [&](int t) {
        const int s = t % V, u = t / V, t2 = u % N, ds = u / N;
        R f[M];
#pragma unroll
        for (int a1 = 0; a1 < M; ++a1) f[a1] = l1[((ds * N + t2) * M + a1) * V + s];
#pragma unroll
        for (int t1 = 0; t1 < N; ++t1) {
            R acc = 0.;
#pragma unroll
            for (int a1 = 0; a1 < M; ++a1) acc = fma(tB<R>(T, a1, t1), f[a1], acc);
            l2[((ds * N + t2) * N + t1) * V + s] = acc;
        }
    });
    sync();
    // 6. expand along the normal directions into the residual
    team_for<NNN * V, TS>(lane, [&](int t) {
        const int s = t % V, b = t / V, i = b % N, j = (b / N) % N, k = b / NN;
        R acc = sR[t];
        // d = 0: r = i, (t1, t2) = (j, k)
        acc += LT.phim[i] * l2[((0 * N + k) * N + j) * V + s] + LT.phip[i] * l2[((1 * N + k) * N + j) * V + s];
        // d = 1: r = j, (t1, t2) = (i, k)
        acc += LT.phim[j] * l2[((2 * N + k) * N + i) * V + s] + LT.phip[j] * l2[((3 * N + k) * N + i) * V + s];
        // d = 2: r = k, (t1, t2) = (i, j)
        acc += LT.phim[k] * l2[((4 * N + j) * N + i) * V + s] + LT.phip[k] * l2[((5 * N + j) * N + i) * V + s];
        sR[t] = acc;
    });
    sync();
}

// ---------------------------------------------------------------------------
// Volume phase: sum-factorised interpolation, point fluxes, weak gradient.
// ---------------------------------------------------------------------------
template <int P, int TS, bool MAPPED = false, class R = double>
__device__ __forceinline__ void volume_phase(const KTables& T, const LaneTabT<R>& LT, const KGeom& G, int lane,
                                             R* sR, R* bankA, R* bankB,
                                             const double* XL = nullptr) {
    using Dm = Dims<P>;
    constexpr int N = Dm::N, M = Dm::M, NN = Dm::NN, MM = Dm::MM, MMM = Dm::MMM;
    const R* sU = bankA;
    R* X1 = bankB;
    R* X2 = bankA;
    R* Uq = bankB;
    R* Qd = bankA;
    R* Cc = bankB;
    R* Bv = bankA;

    // x: X1[k][j][a][s] = sum_i B[a][i] U[k][j][i][s]
    team_for<NN * V, TS>(lane, [&](int t) {
        const int s = t % V, kj = t / V;
        R u[N];
#pragma unroll
        for (int i = 0; i < N; ++i) u[i] = sU[(kj * N + i) * V + s];
#pragma unroll
        for (int a = 0; a < M; ++a) {
            R acc = 0.;
#pragma unroll
            for (int i = 0; i < N; ++i) acc = fma(tB<R>(T, a, i), u[i], acc);
            X1[(kj * M + a) * V + s] = acc;
        }
    });
    sync();
    // y: X2[k][b][a][s] = sum_j B[b][j] X1[k][j][a][s]
    team_for<N * M * V, TS>(lane, [&](int t) {
        const int s = t % V, tt = t / V, a = tt % M, k = tt / M;
        R u[N];
#pragma unroll
        for (int j = 0; j < N; ++j) u[j] = X1[((k * N + j) * M + a) * V + s];
#pragma unroll
        for (int b = 0; b < M; ++b) {
            R acc = 0.;
#pragma unroll
            for (int j = 0; j < N; ++j) acc = fma(tB<R>(T, b, j), u[j], acc);
            X2[((k * M + b) * M + a) * V + s] = acc;
        }
    });
    sync();
    // z: Uq[c][b][a][s] = sum_k B[c][k] X2[k][b][a][s]
    team_for<MM * V, TS>(lane, [&](int t) {
        const int s = t % V, ba = t / V;
        R u[N];
#pragma unroll
        for (int k = 0; k < N; ++k) u[k] = X2[(k * MM + ba) * V + s];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            R acc = 0.;
#pragma unroll
            for (int k = 0; k < N; ++k) acc = fma(tB<R>(T, c, k), u[k], acc);
            Uq[(c * MM + ba) * V + s] = acc;
        }
    });
    sync();
    // point data: (rho, m_x, m_y, m_z, E + p, u, v, w, p); on mapped
    // elements (rho, m, E + p, C_d . u (d = 0..2), p C_d (d = 0..2)) with the
    // cofactor rows C_d = a_{d+1} x a_{d+2} = |J| dxi_d/dx at the point, so
    // that the contravariant flux G_d = sum_j F_j (J^-1)_dj |J| of
    // flux_functors_impl.h:95-102 is W C_d.u + p C_d
    if constexpr (MAPPED) {
        team_for<MMM, TS>(lane, [&](int q) {
            const R* u = Uq + q * V;
            R U[V];
#pragma unroll
            for (int s = 0; s < V; ++s) U[s] = u[s];
            const R r1 = frcp(U[0]);
            const R p = pressure_r(R(G.gamma), U, r1);
            const int a = q % M, b = (q / M) % M, c = q / MM;
            R J[3][3];
            trilinear(XL, LT.xv[a], LT.xv[b], LT.xv[c], J);
            R* o = Qd + q * 17;
            o[0] = U[0];
            o[1] = U[1];
            o[2] = U[2];
            o[3] = U[3];
            o[4] = U[4] + p;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                R Cd[3];
                cross3(J[(d + 1) % 3], J[(d + 2) % 3], Cd);
                o[5 + d] = (Cd[0] * U[1] + Cd[1] * U[2] + Cd[2] * U[3]) * r1;
                o[8 + 3 * d] = p * Cd[0];
                o[9 + 3 * d] = p * Cd[1];
                o[10 + 3 * d] = p * Cd[2];
            }
        });
    } else
    team_for<MMM, TS>(lane, [&](int q) {
        const R* u = Uq + q * V;
        R U[V];
#pragma unroll
        for (int s = 0; s < V; ++s) U[s] = u[s];
        const R r1 = frcp(U[0]);
        const R p = pressure_r(R(G.gamma), U, r1);
        R* o = Qd + q * 9;
        o[0] = U[0];
        o[1] = U[1];
        o[2] = U[2];
        o[3] = U[3];
        o[4] = U[4] + p;
        o[5] = U[1] * r1;
        o[6] = U[2] * r1;
        o[7] = U[3] * r1;
        o[8] = p;
    });
    sync();
    // c-step: contravariant fluxes G_d = F_d (dxi_d/dx_d) |J| w, contracted
    // along z: B for x and y fluxes, D for the z flux
    team_for<MM * V, TS>(lane, [&](int t) {
        const int s = t % V, ba = t / V, a = ba % M, b = ba / M;
        const R wab = LT.w[a] * LT.w[b];
        // pressure enters the momentum component s of flux direction s - 1
        const R e1 = s == 1 ? 1. : 0., e2 = s == 2 ? 1. : 0., e3 = s == 3 ? 1. : 0.;
        R xx[N], yy[N], zz[N];
#pragma unroll
        for (int k = 0; k < N; ++k) xx[k] = yy[k] = zz[k] = 0.;
#pragma unroll
        for (int c = 0; c < M; ++c) {
            const R wq = wab * tw<R>(T, c);
            R gx, gy, gz;
            if constexpr (MAPPED) {
                const R* q = Qd + (c * MM + ba) * 17;
                const R Ws = q[s];
                // p C_d . e_s for the momentum components
                const R px = fma(e1, q[8], fma(e2, q[9], e3 * q[10]));
                const R py = fma(e1, q[11], fma(e2, q[12], e3 * q[13]));
                const R pz = fma(e1, q[14], fma(e2, q[15], e3 * q[16]));
                gx = fma(Ws, q[5], px) * wq;
                gy = fma(Ws, q[6], py) * wq;
                gz = fma(Ws, q[7], pz) * wq;
            } else {
                const R* q = Qd + (c * MM + ba) * 9;
                const R Ws = q[s], p = q[8];
                gx = fma(Ws, q[5], e1 * p) * (R(G.gvol[0]) * wq);
                gy = fma(Ws, q[6], e2 * p) * (R(G.gvol[1]) * wq);
                gz = fma(Ws, q[7], e3 * p) * (R(G.gvol[2]) * wq);
            }
#pragma unroll
            for (int k = 0; k < N; ++k) {
                xx[k] = fma(tB<R>(T, c, k), gx, xx[k]);
                yy[k] = fma(tB<R>(T, c, k), gy, yy[k]);
                zz[k] = fma(tD<R>(T, c, k), gz, zz[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
            Cc[((0 * N + k) * MM + ba) * V + s] = xx[k];
            Cc[((1 * N + k) * MM + ba) * V + s] = yy[k];
            Cc[((2 * N + k) * MM + ba) * V + s] = zz[k];
        }
    });
    sync();
    // b-step: along y (B for x and z fluxes, D for the y flux)
    team_for<N * M * V, TS>(lane, [&](int t) {
        const int s = t % V, tt = t / V, a = tt % M, k = tt / M;
        R cx[M], cy[M], cz[M];
#pragma unroll
        for (int b = 0; b < M; ++b) {
            cx[b] = Cc[((0 * N + k) * MM + b * M + a) * V + s];
            cy[b] = Cc[((1 * N + k) * MM + b * M + a) * V + s];
            cz[b] = Cc[((2 * N + k) * MM + b * M + a) * V + s];
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            R y1 = 0., y2 = 0.;
#pragma unroll
            for (int b = 0; b < M; ++b) {
                y1 = fma(tB<R>(T, b, j), cx[b], y1);
                y2 = fma(tD<R>(T, b, j), cy[b], y2);
                y2 = fma(tB<R>(T, b, j), cz[b], y2);
            }
            Bv[((0 * N + k) * N * M + j * M + a) * V + s] = y1;
            Bv[((1 * N + k) * N * M + j * M + a) * V + s] = y2;
        }
    });
    sync();
    // a-step: along x (D for the x flux, B for the others), into sR
    team_for<NN * V, TS>(lane, [&](int t) {
        const int s = t % V, kj = t / V;
        R y1[M], y2[M];
#pragma unroll
        for (int a = 0; a < M; ++a) {
            y1[a] = Bv[((0 * NN + kj) * M + a) * V + s];
            y2[a] = Bv[((1 * NN + kj) * M + a) * V + s];
        }
#pragma unroll
        for (int i = 0; i < N; ++i) {
            R acc = 0.;
#pragma unroll
            for (int a = 0; a < M; ++a) {
                acc = fma(tD<R>(T, a, i), y1[a], acc);
                acc = fma(tB<R>(T, a, i), y2[a], acc);
            }
            sR[(kj * N + i) * V + s] += acc;
        }
    });
    sync();
}

// ---------------------------------------------------------------------------
// Affine periodic box: the same face and volume work with fewer instructions
// per element.
//   * Each lane's task decode (task t = q V + s, variable innermost) is done
//     once per round at the top of the stage and handed to every step.
//   * The (p+2)-point weights sit in the 1D factors of the projection and
//     the lift (Bw, Dw: w_a B[a][i], w_a D[a][i]), so no weight products
//     are formed per point; the face side's sign sits in the expansion.
//   * The point step writes the three contravariant flux vectors of each
//     volume point (G_d = F_d (dxi_d/dx_d) |J|, 15 values), so the first
//     projection step is nothing but multiply-adds.
//   * The Riemann step picks the left / right trace by address, not by
//     per-value selects.
// ---------------------------------------------------------------------------
// The bf16 matrix-core form of the x-y projection: out[m][n] =
// sum_kk Kproj[m][kk] G[kk][n], m = j (p+1) + i, kk = dir (p+2)^2 + b (p+2) + a,
// n = k V + s, in 16 x 16 x 16 tiles (MT x NT output tiles, KC k-steps); G
// sits in LDS as [n][kk] bf16 rows of KP entries (padded off the bank period).
template <int P>
struct ProjDims {
    static constexpr int N = P + 1, M = P + 2, NN = N * N, MM = M * M;
    static constexpr int MT = (NN + 15) / 16, NT = (N * V + 15) / 16, KC = (3 * MM + 15) / 16;
    static constexpr int KP = KC * 16 + 4;
    static constexpr int GROWS = NT * 16;
    static constexpr int G16 = GROWS * KP;          // bf16 entries of one G part
};

// Row strides (in elements) of the box path's LDS arrays slab, f1, l1, l2
// (face phase) and X1, Bv (volume phase): each array is rows of one line
// of points x V variables.  For p = 3 a row is padded to the smallest
// length >= its own that is = V modulo the slots of one pass of an LDS
// access (32 banks of 4 bytes: 16 lanes over 16 double slots for FP64, 32
// lanes over 32 word slots for FP32): the task of lane t is (row q,
// variable s) with t = q V + s, so the address of a row's element is then
// t + const modulo the slots and the lanes of a pass fall on distinct
// banks.  Measured at p = 3: the f1 row alone at 37 instead of 25 doubles
// takes 75 bank-conflict cycles per element off (635 -> 560); all six rows
// padded, 635 -> 433 in FP64 and 422 -> 357 in FP32.  FP64 pads from
// p = 2 up (the padded rows fit the LDS that the point data and the c-step
// output already take at p = 2 and 4; measured per launch at 64^3: p = 2
// -1.1 %, p = 4 -2.3 %, p = 5 -1.3 %, while p = 1 ran 3 % slower padded);
// FP32 at p = 3 only.
template <int P, class R>
struct BoxPad {
    static constexpr int N = P + 1, M = P + 2;
    static constexpr bool ON = (sizeof(R) == 8 && P >= 2) || P == 3;
    // FP64: 16 lanes per pass over 16 double slots; FP32: 32 lanes per pass
    // over 32 word slots
    static constexpr int SLOTS = sizeof(R) == 8 ? 16 : 32;
    static constexpr int pad(int n) { return ON ? n + ((V - n) % SLOTS + SLOTS) % SLOTS : n; }
    static constexpr int SL = pad(N * V), F1 = pad(M * V), L1 = pad(M * V), L2 = pad(N * V);
    static constexpr int X1 = pad(M * V), BV = pad(M * V);
};

template <int P, int GBF = 0, class PR = float>
struct BoxLDS {
    using Dm = Dims<P>;
    using Pd = BoxPad<P, PR>;
    static constexpr int N = Dm::N, M = Dm::M;
    static constexpr int U = Dm::NNN * V;
    static constexpr int SLAB = 12 * N * Pd::SL;
    static constexpr int F1 = 12 * N * Pd::F1;
    static constexpr int F2 = 12 * Dm::MM * V;
    static constexpr int FH = 6 * Dm::MM * V;
    static constexpr int L1 = 6 * N * Pd::L1;
    static constexpr int L2 = 6 * N * Pd::L2;
    static constexpr int X1 = N * N * Pd::X1;
    static constexpr int X2 = N * M * M * V;
    static constexpr int UQ = Dm::MMM * V;
    // point data: the three flux vectors [d][c][b][a][s] (15 values), or
    // at p = 5, where those would halve the occupancy, 9 values per point
    static constexpr bool PF15 = P <= 4;
    // values per point of the compact form: 9, or 12 for the FP32 form's
    // deviation fluxes
    static constexpr int NQ = sizeof(PR) == 4 ? 12 : 9;
    static constexpr int QD = PF15 ? 3 * Dm::MMM * V : NQ * Dm::MMM;
    static constexpr int C = 3 * N * M * M * V;
    static constexpr int BV = 2 * N * N * Pd::BV;
    static constexpr int A = cmax(cmax(U + cmax(cmax(F1, FH), L2), X2), cmax(QD, BV));
    static constexpr int Bk = cmax(cmax(cmax(cmax(SLAB, F2), L1), cmax(cmax(X1, UQ), C)), GBF);
    static constexpr int R = Dm::NNN * V;
    static constexpr int TOTAL = R + A + Bk;
    static constexpr int RP = ((2 * Dm::MM + WAVE - 1) / WAVE) * WAVE;
};

// bank-B space (in units of R) that the bf16 projection's G operand takes
template <int P, int PROJ, class R>
constexpr int proj_bank() {
    return PROJ == 0 ? 0 : (ProjDims<P>::G16 * (PROJ == 3 ? 2 : 1) * 2 + (int)sizeof(R) - 1) / (int)sizeof(R);
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// bf16 of a float, round to nearest even, and back
__device__ __forceinline__ uint16_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// x-y projection on the matrix cores (v_mfma_f32_16x16x16_bf16, FP32
// accumulation), added into the FP32 residual sR.  Operand layout of the
// 16x16x16 bf16 MFMA (lane l, element q = 0..3): A[l % 16][4 (l / 16) + q],
// B[4 (l / 16) + q][l % 16]; lane l receives D[4 (l / 16) + q][l % 16].
// Each wave takes whole output tiles; PROJ = 3 adds the products of the
// hi part with the lo part and of the lo part with the hi part.
template <int P, int TS, int PROJ>
__device__ __forceinline__ void proj_mfma(const KTables& T, int lane, const uint16_t* Gb, float* sR) {
    using PD = ProjDims<P>;
    constexpr int NW = TS / WAVE;
    const int wv = __builtin_amdgcn_readfirstlane(lane / WAVE);
    const int l = lane % WAVE, c = l % 16, kq = 4 * (l / 16);
    for (int tile = wv; tile < PD::MT * PD::NT; tile += NW) {
        const int mt = tile / PD::NT, nt = tile % PD::NT;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        const uint16_t* arow0 = &T.Kproj[0][mt * 16 + c][kq];
        const uint16_t* arow1 = &T.Kproj[1][mt * 16 + c][kq];
        const uint16_t* brow0 = Gb + (nt * 16 + c) * PD::KP + kq;
        const uint16_t* brow1 = brow0 + PD::G16;
#pragma unroll
        for (int kc = 0; kc < PD::KC; ++kc) {
            const bf16x4 a = *reinterpret_cast<const bf16x4*>(arow0 + 16 * kc);
            const bf16x4 b = *reinterpret_cast<const bf16x4*>(brow0 + 16 * kc);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
            if constexpr (PROJ == 3) {
                const bf16x4 al = *reinterpret_cast<const bf16x4*>(arow1 + 16 * kc);
                const bf16x4 bl = *reinterpret_cast<const bf16x4*>(brow1 + 16 * kc);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bl, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, b, acc, 0, 0, 0);
            }
        }
        const int n = nt * 16 + c;
        if (n < PD::N * V) {
            const int k = n / V, sv = n - V * k;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = mt * 16 + kq + q;
                if (m < PD::NN) sR[(k * PD::NN + m) * V + sv] += acc[q];
            }
        }
    }
}

// per-lane task decode of up to NR rounds of a team of TS threads
template <int TS, int NR>
struct TaskDec {
    int s[NR], q[NR];
    __device__ __forceinline__ explicit TaskDec(int lane) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int t = r * TS + lane;
            q[r] = t / V;
            s[r] = t - V * q[r];
        }
    }
};

template <int P, int TS>
constexpr int box_rounds() {
    constexpr int N = P + 1, M = P + 2;
    constexpr int T = cmax(cmax(12 * M * V, N * N * N * V), 6 * N * N * V);
    return (T + TS - 1) / TS;
}

// rounds over T tasks with the decoded (q, s) of each round
template <int T, int TS, int NR, class F>
__device__ __forceinline__ void team_qs(const TaskDec<TS, NR>& dc, int lane, F&& f) {
    static_assert((T + TS - 1) / TS <= NR, "task decode rounds");
#pragma unroll
    for (int r = 0; r < (T + TS - 1) / TS; ++r) {
        if (r * TS + lane < T) f(dc.q[r], dc.s[r], r * TS + lane);
    }
}

template <int P, int FLUX, int TS, int NR, class R>
__device__ __forceinline__ void face_phase_box(const KTables& T, const LaneTabT<R>& LT, const KGeom& G,
                                               const SplitRef<R>& rf, const TaskDec<TS, NR>& dc, int lane,
                                               const R* sU, R* sR, R* bankA, R* bankB) {
    using Dm = Dims<P>;
    using L = BoxLDS<P>;
    constexpr int N = Dm::N, M = Dm::M, NN = Dm::NN, MM = Dm::MM;
    R* slab = bankB;
    R* f1 = bankA + L::U;
    R* f2 = bankB;
    R* fh = bankA + L::U;
    R* l1 = bankB;
    R* l2 = bankA + L::U;
    using Pd = BoxPad<P, R>;

    // 1. own slabs of both sides (the neighbour slabs are already in `slab`)
    team_qs<3 * NN * V, TS>(dc, lane, [&](int u, int s, int) {
        const int t1 = u % N, u2 = u / N, t2 = u2 % N, d = u2 / N;
        R um = 0., up = 0.;
        const int sr = d == 0 ? V : (d == 1 ? N * V : NN * V);
        const R* src = sU + bidx_d<N>(d, 0, t1, t2) * V + s;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            const R x = src[r * sr];
            um = fma(tphim<R>(T, r), x, um);
            up = fma(tphip<R>(T, r), x, up);
        }
        slab[((d * 4 + 0) * N + t2) * Pd::SL + t1 * V + s] = um;
        slab[((d * 4 + 1) * N + t2) * Pd::SL + t1 * V + s] = up;
    });
    sync();
    // 2. interpolate along t1 to the face quadrature points
    team_qs<12 * N * V, TS>(dc, lane, [&](int u, int s, int) {
        const int t2 = u % N, ds = u / N;
        R x[N];
#pragma unroll
        for (int t1 = 0; t1 < N; ++t1) x[t1] = slab[(ds * N + t2) * Pd::SL + t1 * V + s];
#pragma unroll
        for (int a = 0; a < M; ++a) {
            R acc = 0.;
#pragma unroll
            for (int t1 = 0; t1 < N; ++t1) acc = fma(tB<R>(T, a, t1), x[t1], acc);
            f1[(ds * N + t2) * Pd::F1 + a * V + s] = acc;
        }
    });
    sync();
    // 3. interpolate along t2
    team_qs<12 * M * V, TS>(dc, lane, [&](int u, int s, int) {
        const int a1 = u % M, ds = u / M;
        R x[N];
#pragma unroll
        for (int t2 = 0; t2 < N; ++t2) x[t2] = f1[(ds * N + t2) * Pd::F1 + a1 * V + s];
#pragma unroll
        for (int a2 = 0; a2 < M; ++a2) {
            R acc = 0.;
#pragma unroll
            for (int t2 = 0; t2 < N; ++t2) acc = fma(tB<R>(T, a2, t2), x[t2], acc);
            f2[((ds * M + a2) * M + a1) * V + s] = acc;
        }
    });
    sync();
    // 4. Riemann flux at the face points (the + face's element is the left
    //    one); the weights and the side's sign are applied by the lift.
    //    Where padding each direction to whole waves would leave lanes idle
    //    (all orders but p = 3), the face states are read in the face-normal
    //    frame (momentum components cycled so that the normal one comes
    //    first) and one x-normal solver serves all directions, the flux
    //    cycled back on the store; otherwise each wave takes one direction.
    constexpr bool ROT = 3 * L::RP > ((6 * MM + WAVE - 1) / WAVE) * WAVE;
    if constexpr (ROT) {
        team_for<6 * MM, TS>(lane, [&](int t) {
            const int d = t / (2 * MM), q = t - 2 * MM * d;
            const int pt = q % MM, side = q / MM;
            const R* own = f2 + ((d * 4 + side) * MM + pt) * V;
            const R* nbv = f2 + ((d * 4 + 2 + side) * MM + pt) * V;
            const R* pl = side ? own : nbv;
            const R* pr = side ? nbv : own;
            const int c1 = d == 2 ? 0 : d + 1, c2 = d == 0 ? 2 : d - 1;   // (d+1) % 3, (d+2) % 3
            R UL[V], UR[V], F[V];
            UL[0] = pl[0];
            UL[1] = pl[1 + d];
            UL[2] = pl[1 + c1];
            UL[3] = pl[1 + c2];
            UL[4] = pl[4];
            UR[0] = pr[0];
            UR[1] = pr[1 + d];
            UR[2] = pr[1 + c1];
            UR[3] = pr[1 + c2];
            UR[4] = pr[4];
            const R nmag = R(d == 0 ? G.farea[0] : (d == 1 ? G.farea[1] : G.farea[2]));
            if constexpr (sizeof(R) == 4) riemann_split<FLUX, 0>(R(G.gamma), rf.cycled(d), UL, UR, nmag, F);
            else riemann<FLUX, 0>(R(G.gamma), UL, UR, nmag, F);
            R* o = fh + ((d * 2 + side) * MM + pt) * V;
            o[0] = F[0];
            o[1 + d] = F[1];
            o[1 + c1] = F[2];
            o[1 + c2] = F[3];
            o[4] = F[4];
        });
    } else {
        team_for<3 * L::RP, TS>(lane, [&](int t) {
            const int d = t / L::RP, q = t % L::RP;
            if (q >= 2 * MM) return;
            const int pt = q % MM, side = q / MM;
            const R* own = f2 + ((d * 4 + side) * MM + pt) * V;
            const R* nbv = f2 + ((d * 4 + 2 + side) * MM + pt) * V;
            const R* pl = side ? own : nbv;
            const R* pr = side ? nbv : own;
            R UL[V], UR[V], F[V];
#pragma unroll
            for (int s = 0; s < V; ++s) {
                UL[s] = pl[s];
                UR[s] = pr[s];
            }
            if constexpr (sizeof(R) == 4) {
                if (d == 0) riemann_split<FLUX, 0>(R(G.gamma), rf, UL, UR, R(G.farea[0]), F);
                else if (d == 1) riemann_split<FLUX, 1>(R(G.gamma), rf, UL, UR, R(G.farea[1]), F);
                else riemann_split<FLUX, 2>(R(G.gamma), rf, UL, UR, R(G.farea[2]), F);
            } else {
                if (d == 0) riemann<FLUX, 0>(R(G.gamma), UL, UR, R(G.farea[0]), F);
                else if (d == 1) riemann<FLUX, 1>(R(G.gamma), UL, UR, R(G.farea[1]), F);
                else riemann<FLUX, 2>(R(G.gamma), UL, UR, R(G.farea[2]), F);
            }
            R* o = fh + ((d * 2 + side) * MM + pt) * V;
#pragma unroll
            for (int s = 0; s < V; ++s) o[s] = F[s];
        });
    }
    sync();
    // 5. lift: transpose of the interpolation with the weights, along t2
    //    then t1
    team_qs<6 * M * V, TS>(dc, lane, [&](int u, int s, int) {
        const int a1 = u % M, ds = u / M;
        R f[M];
#pragma unroll
        for (int a2 = 0; a2 < M; ++a2) f[a2] = fh[((ds * M + a2) * M + a1) * V + s];
#pragma unroll
        for (int t2 = 0; t2 < N; ++t2) {
            R acc = 0.;
#pragma unroll
            for (int a2 = 0; a2 < M; ++a2) acc = fma(tBw<R>(T, a2, t2), f[a2], acc);
            l1[(ds * N + t2) * Pd::L1 + a1 * V + s] = acc;
        }
    });
    sync();
    team_qs<6 * N * V, TS>(dc, lane, [&](int u, int s, int) {
        const int t2 = u % N, ds = u / N;
        R f[M];
#pragma unroll
        for (int a1 = 0; a1 < M; ++a1) f[a1] = l1[(ds * N + t2) * Pd::L1 + a1 * V + s];
#pragma unroll
        for (int t1 = 0; t1 < N; ++t1) {
            R acc = 0.;
#pragma unroll
            for (int a1 = 0; a1 < M; ++a1) acc = fma(tBw<R>(T, a1, t1), f[a1], acc);
            l2[(ds * N + t2) * Pd::L2 + t1 * V + s] = acc;
        }
    });
    sync();
    // 6. expand along the normal directions into the residual, its first
    //    contribution (the residual starts here): +F on the - face (the
    //    element is the face's right one), -F on the + face.  One task per
    //    x line (j, k, s), i along it: the x-face values are common to the
    //    line, the y- and z-face rows are read at immediate offsets, and
    //    lanes of one (k, s) or (j, s) read the same words
    team_qs<NN * V, TS>(dc, lane, [&](int q, int s, int) {
        const int j = q % N, k = q / N;
        const R ax = l2[(0 * N + k) * Pd::L2 + j * V + s], bx = l2[(1 * N + k) * Pd::L2 + j * V + s];
        const R mj = LT.phim[j], pj = LT.phip[j], mk = LT.phim[k], pk = LT.phip[k];
        const R* ly = l2 + (2 * N + k) * Pd::L2 + s;
        const R* lz = l2 + (4 * N + j) * Pd::L2 + s;
        R* out = sR + q * N * V + s;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            R acc = tphim<R>(T, i) * ax - tphip<R>(T, i) * bx;
            acc += mj * ly[i * V] - pj * ly[N * Pd::L2 + i * V];
            acc += mk * lz[i * V] - pk * lz[N * Pd::L2 + i * V];
            out[i * V] = acc;
        }
    });
    sync();
}

template <int P, int TS, int NR, class R, int PROJ = 0>
__device__ __forceinline__ void volume_phase_box(const KTables& T, const KGeom& G, const SplitRef<R>& rf,
                                                 const TaskDec<TS, NR>& dc, int lane, R* sR, R* bankA,
                                                 R* bankB) {
    static_assert(PROJ == 0 || sizeof(R) == 4, "the bf16 projection belongs to the FP32 form");
    using Dm = Dims<P>;
    constexpr int N = Dm::N, M = Dm::M, NN = Dm::NN, MM = Dm::MM, MMM = Dm::MMM;
    const R* sU = bankA;
    R* X1 = bankB;
    R* X2 = bankA;
    R* Uq = bankB;
    R* Qd = bankA;
    R* Cc = bankB;
    R* Bv = bankA;
    using Pd = BoxPad<P, R>;

    // x: X1[k][j][a][s] = sum_i B[a][i] U[k][j][i][s]
    team_qs<NN * V, TS>(dc, lane, [&](int q, int s, int) {
        R u[N];
#pragma unroll
        for (int i = 0; i < N; ++i) u[i] = sU[(q * N + i) * V + s];
#pragma unroll
        for (int a = 0; a < M; ++a) {
            R acc = 0.;
#pragma unroll
            for (int i = 0; i < N; ++i) acc = fma(tB<R>(T, a, i), u[i], acc);
            X1[q * Pd::X1 + a * V + s] = acc;
        }
    });
    sync();
    // y: X2[k][b][a][s] = sum_j B[b][j] X1[k][j][a][s]; line q = k M + a
    // (consecutive lanes read consecutive LDS words)
    team_qs<N * M * V, TS>(dc, lane, [&](int q, int s, int) {
        const int a = q % M, k = q / M;
        R u[N];
#pragma unroll
        for (int j = 0; j < N; ++j) u[j] = X1[(k * N + j) * Pd::X1 + a * V + s];
#pragma unroll
        for (int b = 0; b < M; ++b) {
            R acc = 0.;
#pragma unroll
            for (int j = 0; j < N; ++j) acc = fma(tB<R>(T, b, j), u[j], acc);
            X2[((k * M + b) * M + a) * V + s] = acc;
        }
    });
    sync();
    // z: Uq[c][b][a][s] = sum_k B[c][k] X2[k][b][a][s]; line q = b M + a
    team_qs<MM * V, TS>(dc, lane, [&](int q, int s, int) {
        R u[N];
#pragma unroll
        for (int k = 0; k < N; ++k) u[k] = X2[(k * MM + q) * V + s];
#pragma unroll
        for (int c = 0; c < M; ++c) {
            R acc = 0.;
#pragma unroll
            for (int k = 0; k < N; ++k) acc = fma(tB<R>(T, c, k), u[k], acc);
            Uq[(c * MM + q) * V + s] = acc;
        }
    });
    sync();
    // point fluxes: G_d[s] = F_d[s] (dxi_d / dx_d) |J| at each volume point,
    // F_d[s] = U_s u_d + p delta_{s, d+1} (E + p in place of E); the FP32
    // form holds deviations from the reference state and forms
    // F_d(Ur + U') - F_d(Ur) = X_s v_d + Y_s v'_d + p' delta_{s, d+1}, with
    // X = (rho', m', E' + p') and Y = (rho_r, m_r, E_r + p_r)
    constexpr bool PF15 = BoxLDS<P>::PF15;
    constexpr bool SPLIT = sizeof(R) == 4;
    constexpr int NQ = BoxLDS<P, 0, R>::NQ;
    if constexpr (!PF15 && SPLIT) {
        // (X, g_x v_x, g_y v_y, g_z v_z, g_x v'_x, g_y v'_y, g_z v'_z, p')
        team_for<MMM, TS>(lane, [&](int pq) {
            const R* u = Uq + pq * V;
            R U[V];
#pragma unroll
            for (int v = 0; v < V; ++v) U[v] = u[v];
            SplitPt<R> q;
            split_point(R(G.gamma) - R(1), rf, U, q);
            R* o = Qd + pq * NQ;
            o[0] = U[0];
            o[1] = U[1];
            o[2] = U[2];
            o[3] = U[3];
            o[4] = U[4] + q.dp;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                o[5 + d] = q.v[d] * R(G.gvol[d]);
                o[8 + d] = q.dv[d] * R(G.gvol[d]);
            }
            o[11] = q.dp;
        });
    } else if constexpr (!PF15) {
        // (rho, m, E + p, g_x u, g_y v, g_z w, p), g_d = (dxi_d / dx_d) |J|
        team_for<MMM, TS>(lane, [&](int pq) {
            const R* u = Uq + pq * V;
            R U[V];
#pragma unroll
            for (int v = 0; v < V; ++v) U[v] = u[v];
            const R r1 = frcp(U[0]);
            const R p = pressure_r(R(G.gamma), U, r1);
            R* o = Qd + pq * 9;
            o[0] = U[0];
            o[1] = U[1];
            o[2] = U[2];
            o[3] = U[3];
            o[4] = U[4] + p;
#pragma unroll
            for (int d = 0; d < 3; ++d) o[5 + d] = U[1 + d] * r1 * R(G.gvol[d]);
            o[8] = p;
        });
    } else if constexpr (SPLIT) {
        team_for<MMM, TS>(lane, [&](int pq) {
            const R* u = Uq + pq * V;
            R X[V];
#pragma unroll
            for (int v = 0; v < V; ++v) X[v] = u[v];
            SplitPt<R> q;
            split_point(R(G.gamma) - R(1), rf, X, q);
            X[4] += q.dp;
            const R Y[V] = {rf.rho, rf.m[0], rf.m[1], rf.m[2], rf.H};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const R g = R(G.gvol[d]);
                const R vd = q.v[d] * g, dvd = q.dv[d] * g, pd = q.dp * g;
                R* o = Qd + (d * MMM + pq) * V;
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const R f = fma(X[v], vd, Y[v] * dvd);
                    o[v] = (v == 1 + d) ? f + pd : f;
                }
            }
        });
    } else
    team_for<MMM, TS>(lane, [&](int pq) {
        const R* u = Uq + pq * V;
        R U[V];
#pragma unroll
        for (int v = 0; v < V; ++v) U[v] = u[v];
        const R r1 = frcp(U[0]);
        const R p = pressure_r(R(G.gamma), U, r1);
        U[4] += p;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const R g = R(G.gvol[d]);
            const R ud = U[1 + d] * r1 * g, pd = p * g;
            R* o = Qd + (d * MMM + pq) * V;
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = (v == 1 + d) ? fma(U[v], ud, pd) : U[v] * ud;
        }
    });
    sync();
    // c-step: contracted along z with the weighted factors: Bw for the x
    // and y fluxes, Dw for the z flux; line q = b M + a
    team_qs<MM * V, TS>(dc, lane, [&](int q, int s, int) {
        R xx[N], yy[N], zz[N];
#pragma unroll
        for (int k = 0; k < N; ++k) xx[k] = yy[k] = zz[k] = 0.;
        // 9-value points: the pressure enters momentum component s of flux
        // direction s - 1, scaled by that direction's metric
        const R ex = s == 1 ? R(G.gvol[0]) : R(0), ey = s == 2 ? R(G.gvol[1]) : R(0),
                ez = s == 3 ? R(G.gvol[2]) : R(0);
        // split 12-value points: Y_s, the reference part of component s
        const R y0 = rf.rho, y1 = rf.m[0], y2 = rf.m[1], y3 = rf.m[2], y4 = rf.H;
        const R ys = s == 0 ? y0 : (s == 1 ? y1 : (s == 2 ? y2 : (s == 3 ? y3 : y4)));
#pragma unroll
        for (int c = 0; c < M; ++c) {
            R gx, gy, gz;
            if constexpr (PF15) {
                const int o = (c * MM + q) * V + s;
                gx = Qd[o];
                gy = Qd[MMM * V + o];
                gz = Qd[2 * MMM * V + o];
            } else if constexpr (SPLIT) {
                const R* pt = Qd + (c * MM + q) * NQ;
                const R Ws = pt[s], pp = pt[11];
                gx = fma(Ws, pt[5], fma(ys, pt[8], ex * pp));
                gy = fma(Ws, pt[6], fma(ys, pt[9], ey * pp));
                gz = fma(Ws, pt[7], fma(ys, pt[10], ez * pp));
            } else {
                const R* pt = Qd + (c * MM + q) * 9;
                const R Ws = pt[s], pp = pt[8];
                gx = fma(Ws, pt[5], ex * pp);
                gy = fma(Ws, pt[6], ey * pp);
                gz = fma(Ws, pt[7], ez * pp);
            }
#pragma unroll
            for (int k = 0; k < N; ++k) {
                xx[k] = fma(tBw<R>(T, c, k), gx, xx[k]);
                yy[k] = fma(tBw<R>(T, c, k), gy, yy[k]);
                zz[k] = fma(tDw<R>(T, c, k), gz, zz[k]);
            }
        }
        if constexpr (PROJ == 0) {
#pragma unroll
            for (int k = 0; k < N; ++k) {
                Cc[((0 * N + k) * MM + q) * V + s] = xx[k];
                Cc[((1 * N + k) * MM + q) * V + s] = yy[k];
                Cc[((2 * N + k) * MM + q) * V + s] = zz[k];
            }
        } else {
            // the matrix cores' G operand, [n = k V + s][kk = dir MM + q], bf16
            using PD = ProjDims<P>;
            uint16_t* Gb = reinterpret_cast<uint16_t*>(bankB);
#pragma unroll
            for (int k = 0; k < N; ++k) {
                uint16_t* row = Gb + (k * V + s) * PD::KP + q;
                const float g[3] = {float(xx[k]), float(yy[k]), float(zz[k])};
#pragma unroll
                for (int dir = 0; dir < 3; ++dir) {
                    const uint16_t hi = bf16_rne(g[dir]);
                    row[dir * MM] = hi;
                    if constexpr (PROJ == 3) row[PD::G16 + dir * MM] = bf16_rne(g[dir] - bf16_value(hi));
                }
            }
        }
    });
    if constexpr (PROJ != 0) {
        // zero k-padding of the G rows (columns 3 MM .. 16 KC), so that the
        // padded k-steps add 0 x 0
        using PD = ProjDims<P>;
        constexpr int PADC = PD::KC * 16 - 3 * MM;
        if constexpr (PADC > 0) {
            uint16_t* Gb = reinterpret_cast<uint16_t*>(bankB);
            team_for<N * V * PADC, TS>(lane, [&](int t) {
                const int n = t / PADC, kk = 3 * MM + t % PADC;
                Gb[n * PD::KP + kk] = 0;
                if constexpr (PROJ == 3) Gb[PD::G16 + n * PD::KP + kk] = 0;
            });
        }
    }
    sync();
    if constexpr (PROJ != 0) {
        proj_mfma<P, TS, PROJ>(T, lane, reinterpret_cast<const uint16_t*>(bankB), sR);
        sync();
        return;
    }
    // b-step: along y (Bw for x and z fluxes, Dw for the y flux); line q = k M + a
    team_qs<N * M * V, TS>(dc, lane, [&](int q, int s, int) {
        const int a = q % M, k = q / M;
        R cx[M], cy[M], cz[M];
#pragma unroll
        for (int b = 0; b < M; ++b) {
            cx[b] = Cc[((0 * N + k) * MM + b * M + a) * V + s];
            cy[b] = Cc[((1 * N + k) * MM + b * M + a) * V + s];
            cz[b] = Cc[((2 * N + k) * MM + b * M + a) * V + s];
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            R y1 = 0., y2 = 0.;
#pragma unroll
            for (int b = 0; b < M; ++b) {
                y1 = fma(tBw<R>(T, b, j), cx[b], y1);
                y2 = fma(tDw<R>(T, b, j), cy[b], y2);
                y2 = fma(tBw<R>(T, b, j), cz[b], y2);
            }
            Bv[((0 * N + k) * N + j) * Pd::BV + a * V + s] = y1;
            Bv[((1 * N + k) * N + j) * Pd::BV + a * V + s] = y2;
        }
    });
    sync();
    // a-step: along x (Dw for the x flux, Bw for the others), into sR
    team_qs<NN * V, TS>(dc, lane, [&](int q, int s, int) {
        R y1[M], y2[M];
#pragma unroll
        for (int a = 0; a < M; ++a) {
            y1[a] = Bv[(0 * NN + q) * Pd::BV + a * V + s];
            y2[a] = Bv[(1 * NN + q) * Pd::BV + a * V + s];
        }
#pragma unroll
        for (int i = 0; i < N; ++i) {
            R acc = 0.;
#pragma unroll
            for (int a = 0; a < M; ++a) {
                acc = fma(tDw<R>(T, a, i), y1[a], acc);
                acc = fma(tBw<R>(T, a, i), y2[a], acc);
            }
            sR[(q * N + i) * V + s] += acc;
        }
    });
    sync();
}

// Apply M^-1 (tensor product of 1D inverses, scaled by 1/|J|) in place on
// an LDS coefficient block (without the dt and 1/|J| factors for the
// non-diagonal case; the caller scales).
template <int P, int TS>
__device__ __forceinline__ void minv_lines(const KTables& T, int lane, double* sR) {
    using Dm = Dims<P>;
    constexpr int N = Dm::N, NN = Dm::NN;
    // x lines
    team_for<NN * V, TS>(lane, [&](int t) {
        const int s = t % V, kj = t / V;
        double u[N];
#pragma unroll
        for (int i = 0; i < N; ++i) u[i] = sR[(kj * N + i) * V + s];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double acc = 0.;
#pragma unroll
            for (int l = 0; l < N; ++l) acc = fma(T.minv[i][l], u[l], acc);
            sR[(kj * N + i) * V + s] = acc;
        }
    });
    sync();
    // y lines
    team_for<NN * V, TS>(lane, [&](int t) {
        const int s = t % V, ki = t / V, i = ki % N, k = ki / N;
        double u[N];
#pragma unroll
        for (int j = 0; j < N; ++j) u[j] = sR[((k * N + j) * N + i) * V + s];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double acc = 0.;
#pragma unroll
            for (int l = 0; l < N; ++l) acc = fma(T.minv[j][l], u[l], acc);
            sR[((k * N + j) * N + i) * V + s] = acc;
        }
    });
    sync();
    // z lines
    team_for<NN * V, TS>(lane, [&](int t) {
        const int s = t % V, ji = t / V;
        double u[N];
#pragma unroll
        for (int k = 0; k < N; ++k) u[k] = sR[(k * NN + ji) * V + s];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double acc = 0.;
#pragma unroll
            for (int l = 0; l < N; ++l) acc = fma(T.minv[k][l], u[l], acc);
            sR[(k * NN + ji) * V + s] = acc;
        }
    });
    sync();
}

// scale of coefficient b in the diagonal case
template <int P>
__device__ __forceinline__ double minv_diag_scale(const KTables& T, int b) {
    constexpr int N = P + 1;
    const int i = b % N, j = (b / N) % N, k = b / (N * N);
    return T.minv[i][i] * T.minv[j][j] * T.minv[k][k];
}

// the same product from the LDS copy of the diagonal
template <int P, class R>
__device__ __forceinline__ double lane_diag_scale(const LaneTabT<R>& LT, int b) {
    constexpr int N = P + 1;
    const int i = b % N, j = (b / N) % N, k = b / (N * N);
    return LT.mdiag[i] * LT.mdiag[j] * LT.mdiag[k];
}

// ---------------------------------------------------------------------------
// Fused stage kernel: one workgroup per element.
//
// All global loads of the element -- its coefficient block, the lines of
// its six neighbours that feed the neighbour traces, and its RK registers --
// are issued together into registers at the start.  Workgroups are mapped
// to elements XCD-contiguously (block_element), so neighbour blocks read by
// concurrently running workgroups share an XCD's L2.
// ---------------------------------------------------------------------------
template <int P, int TS>
struct StageLoads {
    using Dm = Dims<P>;
    static constexpr int N = Dm::N, NN = Dm::NN, NB = Dm::NNN * V;
    static constexpr int RU = (NB + TS - 1) / TS;          // rounds over an element block
    static constexpr int NT = 6 * NN * V;                  // neighbour-slab tasks
    static constexpr int RN = (NT + TS - 1) / TS;
    double own[RU];
    double nl[RN][N];
    double u0[RU], ac[RU], cm[RU];
    int nside[RN];          // general path: side of the neighbour's face

    __device__ __forceinline__ void load_block(const double* Ust, unsigned e, int lane) {
        const double* src = Ust + (size_t)e * NB;
#pragma unroll
        for (int r = 0; r < RU; ++r) {
            const int t = r * TS + lane;
            own[r] = t < NB ? src[t] : 0.;
        }
    }
    // one neighbour line per task t = ((ds N + t2) N + t1) V + s; the six
    // neighbour elements are wave-uniform and picked per lane by slot ds
    // (selects, no per-lane wrap arithmetic), the line is addressed by a
    // base and the normal stride of the neighbour's face direction
    template <class Pick>
    __device__ __forceinline__ void load_lines(const double* Ust, Pick pick, int lane) {
#pragma unroll
        for (int r = 0; r < RN; ++r) {
            const int t = r * TS + lane;
            if (t < NT) {
                const int s = t % V, u = t / V, t1 = u % N, u2 = u / N, t2 = u2 % N, ds = u2 / N;
                unsigned en;
                int dn;
                pick(ds, en, dn);
                const int sr = dn == 0 ? V : (dn == 1 ? N * V : NN * V);
                const int s1 = dn == 0 ? N * V : V;
                const int s2 = dn == 2 ? N * V : NN * V;
                const double* ln = Ust + (size_t)en * NB + (t1 * s1 + t2 * s2 + s);
#pragma unroll
                for (int q = 0; q < N; ++q) nl[r][q] = ln[q * sr];
            }
        }
    }
    __device__ __forceinline__ void load_neighbours(const double* Ust, const int* ec, const KGeom& G, int lane) {
        const unsigned n0 = elem_at(ec[0] - 1, ec[1], ec[2], G), n1 = elem_at(ec[0] + 1, ec[1], ec[2], G);
        const unsigned n2 = elem_at(ec[0], ec[1] - 1, ec[2], G), n3 = elem_at(ec[0], ec[1] + 1, ec[2], G);
        const unsigned n4 = elem_at(ec[0], ec[1], ec[2] - 1, G), n5 = elem_at(ec[0], ec[1], ec[2] + 1, G);
        load_lines(Ust, [&](int ds, unsigned& en, int& dn) {
            en = n0;
            en = ds == 1 ? n1 : en;
            en = ds == 2 ? n2 : en;
            en = ds == 3 ? n3 : en;
            en = ds == 4 ? n4 : en;
            en = ds == 5 ? n5 : en;
            dn = ds >> 1;
        }, lane);
    }
    // general connectivity: the neighbour and its face come from the face
    // records (read once per element, wave-uniform); the lines run along the
    // neighbour's own normal direction and the slab is kept in the
    // neighbour's tangential frame
    __device__ __forceinline__ void load_neighbours_gen(const double* Ust, unsigned e, const KGeom& G,
                                                        int lane) {
        const unsigned* rec = G.frec + (size_t)e * 12;
        const unsigned n0 = rec[0], n1 = rec[2], n2 = rec[4], n3 = rec[6], n4 = rec[8], n5 = rec[10];
        // neighbour face ids, 3 bits each
        const unsigned fpack = (rec[1] & 7) | ((rec[3] & 7) << 3) | ((rec[5] & 7) << 6) | ((rec[7] & 7) << 9) |
                               ((rec[9] & 7) << 12) | ((rec[11] & 7) << 15);
        load_lines(Ust, [&](int ds, unsigned& en, int& dn) {
            en = n0;
            en = ds == 1 ? n1 : en;
            en = ds == 2 ? n2 : en;
            en = ds == 3 ? n3 : en;
            en = ds == 4 ? n4 : en;
            en = ds == 5 ? n5 : en;
            dn = face_dir((fpack >> (3 * ds)) & 7);
        }, lane);
#pragma unroll
        for (int r = 0; r < RN; ++r) {
            const int t = r * TS + lane;
            const int ds = t / (V * NN);
            nside[r] = ds < 6 ? face_side((fpack >> (3 * ds)) & 7) : 0;
        }
    }
    __device__ __forceinline__ void load_rk(const StageArgs& A, unsigned e, int lane) {
        const size_t base = (size_t)e * NB;
        const bool need_acc = A.mode == MODE_RK_MID || A.mode == MODE_RK_LAST;
        const bool need_cmp = A.cmp && (A.mode == MODE_RK_LAST || A.mode == MODE_FE);
#pragma unroll
        for (int r = 0; r < RU; ++r) {
            const int t = r * TS + lane;
            u0[r] = t < NB ? A.U0[base + t] : 0.;
            ac[r] = (t < NB && need_acc) ? A.acc[base + t] : 0.;
            cm[r] = (t < NB && need_cmp) ? (double)A.cmp[base + t] : 0.;
        }
    }
    __device__ __forceinline__ void clear_rk() {
#pragma unroll
        for (int r = 0; r < RU; ++r) u0[r] = ac[r] = cm[r] = 0.;
    }
    // general path: the trace of the neighbour's face side nside
    __device__ __forceinline__ void store_neighbour_slabs_gen(const LaneTab& LT, double* slab, int lane) const {
#pragma unroll
        for (int r = 0; r < RN; ++r) {
            const int t = r * TS + lane;
            if (t < NT) {
                const int s = t % V, u = t / V, t1 = u % N, u2 = u / N, t2 = u2 % N, ds = u2 / N;
                const int d = ds >> 1, side = ds & 1;
                const double* phi = nside[r] ? LT.phip : LT.phim;
                double acc = 0.;
#pragma unroll
                for (int q = 0; q < N; ++q) acc = fma(phi[q], nl[r][q], acc);
                slab[((d * 4 + 2 + side) * NN + t2 * N + t1) * V + s] = acc;
            }
        }
    }
    // neighbour slabs from the prefetched lines: the neighbour across the -
    // face contributes its + trace and vice versa; same fma chain as the
    // neighbour's own slab of that face
    // (FP32 compute: the line rounded to FP32 first, as the own block is)
    template <class R>
    __device__ __forceinline__ void store_neighbour_slabs(const LaneTabT<R>& LT, int one_kind, R* slab,
                                                          int lane) const {
#pragma unroll
        for (int r = 0; r < RN; ++r) {
            const int t = r * TS + lane;
            if (t < NT) {
                const int s = t % V, u = t / V, t1 = u % N, u2 = u / N, t2 = u2 % N, ds = u2 / N;
                const int d = ds >> 1, side = ds & 1;
                const R* phi = side == 0 ? LT.phip : LT.phim;
                R acc = 0.;
                if constexpr (sizeof(R) == 4) {
                    // the reference state's coefficients come off before the
                    // rounding (see stage_body)
                    const double ur = LT.uref[s];
                    const double u0 = (one_kind != 0 || (t1 == 0 && t2 == 0)) ? ur : 0.;
                    const double uq = one_kind != 0 ? ur : 0.;
#pragma unroll
                    for (int q = 0; q < N; ++q) acc = fma(phi[q], R(nl[r][q] - (q == 0 ? u0 : uq)), acc);
                } else {
#pragma unroll
                    for (int q = 0; q < N; ++q) acc = fma(phi[q], R(nl[r][q]), acc);
                }
                slab[((d * 4 + 2 + side) * N + t2) * BoxPad<P, R>::SL + t1 * V + s] = acc;
            }
        }
    }
};

// Neighbour lines on the box path (p >= 2), whole waves per neighbour: the
// neighbour, its face direction and the line stride are uniform across a
// wave, so the neighbour block's address is scalar, each lane adds one
// 32-bit offset and the points along the line are immediate offsets.  The
// trace that feeds the neighbour slab is the same fma chain as before.
template <int P, int TS>
struct BoxNbr {
    using Dm = Dims<P>;
    static constexpr int N = Dm::N, NN = Dm::NN, NB = Dm::NNN * V;
    static constexpr int LN = NN * V;                      // line tasks per neighbour
    static constexpr int WPN = (LN + WAVE - 1) / WAVE;     // waves per neighbour
    static constexpr int WT = 6 * WPN;                     // waves of work
    static constexpr int WR = TS / WAVE;                   // waves per round
    static constexpr int RB = (WT + WR - 1) / WR;          // rounds
    double nl[RB][N];

    template <int DN>
    __device__ __forceinline__ void line(const double* blk, int l, double* x) const {
        // l = (t2 N + t1) V + s; the line's first point t1 s1 + t2 s2 + s in
        // closed form: l + (N - 1) V (l / V) along x, l + (NN - N) V (l / NV)
        // along y, l along z
        constexpr int sr = DN == 0 ? V : (DN == 1 ? N * V : NN * V);
        const int off = DN == 0 ? l + (N - 1) * V * (l / V) : (DN == 1 ? l + (NN - N) * V * (l / (N * V)) : l);
        const double* ln = blk + off;
#pragma unroll
        for (int q = 0; q < N; ++q) x[q] = ln[q * sr];
    }
    // nbe[ds]: neighbour element across face slot ds = d * 2 + side
    __device__ __forceinline__ void load(const double* Ust, const unsigned* nbe, int lane) {
        const int wv = __builtin_amdgcn_readfirstlane(lane / WAVE);
        const int l0 = lane % WAVE;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int w = r * WR + wv;
            if (w < WT) {
                const int ds = w / WPN, l = (w - ds * WPN) * WAVE + l0;
                if (l < LN) {
                    const double* blk = Ust + (size_t)nbe[ds] * NB;
                    const int dn = ds >> 1;
                    if (dn == 0) line<0>(blk, l, nl[r]);
                    else if (dn == 1) line<1>(blk, l, nl[r]);
                    else line<2>(blk, l, nl[r]);
                }
            }
        }
    }
    // the neighbour across the - face contributes its + trace and vice versa
    template <class R>
    __device__ __forceinline__ void store_slabs(const LaneTabT<R>& LT, int one_kind, R* slab, int lane) const {
        const int wv = __builtin_amdgcn_readfirstlane(lane / WAVE);
        const int l0 = lane % WAVE;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int w = r * WR + wv;
            if (w < WT) {
                const int ds = w / WPN, l = (w - ds * WPN) * WAVE + l0;
                if (l < LN) {
                    const int d = ds >> 1, side = ds & 1;
                    const R* phi = side == 0 ? LT.phip : LT.phim;
                    R acc = 0.;
                    if constexpr (sizeof(R) == 4) {
                        // l = (t2 N + t1) V + s; the reference state's
                        // coefficients come off before the rounding
                        const double ur = LT.uref[l % V];
                        const double u0 = (one_kind != 0 || l < V) ? ur : 0.;
                        const double uq = one_kind != 0 ? ur : 0.;
#pragma unroll
                        for (int q = 0; q < N; ++q) acc = fma(phi[q], R(nl[r][q] - (q == 0 ? u0 : uq)), acc);
                    } else {
#pragma unroll
                        for (int q = 0; q < N; ++q) acc = fma(phi[q], R(nl[r][q]), acc);
                    }
                    // l = (t2 N + t1) V + s: row t2 of the slab, column (t1 V + s)
                    const int t2 = l / (N * V);
                    slab[((d * 4 + 2 + side) * N + t2) * BoxPad<P, R>::SL + (l - t2 * N * V)] = acc;
                }
            }
        }
    }
};

// Dense M^-1 of a mapped element applied in place to an LDS coefficient
// block r[NB][V] (SolverTools::mult_inv_mass_matrix, src/solver/tools.cpp:
// 124-139): r <- imm r with the element's symmetric inverse mass matrix in
// global memory (row b' contiguous, imm[b'][b] = imm[b][b'], so threads of
// consecutive b read consecutive addresses).  The sum over b' is split in
// up to four chunks over the threads; part holds chunks * NB * V doubles.
template <int P, int TS>
constexpr int dense_part_size() {
    constexpr int NB = Dims<P>::NNN;
    constexpr int CH0 = TS / NB;
    constexpr int CH = CH0 < 1 ? 1 : (CH0 > 4 ? 4 : CH0);
    return CH * NB * V;
}

template <int P, int TS>
__device__ __forceinline__ void dense_minv(const double* __restrict__ imm, int lane, double* r, double* part) {
    constexpr int NB = Dims<P>::NNN;
    constexpr int CH0 = TS / NB;
    constexpr int CH = CH0 < 1 ? 1 : (CH0 > 4 ? 4 : CH0);
    constexpr int LEN = (NB + CH - 1) / CH;
    team_for<NB * CH, TS>(lane, [&](int t) {
        const int b = t % NB, ch = t / NB;
        const int b0 = ch * LEN, b1 = b0 + LEN < NB ? b0 + LEN : NB;
        double acc[V] = {0., 0., 0., 0., 0.};
        for (int bp = b0; bp < b1; ++bp) {
            const double m = imm[(size_t)bp * NB + b];
#pragma unroll
            for (int s = 0; s < V; ++s) acc[s] = fma(m, r[bp * V + s], acc[s]);
        }
#pragma unroll
        for (int s = 0; s < V; ++s) part[(ch * NB + b) * V + s] = acc[s];
    });
    sync();
    team_for<NB * V, TS>(lane, [&](int t) {
        const int b = t / V, s = t % V;
        double v = 0.;
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) v += part[(ch * NB + b) * V + s];
        r[t] = v;
    });
    sync();
}

// R = float: the FP32 form -- the face and volume phases (traces, fluxes,
// contractions) in single precision on FP32 copies of the state and the
// tables; the residual is widened to FP64 before M^-1 and the RK update,
// which read and write the FP64 registers.  Periodic box only.
// CORR: the RK update takes dU = dt (M^-1 R + C), C a per-DOF correction in
// FP32 laid out like the state (the learned closure's corrected stage)
template <int P, int FLUX, bool MAPPED, class R, int TS = stage_team<P>(), bool CORR = false, int PROJ = 0,
          int MODE = -1>
__device__ __forceinline__ void stage_body(const StageArgs& A, const KTables& T, const KGeom& G, unsigned e,
                                           const int* ec, const float* __restrict__ Corr = nullptr) {
    static_assert(sizeof(R) == 8 || !MAPPED, "the FP32 form is for the periodic box");
    static_assert(TS % WAVE == 0, "whole waves");
    using Dm = Dims<P>;
    using L = StageLDS<P, MAPPED>;
    using LB = BoxLDS<P, proj_bank<P, PROJ, R>(), R>;
    using PF = StageLoads<P, TS>;
    constexpr int NB = Dm::NNN * V;
    constexpr int LA = MAPPED ? L::A : LB::A;
    constexpr int LTOT = MAPPED ? L::TOTAL : LB::TOTAL;
    static_assert(sizeof(R) == 8 || LA * sizeof(R) >= NB * sizeof(double), "FP64 residual fits bank A");
    __shared__ alignas(16) R lds[LTOT];
    __shared__ LaneTabT<R> LT;
    // general path: node coordinates of the element and of its six face
    // neighbours (slot d*2 + side), for the left elements' normals
    __shared__ double XL[MAPPED ? 168 : 1];
    R* sR = lds;
    R* bankA = lds + L::R;
    R* bankB = bankA + LA;
    const int lane = threadIdx.x;
    const bool rk = (MODE >= 0 ? MODE : A.mode) != MODE_RESIDUAL;
    const TaskDec<TS, box_rounds<P, TS>()> dc(lane);

    // Every global load of the element at once: the lane-table entries
    // (branch-free, clamped indices, into registers), its block and the
    // neighbour lines.  The tables go to LDS only after the element's loads
    // are issued: stored under a lane branch first, each table load was
    // waited for before the element's loads went out (two serialised
    // round trips per workgroup).
    const int tl = lane < PMAX + 2 ? lane : PMAX + 1;
    const int tp = lane < PMAX + 1 ? lane : PMAX;
    const R lt_w = tw<R>(T, tl), lt_xv = R(T.xv[tl]);
    const R lt_phim = tphim<R>(T, tp), lt_phip = tphip<R>(T, tp);
    const double lt_md = T.minv[tp][tp];
    constexpr bool SPLIT = sizeof(R) == 4;
    PF pf;
    BoxNbr<P, TS> bn;
    constexpr bool WAVE_NBR = !MAPPED && P >= 2;
    pf.load_block(A.Ust, e, lane);
    if constexpr (MAPPED) {
        pf.load_neighbours_gen(A.Ust, e, G, lane);
    } else if constexpr (WAVE_NBR) {
        const unsigned nbe[6] = {elem_at(ec[0] - 1, ec[1], ec[2], G), elem_at(ec[0] + 1, ec[1], ec[2], G),
                                 elem_at(ec[0], ec[1] - 1, ec[2], G), elem_at(ec[0], ec[1] + 1, ec[2], G),
                                 elem_at(ec[0], ec[1], ec[2] - 1, G), elem_at(ec[0], ec[1], ec[2] + 1, G)};
        bn.load(A.Ust, nbe, lane);
    } else {
        pf.load_neighbours(A.Ust, ec, G, lane);
    }
    if (lane < PMAX + 2) {
        LT.w[lane] = lt_w;
        LT.xv[lane] = lt_xv;
    }
    if (lane < PMAX + 1) {
        LT.phim[lane] = lt_phim;
        LT.phip[lane] = lt_phip;
        LT.mdiag[lane] = lt_md;
    }
    if constexpr (SPLIT) {
        if (lane < V) LT.uref[lane] = uref_of(G, lane);
    }
    if constexpr (MAPPED) {
        team_for<168, TS>(lane, [&](int t) {
            const int which = t / 24;
            const unsigned el = which == 0 ? e : G.frec[((size_t)e * 6 + (which - 1)) * 2];
            XL[t] = G.X[(size_t)el * 24 + t % 24];
        });
    }
    sync();
    // FP32 form: the reference state's coefficients (uref_s times the
    // constant 1 of the basis) come off in FP64, before the rounding
    const SplitRef<R> rf = split_ref<R>(G);
#pragma unroll
    for (int r = 0; r < PF::RU; ++r) {
        const int t = r * TS + lane;
        if (t < NB) {
            double x = pf.own[r];
            if constexpr (SPLIT) {
                // coefficient b of the constant: 1 for b = 0 only (Legendre)
                // or for every b (Lagrange); the host sets no reference for
                // a basis of another kind
                const double kap = (T.one_kind != 0 || t < V) ? 1.0 : 0.0;
                x = fma(-kap, LT.uref[dc.s[r]], x);
            }
            bankA[t] = R(x);
            if constexpr (MAPPED) sR[t] = R(0);   // the box path's face phase starts it
        }
    }
    if constexpr (MAPPED) pf.store_neighbour_slabs_gen(LT, bankB, lane);
    else if constexpr (WAVE_NBR) bn.store_slabs(LT, T.one_kind, bankB, lane);
    else pf.store_neighbour_slabs(LT, T.one_kind, bankB, lane);
    sync();
    if constexpr (MAPPED) {
        face_phase<P, FLUX, TS, MAPPED, R>(T, LT, G, lane, bankA, sR, bankA, bankB, XL, e);
        volume_phase<P, TS, MAPPED, R>(T, LT, G, lane, sR, bankA, bankB, XL);
    } else {
        face_phase_box<P, FLUX, TS>(T, LT, G, rf, dc, lane, bankA, sR, bankA, bankB);
        volume_phase_box<P, TS, box_rounds<P, TS>(), R, PROJ>(T, G, rf, dc, lane, sR, bankA, bankB);
    }
    // FP32 form: the residual widened to FP64 in LDS (bank A is free again)
    // where M^-1 mixes coefficients; a diagonal M^-1 scales each lane's own
    // entries, which the epilogue widens in registers
    double* sRd;
    if constexpr (sizeof(R) == 8) {
        sRd = sR;
    } else {
        sRd = reinterpret_cast<double*>(bankA);
        if (!T.minv_diag) {
#pragma unroll
            for (int r = 0; r < PF::RU; ++r) {
                const int t = r * TS + lane;
                if (t < NB) sRd[t] = (double)sR[t];
            }
            sync();
        }
    }

    // The RK registers are read only now: held from the start they would
    // cost the whole stage their registers (and a wave per SIMD); other
    // workgroups on the CU compute while these loads are in flight.
    if (rk) pf.load_rk(A, e, lane);
    else pf.clear_rk();
    // Epilogue with one store site per array; the value and the target are
    // picked by selects on the (uniform) mode.  Written as one branch per
    // mode, the backend has merged the stores into one block whose base
    // register was left unset on the forward-Euler path.
    const size_t base = (size_t)e * NB;
    const int mode = MODE >= 0 ? MODE : A.mode;
    if constexpr (MAPPED) {
        if (rk) dense_minv<P, TS>(G.imm + (size_t)e * Dm::NNN * Dm::NNN, lane, sR, bankA);
    } else {
        if (rk && !T.minv_diag) minv_lines<P, TS>(T, lane, sRd);
    }
    const double scale = MAPPED ? A.dt : A.dt * G.jinv;
    double* const out = rk ? A.Uout : A.R;
    const bool wacc = mode == MODE_RK_FIRST || mode == MODE_RK_MID;
    const bool last = mode == MODE_RK_LAST || mode == MODE_FE;
    const bool wcmp = A.cmp && last;
#pragma unroll
    for (int r = 0; r < PF::RU; ++r) {
        const int t = r * TS + lane;
        if (t < NB) {
            double rv;
            if constexpr (sizeof(R) == 4) rv = T.minv_diag ? (double)sR[t] : sRd[t];
            else rv = sRd[t];
            const double sc = (!MAPPED && T.minv_diag) ? scale * lane_diag_scale<P>(LT, t / V) : scale;
            double dU = sc * rv;
            if constexpr (CORR) dU += A.dt * (double)Corr[base + t];
            const double u0 = pf.u0[r], ac = pf.ac[r];
            const double n_rk = u0 + A.alpha * dU;                 // U0 + alpha dU
            // the solution increment: (acc + dU) / 6 (RK4) or dU (FE), plus
            // the carried rounding error of U0 when compensated
            const double inc = (mode == MODE_RK_LAST ? (1.0 / 6.0) * (ac + dU) : dU) + pf.cm[r];
            const double n_up = u0 + inc;
            const double nv = !rk ? rv : (last ? n_up : n_rk);
            out[base + t] = nv;
            if (wacc) A.acc[base + t] = mode == MODE_RK_FIRST ? dU : ac + 2.0 * dU;
            if (wcmp) {
                // TwoSum: n_up + err = u0 + inc exactly
                const double bb = n_up - u0;
                A.cmp[base + t] = (float)((u0 - (n_up - bb)) + (inc - bb));
            }
        }
    }
}

// Waves per SIMD the register allocation leaves room for: the FP32 forms
// are held to 64 VGPRs (8 waves per SIMD) at p <= 3 where their LDS allows
// 8 waves of TS-thread workgroups (not at p >= 4, nor for BF16x3 at p = 3,
// whose matrix-core accumulators then need registers beyond 64); elsewhere
// the compiler's choice.
template <int P, class R, int TS, int PROJ>
constexpr int stage_min_waves() {
    constexpr int lds = BoxLDS<P, proj_bank<P, PROJ, R>(), R>::TOTAL * (int)sizeof(R) + 512;
    constexpr int waves = (160 * 1024 / lds) * (TS / WAVE) / 4;
    return (sizeof(R) == 4 && P <= 3 && !(PROJ == 3 && P == 3) && waves >= 8) ? 8 : 1;
}

template <int P, int FLUX, bool MAPPED, class R = double, int PROJ = 0, int MODE = -1>
__global__ __launch_bounds__(stage_team<P>())
__attribute__((amdgpu_waves_per_eu(stage_min_waves<P, R, stage_team<P>(), PROJ>())))
void stage_kernel(StageArgs A, const KTables* __restrict__ Tp,
                                                                KGeom G) {
    int ec[3];
    const unsigned e = block_element_xyz(G, ec);
    stage_body<P, FLUX, MAPPED, R, stage_team<P>(), false, PROJ, MODE>(A, *Tp, G, e, ec);
}

// The closure's corrected RK stage on an affine FP64 box (RK modes only)
template <int P, int FLUX>
__global__ __launch_bounds__(stage_team<P>()) void stage_corr_kernel(StageArgs A, const KTables* __restrict__ Tp,
                                                                     KGeom G, const float* __restrict__ Corr) {
    int ec[3];
    const unsigned e = block_element_xyz(G, ec);
    stage_body<P, FLUX, false, double, stage_team<P>(), true>(A, *Tp, G, e, ec, Corr);
}

// Batched stage: the elements of many independent box cases of one order
// and flux in one launch.  The grid (a multiple of 8 blocks) is remapped
// so that each XCD walks a contiguous eighth of the concatenated case
// traversals; position q of the sequence is element brick_element(q -
// first_block) of case blk_case[q].  Each case brings its own geometry,
// tables and RK registers (BatchCase), so every element sees the stage
// arithmetic of its case's own launch.  Positions past the last element
// (the padding to a multiple of 8) return at once.
template <int P, int FLUX, class R = double, int TS = stage_team<P>(), int PROJ = 0, int MODE = -1>
__global__ __launch_bounds__(TS) __attribute__((amdgpu_waves_per_eu(stage_min_waves<P, R, TS, PROJ>())))
void stage_batch_kernel(const BatchCase* __restrict__ cases,
                                                                      const unsigned* __restrict__ blk_case,
                                                                      unsigned nreal, int stage,
                                                                      const KTables* __restrict__ Tp,
                                                                      const char* base) {
    const unsigned w = blockIdx.x;
    const unsigned q = (w & 7u) * (gridDim.x >> 3) + (w >> 3);
    if (q >= nreal) return;
    const BatchCase* C = cases + blk_case[q];
    // the RK4 stages of step_once: (Ust, Uout) = (U0, Ua), (Ua, Ub), (Ub, Ua), (Ua, U0)
    const long long ou0 = C->off_U0, oua = C->off_Ua, oub = C->off_Ub;
    const long long ost = stage == 0 ? ou0 : (stage == 2 ? oub : oua);
    const long long oout = stage == 0 ? oua : (stage == 1 ? oub : (stage == 2 ? oua : ou0));
    StageArgs A;
    A.Ust = reinterpret_cast<const double*>(base + ost);
    A.U0 = reinterpret_cast<const double*>(base + ou0);
    A.Uout = reinterpret_cast<double*>(const_cast<char*>(base) + oout);
    A.acc = reinterpret_cast<double*>(const_cast<char*>(base) + C->off_acc);
    A.cmp = C->off_cmp >= 0 ? reinterpret_cast<float*>(const_cast<char*>(base) + C->off_cmp) : nullptr;
    A.R = nullptr;
    A.dt = C->dt;
    A.alpha = stage == 2 ? 1.0 : 0.5;
    A.mode = stage == 0 ? MODE_RK_FIRST : (stage == 3 ? MODE_RK_LAST : MODE_RK_MID);
    // the geometry is read in place (a local copy is indexed at run time in
    // places and would be placed in scratch memory)
    const KGeom& G = C->g;
    int ec[3];
    const unsigned e = brick_element_xyz(G, q - C->first_block, ec);
    stage_body<P, FLUX, false, R, TS, false, PROJ, MODE>(A, *Tp, G, e, ec);
}

// dU = dt M^-1 R
template <int P>
__global__ __launch_bounds__(WAVE) void inv_mass_kernel(double dt, const double* R, double* dU,
                                                        const KTables* __restrict__ Tp, KGeom G) {
    const KTables& T = *Tp;
    constexpr int NB = Dims<P>::NNN * V;
    __shared__ double sR[NB];
    __shared__ double part[dense_part_size<P, WAVE>()];
    const int lane = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * NB;
    team_for<NB>(lane, [&](int t) { sR[t] = R[base + t]; });
    sync();
    if (G.mapped) {
        constexpr int NN3 = Dims<P>::NNN;
        dense_minv<P, WAVE>(G.imm + (size_t)blockIdx.x * NN3 * NN3, lane, sR, part);
        team_for<NB>(lane, [&](int t) { dU[base + t] = dt * sR[t]; });
        return;
    }
    if (!T.minv_diag) minv_lines<P, WAVE>(T, lane, sR);
    const double scale = dt * G.jinv;
    team_for<NB>(lane, [&](int t) {
        const double sc = T.minv_diag ? scale * minv_diag_scale<P>(T, t / V) : scale;
        dU[base + t] = sc * sR[t];
    });
}

// ---------------------------------------------------------------------------
// Point evaluation on the (2p+2)^3 rule, slice by slice in z.
// t2 holds [k][b][a][s] after the x and y contractions of the element.
// ---------------------------------------------------------------------------
template <int P, int NP>
__device__ __forceinline__ void contract_xy(const double (*Bt)[PMAX + 1], int lane, const double* u,
                                            double* t1, double* t2) {
    constexpr int N = P + 1;
    team_for<N * N * V>(lane, [&](int t) {
        const int s = t % V, kj = t / V;
        double v[N];
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = u[(kj * N + i) * V + s];
#pragma unroll
        for (int a = 0; a < NP; ++a) {
            double acc = 0.;
#pragma unroll
            for (int i = 0; i < N; ++i) acc = fma(Bt[a][i], v[i], acc);
            t1[(kj * NP + a) * V + s] = acc;
        }
    });
    sync();
    team_for<N * NP * V>(lane, [&](int t) {
        const int s = t % V, tt = t / V, a = tt % NP, k = tt / NP;
        double v[N];
#pragma unroll
        for (int j = 0; j < N; ++j) v[j] = t1[((k * N + j) * NP + a) * V + s];
#pragma unroll
        for (int b = 0; b < NP; ++b) {
            double acc = 0.;
#pragma unroll
            for (int j = 0; j < N; ++j) acc = fma(Bt[b][j], v[j], acc);
            t2[((k * NP + b) * NP + a) * V + s] = acc;
        }
    });
    sync();
}

// physical coordinate of reference coordinate xi in element index ie along d
__device__ __forceinline__ double phys(const KGeom& G, int d, int ie, double xi) {
    return G.lo[d] + ((double)ie + 0.5 * (xi + 1.0)) * G.h[d];
}

template <int P>
__global__ __launch_bounds__(WAVE) void init_state_kernel(double time, double* U, const KTables* __restrict__ Tp,
                                                          KGeom G, KIC ic) {
    const KTables& T = *Tp;
    using Dm = Dims<P>;
    constexpr int N = Dm::N, NN = Dm::NN, NNN = Dm::NNN, Q = Dm::Q1;
    constexpr int NB = NNN * V;
    __shared__ double rhs[NB];
    __shared__ double fs[Q * Q * V];
    __shared__ double t1[Q * N * V];
    __shared__ double Xe[24];
    __shared__ double part[dense_part_size<P, WAVE>()];
    const int lane = threadIdx.x;
    const unsigned e = blockIdx.x;
    int ec[3];
    elem_ijk(e, G, ec);
    if (G.mapped && lane < 24) Xe[lane] = G.X[(size_t)e * 24 + lane];
    team_for<NB>(lane, [&](int t) { rhs[t] = 0.; });
    sync();
    for (int c = 0; c < Q; ++c) {
        // f w |J| at the slice points (src/solver/tools.cpp:348-352)
        team_for<Q * Q>(lane, [&](int t) {
            const int a = t % Q, b = t / Q;
            double x[3], jd = G.jdet;
            if (G.mapped) {
                double J[3][3], C0[3];
                trilinear(Xe, T.xq[a], T.xq[b], T.xq[c], J, x);
                cross3(J[1], J[2], C0);
                jd = J[0][0] * C0[0] + J[0][1] * C0[1] + J[0][2] * C0[2];
            } else {
                x[0] = phys(G, 0, ec[0], T.xq[a]);
                x[1] = phys(G, 1, ec[1], T.xq[b]);
                x[2] = phys(G, 2, ec[2], T.xq[c]);
            }
            double f[V];
            ic_eval(ic, G, x, time, f);
            const double wd = T.wq[a] * T.wq[b] * T.wq[c] * jd;
#pragma unroll
            for (int s = 0; s < V; ++s) fs[(b * Q + a) * V + s] = wd * f[s];
        });
        sync();
        // contract a -> i
        team_for<Q * V>(lane, [&](int t) {
            const int s = t % V, b = t / V;
            double f[Q];
#pragma unroll
            for (int a = 0; a < Q; ++a) f[a] = fs[(b * Q + a) * V + s];
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double acc = 0.;
#pragma unroll
                for (int a = 0; a < Q; ++a) acc = fma(T.BQ[a][i], f[a], acc);
                t1[(b * N + i) * V + s] = acc;
            }
        });
        sync();
        // contract b -> j, then spread along k with BQ[c][k]
        team_for<N * V>(lane, [&](int t) {
            const int s = t % V, i = t / V;
            double f[Q];
#pragma unroll
            for (int b = 0; b < Q; ++b) f[b] = t1[(b * N + i) * V + s];
#pragma unroll
            for (int j = 0; j < N; ++j) {
                double acc = 0.;
#pragma unroll
                for (int b = 0; b < Q; ++b) acc = fma(T.BQ[b][j], f[b], acc);
#pragma unroll
                for (int k = 0; k < N; ++k) rhs[(k * NN + j * N + i) * V + s] += T.BQ[c][k] * acc;
            }
        });
        sync();
    }
    const size_t base = (size_t)e * NB;
    if (G.mapped) {
        dense_minv<P, WAVE>(G.imm + (size_t)e * NNN * NNN, lane, rhs, part);
        team_for<NB>(lane, [&](int t) { U[base + t] = rhs[t]; });
        return;
    }
    if (!T.minv_diag) minv_lines<P, WAVE>(T, lane, rhs);
    team_for<NB>(lane, [&](int t) {
        const double sc = T.minv_diag ? G.jinv * minv_diag_scale<P>(T, t / V) : G.jinv;
        U[base + t] = sc * rhs[t];
    });
}

// mode 0: squared L2 error per variable; mode 1: integrals
template <int P>
__global__ __launch_bounds__(WAVE) void eval_kernel(int mode, double time, const double* U, double* out,
                                                    const KTables* __restrict__ Tp, KGeom G, KIC ic) {
    const KTables& T = *Tp;
    using Dm = Dims<P>;
    constexpr int N = Dm::N, NNN = Dm::NNN, Q = Dm::Q1;
    constexpr int NB = NNN * V;
    constexpr int NO = 6;
    __shared__ double su[NB];
    __shared__ double t1[N * N * Q * V];
    __shared__ double t2[N * Q * Q * V];
    __shared__ double red[WAVE * NO];
    __shared__ double Xe[24];
    const int lane = threadIdx.x;
    const unsigned e = blockIdx.x;
    int ec[3];
    elem_ijk(e, G, ec);
    const size_t base = (size_t)e * NB;
    if (G.mapped && lane < 24) Xe[lane] = G.X[(size_t)e * 24 + lane];
    team_for<NB>(lane, [&](int t) { su[t] = U[base + t]; });
    sync();
    contract_xy<P, Q>(T.BQ, lane, su, t1, t2);
    double accv[NO] = {0., 0., 0., 0., 0., 0.};
    for (int c = 0; c < Q; ++c) {
        team_for<Q * Q>(lane, [&](int t) {
            const int a = t % Q, b = t / Q;
            double u[V];
#pragma unroll
            for (int s = 0; s < V; ++s) {
                double acc = 0.;
#pragma unroll
                for (int k = 0; k < N; ++k) acc = fma(T.BQ[c][k], t2[((k * Q + b) * Q + a) * V + s], acc);
                u[s] = acc;
            }
            double x[3], jd = G.jdet;
            if (G.mapped) {
                double J[3][3], C0[3];
                trilinear(Xe, T.xq[a], T.xq[b], T.xq[c], J, x);
                cross3(J[1], J[2], C0);
                jd = J[0][0] * C0[0] + J[0][1] * C0[1] + J[0][2] * C0[2];
            } else {
                x[0] = phys(G, 0, ec[0], T.xq[a]);
                x[1] = phys(G, 1, ec[1], T.xq[b]);
                x[2] = phys(G, 2, ec[2], T.xq[c]);
            }
            const double wd = T.wq[a] * T.wq[b] * T.wq[c] * jd;
            if (mode == 0) {
                double ue[V];
                ic_eval(ic, G, x, time, ue);
#pragma unroll
                for (int s = 0; s < V; ++s) {
                    const double d = u[s] - ue[s];
                    accv[s] += d * d * wd;
                }
                accv[5] += wd;   // element volume
            } else {
#pragma unroll
                for (int s = 0; s < V; ++s) accv[s] += u[s] * wd;
                accv[5] += 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0] * wd;
            }
        });
    }
#pragma unroll
    for (int o = 0; o < NO; ++o) red[o * WAVE + lane] = accv[o];
    sync();
    if (lane < NO) {
        double s = 0.;
        for (int l = 0; l < WAVE; ++l) s += red[lane * WAVE + l];
        out[(size_t)e * NO + lane] = s;
    }
}

template <int P>
__global__ __launch_bounds__(WAVE) void wavespeed_kernel(const double* U, double* out, const KTables* __restrict__ Tp, KGeom G) {
    const KTables& T = *Tp;
    using Dm = Dims<P>;
    constexpr int N = Dm::N, NNN = Dm::NNN, M = Dm::M;
    constexpr int NB = NNN * V;
    __shared__ double su[NB];
    __shared__ double t1[N * N * M * V];
    __shared__ double t2[N * M * M * V];
    __shared__ double red[WAVE];
    const int lane = threadIdx.x;
    const unsigned e = blockIdx.x;
    const size_t base = (size_t)e * NB;
    team_for<NB>(lane, [&](int t) { su[t] = U[base + t]; });
    sync();
    contract_xy<P, M>(T.B, lane, su, t1, t2);
    double amax = 0.;
    for (int c = 0; c < M; ++c) {
        team_for<M * M>(lane, [&](int t) {
            const int a = t % M, b = t / M;
            double u[V];
#pragma unroll
            for (int s = 0; s < V; ++s) {
                double acc = 0.;
#pragma unroll
                for (int k = 0; k < N; ++k) acc = fma(T.B[c][k], t2[((k * M + b) * M + a) * V + s], acc);
                u[s] = acc;
            }
            const double w = max_wave_speed(G.gamma, u);
            amax = (w > amax || w != w) ? w : amax;
        });
    }
    red[lane] = amax;
    sync();
    if (lane == 0) {
        double m = 0.;
        for (int l = 0; l < WAVE; ++l) m = (red[l] > m || red[l] != red[l]) ? red[l] : m;
        out[e] = m;
    }
}

// Mass matrix of a mapped element on the (2p+2)^3 rule, M = sum_q w_q |J|_q
// phi(q) phi(q)^T (src/solver/helpers.cpp:53-100,
// src/numerics/basis/tools.cpp:85-110), sum-factorised, then inverted in
// place by Gauss-Jordan elimination (M is symmetric positive definite: no
// pivoting).  One workgroup per element; jmin[e] = smallest |J| at the
// points.
template <int P>
__global__ __launch_bounds__(256) void mass_kernel(double* imm, double* jmin, const KTables* __restrict__ Tp,
                                                   KGeom G) {
    constexpr int TS = 256;
    const KTables& T = *Tp;
    using Dm = Dims<P>;
    constexpr int N = Dm::N, NN = Dm::NN, NB = Dm::NNN, Q = Dm::Q1;
    __shared__ double Xe[24];
    __shared__ double W[Q * Q * Q];          // [c][b][a]
    __shared__ double T1[NN * Q * Q];        // [(i1, i2)][c][b]
    __shared__ double A[NB * NB];
    __shared__ double col[NB];
    __shared__ double red[TS];
    const int lane = threadIdx.x;
    const unsigned e = blockIdx.x;
    if (lane < 24) Xe[lane] = G.X[(size_t)e * 24 + lane];
    sync();
    double jm = 1e300;
    team_for<Q * Q * Q, TS>(lane, [&](int t) {
        const int a = t % Q, b = (t / Q) % Q, c = t / (Q * Q);
        double J[3][3], C0[3];
        trilinear(Xe, T.xq[a], T.xq[b], T.xq[c], J);
        cross3(J[1], J[2], C0);
        const double jd = J[0][0] * C0[0] + J[0][1] * C0[1] + J[0][2] * C0[2];
        jm = fmin(jm, jd);
        W[t] = T.wq[a] * T.wq[b] * T.wq[c] * jd;
    });
    red[lane] = jm;
    sync();
    if (lane == 0) {
        double m = red[0];
        for (int l = 1; l < TS; ++l) m = fmin(m, red[l]);
        jmin[e] = m;
    }
    // T1[(i1, i2)][c][b] = sum_a BQ[a][i1] BQ[a][i2] W[c][b][a]
    team_for<NN * Q * Q, TS>(lane, [&](int t) {
        const int cb = t % (Q * Q), ii = t / (Q * Q), i1 = ii / N, i2 = ii % N;
        double acc = 0.;
#pragma unroll
        for (int a = 0; a < Q; ++a) acc = fma(T.BQ[a][i1] * T.BQ[a][i2], W[cb * Q + a], acc);
        T1[t] = acc;
    });
    sync();
    // M[(k1 j1 i1)][(k2 j2 i2)] = sum_c BQ[c][k1] BQ[c][k2] sum_b BQ[b][j1] BQ[b][j2] T1[(i1, i2)][c][b]
    team_for<NB * NB, TS>(lane, [&](int t) {
        const int r = t / NB, q = t % NB;
        const int i1 = r % N, j1 = (r / N) % N, k1 = r / NN;
        const int i2 = q % N, j2 = (q / N) % N, k2 = q / NN;
        const double* t1 = T1 + (i1 * N + i2) * Q * Q;
        double acc = 0.;
        for (int c = 0; c < Q; ++c) {
            double in = 0.;
#pragma unroll
            for (int b = 0; b < Q; ++b) in = fma(T.BQ[b][j1] * T.BQ[b][j2], t1[c * Q + b], in);
            acc = fma(T.BQ[c][k1] * T.BQ[c][k2], in, acc);
        }
        A[t] = acc;
    });
    sync();
    for (int k = 0; k < NB; ++k) {
        team_for<NB, TS>(lane, [&](int i) { col[i] = A[i * NB + k]; });
        sync();
        const double ip = 1.0 / col[k];
        team_for<NB, TS>(lane, [&](int j) { A[k * NB + j] = (j == k ? 1.0 : A[k * NB + j]) * ip; });
        sync();
        team_for<NB * NB, TS>(lane, [&](int t) {
            const int i = t / NB, j = t % NB;
            if (i != k) A[t] = (j == k ? 0.0 : A[t]) - col[i] * A[k * NB + j];
        });
        sync();
    }
    team_for<NB * NB, TS>(lane, [&](int t) { imm[(size_t)e * NB * NB + t] = A[t]; });
}

#include "dg_kernels_2d.inc"

// ---------------------------------------------------------------------------
// Dispatch
// ---------------------------------------------------------------------------
// Box stage (FP64 or the FP32 form) with the RK mode compiled in (its
// epilogue then has no mode selects); the residual mode keeps the run-time
// form
template <int P, int FLUX, class R>
void stage_box_mode(const StageArgs& a, const KTables* t, const KGeom& g, dim3 grid, dim3 block, hipStream_t st) {
    switch (a.mode) {
    case MODE_RK_FIRST: hipLaunchKernelGGL((stage_kernel<P, FLUX, false, R, 0, MODE_RK_FIRST>), grid, block, 0, st, a, t, g); break;
    case MODE_RK_MID: hipLaunchKernelGGL((stage_kernel<P, FLUX, false, R, 0, MODE_RK_MID>), grid, block, 0, st, a, t, g); break;
    case MODE_RK_LAST: hipLaunchKernelGGL((stage_kernel<P, FLUX, false, R, 0, MODE_RK_LAST>), grid, block, 0, st, a, t, g); break;
    case MODE_FE: hipLaunchKernelGGL((stage_kernel<P, FLUX, false, R, 0, MODE_FE>), grid, block, 0, st, a, t, g); break;
    default: hipLaunchKernelGGL((stage_kernel<P, FLUX, false, R>), grid, block, 0, st, a, t, g); break;
    }
}

template <int P>
hipError_t stage_p(int flux, const StageArgs& a, const KTables* t, const KGeom& g, hipStream_t st) {
    dim3 grid(g.E), block(stage_team<P>());
    if (g.mapped) {
        // non-affine elements: orders up to 4 (the case parser rejects more)
        if constexpr (P <= 4) {
            switch (flux) {
            case 0: hipLaunchKernelGGL((stage_kernel<P, 0, true>), grid, block, 0, st, a, t, g); break;
            case 1: hipLaunchKernelGGL((stage_kernel<P, 1, true>), grid, block, 0, st, a, t, g); break;
            default: hipLaunchKernelGGL((stage_kernel<P, 2, true>), grid, block, 0, st, a, t, g); break;
            }
        } else {
            return hipErrorInvalidValue;
        }
    } else if (g.fp32 && g.proj == 1) {
        switch (flux) {
        case 0: hipLaunchKernelGGL((stage_kernel<P, 0, false, float, 1>), grid, block, 0, st, a, t, g); break;
        case 1: hipLaunchKernelGGL((stage_kernel<P, 1, false, float, 1>), grid, block, 0, st, a, t, g); break;
        default: hipLaunchKernelGGL((stage_kernel<P, 2, false, float, 1>), grid, block, 0, st, a, t, g); break;
        }
    } else if (g.fp32 && g.proj == 3) {
        switch (flux) {
        case 0: hipLaunchKernelGGL((stage_kernel<P, 0, false, float, 3>), grid, block, 0, st, a, t, g); break;
        case 1: hipLaunchKernelGGL((stage_kernel<P, 1, false, float, 3>), grid, block, 0, st, a, t, g); break;
        default: hipLaunchKernelGGL((stage_kernel<P, 2, false, float, 3>), grid, block, 0, st, a, t, g); break;
        }
    } else if (g.fp32) {
        switch (flux) {
        case 0: stage_box_mode<P, 0, float>(a, t, g, grid, block, st); break;
        case 1: stage_box_mode<P, 1, float>(a, t, g, grid, block, st); break;
        default: stage_box_mode<P, 2, float>(a, t, g, grid, block, st); break;
        }
    } else {
        switch (flux) {
        case 0: stage_box_mode<P, 0, double>(a, t, g, grid, block, st); break;
        case 1: stage_box_mode<P, 1, double>(a, t, g, grid, block, st); break;
        default: stage_box_mode<P, 2, double>(a, t, g, grid, block, st); break;
        }
    }
    return hipGetLastError();
}

template <int P, int TS, int PROJ>
hipError_t stage_batch_proj(int flux, const BatchCase* cases, const unsigned* blk_case, unsigned nblocks, int stage,
                            const KTables* t, hipStream_t st) {
    const dim3 grid((nblocks + 7u) & ~7u), block(TS);
    const char* base = reinterpret_cast<const char*>(cases);
    switch (flux) {
    case 0: hipLaunchKernelGGL((stage_batch_kernel<P, 0, float, TS, PROJ>), grid, block, 0, st, cases, blk_case, nblocks, stage, t, base); break;
    case 1: hipLaunchKernelGGL((stage_batch_kernel<P, 1, float, TS, PROJ>), grid, block, 0, st, cases, blk_case, nblocks, stage, t, base); break;
    default: hipLaunchKernelGGL((stage_batch_kernel<P, 2, float, TS, PROJ>), grid, block, 0, st, cases, blk_case, nblocks, stage, t, base); break;
    }
    return hipGetLastError();
}

// batched stage with the RK mode of the stage compiled in (stages 1 and 2
// share the middle mode), as the single-case box stage has it
template <int P, int FLUX, class R, int TS>
void stage_batch_mode(const BatchCase* cases, const unsigned* blk_case, unsigned nblocks, int stage,
                      const KTables* t, dim3 grid, dim3 block, hipStream_t st) {
    const char* base = reinterpret_cast<const char*>(cases);
    if (stage == 0)
        hipLaunchKernelGGL((stage_batch_kernel<P, FLUX, R, TS, 0, MODE_RK_FIRST>), grid, block, 0, st, cases, blk_case,
                           nblocks, stage, t, base);
    else if (stage == 3)
        hipLaunchKernelGGL((stage_batch_kernel<P, FLUX, R, TS, 0, MODE_RK_LAST>), grid, block, 0, st, cases, blk_case,
                           nblocks, stage, t, base);
    else
        hipLaunchKernelGGL((stage_batch_kernel<P, FLUX, R, TS, 0, MODE_RK_MID>), grid, block, 0, st, cases, blk_case,
                           nblocks, stage, t, base);
}

template <int P, int TS>
hipError_t stage_batch_pt(int flux, int fp32, const BatchCase* cases, const unsigned* blk_case, unsigned nblocks,
                          int stage, const KTables* t, hipStream_t st) {
    const dim3 grid((nblocks + 7u) & ~7u), block(TS);
    if constexpr (TS == stage_team<P>()) {
        // the default team: the RK mode compiled in
        if (fp32) {
            switch (flux) {
            case 0: stage_batch_mode<P, 0, float, TS>(cases, blk_case, nblocks, stage, t, grid, block, st); break;
            case 1: stage_batch_mode<P, 1, float, TS>(cases, blk_case, nblocks, stage, t, grid, block, st); break;
            default: stage_batch_mode<P, 2, float, TS>(cases, blk_case, nblocks, stage, t, grid, block, st); break;
            }
        } else {
            switch (flux) {
            case 0: stage_batch_mode<P, 0, double, TS>(cases, blk_case, nblocks, stage, t, grid, block, st); break;
            case 1: stage_batch_mode<P, 1, double, TS>(cases, blk_case, nblocks, stage, t, grid, block, st); break;
            default: stage_batch_mode<P, 2, double, TS>(cases, blk_case, nblocks, stage, t, grid, block, st); break;
            }
        }
        return hipGetLastError();
    } else {
        // the second team (tuning candidate): the run-time mode
        if (fp32) {
            switch (flux) {
            case 0: hipLaunchKernelGGL((stage_batch_kernel<P, 0, float, TS>), grid, block, 0, st, cases, blk_case, nblocks, stage, t, reinterpret_cast<const char*>(cases)); break;
            case 1: hipLaunchKernelGGL((stage_batch_kernel<P, 1, float, TS>), grid, block, 0, st, cases, blk_case, nblocks, stage, t, reinterpret_cast<const char*>(cases)); break;
            default: hipLaunchKernelGGL((stage_batch_kernel<P, 2, float, TS>), grid, block, 0, st, cases, blk_case, nblocks, stage, t, reinterpret_cast<const char*>(cases)); break;
            }
        } else {
            switch (flux) {
            case 0: hipLaunchKernelGGL((stage_batch_kernel<P, 0, double, TS>), grid, block, 0, st, cases, blk_case, nblocks, stage, t, reinterpret_cast<const char*>(cases)); break;
            case 1: hipLaunchKernelGGL((stage_batch_kernel<P, 1, double, TS>), grid, block, 0, st, cases, blk_case, nblocks, stage, t, reinterpret_cast<const char*>(cases)); break;
            default: hipLaunchKernelGGL((stage_batch_kernel<P, 2, double, TS>), grid, block, 0, st, cases, blk_case, nblocks, stage, t, reinterpret_cast<const char*>(cases)); break;
            }
        }
        return hipGetLastError();
    }
}

// the second team size the batched launch can run each order with (the
// first is stage_team<P>); the batch picks per group by timing both
template <int P>
constexpr int alt_team() {
    return P == 1 ? 128 : (P == 2 ? 64 : (P <= 4 ? 128 : 256));
}

template <int P>
hipError_t stage_batch_p(int flux, int fp32, int proj, int variant, const BatchCase* cases, const unsigned* blk_case,
                         unsigned nblocks, int stage, const KTables* t, hipStream_t st) {
    // the bf16 projection forms have the default team only
    if (proj == 1) return stage_batch_proj<P, stage_team<P>(), 1>(flux, cases, blk_case, nblocks, stage, t, st);
    if (proj == 3) return stage_batch_proj<P, stage_team<P>(), 3>(flux, cases, blk_case, nblocks, stage, t, st);
    if (variant == 1)
        return stage_batch_pt<P, alt_team<P>()>(flux, fp32, cases, blk_case, nblocks, stage, t, st);
    return stage_batch_pt<P, stage_team<P>()>(flux, fp32, cases, blk_case, nblocks, stage, t, st);
}

template <int P>
constexpr int team_of(int variant) {
    return variant == 1 ? alt_team<P>() : stage_team<P>();
}

}  // namespace

#define DG_DISPATCH_P(P, CALL)                         \
    switch (P) {                                       \
    case 1: { constexpr int PP = 1; CALL; } break;     \
    case 2: { constexpr int PP = 2; CALL; } break;     \
    case 3: { constexpr int PP = 3; CALL; } break;     \
    case 4: { constexpr int PP = 4; CALL; } break;     \
    case 5: { constexpr int PP = 5; CALL; } break;     \
    default: return hipErrorInvalidValue;              \
    }

hipError_t launch_stage(int P, int flux, const StageArgs& a, const KTables* t, const KGeom& g,
                        hipStream_t st) {
    if (g.dim == 2) {
        DG_DISPATCH_P(P, return stage2d_p<PP>(flux, a, t, g, st));
    }
    DG_DISPATCH_P(P, return stage_p<PP>(flux, a, t, g, st));
    return hipSuccess;
}

hipError_t launch_stage_corrected(int P, int flux, const StageArgs& a, const float* C, const KTables* t,
                                  const KGeom& g, hipStream_t st) {
    if (P != 1 || g.dim != 3 || g.mapped || g.fp32 || a.mode == MODE_RESIDUAL) return hipErrorInvalidValue;
    dim3 grid(g.E), block(stage_team<1>());
    switch (flux) {
    case 0: hipLaunchKernelGGL((stage_corr_kernel<1, 0>), grid, block, 0, st, a, t, g, C); break;
    case 1: hipLaunchKernelGGL((stage_corr_kernel<1, 1>), grid, block, 0, st, a, t, g, C); break;
    default: hipLaunchKernelGGL((stage_corr_kernel<1, 2>), grid, block, 0, st, a, t, g, C); break;
    }
    return hipGetLastError();
}

hipError_t launch_stage_batch(int P, int flux, int fp32, int proj, int variant, const BatchCase* cases,
                              const unsigned* blk_case, unsigned nblocks, int stage, const KTables* t, hipStream_t st) {
    if (nblocks == 0) return hipSuccess;
    DG_DISPATCH_P(P, return stage_batch_p<PP>(flux, fp32, proj, variant, cases, blk_case, nblocks, stage, t, st));
    return hipSuccess;
}

int batch_team(int P, int variant) {
    switch (P) {
    case 1: return team_of<1>(variant);
    case 2: return team_of<2>(variant);
    case 3: return team_of<3>(variant);
    case 4: return team_of<4>(variant);
    case 5: return team_of<5>(variant);
    default: return 0;
    }
}

// Zero n bytes of device memory (n a multiple of 4) with a kernel on the
// stream: a fill kernel, unlike a memset, is a plain kernel node when the
// caller captures the stream into a graph
__global__ void zero_kernel(uint32_t* __restrict__ p, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t st) {
    if (bytes % 4 != 0) return hipErrorInvalidValue;
    const size_t n = bytes / 4;
    if (n == 0) return hipSuccess;
    const size_t want = (n + 255) / 256;
    const unsigned grid = (unsigned)(want < 8192 ? want : 8192);
    hipLaunchKernelGGL(zero_kernel, dim3(grid), dim3(256), 0, st, static_cast<uint32_t*>(p), n);
    return hipGetLastError();
}

hipError_t launch_inv_mass(int P, double dt, const double* R, double* dU, const KTables* t,
                           const KGeom& g, hipStream_t st) {
    if (g.dim == 2) {
        DG_DISPATCH_P(P, hipLaunchKernelGGL((inv_mass2d_kernel<PP>), dim3(g.E), dim3(WAVE), 0, st, dt, R, dU, t, g));
        return hipGetLastError();
    }
    DG_DISPATCH_P(P, hipLaunchKernelGGL((inv_mass_kernel<PP>), dim3(g.E), dim3(WAVE), 0, st, dt, R, dU, t, g));
    return hipGetLastError();
}

hipError_t launch_init_state(int P, double time, double* U, const KTables* t, const KGeom& g,
                             const KIC& ic, hipStream_t st) {
    if (g.dim == 2) {
        DG_DISPATCH_P(P, hipLaunchKernelGGL((init2d_kernel<PP>), dim3(g.E), dim3(WAVE), 0, st, time, U, t, g, ic));
        return hipGetLastError();
    }
    DG_DISPATCH_P(P, hipLaunchKernelGGL((init_state_kernel<PP>), dim3(g.E), dim3(WAVE), 0, st, time, U, t, g, ic));
    return hipGetLastError();
}

hipError_t launch_eval(int P, int mode, double time, const double* U, double* out, const KTables* t,
                       const KGeom& g, const KIC& ic, hipStream_t st) {
    if (g.dim == 2) {
        DG_DISPATCH_P(P, hipLaunchKernelGGL((eval2d_kernel<PP>), dim3(g.E), dim3(WAVE), 0, st, mode, time, U, out, t, g,
                                            ic));
        return hipGetLastError();
    }
    DG_DISPATCH_P(P, hipLaunchKernelGGL((eval_kernel<PP>), dim3(g.E), dim3(WAVE), 0, st, mode, time, U, out, t, g, ic));
    return hipGetLastError();
}

hipError_t launch_wavespeed(int P, const double* U, double* out, const KTables* t, const KGeom& g,
                            hipStream_t st) {
    if (g.dim == 2) {
        DG_DISPATCH_P(P, hipLaunchKernelGGL((wavespeed2d_kernel<PP>), dim3(g.E), dim3(WAVE), 0, st, U, out, t, g));
        return hipGetLastError();
    }
    DG_DISPATCH_P(P, hipLaunchKernelGGL((wavespeed_kernel<PP>), dim3(g.E), dim3(WAVE), 0, st, U, out, t, g));
    return hipGetLastError();
}

hipError_t launch_mass(int P, double* imm, double* jmin, const KTables* t, const KGeom& g, hipStream_t st) {
    if (g.dim == 2) {
        DG_DISPATCH_P(P, hipLaunchKernelGGL((mass2d_kernel<PP>), dim3(g.E), dim3(WAVE), 0, st, imm, jmin, t, g));
        return hipGetLastError();
    }
    const dim3 grid(g.E), block(256);
    switch (P) {
    case 1: hipLaunchKernelGGL((mass_kernel<1>), grid, block, 0, st, imm, jmin, t, g); break;
    case 2: hipLaunchKernelGGL((mass_kernel<2>), grid, block, 0, st, imm, jmin, t, g); break;
    case 3: hipLaunchKernelGGL((mass_kernel<3>), grid, block, 0, st, imm, jmin, t, g); break;
    case 4: hipLaunchKernelGGL((mass_kernel<4>), grid, block, 0, st, imm, jmin, t, g); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace dg
