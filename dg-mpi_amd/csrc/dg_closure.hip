// Learned closure for coarse (p = 1 Legendre) periodic-box runs: per-element
// features and training targets, a small MLP (70 -> 128 -> 128 -> 40, GELU,
// BF16 weights with FP32 accumulation) with its own forward, backward and
// Adam kernels, and RK4 stepping of the coarse solver with the corrected
// tendency dU/dt = M^-1 R(U) + C(U).  No reference counterpart.
#include "dg_amd.h"

#include <hip/hip_runtime.h>

#include <mutex>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dg_solver_impl.hpp"

namespace {

constexpr int NF = 70, NH = 128, NO = 40;      // features, hidden width, outputs
constexpr int NP = NF * NH + NH + NH * NH + NH + NH * NO + NO;
constexpr int OFF_W1 = 0, OFF_B1 = NF * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NO;
constexpr int TB = 256;

struct ClosureError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::ostringstream o;
        o << what << ": " << hipGetErrorString(e);
        throw ClosureError(o.str());
    }
}

unsigned nblk(long n) { return (unsigned)((n + TB - 1) / TB); }

// ---------------------------------------------------------------------------
// bf16 storage (round to nearest even), FP32 arithmetic
// ---------------------------------------------------------------------------
__host__ __device__ inline uint16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// output gradient of the mean squared error in normalised target units:
// G = 2 (O - Y / yscale) / (M NO); per-block partial sums of the squared error
__global__ void loss_grad(const float* __restrict__ O, const float* __restrict__ Y,
                          const float* __restrict__ yscale, float* __restrict__ G, double* __restrict__ part,
                          long M) {
    __shared__ double red[TB];
    const long t = (long)blockIdx.x * TB + threadIdx.x;
    double e2 = 0.;
    if (t < M * NO) {
        const int n = (int)(t % NO);
        const float d = O[t] - Y[t] / yscale[n];
        G[t] = 2.0f * d / (float)(M * NO);
        e2 = (double)d * d;
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = TB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// rows idx[i] of X (NF) and Y (NO) into contiguous mini-batch buffers
__global__ void gather_rows(const float* __restrict__ X, const float* __restrict__ Y, const int* __restrict__ idx,
                            float* __restrict__ Xb, float* __restrict__ Yb, long B, long nrows) {
    const long t = (long)blockIdx.x * TB + threadIdx.x;
    if (t >= B * (NF + NO)) return;
    const long i = t / (NF + NO);
    const int c = (int)(t % (NF + NO));
    long r = idx[i];
    r = r < 0 ? 0 : (r >= nrows ? nrows - 1 : r);
    if (c < NF) Xb[i * NF + c] = X[r * NF + c];
    else Yb[i * NO + (c - NF)] = Y[r * NO + (c - NF)];
}

// Gaussian jitter of the mini-batch features, noise standard deviations in
// normalised units (raw units: noise / scale); counter-based hash RNG
// (splitmix64) and the Box-Muller transform
__device__ inline unsigned long long mix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__global__ void jitter_kernel(float* __restrict__ Xb, const float* __restrict__ scale, float noise, long n,
                              unsigned long long key) {
    const long t = (long)blockIdx.x * TB + threadIdx.x;
    if (t >= n) return;
    const unsigned long long h = mix64(key ^ mix64((unsigned long long)t));
    const float u1 = ((float)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float u2 = ((float)((h >> 16) & 0xffffffu) + 0.5f) * (1.0f / 16777216.0f);
    const float g = sqrtf(-2.0f * logf(u1)) * cosf(6.283185307179586f * u2);
    Xb[t] += noise * g / scale[t % NF];
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters; the BF16 working
// copy follows
__global__ void adam_step(float* __restrict__ p, uint16_t* __restrict__ pb, const float* __restrict__ g,
                          float* __restrict__ m1, float* __restrict__ m2, float lr, float b1, float b2, float eps,
                          float c1, float c2, int n) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    const float q = p[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    p[i] = q;
    pb[i] = to_bf16(q);
}

__global__ void to_bf16_kernel(const float* __restrict__ p, uint16_t* __restrict__ pb, int n) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i < n) pb[i] = to_bf16(p[i]);
}

// ---------------------------------------------------------------------------
// Coarse-state features and targets (p = 1 Legendre periodic box)
// ---------------------------------------------------------------------------

// face mean of element block u (8 x 5) on face (d, side), variable s:
// (1/4) sum_ab w_a w_b U(face point (a, b)) on the (p+2) = 3-point rule,
// factored as sum_b u_b phi_r(+-1) m_t1 m_t2 with the basis means
// m_j = (1/2) sum_a w_a phi_j(x_a)
__device__ double face_mean(const dg::KTables& T, const double* m, const double* u, int d, int side, int s) {
    const double* phi = side ? T.phip : T.phim;
    double acc = 0.;
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) {
                const int b = i + 2 * j + 4 * k;
                const int r = d == 0 ? i : (d == 1 ? j : k);
                const int t1 = d == 0 ? j : i, t2 = d == 2 ? j : k;
                acc = fma(u[b * 5 + s], phi[r] * m[t1] * m[t2], acc);
            }
    return acc;
}

// X[e] = own 40 coefficients, then for the faces x-, x+, y-, y+, z-, z+ the
// jump of the face-mean state (neighbour's matching face minus own), 5
// each.  32 threads per element (32-bit index arithmetic: E 32 < 2^32):
// all copy the 40 coefficients, threads 0..29 then take the 30 jumps
constexpr int FT = 32;
__global__ void features_kernel(const double* __restrict__ U, float* __restrict__ X,
                                const dg::KTables* __restrict__ Tp, dg::KGeom G) {
    const unsigned t = blockIdx.x * TB + threadIdx.x;
    const unsigned e = t / FT, q = t % FT;
    if (e >= G.E) return;
    const double* u = U + (size_t)e * 40;
    float* x = X + (size_t)e * NF;
    x[q] = (float)u[q];
    if (q < 40 - FT) x[FT + q] = (float)u[FT + q];
    if (q >= 30) return;
    const int c = 40 + (int)q;
    const dg::KTables& T = *Tp;
    double m[2];
    for (int jj = 0; jj < 2; ++jj) {
        double a = 0.;
        for (int g = 0; g < 3; ++g) a += T.w[g] * T.B[g][jj];
        m[jj] = 0.5 * a;
    }
    const int f = (c - 40) / 5, s = (c - 40) % 5, d = f >> 1, side = f & 1;
    const unsigned nx = G.ne[0], ny = G.ne[1], nz = G.ne[2];
    const unsigned eyz = e / nx;
    unsigned cc[3] = {e - eyz * nx, eyz % ny, eyz / ny};
    const unsigned nd = d == 0 ? nx : (d == 1 ? ny : nz);
    cc[d] = side ? (cc[d] + 1 == nd ? 0u : cc[d] + 1) : (cc[d] == 0 ? nd - 1 : cc[d] - 1);
    const unsigned en = cc[0] + nx * (cc[1] + ny * cc[2]);
    x[c] = (float)(face_mean(T, m, U + (size_t)en * 40, d, 1 - side, s) - face_mean(T, m, u, d, side, s));
}

// coarse state = Legendre modes (i, j, k) <= 1 of the fine state (the L2
// projection onto p = 1 for an orthogonal basis on affine elements)
__global__ void truncate_kernel(const double* __restrict__ Uf, double* __restrict__ Uc, int Pf, long E) {
    const long t = (long)blockIdx.x * TB + threadIdx.x;
    if (t >= E * 40) return;
    const long e = t / 40;
    const int r = (int)(t % 40), b = r / 5, s = r % 5;
    const int i = b & 1, j = (b >> 1) & 1, k = b >> 2, n = Pf + 1;
    Uc[t] = Uf[(e * n * n * n + i + n * j + n * n * k) * 5 + s];
}

// Y = trunc(Df) - Dc, in FP32
__global__ void target_kernel(const double* __restrict__ Df, const double* __restrict__ Dc, float* __restrict__ Y,
                              int Pf, long E) {
    const long t = (long)blockIdx.x * TB + threadIdx.x;
    if (t >= E * 40) return;
    const long e = t / 40;
    const int r = (int)(t % 40), b = r / 5, s = r % 5;
    const int i = b & 1, j = (b >> 1) & 1, k = b >> 2, n = Pf + 1;
    Y[t] = (float)(Df[(e * n * n * n + i + n * j + n * n * k) * 5 + s] - Dc[t]);
}

// Fused inference for the closure step and dg_mlp_forward, on the matrix
// cores: v_mfma_f32_16x16x16_bf16 tiles (bf16 operands, FP32 accumulation)
// for the three layers, with the normalisation, GELU, de-normalisation and
// the fallback bound around them.  The grid is persistent (at most one
// block per CU): a block stages the three weight matrices once, transposed
// to [output][input] in LDS, then walks 64-row tiles; wave w owns rows
// 16w .. 16w+15 of a tile through all three layers.  Activations enter each
// layer rounded to bf16.
//
// Operand layout of the 16x16x16 bf16 MFMA (lane l, element q = 0..3):
// A[l % 16][4 (l / 16) + q], B[4 (l / 16) + q][l % 16]; the result lane l
// holds D[4 (l / 16) + q][l % 16].
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int MT = 256, MR = 64;               // threads, rows per tile
constexpr int K1 = 80;                         // layer-1 inputs padded to whole 16-wide k steps
constexpr int L1S = K1 + 4, L2S = NH + 4;      // LDS row strides (bf16), padded off the bank period
constexpr int N3 = 48;                         // layer-3 outputs padded to whole 16-column tiles
constexpr int MAX_INFER_BLOCKS = 256;
constexpr int GCHUNKS = 32;                   // most row chunks of one weight-gradient pass

struct InferLDS {
    uint16_t w1[NH][L1S];                      // W1^T; inputs k >= NF are zero
    uint16_t w2[NH][L2S];                      // W2^T
    uint16_t w3[N3][L2S];                      // W3^T; outputs n >= NO are zero
    uint16_t a0[MR][L1S];                      // normalised features
    uint16_t h1[MR][L2S];
    uint16_t h2[MR][L2S];
    float o[MR][NO + 1];                       // de-normalised outputs
};

__device__ inline bf16x4 ld_bf16x4(const uint16_t* q) { return *reinterpret_cast<const bf16x4*>(q); }

// The weights' LDS image (w1, w2, w3 of InferLDS, contiguous) in global
// memory, built once per change of the bf16 weights: element i of the image
// is W1^T / W2^T / W3^T at (row, column) with the padding zero
constexpr int IMG1 = NH * L1S, IMG2 = NH * L2S, IMG3 = N3 * L2S, IMG = IMG1 + IMG2 + IMG3;

__global__ void pack_infer_weights(const uint16_t* __restrict__ W, uint16_t* __restrict__ img) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= IMG) return;
    uint16_t v = 0;
    if (i < IMG1) {
        const int n = i / L1S, k = i % L1S;
        if (k < NF) v = W[OFF_W1 + k * NH + n];
    } else if (i < IMG1 + IMG2) {
        const int n = (i - IMG1) / L2S, k = (i - IMG1) % L2S;
        if (k < NH) v = W[OFF_W2 + k * NH + n];
    } else {
        const int n = (i - IMG1 - IMG2) / L2S, k = (i - IMG1 - IMG2) % L2S;
        if (n < NO && k < NH) v = W[OFF_W3 + k * NO + n];
    }
    img[i] = v;
}

// acc[j] += A[16 rows][K] x W^T[16 j .. 16 j + 15][K] for j < NT
template <int NT, int K, int LA, int LW>
__device__ __forceinline__ void mfma_layer(const uint16_t (*A)[LA], const uint16_t (*Wt)[LW], int rw, int lane,
                                           f32x4* acc) {
    const int c = lane & 15, kq = 4 * (lane >> 4);
#pragma unroll
    for (int kk = 0; kk < K; kk += 16) {
        const bf16x4 a = ld_bf16x4(&A[rw + c][kk + kq]);
#pragma unroll
        for (int j = 0; j < NT; ++j)
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, ld_bf16x4(&Wt[16 * j + c][kk + kq]), acc[j], 0, 0,
                                                              0);
    }
}

__global__ __launch_bounds__(MT) void mlp_infer_kernel(const float* __restrict__ X,
                                                       const uint16_t* __restrict__ Wimg,
                                                       const float* __restrict__ P, const float* __restrict__ xs,
                                                       float* __restrict__ C, double bound, long rows,
                                                       unsigned long long* __restrict__ count) {
    __shared__ alignas(16) InferLDS L;
    const int t = threadIdx.x, lane = t & 63, rw = 16 * (t >> 6);
    const int c = lane & 15, rq = 4 * (lane >> 4);
    // the weights' LDS image, a straight 16-byte copy
    static_assert(offsetof(InferLDS, w2) == IMG1 * 2 && offsetof(InferLDS, w3) == (IMG1 + IMG2) * 2 &&
                      IMG * 2 % 16 == 0,
                  "contiguous weight image");
    {
        constexpr int NV = IMG * 2 / 16;
        const uint4* src = reinterpret_cast<const uint4*>(Wimg);
        uint4* dst = reinterpret_cast<uint4*>(&L.w1[0][0]);
        for (int i = t; i < NV; i += 4 * MT) {
            // four loads in flight before the stores
            const int i1 = i + MT, i2 = i + 2 * MT, i3 = i + 3 * MT;
            const uint4 v0 = src[i];
            const uint4 v1 = i1 < NV ? src[i1] : v0, v2 = i2 < NV ? src[i2] : v0, v3 = i3 < NV ? src[i3] : v0;
            dst[i] = v0;
            if (i1 < NV) dst[i1] = v1;
            if (i2 < NV) dst[i2] = v2;
            if (i3 < NV) dst[i3] = v3;
        }
    }
    const long ntile = (rows + MR - 1) / MR;
    for (long tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const long r0 = tile * MR;
        const int nr = (int)(rows - r0 < MR ? rows - r0 : MR);
        __syncthreads();                       // the staged weights; the previous tile's a0 and o are read
        {
            // the tile's features: all global loads first, then the stores
            constexpr int PER = (MR * K1 + MT - 1) / MT;
            float xv[PER];
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int i = t + u * MT, r = i / K1, k = i % K1;
                xv[u] = (i < MR * K1 && r < nr && k < NF) ? X[(r0 + r) * NF + k] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int i = t + u * MT, r = i / K1, k = i % K1;
                if (i < MR * K1) L.a0[r][k] = (r < nr && k < NF) ? to_bf16((xv[u] - xs[k]) * xs[NF + k]) : (uint16_t)0;
            }
        }
        __syncthreads();
        f32x4 acc[8];
        // layer 1: NF -> NH, GELU
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float b = P[OFF_B1 + 16 * j + c];
            acc[j] = f32x4{b, b, b, b};
        }
        mfma_layer<8, K1, L1S, L1S>(L.a0, L.w1, rw, lane, acc);
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) L.h1[rw + rq + q][16 * j + c] = to_bf16(gelu(acc[j][q]));
        __syncthreads();
        // layer 2: NH -> NH, GELU
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float b = P[OFF_B2 + 16 * j + c];
            acc[j] = f32x4{b, b, b, b};
        }
        mfma_layer<8, NH, L2S, L2S>(L.h1, L.w2, rw, lane, acc);
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) L.h2[rw + rq + q][16 * j + c] = to_bf16(gelu(acc[j][q]));
        __syncthreads();
        // layer 3: NH -> NO, de-normalised
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int n = 16 * j + c;
            const float b = n < NO ? P[OFF_B3 + n] : 0.f;
            acc[j] = f32x4{b, b, b, b};
        }
        mfma_layer<3, NH, L2S, L2S>(L.h2, L.w3, rw, lane, acc);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int n = 16 * j + c;
            if (n < NO) {
                const float ys = xs[2 * NF + n];
#pragma unroll
                for (int q = 0; q < 4; ++q) L.o[rw + rq + q][n] = acc[j][q] * ys;
            }
        }
        __syncthreads();
        // fallback bound per row (count null: no bound), then the store
        if (count && t < MR) {
            double n2 = 0.;
            for (int i = 0; i < NO; ++i) n2 += (double)L.o[t][i] * L.o[t][i];
            if (t < nr && !(sqrt(n2) <= bound)) {
                for (int i = 0; i < NO; ++i) L.o[t][i] = 0.f;
                atomicAdd(count, 1ull);
            }
        }
        __syncthreads();
        for (int i = t; i < nr * NO; i += MT) C[r0 * NO + i] = L.o[i / NO][i % NO];
    }
}

// ---------------------------------------------------------------------------
// Training on the matrix cores: FP32 GEMM tiles (v_mfma_f32_16x16x4f32,
// FP32 operands and accumulation).  In one 16x16x4 step lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[4 (l / 16) + r][l % 16],
// r = 0..3, of the 16x16 result.  The layer inputs and the gradients stay
// FP32 and the weights are the bf16 working copy widened exactly, so the
// arithmetic is that of a plain FP32 network with bf16 weights up to the
// order of the FP32 sums.
// ---------------------------------------------------------------------------
constexpr int GW = 4;                          // waves per training workgroup
constexpr int WV = 64;                         // lanes per wave
constexpr int NJH = 2;                         // 16-column tiles per wave of the 128-wide layers

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Z[m][n] = sum_k A'[m][k] W[k][n] + b[n] and H = act(Z) (Z kept when
// non-null), A' = (A - mean) scale when mean is non-null.  Each wave takes
// 16 rows and NJ 16-column tiles of the N outputs, the tiles of column
// group blockIdx.y (columns >= N masked).
template <bool ACT, int NJ>
__global__ __launch_bounds__(GW * WV) void mfma_fwd(const float* __restrict__ A, const float* __restrict__ mean,
                                                    const float* __restrict__ scale, const uint16_t* __restrict__ W,
                                                    const float* __restrict__ b, float* __restrict__ Z,
                                                    float* __restrict__ H, long M, int K, int N) {
    const int lane = threadIdx.x & (WV - 1), c = lane & 15, kq = lane >> 4;
    const long m0 = ((long)blockIdx.x * GW + (threadIdx.x / WV)) * 16;
    if (m0 >= M) return;                       // whole waves; no barrier follows
    const long ma = m0 + c;
    const bool rowok = ma < M;
    const int n0 = blockIdx.y * 16 * NJ;
    f32x4 acc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int n = n0 + 16 * j + c;
        const float bn = n < N ? b[n] : 0.f;
        acc[j] = f32x4{bn, bn, bn, bn};
    }
    for (int k0 = 0; k0 < K; k0 += 4) {
        const int k = k0 + kq;
        float a = 0.f;
        if (rowok && k < K) {
            a = A[ma * K + k];
            if (mean) a = (a - mean[k]) * scale[k];
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int n = n0 + 16 * j + c;
            const float w = (k < K && n < N) ? from_bf16(W[(size_t)k * N + n]) : 0.f;
            acc[j] = mfma4(a, w, acc[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int n = n0 + 16 * j + c;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const long m = m0 + 4 * kq + r;
            if (m < M && n < N) {
                const float z = acc[j][r];
                if (Z) Z[m * N + n] = z;
                H[m * N + n] = ACT ? gelu(z) : z;
            }
        }
    }
}

// Gin[m][k] = (sum_n G[m][n] W[k][n]) gelu'(Z[m][k]) for the KO inputs k
// of a layer with N outputs; each wave takes 16 rows and the NJ 16-column
// tiles of column group blockIdx.y
template <int KO, int NJ>
__global__ __launch_bounds__(GW * WV) void mfma_bwd_in(const float* __restrict__ G, const uint16_t* __restrict__ W,
                                                       const float* __restrict__ Z, float* __restrict__ Gin, long M,
                                                       int N) {
    static_assert(KO % (16 * NJ) == 0, "whole column groups");
    const int lane = threadIdx.x & (WV - 1), c = lane & 15, kq = lane >> 4;
    const int k0 = blockIdx.y * 16 * NJ;
    const long m0 = ((long)blockIdx.x * GW + (threadIdx.x / WV)) * 16;
    if (m0 >= M) return;
    const long ma = m0 + c;
    const bool rowok = ma < M;
    f32x4 acc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int n0 = 0; n0 < N; n0 += 4) {
        const int n = n0 + kq;
        const float g = (rowok && n < N) ? G[ma * N + n] : 0.f;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int k = k0 + 16 * j + c;
            const float w = n < N ? from_bf16(W[(size_t)k * N + n]) : 0.f;
            acc[j] = mfma4(g, w, acc[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int k = k0 + 16 * j + c;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const long m = m0 + 4 * kq + r;
            if (m < M) Gin[m * KO + k] = acc[j][r] * gelu_d(Z[m * KO + k]);
        }
    }
}

// Weight and bias gradients of a layer with K inputs and N outputs, partial
// over one chunk of rows: part[s][k][n] = sum_{m in chunk s} A'[m][k]
// G[m][n], with the row k = K the bias (A' = 1); A' normalised as in
// mfma_fwd when mean is non-null.  Workgroup (16-row k tile, chunk s,
// column group of NJ 16-column tiles; columns >= N masked); its waves split
// the chunk's rows and add their tiles in LDS, in a fixed order.
template <int NJ>
__global__ __launch_bounds__(GW * WV) void mfma_wgrad(const float* __restrict__ A, const float* __restrict__ mean,
                                                      const float* __restrict__ scale, const float* __restrict__ G,
                                                      float* __restrict__ part, long M, int K, int N, long chunk) {
    __shared__ f32x4 red[GW][NJ][WV];
    const int wv = threadIdx.x / WV, lane = threadIdx.x & (WV - 1), c = lane & 15, kq = lane >> 4;
    const int k0 = blockIdx.x * 16, k = k0 + c, n0 = blockIdx.z * 16 * NJ;
    const long mb = (long)blockIdx.y * chunk, me = mb + chunk < M ? mb + chunk : M;
    f32x4 acc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long m4 = mb + 4 * wv; m4 < me; m4 += 4 * GW) {
        const long m = m4 + kq;
        const bool ok = m < me;
        float a = 0.f;
        if (ok && k <= K) {
            if (k == K) {
                a = 1.f;
            } else {
                a = A[m * K + k];
                if (mean) a = (a - mean[k]) * scale[k];
            }
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int n = n0 + 16 * j + c;
            const float g = (ok && n < N) ? G[m * N + n] : 0.f;
            acc[j] = mfma4(a, g, acc[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) red[wv][j][lane] = acc[j];
    __syncthreads();
    for (int i = threadIdx.x; i < NJ * WV; i += GW * WV) {
        const int j = i / WV, l = i % WV;
        f32x4 v = red[0][j][l];
#pragma unroll
        for (int w = 1; w < GW; ++w) v += red[w][j][l];
        const int n = n0 + 16 * j + (l & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int kk = k0 + 4 * (l >> 4) + r;
            if (kk <= K && n < N) part[(size_t)blockIdx.y * NP + (size_t)kk * N + n] = v[r];
        }
    }
}

// g[i] = sum_s part[s][i] over the row chunks (fixed order)
__global__ void grad_reduce(const float* __restrict__ part, float* __restrict__ g, int nchunk) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= NP) return;
    float v = 0.f;
    for (int s = 0; s < nchunk; ++s) v += part[(size_t)s * NP + i];
    g[i] = v;
}

// acc += (sum of the nb partial squared errors) * inv, one block, fixed order
__global__ void loss_add(const double* __restrict__ part, long nb, double inv, double* __restrict__ acc) {
    __shared__ double red[TB];
    double v = 0.;
    for (long i = threadIdx.x; i < nb; i += TB) v += part[i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = TB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += red[0] * inv;
}

__global__ void clear_f64_kernel(double* p) { p[threadIdx.x] = 0.; }

// Y = trunc(Df) - Dc + (trunc(Uf) - Uc) / tau, for a coarse state Uc of
// its own (a coarse run beside the fine one): the tendency mismatch at Uc
// plus a relaxation of Uc toward the fine state's projection
__global__ void target_at_kernel(const double* __restrict__ Df, const double* __restrict__ Dc,
                                 const double* __restrict__ Uf, const double* __restrict__ Uc, double rtau,
                                 float* __restrict__ Y, int Pf, long E) {
    const long t = (long)blockIdx.x * TB + threadIdx.x;
    if (t >= E * 40) return;
    const long e = t / 40;
    const int r = (int)(t % 40), b = r / 5, s = r % 5;
    const int i = b & 1, j = (b >> 1) & 1, k = b >> 2, n = Pf + 1;
    const long f = (e * n * n * n + i + n * j + n * n * k) * 5 + s;
    Y[t] = (float)(Df[f] - Dc[t] + (Uf[f] - Uc[t]) * rtau);
}

// Device scratch of dg_closure_step_noalloc: the feature rows [E][NF], the
// corrections [E][NO] and the 64-bit fallback counter of up to
// NOALLOC_MAX_ELEMS elements (32^3, the closure's box), 14.4 MB, in a
// file-scope array that each GPU holds until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is a stand-in,
// not a device address): it uses the address hipGetSymbolAddress returns
// for the current device (noalloc_scratch).
constexpr long NOALLOC_MAX_ELEMS = 32768;
constexpr size_t NOALLOC_X = (size_t)NOALLOC_MAX_ELEMS * NF;        // floats
constexpr size_t NOALLOC_C = (size_t)NOALLOC_MAX_ELEMS * NO;        // floats
constexpr size_t NOALLOC_FLOATS = NOALLOC_X + NOALLOC_C + 2;        // + the counter (8 B, 8-aligned)
static_assert(((NOALLOC_X + NOALLOC_C) * sizeof(float)) % 8 == 0, "counter alignment");
static __device__ __align__(256) float scratch[NOALLOC_FLOATS];

// one thread clears *p (a vector store: the address is lane-indexed)
__global__ void clear_u64_kernel(unsigned long long* p) { p[threadIdx.x] = 0ull; }

thread_local std::string g_mlp_error;

}  // namespace

// ---------------------------------------------------------------------------
// MLP object
// ---------------------------------------------------------------------------
struct dg_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float* p = nullptr;          // master parameters [NP]
    uint16_t* pb = nullptr;      // BF16 working copy
    uint16_t* img = nullptr;     // the inference kernel's LDS image of the bf16 weights
    bool img_stale = true;       // pb changed since img was built
    float* g = nullptr;          // gradient
    float* m1 = nullptr;         // Adam moments
    float* m2 = nullptr;
    float* xs = nullptr;         // feature mean [NF], scale [NF], target scale [NO]
    long cap = 0;                // rows the activation buffers hold
    float *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *O = nullptr, *G3 = nullptr,
          *G2 = nullptr, *G1 = nullptr, *Xb = nullptr, *Yb = nullptr;
    double* part = nullptr;
    float* gpart = nullptr;      // per-chunk weight-gradient partials [GCHUNKS][NP]
    double* lacc = nullptr;      // training loss accumulator (device)
    long step = 0;
    std::string err;
};

namespace {

void free_acts(dg_mlp* m) {
    for (float* q : {m->Z1, m->H1, m->Z2, m->H2, m->O, m->G3, m->G2, m->G1, m->Xb, m->Yb})
        if (q) (void)hipFree(q);
    if (m->part) (void)hipFree(m->part);
    if (m->gpart) (void)hipFree(m->gpart);
    if (m->lacc) (void)hipFree(m->lacc);
    m->Z1 = m->H1 = m->Z2 = m->H2 = m->O = m->G3 = m->G2 = m->G1 = m->Xb = m->Yb = nullptr;
    m->part = nullptr;
    m->gpart = nullptr;
    m->lacc = nullptr;
    m->cap = 0;
}

// the inference kernel on a persistent grid of at most one block per CU
// (the weight image is rebuilt first, on the same stream, when pb changed)
void launch_infer(dg_mlp* m, const float* X, float* C, double bound, long rows, unsigned long long* count,
                  hipStream_t st) {
    if (m->img_stale) {
        hipLaunchKernelGGL(pack_infer_weights, dim3(nblk(IMG)), dim3(TB), 0, st, m->pb, m->img);
        hcheck(hipGetLastError(), "weight image kernel");
        m->img_stale = false;
    }
    const long ntile = (rows + MR - 1) / MR;
    const unsigned grid = (unsigned)(ntile < MAX_INFER_BLOCKS ? ntile : MAX_INFER_BLOCKS);
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid), dim3(MT), 0, st, X, m->img, m->p, m->xs, C, bound, rows,
                       count);
    hcheck(hipGetLastError(), "inference kernel");
}

void ensure_acts(dg_mlp* m, long rows) {
    if (rows <= m->cap) return;
    free_acts(m);
    auto al = [&](float** q, long n) { hcheck(hipMalloc(q, (size_t)n * sizeof(float)), "hipMalloc(mlp)"); };
    al(&m->Z1, rows * NH);
    al(&m->H1, rows * NH);
    al(&m->Z2, rows * NH);
    al(&m->H2, rows * NH);
    al(&m->O, rows * NO);
    al(&m->G3, rows * NO);
    al(&m->G2, rows * NH);
    al(&m->G1, rows * NH);
    al(&m->Xb, rows * NF);
    al(&m->Yb, rows * NO);
    hcheck(hipMalloc(&m->part, (size_t)nblk(rows * NO) * sizeof(double)), "hipMalloc(mlp)");
    hcheck(hipMalloc(&m->gpart, (size_t)GCHUNKS * NP * sizeof(float)), "hipMalloc(mlp)");
    hcheck(hipMalloc(&m->lacc, sizeof(double)), "hipMalloc(mlp)");
    m->cap = rows;
}

// the weight-gradient row chunks of a batch of M rows: at most GCHUNKS
// chunks of at least 64 rows
long grad_chunk(long M) {
    const long c = (M + GCHUNKS - 1) / GCHUNKS;
    return c < 64 ? 64 : c;
}

// forward pass of rows X (raw features) into m->O (normalised outputs)
void forward(dg_mlp* m, const float* X, long rows) {
    hipStream_t st = m->stream;
    const float* mean = m->xs;
    const float* scale = m->xs + NF;
    const unsigned rt = (unsigned)((rows + 16 * GW - 1) / (16 * GW));
    const dim3 gh(rt, NH / (16 * NJH)), go(rt, (NO + 15) / 16), block(GW * WV);
    hipLaunchKernelGGL((mfma_fwd<true, NJH>), gh, block, 0, st, X, mean, scale, m->pb + OFF_W1, m->p + OFF_B1,
                       m->Z1, m->H1, rows, NF, NH);
    hipLaunchKernelGGL((mfma_fwd<true, NJH>), gh, block, 0, st, m->H1, nullptr, nullptr, m->pb + OFF_W2,
                       m->p + OFF_B2, m->Z2, m->H2, rows, NH, NH);
    hipLaunchKernelGGL((mfma_fwd<false, 1>), go, block, 0, st, m->H2, nullptr, nullptr,
                       m->pb + OFF_W3, m->p + OFF_B3, nullptr, m->O, rows, NH, NO);
    hcheck(hipGetLastError(), "mlp forward");
}

// loss and gradient (into m->g) of rows X, Y: the batch's mean squared
// error is added to m->lacc on the device (no host synchronisation)
void backward(dg_mlp* m, const float* X, const float* Y, long rows) {
    hipStream_t st = m->stream;
    const float* mean = m->xs;
    const float* scale = m->xs + NF;
    const float* ys = m->xs + 2 * NF;
    forward(m, X, rows);
    const unsigned nb = nblk(rows * NO);
    hipLaunchKernelGGL(loss_grad, dim3(nb), dim3(TB), 0, st, m->O, Y, ys, m->G3, m->part, rows);
    hipLaunchKernelGGL(loss_add, dim3(1), dim3(TB), 0, st, m->part, (long)nb, 1.0 / (double)(rows * NO), m->lacc);
    const long chunk = grad_chunk(rows);
    const unsigned nch = (unsigned)((rows + chunk - 1) / chunk);
    const dim3 gh((unsigned)((rows + 16 * GW - 1) / (16 * GW)), NH / (16 * NJH)), block(GW * WV);
    hipLaunchKernelGGL((mfma_wgrad<1>), dim3((NH + 1 + 15) / 16, nch, (NO + 15) / 16), block, 0, st, m->H2,
                       nullptr, nullptr, m->G3, m->gpart + OFF_W3, rows, NH, NO, chunk);
    hipLaunchKernelGGL((mfma_bwd_in<NH, NJH>), gh, block, 0, st, m->G3, m->pb + OFF_W3, m->Z2, m->G2, rows, NO);
    hipLaunchKernelGGL((mfma_wgrad<NJH>), dim3((NH + 1 + 15) / 16, nch, NH / (16 * NJH)), block, 0, st, m->H1,
                       nullptr, nullptr, m->G2, m->gpart + OFF_W2, rows, NH, NH, chunk);
    hipLaunchKernelGGL((mfma_bwd_in<NH, NJH>), gh, block, 0, st, m->G2, m->pb + OFF_W2, m->Z1, m->G1, rows, NH);
    hipLaunchKernelGGL((mfma_wgrad<NJH>), dim3((NF + 1 + 15) / 16, nch, NH / (16 * NJH)), block, 0, st, X, mean,
                       scale, m->G1, m->gpart + OFF_W1, rows, NF, NH, chunk);
    hipLaunchKernelGGL(grad_reduce, dim3(nblk(NP)), dim3(TB), 0, st, m->gpart, m->g, (int)nch);
    hcheck(hipGetLastError(), "mlp backward");
}

// m->lacc = 0 (a kernel: no memset node on a captured stream)
void clear_loss(dg_mlp* m) {
    hipLaunchKernelGGL(clear_f64_kernel, dim3(1), dim3(1), 0, m->stream, m->lacc);
    hcheck(hipGetLastError(), "clear kernel");
}

double read_loss(dg_mlp* m) {
    double v = 0.;
    hcheck(hipMemcpyAsync(&v, m->lacc, sizeof(double), hipMemcpyDeviceToHost, m->stream), "hipMemcpy");
    hcheck(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
    return v;
}

template <class F>
int mlp_guard(std::string& err, F&& f) {
    try {
        f();
        err.clear();
        return DG_OK;
    } catch (const std::invalid_argument& e) {
        err = e.what();
        return DG_ERR_ARGUMENT;
    } catch (const dg::UnsupportedError& e) {
        err = e.what();
        return DG_ERR_UNSUPPORTED;
    } catch (const std::domain_error& e) {
        err = e.what();
        return DG_ERR_NUMERICS;
    } catch (const std::exception& e) {
        err = e.what();
        return DG_ERR_DEVICE;
    }
}

void need_ptr(const void* p, const char* what) {
    if (!p) throw std::invalid_argument(std::string(what) + " is NULL");
}

// the coarse / fine solvers this module works with: 3D Legendre periodic boxes
void check_box(const dg_solver* s, const char* what, int order) {
    need_ptr(s, what);
    if ((unsigned long long)s->geo.E * NF >= (1ull << 32))
        throw std::invalid_argument(std::string(what) + ": too many elements for the closure features");
    if (s->cfg.dim != 3 || s->geo.mapped || s->cfg.basis != dg::Basis::LegendreHex || s->geo.fp32)
        throw dg::UnsupportedError(std::string(what) + ": a 3D affine FP64 LegendreHex periodic box is required");
    if (order > 0 && s->P != order)
        throw dg::UnsupportedError(std::string(what) + ": order " + std::to_string(order) + " is required");
}

// dD = M^-1 R(U) of solver s (dt = 1), with its own stage registers as scratch
void tendency(dg_solver* s, const double* U, double* D) {
    dg::StageArgs a{};
    a.Ust = U;
    a.R = s->Ua;
    a.mode = dg::MODE_RESIDUAL;
    hcheck(dg::launch_stage(s->P, (int)s->cfg.flux, a, s->dtab, s->geo, s->stream), "stage kernel");
    hcheck(dg::launch_inv_mass(s->P, 1.0, s->Ua, D, s->dtab, s->geo, s->stream), "inverse mass kernel");
}

// The device address of `scratch` on the current device, looked up once
// per device (under a lock) and kept; no heap memory is involved.
constexpr int NOALLOC_MAX_DEVICES = 64;
float* noalloc_scratch(int device) {
    static std::mutex mu;
    static float* addr[NOALLOC_MAX_DEVICES] = {};
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw std::invalid_argument("device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        void* ptr = nullptr;
        hcheck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) throw std::runtime_error("hipGetSymbolAddress(scratch) returned NULL");
        addr[device] = static_cast<float*>(ptr);
    }
    return addr[device];
}

// argument checks of the closure step; returns the time step
double closure_check(dg_solver* s, dg_mlp* m, double dt_step, int nsteps) {
    check_box(s, "solver", 1);
    need_ptr(m, "mlp");
    if (nsteps < 0) throw std::invalid_argument("nsteps < 0");
    const double dt = dt_step > 0.0 ? dt_step : s->dt;
    if (!(dt > 0.0)) throw std::invalid_argument("no time step");
    if (s->cfg.stepper != dg::Stepper::RK4) throw dg::UnsupportedError("the closure steps RK4 only");
    hcheck(hipSetDevice(s->device), "hipSetDevice");
    return dt;
}

// nsteps corrected RK4 steps with the given feature rows X [E][NF],
// corrections C [E][NO] and fallback counter cnt (device memory)
void closure_steps(dg_solver* s, dg_mlp* m, double dt, double bound, int nsteps, float* X, float* C,
                   unsigned long long* cnt, long* n_fallback) {
    const long E = s->geo.E;
    // the counter is cleared by a kernel (no memset node if the caller
    // captures the solver's stream)
    hipLaunchKernelGGL(clear_u64_kernel, dim3(1), dim3(1), 0, s->stream, cnt);
    hcheck(hipGetLastError(), "clear kernel");
    // the network's parameters are written by its own stream (training)
    hcheck(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
    auto stage = [&](const double* Ust, double* Uout, int mode, double alpha) {
        // everything on the solver's stream: features, the network with
        // its fallback, and the stage kernel whose RK epilogue takes
        // dU = dt (M^-1 R + C)
        hipLaunchKernelGGL(features_kernel, dim3(nblk(E * FT)), dim3(TB), 0, s->stream, Ust, X, s->dtab, s->geo);
        launch_infer(m, X, C, bound, E, cnt, s->stream);
        dg::StageArgs a{};
        a.Ust = Ust;
        a.U0 = s->U0;
        a.Uout = Uout;
        a.acc = s->acc;
        a.cmp = s->cmp;
        a.dt = dt;
        a.alpha = alpha;
        a.mode = mode;
        hcheck(dg::launch_stage_corrected(s->P, (int)s->cfg.flux, a, C, s->dtab, s->geo, s->stream),
               "corrected stage kernel");
    };
    for (int k = 0; k < nsteps; ++k) {
        // the stage kernel reads the neighbours of Ust: stages ping-pong
        // U0 -> Ua -> Ub -> Ua -> U0, as in dg_step
        stage(s->U0, s->Ua, dg::MODE_RK_FIRST, 0.5);
        s->time += dt / 2.0;
        stage(s->Ua, s->Ub, dg::MODE_RK_MID, 0.5);
        stage(s->Ub, s->Ua, dg::MODE_RK_MID, 1.0);
        s->time += dt / 2.0;
        stage(s->Ua, s->U0, dg::MODE_RK_LAST, 0.0);
    }
    unsigned long long c = 0;
    hcheck(hipMemcpyAsync(&c, cnt, sizeof(c), hipMemcpyDeviceToHost, s->stream), "hipMemcpyAsync");
    hcheck(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    if (n_fallback) *n_fallback = (long)c;
}

}  // namespace

extern "C" {

int dg_mlp_create(unsigned long long seed, dg_mlp** out) {
    if (out) *out = nullptr;
    dg_mlp* m = nullptr;
    const int rc = mlp_guard(g_mlp_error, [&] {
        need_ptr(out, "out");
        m = new dg_mlp();
        hcheck(hipGetDevice(&m->device), "hipGetDevice");
        hcheck(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "hipStreamCreate");
        hcheck(hipMalloc(&m->p, NP * sizeof(float)), "hipMalloc(mlp)");
        hcheck(hipMalloc(&m->pb, NP * sizeof(uint16_t)), "hipMalloc(mlp)");
        hcheck(hipMalloc(&m->img, IMG * sizeof(uint16_t)), "hipMalloc(mlp)");
        hcheck(hipMalloc(&m->g, NP * sizeof(float)), "hipMalloc(mlp)");
        hcheck(hipMalloc(&m->m1, NP * sizeof(float)), "hipMalloc(mlp)");
        hcheck(hipMalloc(&m->m2, NP * sizeof(float)), "hipMalloc(mlp)");
        hcheck(hipMalloc(&m->xs, (2 * NF + NO) * sizeof(float)), "hipMalloc(mlp)");
        // He-normal hidden layers, a small output layer (the correction
        // starts near zero), zero biases; identity scaling
        std::mt19937_64 rng(seed);
        std::normal_distribution<float> nd(0.f, 1.f);
        std::vector<float> h(NP, 0.f);
        for (int i = 0; i < NF * NH; ++i) h[OFF_W1 + i] = nd(rng) * std::sqrt(2.f / NF);
        for (int i = 0; i < NH * NH; ++i) h[OFF_W2 + i] = nd(rng) * std::sqrt(2.f / NH);
        for (int i = 0; i < NH * NO; ++i) h[OFF_W3 + i] = nd(rng) * 0.1f * std::sqrt(1.f / NH);
        hcheck(hipMemcpy(m->p, h.data(), NP * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        std::vector<float> z(NP, 0.f);
        hcheck(hipMemcpy(m->m1, z.data(), NP * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        hcheck(hipMemcpy(m->m2, z.data(), NP * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        std::vector<float> xs(2 * NF + NO, 1.f);
        for (int i = 0; i < NF; ++i) xs[i] = 0.f;
        hcheck(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
        hcheck(hipMemcpy(m->xs, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk(NP)), dim3(TB), 0, m->stream, m->p, m->pb, NP);
        m->img_stale = true;
        hcheck(hipGetLastError(), "to_bf16");
        hcheck(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
    });
    if (rc != DG_OK) {
        if (m) dg_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return DG_OK;
}

int dg_mlp_destroy(dg_mlp* m) {
    if (!m) return DG_OK;
    (void)hipSetDevice(m->device);
    free_acts(m);
    for (float* q : {m->p, m->g, m->m1, m->m2, m->xs})
        if (q) (void)hipFree(q);
    if (m->pb) (void)hipFree(m->pb);
    if (m->img) (void)hipFree(m->img);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return DG_OK;
}

const char* dg_mlp_last_error(const dg_mlp* m) { return m ? m->err.c_str() : g_mlp_error.c_str(); }

int dg_mlp_num_params(long* n) {
    if (!n) return DG_ERR_ARGUMENT;
    *n = NP;
    return DG_OK;
}

int dg_mlp_get_params(const dg_mlp* m, float* host) {
    return mlp_guard(m ? const_cast<dg_mlp*>(m)->err : g_mlp_error, [&] {
        need_ptr(m, "mlp");
        need_ptr(host, "host");
        hcheck(hipMemcpy(host, m->p, NP * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

int dg_mlp_set_params(dg_mlp* m, const float* host) {
    return mlp_guard(m ? m->err : g_mlp_error, [&] {
        need_ptr(m, "mlp");
        need_ptr(host, "host");
        // work already queued on the network's stream may read the parameters
        hcheck(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
        hcheck(hipMemcpy(m->p, host, NP * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk(NP)), dim3(TB), 0, m->stream, m->p, m->pb, NP);
        m->img_stale = true;
        hcheck(hipGetLastError(), "to_bf16");
        hcheck(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
    });
}

int dg_mlp_set_scaling(dg_mlp* m, const float* x_mean, const float* x_scale, const float* y_scale) {
    return mlp_guard(m ? m->err : g_mlp_error, [&] {
        need_ptr(m, "mlp");
        need_ptr(x_mean, "x_mean");
        need_ptr(x_scale, "x_scale");
        need_ptr(y_scale, "y_scale");
        std::vector<float> xs(2 * NF + NO);
        for (int i = 0; i < NF; ++i) {
            xs[i] = x_mean[i];
            xs[NF + i] = x_scale[i];
        }
        for (int i = 0; i < NO; ++i) {
            if (!(y_scale[i] > 0.f)) throw std::invalid_argument("y_scale entries must be > 0");
            xs[2 * NF + i] = y_scale[i];
        }
        hcheck(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
        hcheck(hipMemcpy(m->xs, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    });
}

int dg_mlp_forward(dg_mlp* m, const float* X_dev, float* Y_dev, long rows) {
    return mlp_guard(m ? m->err : g_mlp_error, [&] {
        need_ptr(m, "mlp");
        need_ptr(X_dev, "X");
        need_ptr(Y_dev, "Y");
        if (rows < 1) throw std::invalid_argument("rows < 1");
        hcheck(hipSetDevice(m->device), "hipSetDevice");
        // the closure step's inference kernel, without the fallback bound
        launch_infer(m, X_dev, Y_dev, 0.0, rows, nullptr, m->stream);
        hcheck(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
    });
}

int dg_mlp_gradient(dg_mlp* m, const float* X_dev, const float* Y_dev, long rows, float* grad_host, double* loss) {
    return mlp_guard(m ? m->err : g_mlp_error, [&] {
        need_ptr(m, "mlp");
        need_ptr(X_dev, "X");
        need_ptr(Y_dev, "Y");
        if (rows < 1) throw std::invalid_argument("rows < 1");
        hcheck(hipSetDevice(m->device), "hipSetDevice");
        ensure_acts(m, rows);
        clear_loss(m);
        backward(m, X_dev, Y_dev, rows);
        const double l = read_loss(m);
        if (loss) *loss = l;
        if (grad_host) hcheck(hipMemcpy(grad_host, m->g, NP * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

int dg_mlp_train(dg_mlp* m, const float* X_dev, const float* Y_dev, long nrows, const int* idx_dev, long nidx,
                 int batch, float lr, float noise, double* mean_loss) {
    return mlp_guard(m ? m->err : g_mlp_error, [&] {
        need_ptr(m, "mlp");
        need_ptr(X_dev, "X");
        need_ptr(Y_dev, "Y");
        need_ptr(idx_dev, "idx");
        if (nrows < 1 || nidx < 1 || batch < 1) throw std::invalid_argument("empty training set or batch");
        if (nidx < batch) throw std::invalid_argument("fewer row indices than one mini-batch (nidx < batch)");
        hcheck(hipSetDevice(m->device), "hipSetDevice");
        ensure_acts(m, batch);
        clear_loss(m);
        long nb = 0;
        for (long i0 = 0; i0 + batch <= nidx; i0 += batch) {
            hipLaunchKernelGGL(gather_rows, dim3(nblk((long)batch * (NF + NO))), dim3(TB), 0, m->stream, X_dev,
                               Y_dev, idx_dev + i0, m->Xb, m->Yb, (long)batch, nrows);
            if (noise > 0.f)
                hipLaunchKernelGGL(jitter_kernel, dim3(nblk((long)batch * NF)), dim3(TB), 0, m->stream, m->Xb,
                                   m->xs + NF, noise, (long)batch * NF, (unsigned long long)m->step * 0x100000001ull);
            backward(m, m->Xb, m->Yb, batch);
            ++nb;
            ++m->step;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
            m->img_stale = true;
            hipLaunchKernelGGL(adam_step, dim3(nblk(NP)), dim3(TB), 0, m->stream, m->p, m->pb, m->g, m->m1, m->m2, lr,
                               b1, b2, 1e-8f, c1, c2, NP);
            hcheck(hipGetLastError(), "adam");
        }
        const double sum = read_loss(m);
        if (mean_loss) *mean_loss = nb ? sum / nb : 0.;
    });
}

int dg_closure_features(dg_solver* s, const double* U_dev, float* X_dev) {
    return mlp_guard(g_mlp_error, [&] {
        check_box(s, "solver", 1);
        need_ptr(U_dev, "U");
        need_ptr(X_dev, "X");
        hcheck(hipSetDevice(s->device), "hipSetDevice");
        hipLaunchKernelGGL(features_kernel, dim3(nblk((long)s->geo.E * FT)), dim3(TB), 0, s->stream, U_dev, X_dev,
                           s->dtab, s->geo);
        hcheck(hipGetLastError(), "features kernel");
        hcheck(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    });
}

int dg_closure_target(dg_solver* fine, const double* Uf_dev, dg_solver* coarse, double* Uc_dev, float* Y_dev) {
    return mlp_guard(g_mlp_error, [&] {
        check_box(fine, "fine solver", 0);
        check_box(coarse, "coarse solver", 1);
        need_ptr(Uf_dev, "Uf");
        need_ptr(Uc_dev, "Uc");
        need_ptr(Y_dev, "Y");
        if (fine->P < 2) throw std::invalid_argument("the fine solver's order must exceed 1");
        for (int d = 0; d < 3; ++d)
            if (fine->geo.ne[d] != coarse->geo.ne[d] || fine->geo.h[d] != coarse->geo.h[d] ||
                fine->geo.lo[d] != coarse->geo.lo[d])
                throw std::invalid_argument("fine and coarse solvers must share the mesh");
        if (fine->device != coarse->device) throw std::invalid_argument("the solvers live on different devices");
        hcheck(hipSetDevice(fine->device), "hipSetDevice");
        hcheck(hipStreamSynchronize(coarse->stream), "hipStreamSynchronize");
        const long E = fine->geo.E;
        // fine tendency into the fine handle's Ub, then the coarse state and its tendency
        tendency(fine, Uf_dev, fine->Ub);
        hipLaunchKernelGGL(truncate_kernel, dim3(nblk(E * 40)), dim3(TB), 0, fine->stream, Uf_dev, Uc_dev, fine->P, E);
        hcheck(hipGetLastError(), "truncate kernel");
        hcheck(hipStreamSynchronize(fine->stream), "hipStreamSynchronize");
        tendency(coarse, Uc_dev, coarse->Ub);
        hcheck(hipStreamSynchronize(coarse->stream), "hipStreamSynchronize");
        hipLaunchKernelGGL(target_kernel, dim3(nblk(E * 40)), dim3(TB), 0, fine->stream, fine->Ub, coarse->Ub, Y_dev,
                           fine->P, E);
        hcheck(hipGetLastError(), "target kernel");
        hcheck(hipStreamSynchronize(fine->stream), "hipStreamSynchronize");
    });
}

int dg_closure_target_at(dg_solver* fine, const double* Uf_dev, dg_solver* coarse, const double* Uc_dev, double tau,
                         float* Y_dev) {
    return mlp_guard(g_mlp_error, [&] {
        check_box(fine, "fine solver", 0);
        check_box(coarse, "coarse solver", 1);
        need_ptr(Uf_dev, "Uf");
        need_ptr(Uc_dev, "Uc");
        need_ptr(Y_dev, "Y");
        if (fine->P < 2) throw std::invalid_argument("the fine solver's order must exceed 1");
        if (!(tau > 0.0)) throw std::invalid_argument("tau must be > 0");
        for (int d = 0; d < 3; ++d)
            if (fine->geo.ne[d] != coarse->geo.ne[d] || fine->geo.h[d] != coarse->geo.h[d] ||
                fine->geo.lo[d] != coarse->geo.lo[d])
                throw std::invalid_argument("fine and coarse solvers must share the mesh");
        if (fine->device != coarse->device) throw std::invalid_argument("the solvers live on different devices");
        hcheck(hipSetDevice(fine->device), "hipSetDevice");
        const long E = fine->geo.E;
        tendency(fine, Uf_dev, fine->Ub);
        hcheck(hipStreamSynchronize(fine->stream), "hipStreamSynchronize");
        tendency(coarse, Uc_dev, coarse->Ub);
        hcheck(hipStreamSynchronize(coarse->stream), "hipStreamSynchronize");
        hipLaunchKernelGGL(target_at_kernel, dim3(nblk(E * 40)), dim3(TB), 0, fine->stream, fine->Ub, coarse->Ub,
                           Uf_dev, Uc_dev, 1.0 / tau, Y_dev, fine->P, E);
        hcheck(hipGetLastError(), "target kernel");
        hcheck(hipStreamSynchronize(fine->stream), "hipStreamSynchronize");
    });
}

int dg_closure_step(dg_solver* s, dg_mlp* m, double dt_step, double bound, int nsteps, long* n_fallback) {
    return mlp_guard(g_mlp_error, [&] {
        const double dt = closure_check(s, m, dt_step, nsteps);
        const long E = s->geo.E;
        // the feature rows, corrections and counter live on the solver
        // handle from the first call on (freed by dg_destroy)
        if (!s->cl_X) hcheck(hipMalloc(&s->cl_X, (size_t)E * NF * sizeof(float)), "hipMalloc");
        if (!s->cl_C) hcheck(hipMalloc(&s->cl_C, (size_t)E * NO * sizeof(float)), "hipMalloc");
        if (!s->cl_cnt) hcheck(hipMalloc(&s->cl_cnt, sizeof(unsigned long long)), "hipMalloc");
        closure_steps(s, m, dt, bound, nsteps, s->cl_X, s->cl_C, s->cl_cnt, n_fallback);
    });
}

int dg_closure_step_noalloc(dg_solver* s, dg_mlp* m, double dt_step, double bound, int nsteps, long* n_fallback) {
    return mlp_guard(g_mlp_error, [&] {
        const double dt = closure_check(s, m, dt_step, nsteps);
        const long E = s->geo.E;
        if (E > NOALLOC_MAX_ELEMS)
            throw std::invalid_argument("dg_closure_step_noalloc: " + std::to_string(E) + " elements exceed the " +
                                        std::to_string(NOALLOC_MAX_ELEMS) + " of its static scratch (use dg_closure_step)");
        // the device is current (closure_check): its own copy of the array
        float* base = noalloc_scratch(s->device);
        float* X = base;
        float* C = base + NOALLOC_X;
        auto* cnt = reinterpret_cast<unsigned long long*>(base + NOALLOC_X + NOALLOC_C);
        closure_steps(s, m, dt, bound, nsteps, X, C, cnt, n_fallback);
    });
}

}  // extern "C"
