// fj_mlp.hpp -- the action surrogate: a 5 -> 128 -> 128 -> 128 -> 2 tanh MLP
// with BF16 weights and activations on the matrix cores (FP32 accumulation),
// trained on the GPU on actions the Staeckel fudge computed.
//
// No counterpart in the reference: it stands in for uvOrb (uvOrb.cpp:15-262)
// at the quadrature nodes of intermediate SCF iterations.
//
// Building blocks:
//   k_gemm_bf16   C[M x N] (fp32) = op(A)[M x K] op(B)[K x N], A and B bf16,
//                 64 x 64 block tiles, one 32 x 32 tile per wave, built from
//                 v_mfma_f32_32x32x8_bf16 (operands staged in LDS)
//   elementwise   bias + tanh, output bias, MSE gradient, tanh gradient,
//                 column sums (bias gradients), Adam, fp32 -> bf16 copies
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fj {

typedef uint16_t bf16_t;
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

// float -> bf16, round to nearest even (NaN kept quiet)
__host__ __device__ inline bf16_t f2bf(float x) {
    union { float f; uint32_t u; } v;
    v.f = x;
    if ((v.u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((v.u >> 16) | 0x40);
    const uint32_t lsb = (v.u >> 16) & 1u;
    return (bf16_t)((v.u + 0x7fffu + lsb) >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h) {
    union { float f; uint32_t u; } v;
    v.u = (uint32_t)h << 16;
    return v.f;
}

// ---------------------------------------------------------------------------
// GEMM: C = op(A) op(B).  op(A)(i,k) = TA ? A[k*lda + i] : A[i*lda + k];
// op(B)(k,n) = TB ? B[n*ldb + k] : B[k*ldb + n].  Any M, N, K (tails are
// zero-filled in LDS).  256 threads: waves (0,1,2,3) own the 32 x 32 tiles
// (0,0), (0,32), (32,0), (32,32) of the block's 64 x 64 tile.
// ---------------------------------------------------------------------------
constexpr int GT = 64;    // block tile (rows and columns of C)
constexpr int GK = 32;    // K per LDS stage

// Split K: block z of gridDim.z handles K range [z kc, (z+1) kc) and writes
// its partial product to C + z * cz (a reduction kernel adds the partials).
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm_bf16(int M, int N, int K, const bf16_t *__restrict__ A, int lda,
                                                   const bf16_t *__restrict__ B, int ldb, float *__restrict__ C,
                                                   int ldc, int accumulate, int kc, long long cz) {
    // As[row][k] and Bt[col][k]: every lane reads 4 consecutive k (8 bytes)
    __shared__ bf16_t As[GT][GK + 4];
    __shared__ bf16_t Bt[GT][GK + 4];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const long long row0 = (long long)blockIdx.y * GT, col0 = (long long)blockIdx.x * GT;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    v16f acc;
    for (int r = 0; r < 16; r++) acc[r] = 0.f;
    const int kbeg = blockIdx.z * kc, kend = min(K, kbeg + kc);
    C += blockIdx.z * cz;
    for (int k0 = kbeg; k0 < kend; k0 += GK) {
        // stage: 64 x 32 of A and 32 x 64 of B, 8 elements each per thread
        for (int e = t; e < GT * GK; e += 256) {
            int i, k;
            if (TA) { i = e % GT; k = e / GT; } else { i = e / GK; k = e % GK; }   // coalesced along memory
            const long long gi = row0 + i;
            const int gk = k0 + k;
            bf16_t v = 0;
            if (gi < M && gk < kend) v = TA ? A[(long long)gk * lda + gi] : A[gi * lda + gk];
            As[i][k] = v;
        }
        for (int e = t; e < GT * GK; e += 256) {
            int n, k;
            if (TB) { n = e / GK; k = e % GK; } else { n = e % GT; k = e / GT; }
            const long long gn = col0 + n;
            const int gk = k0 + k;
            bf16_t v = 0;
            if (gn < N && gk < kend) v = TB ? B[gn * ldb + gk] : B[(long long)gk * ldb + gn];
            Bt[n][k] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; kk += 8) {
            const int ka = kk + 4 * (lane >> 5);
            v4s a, b;
            const bf16_t *pa = &As[wr + (lane & 31)][ka];
            const bf16_t *pb = &Bt[wc + (lane & 31)][ka];
            for (int j = 0; j < 4; j++) { a[j] = (short)pa[j]; b[j] = (short)pb[j]; }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout: lane holds column wc + lane%32, rows wr + 8(r/4) + 4(lane/32) + r%4
    const long long gc = col0 + wc + (lane & 31);
    if (gc < N) {
        for (int r = 0; r < 16; r++) {
            const long long gr = row0 + wr + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (gr < M) {
                float *p = &C[gr * ldc + gc];
                *p = accumulate ? *p + acc[r] : acc[r];
            }
        }
    }
}

// h = tanh(pre + b) -> fp32 and bf16 copies; pre [M x N] row-major
__global__ void k_bias_tanh(long long M, int N, const float *pre, const float *b, float *h32, bf16_t *h16) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * N) return;
    const float v = tanhf(pre[i] + b[i % N]);
    if (h32) h32[i] = v;
    h16[i] = f2bf(v);
}

// y = pre + b on the first nout columns of an ldp-wide pre
__global__ void k_bias_out(long long M, int nout, int ldp, const float *pre, const float *b, float *y) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * nout) return;
    const long long r = i / nout;
    const int c = (int)(i % nout);
    y[i] = pre[r * ldp + c] + b[c];
}

// Mean-squared error on nout outputs, row weights wt (null: 1):
// dY = 2 wt (y - t) / (M nout) into an ldd-wide bf16 buffer (other columns
// 0); per-block partial sums of wt (y - t)^2.
__global__ void k_mse_grad(long long M, int nout, int ldd, const float *y, const float *tgt, const float *wt,
                           bf16_t *dY, float *loss_part) {
    __shared__ float red[256];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < M * ldd) {
        const long long r = i / ldd;
        const int c = (int)(i % ldd);
        float g = 0.f;
        if (c < nout) {
            const float d = y[r * nout + c] - tgt[r * nout + c];
            const float wr = wt ? wt[r] : 1.f;
            l = wr * d * d;
            g = 2.f * wr * d / (float)(M * nout);
        }
        dY[i] = f2bf(g);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// dpre = dH (1 - h^2) -> fp32 and bf16
__global__ void k_tanh_grad(long long n, const float *dH, const float *h, float *dp32, bf16_t *dp16) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = dH[i] * (1.f - h[i] * h[i]);
    if (dp32) dp32[i] = v;
    dp16[i] = f2bf(v);
}

// Column sums db[c] = sum_r X[r][c] of an M x N row-major matrix (ld ldx,
// N <= 256), in two fixed-order stages.  Stage 1: block b sums rows
// [b R, (b+1) R) with row-coalesced reads (thread t takes column t % N of
// every (256 / N)-th row from t / N) into part[b][0..N).  Stage 2: one block
// per column adds the nb partials (a 256-wide tree).
__device__ __forceinline__ float as_f32(float x) { return x; }
__device__ __forceinline__ float as_f32(bf16_t x) { return bf2f(x); }
template <class T>
__global__ void __launch_bounds__(256) k_colsum_part(long long M, int N, int ldx, long long R, const T *X,
                                                     float *part) {
    __shared__ float red[256];
    const int t = threadIdx.x, rstep = 256 / N, c = t % N, r0 = t / N;
    float s = 0.f;
    if (r0 < rstep) {
        const long long rb = (long long)blockIdx.x * R, re = rb + R < M ? rb + R : M;
        for (long long r = rb + r0; r < re; r += rstep) s += as_f32(X[r * ldx + c]);
    }
    red[t] = s;
    __syncthreads();
    if (t < N) {
        float v = 0.f;
        for (int k = 0; k < rstep; k++) v += red[k * N + t];
        part[(long long)blockIdx.x * N + t] = v;
    }
}
__global__ void __launch_bounds__(256) k_colsum_fin(int nb, int N, const float *part, float *db) {
    __shared__ float red[256];
    const int c = blockIdx.x, t = threadIdx.x;
    float s = 0.f;
    for (int b = t; b < nb; b += 256) s += part[(long long)b * N + c];
    red[t] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st) red[t] += red[t + st];
        __syncthreads();
    }
    if (t == 0) db[c] = red[0];
}

// Adam step on n fp32 parameters; writes the bf16 copy used by the GEMMs.
__global__ void k_adam(int n, float *p, const float *g, float *m, float *v, bf16_t *p16, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    const float pi = p[i] - upd;
    p[i] = pi;
    if (p16) p16[i] = f2bf(pi);
}

__global__ void k_to_bf16(long long n, const float *x, bf16_t *y) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

}  // namespace fj

// ===========================================================================
// Host side: parameters, forward, backward and Adam on one stream.
// ===========================================================================
#include <math.h>
#include <string.h>

#include <vector>

namespace fj {

constexpr int MLP_IN = 5, MLP_H = 128, MLP_OUT = 2;

// Parameter offsets in the flat fp32 vector: W1 [5][128], b1, W2 [128][128],
// b2, W3 [128][128], b3, W4 [128][2], b4 (W[in][out], y = x W + b).
struct MlpLayout {
    static constexpr int W1 = 0, B1 = W1 + MLP_IN * MLP_H, W2 = B1 + MLP_H, B2 = W2 + MLP_H * MLP_H,
                         W3 = B2 + MLP_H, B3 = W3 + MLP_H * MLP_H, W4 = B3 + MLP_H, B4 = W4 + MLP_H * MLP_OUT,
                         N = B4 + MLP_OUT;
};

// Partial-product buffer for split-K GEMMs.
struct SplitK {
    float *part = nullptr;
    long long cap = 0;   // floats
};

struct Mlp {
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // [N] fp32
    bf16_t *p16 = nullptr;                                          // [N] bf16 copy
    int step = 0;
    // normalisation of inputs and targets
    float mu[MLP_IN], sig[MLP_IN], tmu[MLP_OUT], tsig[MLP_OUT];
    // work buffers for batches of up to cap rows
    long long cap = 0;
    bf16_t *x16 = nullptr, *h16[3] = {nullptr, nullptr, nullptr}, *d16 = nullptr, *dy16 = nullptr;
    float *pre = nullptr, *h32[3] = {nullptr, nullptr, nullptr}, *y = nullptr, *dh = nullptr, *dp32 = nullptr,
          *loss_part = nullptr;
    SplitK sk;
};

inline int mlp_blocks(long long n, int t = 256) { return (int)((n + t - 1) / t); }

// db[0..N) = column sums of X [M][N] (ld ldx), through the split-K buffer.
template <class T>
inline void colsum(hipStream_t s, long long M, int N, int ldx, const T *X, float *db, SplitK &sk) {
    long long nb = (M + 127) / 128;
    if (nb > 256) nb = 256;
    if (nb * N > sk.cap) nb = sk.cap / N;
    if (nb < 1) nb = 1;
    const long long R = (M + nb - 1) / nb;
    k_colsum_part<T><<<(unsigned)nb, 256, 0, s>>>(M, N, ldx, R, X, sk.part);
    k_colsum_fin<<<N, 256, 0, s>>>((int)nb, N, sk.part, db);
}

// C[i][j] = sum_z part[z][i][j] (fixed order), C with leading dimension
// ldc.  A block takes 64 consecutive outputs; its four waves sum the z in
// four interleaved groups (wave w: z = w, w + 4, ...), which then add in
// wave order.
__global__ void __launch_bounds__(256) k_splitk_sum(int M, int N, int ldc, int nz, const float *part, float *C) {
    __shared__ float red[4][64];
    const int t = threadIdx.x, o = t & 63, w = t >> 6;
    const long long i = (long long)blockIdx.x * 64 + o, MN = (long long)M * N;
    float s = 0.f;
    if (i < MN)
        for (int z = w; z < nz; z += 4) s += part[(long long)z * MN + i];
    red[w][o] = s;
    __syncthreads();
    if (w == 0 && i < MN) {
        const long long r = i / N;
        const int c = (int)(i % N);
        C[r * ldc + c] = ((red[0][o] + red[1][o]) + red[2][o]) + red[3][o];
    }
}

template <bool TA, bool TB>
inline void gemm(hipStream_t s, int M, int N, int K, const bf16_t *A, int lda, const bf16_t *B, int ldb, float *C,
                 int ldc, bool acc = false, SplitK *sk = nullptr) {
    dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
    // few output tiles and a long K (weight gradients): split K over blocks
    const long long tiles = (long long)grid.x * grid.y;
    int nz = 1;
    if (sk && !acc && tiles < 1024 && K >= 2048) {
        nz = (int)((1024 + tiles - 1) / tiles);   // about four blocks per CU
        const int maxz = K / 256;
        if (nz > maxz) nz = maxz;
        while (nz > 1 && (long long)nz * M * N > sk->cap) nz--;
    }
    if (nz <= 1) {
        k_gemm_bf16<TA, TB><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, C, ldc, acc ? 1 : 0, K, 0);
        return;
    }
    const int kc = ((K + nz - 1) / nz + GK - 1) / GK * GK;
    grid.z = nz;
    k_gemm_bf16<TA, TB><<<grid, 256, 0, s>>>(M, N, K, A, lda, B, ldb, sk->part, N, 0, kc, (long long)M * N);
    k_splitk_sum<<<(unsigned)(((long long)M * N + 63) / 64), 256, 0, s>>>(M, N, ldc, nz, sk->part, C);
}

// Forward of rows x16 [M][5] (normalised, bf16) -> y [M][2] (normalised
// outputs), M <= cap.  Keeps h32/h16 of the three hidden layers for a
// backward pass on n.x16 / n.y.
inline void mlp_forward(hipStream_t s, Mlp &n, long long M, const bf16_t *x16, float *y) {
    typedef MlpLayout L;
    const bf16_t *W[4] = {n.p16 + L::W1, n.p16 + L::W2, n.p16 + L::W3, n.p16 + L::W4};
    const float *b[4] = {n.p + L::B1, n.p + L::B2, n.p + L::B3, n.p + L::B4};
    const bf16_t *in = x16;
    int K = MLP_IN;
    for (int l = 0; l < 3; l++) {
        gemm<false, false>(s, (int)M, MLP_H, K, in, K, W[l], MLP_H, n.pre, MLP_H);
        k_bias_tanh<<<mlp_blocks(M * MLP_H), 256, 0, s>>>(M, MLP_H, n.pre, b[l], n.h32[l], n.h16[l]);
        in = n.h16[l];
        K = MLP_H;
    }
    gemm<false, false>(s, (int)M, MLP_OUT, MLP_H, in, MLP_H, W[3], MLP_OUT, n.pre, MLP_OUT);
    k_bias_out<<<mlp_blocks(M * MLP_OUT), 256, 0, s>>>(M, MLP_OUT, MLP_OUT, n.pre, b[3], y);
}

// Gradient of the (row-weighted) mean squared error against tgt [M][2]
// (normalised) into n.g;
// n.loss_part gets per-block partial sums of squared errors.
inline void mlp_backward(hipStream_t s, Mlp &n, long long M, const float *tgt, const float *wt = nullptr) {
    typedef MlpLayout L;
    k_mse_grad<<<mlp_blocks(M * MLP_OUT), 256, 0, s>>>(M, MLP_OUT, MLP_OUT, n.y, tgt, wt, n.dy16, n.loss_part);
    // output layer
    gemm<true, false>(s, MLP_H, MLP_OUT, (int)M, n.h16[2], MLP_H, n.dy16, MLP_OUT, n.g + L::W4, MLP_OUT, false,
                      &n.sk);
    colsum(s, M, MLP_OUT, MLP_OUT, n.dy16, n.g + L::B4, n.sk);
    gemm<false, true>(s, (int)M, MLP_H, MLP_OUT, n.dy16, MLP_OUT, n.p16 + L::W4, MLP_OUT, n.dh, MLP_H);
    const int Wo[3] = {L::W1, L::W2, L::W3}, Bo[3] = {L::B1, L::B2, L::B3};
    for (int l = 2; l >= 0; l--) {
        k_tanh_grad<<<mlp_blocks(M * MLP_H), 256, 0, s>>>(M * MLP_H, n.dh, n.h32[l], n.dp32, n.d16);
        colsum(s, M, MLP_H, MLP_H, n.dp32, n.g + Bo[l], n.sk);
        const bf16_t *in = l > 0 ? n.h16[l - 1] : n.x16;
        const int K = l > 0 ? MLP_H : MLP_IN;
        gemm<true, false>(s, K, MLP_H, (int)M, in, K, n.d16, MLP_H, n.g + Wo[l], MLP_H, false, &n.sk);
        if (l > 0) gemm<false, true>(s, (int)M, MLP_H, MLP_H, n.d16, MLP_H, n.p16 + Wo[l], MLP_H, n.dh, MLP_H);
    }
}

// acc[0] += sum of part[0 .. n) (one block, fixed order)
__global__ void k_sum_into(int n, const float *part, double *acc) {
    __shared__ double red[256];
    double v = 0;
    for (int i = threadIdx.x; i < n; i += 256) v += part[i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += red[0];
}

inline void mlp_adam(hipStream_t s, Mlp &n, float lr) {
    const float b1 = 0.9f, b2 = 0.999f;
    n.step++;
    const float c1 = 1.f - powf(b1, (float)n.step), c2 = 1.f - powf(b2, (float)n.step);
    k_adam<<<mlp_blocks(MlpLayout::N), 256, 0, s>>>(MlpLayout::N, n.p, n.g, n.m, n.v, n.p16, lr, b1, b2, c1, c2,
                                                   1e-8f);
}

// Glorot-uniform weights, zero biases, from a 64-bit LCG (host side).
inline void mlp_init_params(std::vector<float> &p, unsigned long long seed) {
    typedef MlpLayout L;
    p.assign(L::N, 0.f);
    unsigned long long st = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    const int off[4] = {L::W1, L::W2, L::W3, L::W4};
    const int fin[4] = {MLP_IN, MLP_H, MLP_H, MLP_H}, fout[4] = {MLP_H, MLP_H, MLP_H, MLP_OUT};
    for (int l = 0; l < 4; l++) {
        const double a = sqrt(6.0 / (fin[l] + fout[l]));
        for (int i = 0; i < fin[l] * fout[l]; i++) p[off[l] + i] = (float)((2 * uni() - 1) * a);
    }
}

}  // namespace fj

// ===========================================================================
// Fused inference.  Nodes are the N dimension (columns) of every product:
// H_l^T [128 x 32 nodes] = W_l^T [128 x 128] H_{l-1}^T, one 32-node tile per
// wave.  The 32 x 32 accumulator of output block ob holds, in lane l and
// element r, neuron 32 ob + 8 (r/4) + 4 (l/32) + r%4 of node l%32; the next
// product's B operand for K chunk kc (neurons 8 kc .. 8 kc + 7) wants in lane
// l neurons 8 kc + 4 (l/32) + j, j < 4: elements 4 (kc%4) + j of block kc/4.
// So each layer's output feeds the next layer in registers.  The A operands
// (weights) are pre-arranged so that each lane reads its 4 values at once.
// ===========================================================================
namespace fj {

// Packed weights (bf16): L1 [4 ob][64 lanes][4], L2 and L3 [4 ob][16 kc][64][4],
// L4 [16 kc][64][4]; then the biases as fp32.
struct MlpPacked {
    static constexpr int L1 = 0, L2 = L1 + 4 * 64 * 4, L3 = L2 + 4 * 16 * 64 * 4, L4 = L3 + 4 * 16 * 64 * 4,
                         N16 = L4 + 16 * 64 * 4;
};

__global__ void k_mlp_pack(const float *p, bf16_t *w, float *bias) {
    typedef MlpLayout L;
    typedef MlpPacked K;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4 * 64 * 4) {   // layer 1
        const int ob = i / 256, l = (i / 4) % 64, j = i % 4, k = 4 * (l / 32) + j;
        w[K::L1 + i] = k < MLP_IN ? f2bf(p[L::W1 + k * MLP_H + 32 * ob + l % 32]) : (bf16_t)0;
    }
    if (i < 4 * 16 * 64 * 4) {   // layers 2 and 3
        const int ob = i / 4096, kc = (i / 256) % 16, l = (i / 4) % 64, j = i % 4;
        const int k = 8 * kc + 4 * (l / 32) + j, n = 32 * ob + l % 32;
        w[K::L2 + i] = f2bf(p[L::W2 + k * MLP_H + n]);
        w[K::L3 + i] = f2bf(p[L::W3 + k * MLP_H + n]);
    }
    if (i < 16 * 64 * 4) {   // layer 4: rows 0, 1 of W4^T
        const int kc = i / 256, l = (i / 4) % 64, j = i % 4, k = 8 * kc + 4 * (l / 32) + j, n = l % 32;
        w[K::L4 + i] = n < MLP_OUT ? f2bf(p[L::W4 + k * MLP_OUT + n]) : (bf16_t)0;
    }
    if (i < 3 * MLP_H) bias[i] = p[(i / MLP_H == 0 ? L::B1 : i / MLP_H == 1 ? L::B2 : L::B3) + i % MLP_H];
    if (i < MLP_OUT) bias[3 * MLP_H + i] = p[L::B4 + i];
}

__device__ __forceinline__ float tanh_fast(float x) {
    // 1 - 2 / (exp(2x) + 1); saturates cleanly for large |x|
    const float e = __expf(2.f * x);
    return 1.f - 2.f / (e + 1.f);
}

__device__ __forceinline__ v4s pack4(float a, float b, float c, float d) {
    v4s v;
    v[0] = (short)f2bf(a); v[1] = (short)f2bf(b); v[2] = (short)f2bf(c); v[3] = (short)f2bf(d);
    return v;
}

// One 32-node tile in a wave: x4 holds this lane's 4 inputs (features
// 4 (lane/32) .. +3 of node lane%32, normalised).  Returns the two outputs in
// lanes 0..31 (y0, y1).  wl: LDS copy of the L2/L3 packed weights.
__device__ __forceinline__ void mlp_tile(const bf16_t *__restrict__ wg, const bf16_t *wl, const float *__restrict__ bias,
                                         v4s x4, float &y0, float &y1) {
    typedef MlpPacked K;
    const int lane = __lane_id();
    v16f acc[4];
    v4s h[16];   // B operands of the next layer: chunk kc = 4 ob + q
    // layer 1 (K = 8)
    for (int ob = 0; ob < 4; ob++) {
        for (int r = 0; r < 16; r++) acc[ob][r] = 0.f;
        const v4s a = *reinterpret_cast<const v4s *>(wg + K::L1 + (ob * 64 + lane) * 4);
        acc[ob] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, x4, acc[ob], 0, 0, 0);
    }
    for (int layer = 0; layer < 3; layer++) {
        // bias + tanh of the previous products -> next B operands
        const float *b = bias + layer * MLP_H;
        for (int ob = 0; ob < 4; ob++)
            for (int q = 0; q < 4; q++) {
                const int n0 = 32 * ob + 8 * q + 4 * (lane >> 5);
                h[4 * ob + q] = pack4(tanh_fast(acc[ob][4 * q] + b[n0]), tanh_fast(acc[ob][4 * q + 1] + b[n0 + 1]),
                                      tanh_fast(acc[ob][4 * q + 2] + b[n0 + 2]),
                                      tanh_fast(acc[ob][4 * q + 3] + b[n0 + 3]));
            }
        if (layer == 2) break;
        const bf16_t *W = wl + (layer == 0 ? 0 : K::L3 - K::L2);
        for (int ob = 0; ob < 4; ob++) {
            for (int r = 0; r < 16; r++) acc[ob][r] = 0.f;
            for (int kc = 0; kc < 16; kc++) {
                const v4s a = *reinterpret_cast<const v4s *>(W + ((ob * 16 + kc) * 64 + lane) * 4);
                acc[ob] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, h[kc], acc[ob], 0, 0, 0);
            }
        }
    }
    // layer 4: rows 0, 1 of the output block
    v16f o;
    for (int r = 0; r < 16; r++) o[r] = 0.f;
    for (int kc = 0; kc < 16; kc++) {
        const v4s a = *reinterpret_cast<const v4s *>(wg + K::L4 + (kc * 64 + lane) * 4);
        o = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, h[kc], o, 0, 0, 0);
    }
    y0 = o[0] + bias[3 * MLP_H];
    y1 = o[1] + bias[3 * MLP_H + 1];
}

}  // namespace fj
