// mlp.hip -- the proposal network: a 64 -> 256 -> 256 -> 256 -> 1 ReLU MLP
// trained and evaluated with hand-written BF16 MFMA kernels (gfx950
// v_mfma_f32_16x16x32_bf16, FP32 accumulation) and FP32 Adam master weights.
//
// GEMM: C[M x N] = op(A)[M x K] * op(B)[K x N], A and B bf16, C fp32 in
// registers.  A wave owns a 32x32 tile of C (2x2 MFMA tiles), a 256-thread
// block a 64x64 tile; K advances 32 per MFMA.  Operand lane layout of the
// 16x16x32 instruction: lane l supplies A[row = l%16][k = 8*(l/16) .. +8] and
// B[k = 8*(l/16) .. +8][col = l%16]; it receives C[rows 4*(l/16) .. +4][col
// = l%16].  Operands come straight from global memory (L2-resident weights,
// streamed activations): 16-byte loads where the eight k-values are
// contiguous, eight 2-byte loads where they are strided (transposed
// operands).  Epilogues fuse bias, ReLU and the ReLU mask of the backward
// pass.  Weight gradients reduce over the batch in fixed-size chunks whose
// partial sums are added in a fixed order (deterministic training).
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "walk_common.hpp"

namespace cmfw {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

union Pack8 {
    bf16x8 v;
    uint16_t h[8];
    uint4 q;
};

enum Epi { EPI_STORE = 0, EPI_BIAS_RELU_BF16 = 1, EPI_MASK_BF16 = 2, EPI_PARTIAL = 3 };

struct GemmArgs {
    int M, N, K;
    const uint16_t *A;  // TA=0: [M][lda] (row m, col k); TA=1: [K][lda] (row k, col m)
    int lda;
    const uint16_t *B;  // TB=0: [K][ldb] (row k, col n); TB=1: [N][ldb] (row n, col k)
    int ldb;
    float *C;           // EPI_STORE: [M][ldc] fp32; EPI_PARTIAL: [chunk][M][N]
    int ldc;
    uint16_t *Cb;       // bf16 output [M][ldc] for the bf16 epilogues
    const float *bias;  // [N]
    const uint16_t *mask;  // [M][ldc]: the forward activation; grad passes where > 0
    int kchunk;         // EPI_PARTIAL: K range per blockIdx.z
    int vecA, vecB;     // 16-byte loads allowed: base 16-byte aligned and ld % 8 == 0
};

static inline int vec_ok(const void *p, int ld) {
    return ((reinterpret_cast<uintptr_t>(p) & 15) == 0) && ((ld & 7) == 0);
}

template <int TA, int TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;           // 4 waves: 2 x 2 of 32x32
    const int m0 = blockIdx.x * 64 + (wave >> 1) * 32;
    const int n0 = blockIdx.y * 64 + (wave & 1) * 32;
    int kb = 0, ke = g.K;
    if (EPI == EPI_PARTIAL) {
        kb = blockIdx.z * g.kchunk;
        ke = min(g.K, kb + g.kchunk);
    }
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int r = lane & 15;       // row of A / col of B this lane feeds
    const int kg = (lane >> 4) * 8;  // first of its eight k values
    for (int k0 = kb; k0 < ke; k0 += 32) {
        Pack8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int m = m0 + i * 16 + r;
#pragma unroll
            for (int e = 0; e < 8; ++e) a[i].h[e] = 0;
            if (m < g.M) {
                if (TA == 0) {
                    const int k = k0 + kg;
                    if (k + 8 <= ke && g.vecA) {
                        a[i].q = *reinterpret_cast<const uint4 *>(g.A + (size_t)m * g.lda + k);
                    } else {
#pragma unroll
                        for (int e = 0; e < 8; ++e)
                            if (k + e < ke) a[i].h[e] = g.A[(size_t)m * g.lda + k + e];
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int k = k0 + kg + e;
                        if (k < ke) a[i].h[e] = g.A[(size_t)k * g.lda + m];
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int n = n0 + j * 16 + r;
#pragma unroll
            for (int e = 0; e < 8; ++e) b[j].h[e] = 0;
            if (n < g.N) {
                if (TB == 1) {
                    const int k = k0 + kg;
                    if (k + 8 <= ke && g.vecB) {
                        b[j].q = *reinterpret_cast<const uint4 *>(g.B + (size_t)n * g.ldb + k);
                    } else {
#pragma unroll
                        for (int e = 0; e < 8; ++e)
                            if (k + e < ke) b[j].h[e] = g.B[(size_t)n * g.ldb + k + e];
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int k = k0 + kg + e;
                        if (k < ke) b[j].h[e] = g.B[(size_t)k * g.ldb + n];
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i].v, b[j].v, acc[i][j], 0, 0, 0);
    }
    // epilogue: this lane holds C[m0+16i + 4*(lane/16) + e][n0 + 16j + lane%16]
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + i * 16 + (lane >> 4) * 4 + e;
                const int n = n0 + j * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                const float v = acc[i][j][e];
                if (EPI == EPI_STORE) {
                    g.C[(size_t)m * g.ldc + n] = v;
                } else if (EPI == EPI_BIAS_RELU_BF16) {
                    g.Cb[(size_t)m * g.ldc + n] = f2bf(fmaxf(v + g.bias[n], 0.f));
                } else if (EPI == EPI_MASK_BF16) {
                    const bool on = bf2f(g.mask[(size_t)m * g.ldc + n]) > 0.f;
                    g.Cb[(size_t)m * g.ldc + n] = f2bf(on ? v : 0.f);
                } else {
                    g.C[((size_t)blockIdx.z * g.M + m) * g.N + n] = v;
                }
            }
}

// sum of partial weight gradients over chunks, in chunk order
__global__ void k_sum_partials(const float *part, int nchunk, int count, float *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * count + i];
    out[i] = s;
}

// last layer: y[b] = sum_k H[b][k] * w[k] + bias (fp32 weights), one lane per row
__global__ void k_head_forward(const uint16_t *H, int B, int K, const float *w, const float *bias,
                               float *y) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float s = bias[0];
    for (int k = 0; k < K; ++k) s = fmaf(bf2f(H[(size_t)b * K + k]), w[k], s);
    y[b] = s;
}

// loss gradient: dy = 2 (y - t) / B for rows with weight; writes the
// per-row squared error for the loss
__global__ void k_mse_grad(const float *y, const float *t, int B, float *dy, float *sq) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float d = y[b] - t[b];
    dy[b] = 2.f * d / (float)B;
    sq[b] = d * d;
}

// dH3 = dy (outer) w, masked by H3 > 0, bf16 out
__global__ void k_head_backward(const float *dy, const float *w, const uint16_t *H, int B, int K,
                                uint16_t *dH) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)B * K) return;
    const int b = (int)(i / K), k = (int)(i % K);
    const bool on = bf2f(H[i]) > 0.f;
    dH[i] = f2bf(on ? dy[b] * w[k] : 0.f);
}

// column sums over the batch in fixed chunks (deterministic):
// part[c][j] = sum over rows of chunk c of X[row][j] (X fp32 [B][ld] or bf16)
__global__ void k_colsum_partial_f32(const float *X, int B, int cols, int ld, int chunk, float *part) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (j >= cols) return;
    const int b0 = c * chunk, b1 = min(B, b0 + chunk);
    float s = 0.f;
    for (int b = b0; b < b1; ++b) s += X[(size_t)b * ld + j];
    part[(size_t)c * cols + j] = s;
}
__global__ void k_colsum_partial_bf16(const uint16_t *X, int B, int cols, int ld, int chunk,
                                      float *part) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (j >= cols) return;
    const int b0 = c * chunk, b1 = min(B, b0 + chunk);
    float s = 0.f;
    for (int b = b0; b < b1; ++b) s += bf2f(X[(size_t)b * ld + j]);
    part[(size_t)c * cols + j] = s;
}
// dW4[k] partials: sum_b H[b][k] * dy[b]
__global__ void k_head_wgrad_partial(const uint16_t *H, const float *dy, int B, int K, int chunk,
                                     float *part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (k >= K) return;
    const int b0 = c * chunk, b1 = min(B, b0 + chunk);
    float s = 0.f;
    for (int b = b0; b < b1; ++b) s = fmaf(bf2f(H[(size_t)b * K + k]), dy[b], s);
    part[(size_t)c * K + k] = s;
}

// Adam on FP32 master weights; refresh the bf16 copy of the matrices
__global__ void k_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2, uint16_t *wb, int nb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (i < nb) wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float *x, int n, uint16_t *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(x[i]);
}

// ---------------------------------------------------------------- host --

template <int TA, int TB, int EPI>
static hipError_t gemm(GemmArgs g, int nchunk, hipStream_t st) {
    g.vecA = vec_ok(g.A, g.lda);
    g.vecB = vec_ok(g.B, g.ldb);
    dim3 grid((g.M + 63) / 64, (g.N + 63) / 64, EPI == EPI_PARTIAL ? nchunk : 1);
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, st, g);
    return hipGetLastError();
}

hipError_t mlp_gemm_test(int mode, int M, int N, int K, const uint16_t *A, int lda,
                         const uint16_t *B, int ldb, float *C, int ldc, hipStream_t st) {
    GemmArgs g{};
    g.M = M;
    g.N = N;
    g.K = K;
    g.A = A;
    g.lda = lda;
    g.B = B;
    g.ldb = ldb;
    g.C = C;
    g.ldc = ldc;
    switch (mode) {
        case 0: return gemm<0, 0, EPI_STORE>(g, 1, st);
        case 1: return gemm<0, 1, EPI_STORE>(g, 1, st);
        case 2: return gemm<1, 0, EPI_STORE>(g, 1, st);
        case 3: return gemm<1, 1, EPI_STORE>(g, 1, st);
        default: return hipErrorInvalidValue;
    }
}

// Parameter layout (fp32 master): W1[64][256], W2[256][256], W3[256][256],
// w4[256], b1[256], b2[256], b3[256], b4[1]; the bf16 copy holds W1..W3.
constexpr int kIn = 64, kHid = 256;
constexpr size_t oW1 = 0, oW2 = oW1 + (size_t)kIn * kHid, oW3 = oW2 + (size_t)kHid * kHid,
                 oW4 = oW3 + (size_t)kHid * kHid, oB1 = oW4 + kHid, oB2 = oB1 + kHid,
                 oB3 = oB2 + kHid, oB4 = oB3 + kHid, kParams = oB4 + 1;
constexpr size_t kParamsBf16 = oW4;
// Weight gradients reduce the batch in split-K chunks of at least
// kMinChunk rows (a multiple of 32), at most kMaxChunks of them: enough
// workgroups to fill the chip at training batch sizes, a bounded partial
// buffer at any batch size.
constexpr int kMinChunk = 256, kMaxChunks = 64;

static int chunk_rows(int B) {
    int c = (B + kMaxChunks - 1) / kMaxChunks;
    c = (c + 31) / 32 * 32;
    return c < kMinChunk ? kMinChunk : c;
}

size_t mlp_param_count() { return kParams; }
size_t mlp_param_bf16_count() { return kParamsBf16; }

static size_t a256(size_t b) { return align256(b); }

size_t mlp_workspace_bytes(int B) {
    const size_t b = (size_t)B;
    const int nchunk = (B + chunk_rows(B) - 1) / chunk_rows(B);
    return 3 * a256(b * kHid * 2) + 2 * a256(b * kHid * 2) + 3 * a256(b * 4) +
           a256((size_t)nchunk * kHid * kHid * 4);
}

struct MlpWs {
    uint16_t *H1, *H2, *H3, *G0, *G1;
    float *y, *dy, *sq, *part;
    int chunk, nchunk;
};

static MlpWs carve(void *ws, int B) {
    MlpWs w;
    char *p = (char *)ws;
    const size_t b = (size_t)B;
    w.H1 = (uint16_t *)p; p += a256(b * kHid * 2);
    w.H2 = (uint16_t *)p; p += a256(b * kHid * 2);
    w.H3 = (uint16_t *)p; p += a256(b * kHid * 2);
    w.G0 = (uint16_t *)p; p += a256(b * kHid * 2);
    w.G1 = (uint16_t *)p; p += a256(b * kHid * 2);
    w.y = (float *)p; p += a256(b * 4);
    w.dy = (float *)p; p += a256(b * 4);
    w.sq = (float *)p; p += a256(b * 4);
    w.part = (float *)p;
    w.chunk = chunk_rows(B);
    w.nchunk = (B + w.chunk - 1) / w.chunk;
    return w;
}

static hipError_t layer_fwd(const uint16_t *X, int B, int K, const uint16_t *W, const float *bias,
                            uint16_t *H, hipStream_t st) {
    GemmArgs g{};
    g.M = B;
    g.N = kHid;
    g.K = K;
    g.A = X;
    g.lda = K;
    g.B = W;
    g.ldb = kHid;
    g.Cb = H;
    g.ldc = kHid;
    g.bias = bias;
    return gemm<0, 0, EPI_BIAS_RELU_BF16>(g, 1, st);
}

hipError_t mlp_forward(const float *P, const uint16_t *Pb, const uint16_t *X, int B, float *y,
                       void *ws, hipStream_t st) {
    MlpWs w = carve(ws, B);
    hipError_t e;
    if ((e = layer_fwd(X, B, kIn, Pb + oW1, P + oB1, w.H1, st))) return e;
    if ((e = layer_fwd(w.H1, B, kHid, Pb + oW2, P + oB2, w.H2, st))) return e;
    if ((e = layer_fwd(w.H2, B, kHid, Pb + oW3, P + oB3, w.H3, st))) return e;
    hipLaunchKernelGGL(k_head_forward, dim3((B + 255) / 256), dim3(256), 0, st, w.H3, B, kHid,
                       P + oW4, P + oB4, y);
    return hipGetLastError();
}

// weight gradient dW[Kin][256] = X^T * dZ over the batch, deterministic
static hipError_t wgrad(const uint16_t *X, int B, int Kin, const uint16_t *dZ, float *dW, float *part,
                        int chunk, int nchunk, hipStream_t st) {
    GemmArgs g{};
    g.M = Kin;
    g.N = kHid;
    g.K = B;
    g.A = X;
    g.lda = Kin;
    g.B = dZ;
    g.ldb = kHid;
    g.C = part;
    g.kchunk = chunk;
    hipError_t e = gemm<1, 0, EPI_PARTIAL>(g, nchunk, st);
    if (e) return e;
    const int cnt = Kin * kHid;
    hipLaunchKernelGGL(k_sum_partials, dim3((cnt + 255) / 256), dim3(256), 0, st, part, nchunk, cnt, dW);
    return hipGetLastError();
}

static hipError_t bgrad(const uint16_t *dZ, int B, float *db, float *part, int chunk, int nchunk,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_colsum_partial_bf16, dim3(1, nchunk), dim3(kHid), 0, st, dZ, B, kHid, kHid,
                       chunk, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kHid), 0, st, part, nchunk, kHid, db);
    return hipGetLastError();
}

// dZ_prev = (dZ * W^T) masked by H_prev > 0, W stored [Kin][256]
static hipError_t dgrad(const uint16_t *dZ, int B, const uint16_t *W, int Kin, const uint16_t *Hprev,
                        uint16_t *out, hipStream_t st) {
    GemmArgs g{};
    g.M = B;
    g.N = Kin;
    g.K = kHid;
    g.A = dZ;
    g.lda = kHid;
    g.B = W;
    g.ldb = kHid;
    g.Cb = out;
    g.ldc = Kin;
    g.mask = Hprev;
    return gemm<0, 1, EPI_MASK_BF16>(g, 1, st);
}

hipError_t mlp_grad(const float *P, const uint16_t *Pb, const uint16_t *X, const float *target,
                    int B, float *grad, float *sq_out, void *ws, hipStream_t st) {
    MlpWs w = carve(ws, B);
    hipError_t e;
    if ((e = mlp_forward(P, Pb, X, B, w.y, ws, st))) return e;
    hipLaunchKernelGGL(k_mse_grad, dim3((B + 255) / 256), dim3(256), 0, st, w.y, target, B, w.dy,
                       sq_out ? sq_out : w.sq);
    // head
    hipLaunchKernelGGL(k_head_wgrad_partial, dim3(1, w.nchunk), dim3(kHid), 0, st, w.H3, w.dy, B,
                       kHid, w.chunk, w.part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kHid), 0, st, w.part, w.nchunk, kHid, grad + oW4);
    hipLaunchKernelGGL(k_colsum_partial_f32, dim3(1, w.nchunk), dim3(1), 0, st, w.dy, B, 1, 1, w.chunk,
                       w.part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1), 0, st, w.part, w.nchunk, 1, grad + oB4);
    {
        const size_t cnt = (size_t)B * kHid;
        hipLaunchKernelGGL(k_head_backward, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                           w.dy, P + oW4, w.H3, B, kHid, w.G0);
    }
    // layer 3: dZ3 = G0
    if ((e = wgrad(w.H2, B, kHid, w.G0, grad + oW3, w.part, w.chunk, w.nchunk, st))) return e;
    if ((e = bgrad(w.G0, B, grad + oB3, w.part, w.chunk, w.nchunk, st))) return e;
    if ((e = dgrad(w.G0, B, Pb + oW3, kHid, w.H2, w.G1, st))) return e;
    // layer 2: dZ2 = G1
    if ((e = wgrad(w.H1, B, kHid, w.G1, grad + oW2, w.part, w.chunk, w.nchunk, st))) return e;
    if ((e = bgrad(w.G1, B, grad + oB2, w.part, w.chunk, w.nchunk, st))) return e;
    if ((e = dgrad(w.G1, B, Pb + oW2, kHid, w.H1, w.G0, st))) return e;
    // layer 1: dZ1 = G0
    if ((e = wgrad(X, B, kIn, w.G0, grad + oW1, w.part, w.chunk, w.nchunk, st))) return e;
    if ((e = bgrad(w.G0, B, grad + oB1, w.part, w.chunk, w.nchunk, st))) return e;
    return hipGetLastError();
}

hipError_t mlp_adam(float *P, uint16_t *Pb, const float *grad, float *m, float *v, int step,
                    float lr, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    const int n = (int)kParams;
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, P, grad, m, v, n, lr, b1, b2,
                       eps, c1, c2, Pb, (int)kParamsBf16);
    return hipGetLastError();
}

hipError_t mlp_sync_bf16(const float *P, uint16_t *Pb, hipStream_t st) {
    const int n = (int)kParamsBf16;
    hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, st, P, n, Pb);
    return hipGetLastError();
}

// Proposal features of each candidate (64 bf16 per row).  The pFq family is
// symmetric under permutations of its x parameters and, separately, of its
// y parameters, so each group's (direction, start) pairs are sorted
// (direction descending, then start descending) before they are encoded:
// permuted candidates get identical rows.  Slots:
//   0..7   x starts / 16 (sorted order), 8..15 y starts / 16
//   16..23 x directions / 4,             24..31 y directions / 4
//   32 z; then the outcome of the K-step feature walk:
//   33 valid, 34 fraction of the prefix's axis steps executed,
//   35 log scale / K, 36 delta clipped to [-4, 4] / 4, 37 finite-delta flag,
//   38 precision (min 32) / 16 (2 if infinite), 39 tanh(limit),
//   40 n_limits / 256, 41 sum |dir| / 64,
//   42 log|q_K| + log scale (the growth of the convergent's denominator,
//      which decides delta's scale) / 16 clipped to [-4, 4], 43 its sign > 0,
//   44 log|p_K| + log scale / 16 clipped to [-4, 4].
template <int K>
__device__ __forceinline__ void sort_pairs(int (&d)[K], double (&v)[K]) {
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int j = 0; j + 1 < K - i; ++j) {
            const bool sw = d[j] < d[j + 1] || (d[j] == d[j + 1] && v[j] < v[j + 1]);
            const int td = sw ? d[j + 1] : d[j];
            const int tu = sw ? d[j] : d[j + 1];
            const double vd = sw ? v[j + 1] : v[j];
            const double vu = sw ? v[j] : v[j + 1];
            d[j] = td;
            d[j + 1] = tu;
            v[j] = vd;
            v[j + 1] = vu;
        }
}

__global__ void k_features(WalkArgs A, int K, uint16_t *X) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.n) return;
    int dx[8], dy[8];
    double sx[8], sy[8];
    int sdir = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const bool hx = j < A.p, hy = j < A.q;
        // absent axes sort last
        dx[j] = hx ? A.dir[(size_t)j * A.ld + t] : INT_MIN;
        sx[j] = hx ? A.start[(size_t)j * A.ld + t] : 0.0;
        dy[j] = hy ? A.dir[(size_t)(A.p + j) * A.ld + t] : INT_MIN;
        sy[j] = hy ? A.start[(size_t)(A.p + j) * A.ld + t] : 0.0;
        sdir += (hx ? abs(dx[j]) : 0) + (hy ? abs(dy[j]) : 0);
    }
    sort_pairs(dx, sx);
    sort_pairs(dy, sy);
    float f[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) f[i] = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j < A.p) {
            f[j] = (float)(sx[j] / 16.0);
            f[16 + j] = dx[j] / 4.f;
        }
        if (j < A.q) {
            f[8 + j] = (float)(sy[j] / 16.0);
            f[24 + j] = dy[j] / 4.f;
        }
    }
    f[32] = (float)A.z;
    const bool v = A.valid[t] != 0;
    f[33] = v ? 1.f : 0.f;
    const float steps = (float)(A.n_fwd[t] + A.n_inv[t]);
    f[34] = sdir > 0 ? steps / ((float)K * sdir) : 0.f;
    if (v) {
        const double ls = A.log_scale[t], dl = A.delta[t], pr = A.precision[t], lm = A.limit[t];
        f[35] = isfinite(ls) ? (float)(ls / K) : 0.f;
        const bool fd = isfinite(dl);
        f[36] = fd ? (float)(fmin(fmax(dl, -4.0), 4.0) / 4.0) : 0.f;
        f[37] = fd ? 1.f : 0.f;
        f[38] = isfinite(pr) ? (float)(fmin(pr, 32.0) / 16.0) : 2.f;
        f[39] = isfinite(lm) ? tanhf((float)lm) : 0.f;
        f[40] = A.n_limits[t] / 256.f;
        const double lq = log(fabs(A.rq[t])) + ls, lp = log(fabs(A.rp[t])) + ls;
        if (isfinite(lq)) {
            f[42] = (float)(fmin(fmax(lq / 16.0, -4.0), 4.0));
            f[43] = lq > 0.0 ? 1.f : 0.f;
        }
        if (isfinite(lp)) f[44] = (float)(fmin(fmax(lp / 16.0, -4.0), 4.0));
    }
    f[41] = sdir / 64.f;
    uint16_t *row = X + (size_t)t * kIn;
#pragma unroll
    for (int i = 0; i < 64; ++i) row[i] = f2bf(f[i]);
}

hipError_t mlp_features(const WalkArgs &A, int K, uint16_t *X, hipStream_t st) {
    hipLaunchKernelGGL(k_features, dim3((A.n + 255) / 256), dim3(256), 0, st, A, K, X);
    return hipGetLastError();
}

}  // namespace cmfw
