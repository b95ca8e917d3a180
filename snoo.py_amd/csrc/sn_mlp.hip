// Kernels of the learned warm start: BF16 GEMM on the gfx950 matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation) with fused bias / GELU /
// GELU-derivative epilogues, plus the element-wise pieces of training (MSE
// gradient, bias gradient, Adam) and of inference (POD expansion, mesh-to-
// mesh transfer of edge DOFs).  No counterpart in the reference: its solve
// always starts from A = 0 or a caller-supplied x_0
// (snoopy/red_mvp_solver.py:146-147,174); this fills x_0.
#include <algorithm>
#include <cstdint>

#include "sn_common.hpp"

namespace sn {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;

__device__ __forceinline__ __bf16 to_bf16(float v) { return (__bf16)v; }
__device__ __forceinline__ float from_bf16(__bf16 v) { return (float)v; }

__device__ __forceinline__ float gelu(float z) {
    return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
    // Phi(z) + z phi(z)
    const float Phi = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float phi = 0.39894228040143268f * expf(-0.5f * z * z);
    return Phi + z * phi;
}

// C[m][n] = sum_z part[z][m][n] over the split-K layers, in layer order
__global__ void splitk_sum_kernel(int M, int N, int ldc, int S, const float *__restrict__ part,
                                  float *__restrict__ C) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)M * N) return;
    const size_t o = (size_t)(t / N) * ldc + (size_t)(t % N);
    float s = 0.0f;
    for (int z = 0; z < S; ++z) s += part[(size_t)z * M * ldc + o];
    C[o] = s;
}

// Large-tile variant: 128x128 tile, BK = 32, 4 waves in 2x2, each wave a
// 64x64 sub-tile of 2x2 MFMA 32x32 tiles (8 MFMAs per wave per k-tile);
// global -> register -> LDS staging with 16-byte loads along the contiguous
// dimension, double-buffered LDS so one barrier per k-tile separates the
// writes of the next tile from the reads of the current one.  VEC = 16-byte
// loads allowed (leading dimensions multiples of 8, 16-byte aligned bases);
// vectors that cross the matrix edge are loaded element by element.
constexpr int tBM = 128, tBN = 128, tBK = 32, tPad = 8;

template <bool VEC>
__device__ __forceinline__ bf16x8 load8(const __bf16 *__restrict__ p, size_t off, int n_ok) {
    bf16x8 v;
    if (VEC && n_ok >= 8) {
        v = *reinterpret_cast<const bf16x8 *>(p + off);
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = (i < n_ok) ? p[off + i] : to_bf16(0.0f);
    }
    return v;
}

template <int TA, int TB, int EPI, bool VEC>
__global__ __launch_bounds__(256) void gemm_bf16_big_kernel(
    int M, int N, int K, const __bf16 *__restrict__ A, int lda, const __bf16 *__restrict__ B,
    int ldb, const float *__restrict__ bias, const float *__restrict__ Zin,
    float *__restrict__ Zout, float *__restrict__ Cf, __bf16 *__restrict__ Cb, int ldc) {
    __shared__ __attribute__((aligned(16))) __bf16 As[2][tBM][tBK + tPad];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][tBN][tBK + tPad];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int m0 = blockIdx.y * tBM, n0 = blockIdx.x * tBN;
    f32x16 acc[2][2] = {};
    bf16x8 ra[2], rb[2];
    // 512 vectors of 8 per operand tile, two per thread
    auto load_tile = [&](int k0) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int v = tid + 256 * s;
            if (TA == 0) {           // rows m, k contiguous
                const int m = v >> 2, kv = (v & 3) * 8, gm = m0 + m, gk = k0 + kv;
                const int ok = (gm < M) ? (K - gk < 8 ? (K - gk > 0 ? K - gk : 0) : 8) : 0;
                ra[s] = load8<VEC>(A, (size_t)gm * lda + gk, ok);
            } else {                 // rows k, m contiguous
                const int k = v >> 4, mv = (v & 15) * 8, gk = k0 + k, gm = m0 + mv;
                const int ok = (gk < K) ? (M - gm < 8 ? (M - gm > 0 ? M - gm : 0) : 8) : 0;
                ra[s] = load8<VEC>(A, (size_t)gk * lda + gm, ok);
            }
            if (TB == 1) {           // rows n, k contiguous
                const int n = v >> 2, kv = (v & 3) * 8, gn = n0 + n, gk = k0 + kv;
                const int ok = (gn < N) ? (K - gk < 8 ? (K - gk > 0 ? K - gk : 0) : 8) : 0;
                rb[s] = load8<VEC>(B, (size_t)gn * ldb + gk, ok);
            } else {                 // rows k, n contiguous
                const int k = v >> 4, nv = (v & 15) * 8, gk = k0 + k, gn = n0 + nv;
                const int ok = (gk < K) ? (N - gn < 8 ? (N - gn > 0 ? N - gn : 0) : 8) : 0;
                rb[s] = load8<VEC>(B, (size_t)gk * ldb + gn, ok);
            }
        }
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int v = tid + 256 * s;
            if (TA == 0) {
                *reinterpret_cast<bf16x8 *>(&As[buf][v >> 2][(v & 3) * 8]) = ra[s];
            } else {
                const int k = v >> 4, mv = (v & 15) * 8;
#pragma unroll
                for (int i = 0; i < 8; ++i) As[buf][mv + i][k] = ra[s][i];
            }
            if (TB == 1) {
                *reinterpret_cast<bf16x8 *>(&Bs[buf][v >> 2][(v & 3) * 8]) = rb[s];
            } else {
                const int k = v >> 4, nv = (v & 15) * 8;
#pragma unroll
                for (int i = 0; i < 8; ++i) Bs[buf][nv + i][k] = rb[s][i];
            }
        }
    };
    const int nk = (K + tBK - 1) / tBK;
    load_tile(0);
    store_tile(0);
    __syncthreads();
    const int r = lane & 31, kh = (lane >> 5) * 8;
    for (int t = 0; t < nk; ++t) {
        const int cur = t & 1;
        if (t + 1 < nk) load_tile((t + 1) * tBK);
#pragma unroll
        for (int kk = 0; kk < tBK; kk += 16) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                a[i] = *reinterpret_cast<const bf16x8 *>(&As[cur][wm + 32 * i + r][kk + kh]);
                b[i] = *reinterpret_cast<const bf16x8 *>(&Bs[cur][wn + 32 * i + r][kk + kh]);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0,
                                                                          0, 0);
        }
        if (t + 1 < nk) store_tile(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + 32 * j + (lane & 31);
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int row = m0 + wm + 32 * i + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                if (row >= M || col >= N) continue;
                float v = acc[i][j][q];
                if (bias) v += bias[col];
                const size_t o = (size_t)row * ldc + col;
                if (EPI == 1) {
                    if (Zout) Zout[o] = v;
                    v = gelu(v);
                } else if (EPI == 2) {
                    v *= gelu_grad(Zin[o]);
                }
                if (Cf) Cf[o] = v;
                if (Cb) Cb[o] = to_bf16(v);
            }
        }
}

// C[M,N] = op(A)[M,K] op(B)[K,N]
//   TA = 0: A stored [M][lda] (row m, k contiguous); TA = 1: A stored [K][lda] (A^T)
//   TB = 0: B stored [K][ldb];                       TB = 1: B stored [N][ldb] (B^T)
// epilogue: v = acc (+ bias[n]); EPI 1: Z = v stored (fp32), v = gelu(v);
//           EPI 2: v *= gelu'(Zin[m][n]) (backward through GELU).
// Outputs Cf (fp32) and/or Cb (bf16), both [M][ldc].
// Small-tile form for products with few output tiles (the network's own
// layers, M <= a few thousand, N, K <= 512): 64x64 tile, BK = 64, 4 waves in
// 2x2, each wave a 32x32 sub-tile (4 MFMAs 32x32x16 per k-tile); 16-byte
// global loads along the contiguous dimension staged through registers into
// double-buffered LDS, the next k-tile's loads in flight during the current
// tile's MFMAs, one barrier per k-tile.  Split K: workgroup layer z sums k
// in [z kslice, (z+1) kslice) into its own partial output Cf + z M ldc.
constexpr int kBM = 64, kBN = 64, kBK = 64, kPad = 8;

template <int TA, int TB, int EPI, bool VEC>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(
    int M, int N, int K, const __bf16 *__restrict__ A, int lda, const __bf16 *__restrict__ B,
    int ldb, const float *__restrict__ bias, const float *__restrict__ Zin,
    float *__restrict__ Zout, float *__restrict__ Cf, __bf16 *__restrict__ Cb, int ldc,
    int kslice) {
    __shared__ __attribute__((aligned(16))) __bf16 As[2][kBM][kBK + kPad];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][kBN][kBK + kPad];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
    const int kz0 = blockIdx.z * kslice, kz1 = min(K, kz0 + kslice);
    if (Cf) Cf += (size_t)blockIdx.z * M * ldc;
    f32x16 acc = {};
    bf16x8 ra[2], rb[2];
    auto clamp8 = [](int n) { return n < 0 ? 0 : (n > 8 ? 8 : n); };
    // 512 vectors of 8 per 64x64 operand tile, two per thread
    auto load_tile = [&](int k0) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int v = tid + 256 * s;
            if (TA == 0) {           // rows m, k contiguous
                const int m = v >> 3, kv = (v & 7) * 8, gm = m0 + m, gk = k0 + kv;
                ra[s] = load8<VEC>(A, (size_t)gm * lda + gk, gm < M ? clamp8(kz1 - gk) : 0);
            } else {                 // rows k, m contiguous
                const int k = v >> 3, mv = (v & 7) * 8, gk = k0 + k, gm = m0 + mv;
                ra[s] = load8<VEC>(A, (size_t)gk * lda + gm, gk < kz1 ? clamp8(M - gm) : 0);
            }
            if (TB == 1) {           // rows n, k contiguous
                const int n = v >> 3, kv = (v & 7) * 8, gn = n0 + n, gk = k0 + kv;
                rb[s] = load8<VEC>(B, (size_t)gn * ldb + gk, gn < N ? clamp8(kz1 - gk) : 0);
            } else {                 // rows k, n contiguous
                const int k = v >> 3, nv = (v & 7) * 8, gk = k0 + k, gn = n0 + nv;
                rb[s] = load8<VEC>(B, (size_t)gk * ldb + gn, gk < kz1 ? clamp8(N - gn) : 0);
            }
        }
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int v = tid + 256 * s;
            if (TA == 0) {
                *reinterpret_cast<bf16x8 *>(&As[buf][v >> 3][(v & 7) * 8]) = ra[s];
            } else {
                const int k = v >> 3, mv = (v & 7) * 8;
#pragma unroll
                for (int i = 0; i < 8; ++i) As[buf][mv + i][k] = ra[s][i];
            }
            if (TB == 1) {
                *reinterpret_cast<bf16x8 *>(&Bs[buf][v >> 3][(v & 7) * 8]) = rb[s];
            } else {
                const int k = v >> 3, nv = (v & 7) * 8;
#pragma unroll
                for (int i = 0; i < 8; ++i) Bs[buf][nv + i][k] = rb[s][i];
            }
        }
    };
    const int nk = kz1 > kz0 ? (kz1 - kz0 + kBK - 1) / kBK : 0;
    if (nk > 0) {
        load_tile(kz0);
        store_tile(0);
    }
    __syncthreads();
    // lane l: row/column l % 32, k = 8 * (l / 32) .. +7 of each 16-wide step
    const int r = lane & 31, kh = (lane >> 5) * 8;
    for (int t = 0; t < nk; ++t) {
        const int cur = t & 1;
        if (t + 1 < nk) load_tile(kz0 + (t + 1) * kBK);
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 16) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[cur][wm + r][kk + kh]);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bs[cur][wn + r][kk + kh]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        if (t + 1 < nk) store_tile(cur ^ 1);
        __syncthreads();
    }
    // D layout: lane l, register i -> column l % 32, row 8*(i/4) + 4*(l/32) + i%4
    const int col = n0 + wn + (lane & 31);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        if (row >= M || col >= N) continue;
        float v = acc[i];
        if (bias) v += bias[col];
        const size_t o = (size_t)row * ldc + col;
        if (EPI == 1) {
            if (Zout) Zout[o] = v;
            v = gelu(v);
        } else if (EPI == 2) {
            v *= gelu_grad(Zin[o]);
        }
        if (Cf) Cf[o] = v;
        if (Cb) Cb[o] = to_bf16(v);
    }
}

// dY = 2 (Y - T) * scale, loss partial sums of (Y - T)^2 per block
__global__ void mse_grad_kernel(int n, const float *__restrict__ Y, const float *__restrict__ T,
                                float scale, float *__restrict__ dYf, __bf16 *__restrict__ dYb,
                                double *__restrict__ partial) {
    __shared__ double lds[256 / kWave];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        const float g = 2.0f * d * scale;
        dYf[i] = g;
        dYb[i] = to_bf16(g);
        s += (double)d * (double)d;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// db[n] = sum_m dZ[m][n]: 8 columns per workgroup, each thread four
// adjacent columns of every 128th row (row m goes to group m % 128, summed
// in row order), the 128 group sums combined by a fixed pairwise tree --
// deterministic.  The layers' N of a few hundred gives N / 8 workgroups,
// each thread a short run of independent loads.
constexpr int kColW = 8, kColG = 128;
template <bool VEC>
__global__ __launch_bounds__(256) void colsum_kernel(int M, int N, const float *__restrict__ dZ,
                                                     float *__restrict__ db) {
    __shared__ float part[kColG][kColW];
    const int q = threadIdx.x & 1, g = threadIdx.x >> 1;
    const int n = blockIdx.x * kColW + 4 * q;
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (VEC && n + 4 <= N) {
        for (int m = g; m < M; m += kColG) {
            const float4 v = *reinterpret_cast<const float4 *>(dZ + (size_t)m * N + n);
            s[0] += v.x;
            s[1] += v.y;
            s[2] += v.z;
            s[3] += v.w;
        }
    } else {
        for (int m = g; m < M; m += kColG)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (n + j < N) s[j] += dZ[(size_t)m * N + n + j];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) part[g][4 * q + j] = s[j];
    __syncthreads();
    for (int h = kColG / 2; h > 0; h >>= 1) {
        if (g < h)
#pragma unroll
            for (int j = 0; j < 4; ++j) part[g][4 * q + j] += part[g + h][4 * q + j];
        __syncthreads();
    }
    if (g == 0)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (n + j < N) db[n + j] = part[0][4 * q + j];
}

// Adam on FP32 master weights; refreshes the BF16 copy
__global__ void adam_kernel(int n, float *__restrict__ p, const float *__restrict__ g,
                            float *__restrict__ m1, float *__restrict__ m2, __bf16 *__restrict__ pb,
                            float lr, float b1, float b2, float eps, float c1, float c2,
                            float wd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = v;
    if (pb) pb[i] = to_bf16(v);
}

__global__ void f32_to_bf16_kernel(int n, const float *__restrict__ x, __bf16 *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// out[i] = mean[i] + sum_j U[j*n + i] c[j]   (U column-major: mode j contiguous)
__global__ void pod_expand_kernel(int n, int r, const float *__restrict__ U,
                                  const float *__restrict__ c, const double *__restrict__ mean,
                                  double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = mean ? mean[i] : 0.0;
    // eight modes' loads issued together, summed in mode order
    int j = 0;
    for (; j + 8 <= r; j += 8) {
        float u[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) u[m] = U[(size_t)(j + m) * n + i];
#pragma unroll
        for (int m = 0; m < 8; ++m) s += (double)u[m] * (double)c[j + m];
    }
    for (; j < r; ++j) s += (double)U[(size_t)j * n + i] * (double)c[j];
    out[i] = s;
}

// pod_expand_kernel for nvec coefficient vectors c[w*r + j] at once, out[w*n
// + i]: one pass over the modes serves every vector (the coefficients staged
// in LDS); each output summed in the order of pod_expand_kernel
constexpr int kExpandMaxVec = 8, kExpandMaxModes = 256;
template <int NVEC>
__global__ __launch_bounds__(256) void pod_expand_many_kernel(int n, int r,
                                                              const float *__restrict__ U,
                                                              const float *__restrict__ c,
                                                              const double *__restrict__ mean,
                                                              double *__restrict__ out) {
    __shared__ float cs[NVEC][kExpandMaxModes];
    for (int t = threadIdx.x; t < NVEC * r; t += blockDim.x) cs[t / r][t % r] = c[t];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s[NVEC];
    const double m0 = mean ? mean[i] : 0.0;
#pragma unroll
    for (int w = 0; w < NVEC; ++w) s[w] = m0;
    int j = 0;
    for (; j + 8 <= r; j += 8) {
        float u[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) u[m] = U[(size_t)(j + m) * n + i];
#pragma unroll
        for (int w = 0; w < NVEC; ++w)
#pragma unroll
            for (int m = 0; m < 8; ++m) s[w] += (double)u[m] * (double)cs[w][j + m];
    }
    for (; j < r; ++j) {
        const float u = U[(size_t)j * n + i];
#pragma unroll
        for (int w = 0; w < NVEC; ++w) s[w] += (double)u * (double)cs[w][j];
    }
#pragma unroll
    for (int w = 0; w < NVEC; ++w) out[(size_t)w * n + i] = s[w];
}

// x_f[i] = sum_k w[6i+k] x_c[idx[6i+k]]  (idx < 0: no contribution)
__global__ void edge_transfer_kernel(int nf, const int *__restrict__ idx,
                                     const double *__restrict__ w, const double *__restrict__ xc,
                                     double *__restrict__ xf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int j = idx[6 * i + k];
        if (j >= 0) s += w[6 * i + k] * xc[j];
    }
    xf[i] = s;
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_gemm_bf16(int ta, int tb, int epi, int M, int N, int K, const void *A, int lda,
                 const void *B, int ldb, const float *bias, const float *Zin, float *Zout,
                 float *Cf, void *Cb, int ldc, void *stream) {
    SN_REQUIRE(M > 0 && N > 0 && K > 0, "sn_gemm_bf16: bad shape %dx%dx%d", M, N, K);
    SN_REQUIRE(A && B && (Cf || Cb), "sn_gemm_bf16: null argument");
    SN_REQUIRE(epi >= 0 && epi <= 2 && ta >= 0 && ta <= 1 && tb >= 0 && tb <= 1,
               "sn_gemm_bf16: bad flags");
    SN_REQUIRE(epi != 2 || Zin, "sn_gemm_bf16: GELU backward needs Zin");
    SN_REQUIRE(lda >= (ta ? M : K) && ldb >= (tb ? K : N) && ldc >= N,
               "sn_gemm_bf16: leading dimension too small");
    hipStream_t s = as_stream(stream);
    const __bf16 *a = (const __bf16 *)A;
    const __bf16 *b = (const __bf16 *)B;
    __bf16 *cb = (__bf16 *)Cb;
    // large tiles when they still give every CU a workgroup
    const long long big_tiles = (long long)((N + tBN - 1) / tBN) * ((M + tBM - 1) / tBM);
    const bool big = big_tiles >= num_cus();
    const bool vec = (lda % 8 == 0) && (ldb % 8 == 0) &&
                     (reinterpret_cast<uintptr_t>(A) % 16 == 0) &&
                     (reinterpret_cast<uintptr_t>(B) % 16 == 0);
    const dim3 grid = big ? dim3((N + tBN - 1) / tBN, (M + tBM - 1) / tBM)
                          : dim3((N + kBN - 1) / kBN, (M + kBM - 1) / kBM);
    const dim3 block(256);
    // a plain product with few output tiles and a long K (the weight
    // gradients A^T dZ of the network) is split along K over several
    // workgroup layers, summed afterwards in layer order
    const int tiles = (int)grid.x * (int)grid.y;
    int S = 1, kslice = K;
    float *Cdst = Cf;
    float *part = nullptr;
    if (!big && epi == 0 && Cf && !Cb && !bias && 2 * tiles <= num_cus() && K >= 512) {
        S = std::min(K / 256, std::max(1, num_cus() / tiles));
        kslice = ((K + S - 1) / S + kBK - 1) / kBK * kBK;
        S = (K + kslice - 1) / kslice;
    }
    if (S > 1) {
        SN_HIP(hipMallocAsync((void **)&part, (size_t)S * M * ldc * sizeof(float), s));
        Cdst = part;
    }
    const dim3 gridz(grid.x, grid.y, S);
#define SN_G(TA, TB, E)                                                                          \
    do {                                                                                         \
        if (!big && vec)                                                                         \
            hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, E, true>), gridz, block, 0, s, M, N, K,  \
                               a, lda, b, ldb, bias, Zin, Zout, Cdst, cb, ldc, kslice);          \
        else if (!big)                                                                           \
            hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, E, false>), gridz, block, 0, s, M, N, K, \
                               a, lda, b, ldb, bias, Zin, Zout, Cdst, cb, ldc, kslice);          \
        else if (vec)                                                                            \
            hipLaunchKernelGGL((gemm_bf16_big_kernel<TA, TB, E, true>), grid, block, 0, s, M, N,  \
                               K, a, lda, b, ldb, bias, Zin, Zout, Cf, cb, ldc);                 \
        else                                                                                     \
            hipLaunchKernelGGL((gemm_bf16_big_kernel<TA, TB, E, false>), grid, block, 0, s, M, N, \
                               K, a, lda, b, ldb, bias, Zin, Zout, Cf, cb, ldc);                 \
    } while (0)
#define SN_GE(TA, TB) \
    if (epi == 0) SN_G(TA, TB, 0); else if (epi == 1) SN_G(TA, TB, 1); else SN_G(TA, TB, 2)
    if (ta == 0 && tb == 0) { SN_GE(0, 0); }
    else if (ta == 0 && tb == 1) { SN_GE(0, 1); }
    else if (ta == 1 && tb == 0) { SN_GE(1, 0); }
    else { SN_GE(1, 1); }
#undef SN_GE
#undef SN_G
    SN_CHECK_LAUNCH();
    if (S > 1) {
        const long long mn = (long long)M * N;
        hipLaunchKernelGGL(splitk_sum_kernel, dim3(grid_for(mn, 256, 1 << 30)), dim3(256), 0, s, M,
                           N, ldc, S, part, Cf);
        SN_CHECK_LAUNCH();
        SN_HIP(hipFreeAsync(part, s));
    }
    return SN_OK;
}

int sn_mse_grad(int n, const float *Y, const float *T, float scale, float *dYf, void *dYb,
                double *loss_host, void *stream) {
    SN_REQUIRE(n > 0 && Y && T && dYf && dYb, "sn_mse_grad: bad argument");
    hipStream_t s = as_stream(stream);
    const int G = grid_for(n, 256, 256);
    double *part = nullptr;
    SN_HIP(hipMallocAsync(&part, G * sizeof(double), s));
    hipLaunchKernelGGL(mse_grad_kernel, dim3(G), dim3(256), 0, s, n, Y, T, scale, dYf,
                       (__bf16 *)dYb, part);
    SN_CHECK_LAUNCH();
    if (loss_host) {
        double h[256];
        SN_HIP(hipMemcpyAsync(h, part, G * sizeof(double), hipMemcpyDeviceToHost, s));
        SN_HIP(hipStreamSynchronize(s));
        double t = 0.0;
        for (int k = 0; k < G; ++k) t += h[k];
        *loss_host = t / n;
    }
    SN_HIP(hipFreeAsync(part, s));
    return SN_OK;
}

int sn_colsum_f32(int M, int N, const float *dZ, float *db, void *stream) {
    SN_REQUIRE(M > 0 && N > 0 && dZ && db, "sn_colsum_f32: bad argument");
    const dim3 grid((N + kColW - 1) / kColW);
    if (N % 4 == 0 && reinterpret_cast<uintptr_t>(dZ) % 16 == 0)
        hipLaunchKernelGGL(colsum_kernel<true>, grid, dim3(256), 0, as_stream(stream), M, N, dZ,
                           db);
    else
        hipLaunchKernelGGL(colsum_kernel<false>, grid, dim3(256), 0, as_stream(stream), M, N, dZ,
                           db);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_adam_f32(int n, float *p, const float *g, float *m1, float *m2, void *pb, float lr,
                float b1, float b2, float eps, int step, float wd, void *stream) {
    SN_REQUIRE(n > 0 && p && g && m1 && m2 && step >= 1, "sn_adam_f32: bad argument");
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), n, p,
                       g, m1, m2, (__bf16 *)pb, lr, b1, b2, eps, c1, c2, wd);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_f32_to_bf16(int n, const float *x, void *y, void *stream) {
    SN_REQUIRE(n >= 0 && (n == 0 || (x && y)), "sn_f32_to_bf16: bad argument");
    if (n == 0) return SN_OK;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream),
                       n, x, (__bf16 *)y);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_pod_expand(int n, int r, const float *U, const float *c, const double *mean, double *out,
                  void *stream) {
    SN_REQUIRE(n > 0 && r >= 0 && out && (r == 0 || (U && c)), "sn_pod_expand: bad argument");
    hipLaunchKernelGGL(pod_expand_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream),
                       n, r, U, c, mean, out);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_pod_expand_many(int n, int r, int nvec, const float *U, const float *c, const double *mean,
                       double *out, void *stream) {
    SN_REQUIRE(n > 0 && r >= 0 && out && (r == 0 || (U && c)), "sn_pod_expand_many: bad argument");
    SN_REQUIRE(nvec == 1 || nvec == 2 || nvec == 4 || nvec == 8,
               "sn_pod_expand_many: nvec must be 1, 2, 4 or 8 (%d)", nvec);
    SN_REQUIRE(r <= kExpandMaxModes, "sn_pod_expand_many: at most %d modes (%d)", kExpandMaxModes,
               r);
    const dim3 grid((n + 255) / 256);
    hipStream_t s = as_stream(stream);
    switch (nvec) {
        case 1: hipLaunchKernelGGL(pod_expand_many_kernel<1>, grid, dim3(256), 0, s, n, r, U, c, mean, out); break;
        case 2: hipLaunchKernelGGL(pod_expand_many_kernel<2>, grid, dim3(256), 0, s, n, r, U, c, mean, out); break;
        case 4: hipLaunchKernelGGL(pod_expand_many_kernel<4>, grid, dim3(256), 0, s, n, r, U, c, mean, out); break;
        default: hipLaunchKernelGGL(pod_expand_many_kernel<8>, grid, dim3(256), 0, s, n, r, U, c, mean, out); break;
    }
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_edge_transfer(int nf, const int *idx, const double *w, const double *xc, double *xf,
                     void *stream) {
    SN_REQUIRE(nf >= 0 && (nf == 0 || (idx && w && xc && xf)), "sn_edge_transfer: bad argument");
    if (nf == 0) return SN_OK;
    hipLaunchKernelGGL(edge_transfer_kernel, dim3((nf + 255) / 256), dim3(256), 0,
                       as_stream(stream), nf, idx, w, xc, xf);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

}  // extern "C"
