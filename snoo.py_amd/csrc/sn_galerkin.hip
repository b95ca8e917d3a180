// Reduced (POD-Galerkin) Newton step of the learned warm start.  No
// counterpart in the reference, whose Newton starts from A = 0 or a given x_0
// (snoopy/red_mvp_solver.py:146-147,174).
//
// With U the first r POD modes (U[k*n + i], mode k contiguous) and the
// Jacobian in the reference-curl form of the element kernel (J_e[j][l] =
// s_j s_l c_j . Mr_e c_l, Mr_e a symmetric 3x3 per element), the reduced
// Jacobian is
//     H = U^T J U = sum_e W_e^T Mr_e W_e,   W_e[:, k] = sum_j s_j c_j U[k][g_ej],
// where W_e (3 x r) are the face circulations of the modes on element e.  W
// depends on the mesh topology and the modes only, so it is formed once per
// trained model; every correction then streams W and the 6 Mr values per
// element once.  Arithmetic: FP32 products with FP32 accumulation inside a
// workgroup's element range, partial sums combined in FP64 in a fixed order.
#include "sn_common.hpp"

namespace sn {

// rows of the integer reference curls c_j of the local edges (0,1), (0,2),
// (0,3), (1,2), (1,3), (2,3), the convention of the element kernels
__constant__ int kCurl[6][3] = {{0, -1, 1}, {1, 0, -1}, {-1, 1, 0},
                                {0, 0, 1},  {0, -1, 0}, {1, 0, 0}};

__global__ void pod_element_modes_kernel(int num_el, int n, int r, const float *__restrict__ U,
                                         const int *__restrict__ gids,
                                         const int *__restrict__ orient, float *__restrict__ W) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)num_el * r) return;
    const int e = (int)(t % num_el);     // element fastest: neighbouring DOFs per wave
    const int k = (int)(t / num_el);
    const int o = orient[e];
    float w0 = 0.0f, w1 = 0.0f, w2 = 0.0f;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const float u = U[(size_t)k * n + gids[6 * e + j]];
        const float su = ((o >> j) & 1) ? -u : u;
        w0 += kCurl[j][0] * su;
        w1 += kCurl[j][1] * su;
        w2 += kCurl[j][2] * su;
    }
    float *dst = W + (size_t)e * 3 * r;
    dst[k] = w0;
    dst[r + k] = w1;
    dst[2 * r + k] = w2;
}

constexpr int kGalBlock = 256;
constexpr int kGalEB = 16;   // elements staged in LDS per step

// Partial H of a contiguous element range per workgroup; thread t owns the
// TK x TK block of rows (t / 16) * TK.. and columns (t % 16) * TK.. of H.
// W and Mr start at the first element of the range; Mr value u of element e
// is Mr[u * mr_stride + e] (the element arrays' own stride).
template <int R>
__global__ __launch_bounds__(kGalBlock) void pod_reduced_jacobian_kernel(
    int num_el, int mr_stride, const float *__restrict__ W, const float *__restrict__ Mr,
    float *__restrict__ part) {
    constexpr int TK = R / 16;
    __shared__ __align__(16) float sW[kGalEB][3][R];
    __shared__ __align__(16) float sZ[kGalEB][3][R];
    __shared__ float sM[kGalEB][6];
    const int tid = threadIdx.x;
    const int kb = (tid / 16) * TK, lb = (tid % 16) * TK;
    float acc[TK][TK];
#pragma unroll
    for (int a = 0; a < TK; ++a)
#pragma unroll
        for (int b = 0; b < TK; ++b) acc[a][b] = 0.0f;
    const int nb = (num_el + kGalEB - 1) / kGalEB;
    const int per = (nb + gridDim.x - 1) / gridDim.x;
    const int b0 = blockIdx.x * per, b1 = min(nb, b0 + per);
    for (int bt = b0; bt < b1; ++bt) {
        const int e0 = bt * kGalEB;
        // the batch's W rows are contiguous: float4 loads (3 R is a multiple of 4)
        const float4 *src = reinterpret_cast<const float4 *>(W + (size_t)e0 * 3 * R);
        float4 *dstW = reinterpret_cast<float4 *>(&sW[0][0][0]);
        for (int v = tid; v < kGalEB * 3 * R / 4; v += kGalBlock) {
            const int e = e0 + (4 * v) / (3 * R);
            dstW[v] = (e < num_el) ? src[v] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        if (tid < kGalEB * 6) {
            const int le = tid / 6, u = tid % 6, e = e0 + le;
            sM[le][u] = (e < num_el) ? Mr[(size_t)u * mr_stride + e] : 0.0f;
        }
        __syncthreads();
        // Z_e = Mr_e W_e (Mr upper triangle 00 01 02 11 12 22)
        for (int v = tid; v < kGalEB * R; v += kGalBlock) {
            const int le = v / R, l = v % R;
            const float *m = sM[le];
            const float x0 = sW[le][0][l], x1 = sW[le][1][l], x2 = sW[le][2][l];
            sZ[le][0][l] = m[0] * x0 + m[1] * x1 + m[2] * x2;
            sZ[le][1][l] = m[1] * x0 + m[3] * x1 + m[4] * x2;
            sZ[le][2][l] = m[2] * x0 + m[4] * x1 + m[5] * x2;
        }
        __syncthreads();
        for (int le = 0; le < kGalEB; ++le)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float wk[TK], zl[TK];
#pragma unroll
                for (int q = 0; q < TK; q += 2) {     // 8-byte LDS reads (TK is even)
                    const float2 w2 = *reinterpret_cast<const float2 *>(&sW[le][a][kb + q]);
                    const float2 z2 = *reinterpret_cast<const float2 *>(&sZ[le][a][lb + q]);
                    wk[q] = w2.x;
                    wk[q + 1] = w2.y;
                    zl[q] = z2.x;
                    zl[q + 1] = z2.y;
                }
#pragma unroll
                for (int p = 0; p < TK; ++p)
#pragma unroll
                    for (int q = 0; q < TK; ++q) acc[p][q] += wk[p] * zl[q];
            }
        __syncthreads();
    }
    float *dst = part + (size_t)blockIdx.x * R * R;
#pragma unroll
    for (int p = 0; p < TK; ++p)
#pragma unroll
        for (int q = 0; q < TK; ++q) dst[(kb + p) * R + lb + q] = acc[p][q];
}

// The same sum for r = 128 on the matrix cores.  Over one element chunk the
// K = 3 EB rows k = 3 le + a of W and of Z = Mr W are staged in LDS, and
//     H[p][q] += sum_k W[k][p] Z[k][q]
// is a 128 x 128 x K product.  H is symmetric, so only its 10 upper 32 x 32
// tiles are formed, two per wave (five waves), with v_mfma_f32_32x32x2_f32
// (FP32 operands, FP32 accumulation); the partial writes each off-diagonal
// tile twice, as itself and transposed.  Lane layouts of the 32x32x2 form:
// A[i][kk] and B[kk][j] in lane i + 32 kk (resp. j + 32 kk); C[i][j] in lane
// j, register 8 (q / 4) + 4 (lane / 32) + q % 4 = i.
typedef __attribute__((ext_vector_type(16))) float gal_f32x16;
constexpr int kGalMfmaEB = 16;
static_assert(kGalMfmaEB == kGalEB, "the launch sizes its grid with the chunk of kGalEB elements");
constexpr int kGalMfmaWaves = 5;
constexpr int kGalMfmaBlock = 64 * kGalMfmaWaves;
__constant__ int kGalUpperTiles[10][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 1},
                                          {1, 2}, {1, 3}, {2, 2}, {2, 3}, {3, 3}};

__global__ __launch_bounds__(kGalMfmaBlock) void pod_reduced_jacobian_mfma_kernel(
    int num_el, int mr_stride, const float *__restrict__ W, const float *__restrict__ Mr,
    float *__restrict__ part) {
    constexpr int R = 128, EB = kGalMfmaEB, K = 3 * EB;
    // staging items: (element of the chunk, group of 4 modes), each a thread's
    // 3 float4 of W and the element's 6 Mr values, loaded one chunk ahead
    constexpr int ITEMS = EB * R / 4;
    constexpr int PER = (ITEMS + kGalMfmaBlock - 1) / kGalMfmaBlock;
    __shared__ __align__(16) float sW[K][R];
    __shared__ __align__(16) float sZ[K][R];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ti0 = kGalUpperTiles[2 * wave][0], tj0 = kGalUpperTiles[2 * wave][1];
    const int ti1 = kGalUpperTiles[2 * wave + 1][0], tj1 = kGalUpperTiles[2 * wave + 1][1];
    const int kh = lane >> 5, c = lane & 31;
    gal_f32x16 acc0, acc1;
#pragma unroll
    for (int q = 0; q < 16; ++q) acc0[q] = acc1[q] = 0.0f;
    const int nb = (num_el + EB - 1) / EB;
    const int per = (nb + gridDim.x - 1) / gridDim.x;
    const int b0 = blockIdx.x * per, b1 = min(nb, b0 + per);
    float4 w[PER][3];
    float m[PER][6];
    auto load = [&](int bt) {
        const int e0 = bt * EB;
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int it = tid + t * kGalMfmaBlock;
            const int le = it / (R / 4), l4 = it % (R / 4), e = e0 + le;
            const bool ok = it < ITEMS && e < num_el;
            const float4 *src = reinterpret_cast<const float4 *>(W + (size_t)e * 3 * R) + l4;
#pragma unroll
            for (int a = 0; a < 3; ++a)
                w[t][a] = ok ? src[a * (R / 4)] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
            for (int u = 0; u < 6; ++u) m[t][u] = ok ? Mr[(size_t)u * mr_stride + e] : 0.0f;
        }
    };
    if (b0 < b1) load(b0);
    for (int bt = b0; bt < b1; ++bt) {
        // W and Z = Mr W of the chunk into LDS
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int it = tid + t * kGalMfmaBlock;
            if (it >= ITEMS) continue;
            const int le = it / (R / 4), l4 = it % (R / 4);
            const float *mm = m[t];
            float4 z[3];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float x0 = w[t][0][q], x1 = w[t][1][q], x2 = w[t][2][q];
                z[0][q] = mm[0] * x0 + mm[1] * x1 + mm[2] * x2;
                z[1][q] = mm[1] * x0 + mm[3] * x1 + mm[4] * x2;
                z[2][q] = mm[2] * x0 + mm[4] * x1 + mm[5] * x2;
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                *reinterpret_cast<float4 *>(&sW[3 * le + a][4 * l4]) = w[t][a];
                *reinterpret_cast<float4 *>(&sZ[3 * le + a][4 * l4]) = z[a];
            }
        }
        __syncthreads();
        if (bt + 1 < b1) load(bt + 1);       // in flight during the products
#pragma unroll 4
        for (int k = 0; k < K; k += 2) {
            const float a0 = sW[k + kh][32 * ti0 + c], bz0 = sZ[k + kh][32 * tj0 + c];
            const float a1 = sW[k + kh][32 * ti1 + c], bz1 = sZ[k + kh][32 * tj1 + c];
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, bz0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bz1, acc1, 0, 0, 0);
        }
        __syncthreads();
    }
    float *dst = part + (size_t)blockIdx.x * R * R;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = 8 * (q >> 2) + 4 * kh + (q & 3);
        const int r0 = 32 * ti0 + i, c0 = 32 * tj0 + c;
        dst[r0 * R + c0] = acc0[q];
        if (ti0 != tj0) dst[c0 * R + r0] = acc0[q];
        const int r1 = 32 * ti1 + i, c1 = 32 * tj1 + c;
        dst[r1 * R + c1] = acc1[q];
        if (ti1 != tj1) dst[c1 * R + r1] = acc1[q];
    }
}

// out[i] = sum_g part[g*m + i] (float partials, double sums): 64 entries per
// workgroup, 4 groups of 64 lanes (partial g goes to group g % 4, each group
// sums in order), the group sums added in a fixed order
template <typename T>
__global__ __launch_bounds__(256) void sum_partials_kernel(int G, int m,
                                                           const T *__restrict__ part,
                                                           double *__restrict__ out) {
    __shared__ double acc[4][64];
    const int c = threadIdx.x % 64, grp = threadIdx.x / 64;
    const int i = blockIdx.x * 64 + c;
    double s = 0.0;
    if (i < m)
        for (int g = grp; g < G; g += 4) s += (double)part[(size_t)g * m + i];
    acc[grp][c] = s;
    __syncthreads();
    if (grp == 0 && i < m) out[i] = ((acc[0][c] + acc[1][c]) + acc[2][c]) + acc[3][c];
}

// partial g[k] over one chunk of kProjPer * 256 entries and one group of
// kProjModes modes per workgroup (grid: chunks x mode groups): the chunk's v
// values stay in registers, every mode k of the group streams its slice of U
// once (coalesced), wave sums go through LDS; part[b*r + k]
constexpr int kProjPer = 8, kProjModes = 32;
__global__ __launch_bounds__(256) void pod_project_kernel(int n, int r,
                                                          const float *__restrict__ U,
                                                          const double *__restrict__ v,
                                                          double *__restrict__ part) {
    __shared__ double ws[4][kProjModes];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const long long base = (long long)blockIdx.x * 256 * kProjPer + threadIdx.x;
    const int k0 = blockIdx.y * kProjModes, kc = min(kProjModes, r - k0);
    double vv[kProjPer];
#pragma unroll
    for (int j = 0; j < kProjPer; ++j) {
        const long long i = base + 256LL * j;
        vv[j] = i < n ? v[i] : 0.0;
    }
    // two modes per trip: both modes' loads in flight before either sum
    for (int kk = 0; kk < kc; kk += 2) {
        const bool two = kk + 1 < kc;
        const float *u0 = U + (size_t)(k0 + kk) * n;
        const float *u1 = U + (size_t)(k0 + kk + (two ? 1 : 0)) * n;
        float a0[kProjPer], a1[kProjPer];
#pragma unroll
        for (int j = 0; j < kProjPer; ++j) {
            const long long i = base + 256LL * j;
            a0[j] = i < n ? u0[i] : 0.0f;
            a1[j] = i < n ? u1[i] : 0.0f;
        }
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int j = 0; j < kProjPer; ++j) {
            const long long i = base + 256LL * j;
            if (i < n) {
                s0 += (double)a0[j] * vv[j];
                s1 += (double)a1[j] * vv[j];
            }
        }
        s0 = wave_sum(s0);
        s1 = wave_sum(s1);
        if (lane == 0) {
            ws[wid][kk] = s0;
            if (two) ws[wid][kk + 1] = s1;
        }
    }
    __syncthreads();
    if (threadIdx.x < kc)
        part[(size_t)blockIdx.x * r + k0 + threadIdx.x] =
            ((ws[0][threadIdx.x] + ws[1][threadIdx.x]) + ws[2][threadIdx.x]) + ws[3][threadIdx.x];
}

// pod_project_kernel for nv vectors v[w*n + i] at once: one pass over the
// modes serves every vector; part[(b*nv + w)*r + k], each vector's partial
// formed in the order of pod_project_kernel (bit-identical per vector)
constexpr int kProjMaxVec = 8;
template <int NVEC>
__global__ __launch_bounds__(256) void pod_project_many_kernel(int n, int r,
                                                               const float *__restrict__ U,
                                                               const double *__restrict__ v,
                                                               double *__restrict__ part) {
    __shared__ double ws[4][kProjModes][NVEC];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const long long base = (long long)blockIdx.x * 256 * kProjPer + threadIdx.x;
    const int k0 = blockIdx.y * kProjModes, kc = min(kProjModes, r - k0);
    double vv[NVEC][kProjPer];
#pragma unroll
    for (int w = 0; w < NVEC; ++w)
#pragma unroll
        for (int j = 0; j < kProjPer; ++j) {
            const long long i = base + 256LL * j;
            vv[w][j] = i < n ? v[(long long)w * n + i] : 0.0;
        }
    // two modes per trip, the next trip's two modes loaded during this one's
    // sums
    float a0[kProjPer], a1[kProjPer];
    auto load2 = [&](int kk, float *x0, float *x1) {
        const bool two = kk + 1 < kc;
        const float *u0 = U + (size_t)(k0 + kk) * n;
        const float *u1 = U + (size_t)(k0 + kk + (two ? 1 : 0)) * n;
#pragma unroll
        for (int j = 0; j < kProjPer; ++j) {
            const long long i = base + 256LL * j;
            x0[j] = i < n ? u0[i] : 0.0f;
            x1[j] = i < n ? u1[i] : 0.0f;
        }
    };
    if (kc > 0) load2(0, a0, a1);
    for (int kk = 0; kk < kc; kk += 2) {
        const bool two = kk + 1 < kc;
        float n0[kProjPer], n1[kProjPer];
        if (kk + 2 < kc) load2(kk + 2, n0, n1);
        // per mode, the NVEC vectors' sums reduced over the wave together
#pragma unroll
        for (int which = 0; which < 2; ++which) {
            const float *am = which ? a1 : a0;
            double a[NVEC];
#pragma unroll
            for (int w = 0; w < NVEC; ++w) {
                double sw = 0.0;
#pragma unroll
                for (int j = 0; j < kProjPer; ++j) {
                    const long long i = base + 256LL * j;
                    if (i < n) sw += (double)am[j] * vv[w][j];
                }
                a[w] = sw;
            }
            // Butterfly over the wave, lane masks 32, 16, ..., 1 in the order
            // of wave_sum, so every sum has wave_sum's bits: while more than
            // one value is left, each lane keeps half of them (by its mask
            // bit) and adds its partner's partial of that half; then lane l
            // holds vector l >> (6 - log2 NVEC), summed over the remaining
            // masks like one value.  NVEC - 1 + 6 - log2 NVEC shuffles
            // instead of 6 NVEC.
            int m = 32;
#pragma unroll
            for (int c = NVEC; c > 1; c >>= 1, m >>= 1) {
                const bool hi = (lane & m) != 0;
#pragma unroll
                for (int q = 0; q < c / 2; ++q) {
                    const double keep = hi ? a[q + c / 2] : a[q];
                    const double send = hi ? a[q] : a[q + c / 2];
                    a[q] = keep + __shfl_xor(send, m, kWave);
                }
            }
            double t = a[0];
            for (; m > 0; m >>= 1) t += __shfl_xor(t, m, kWave);
            constexpr int LOGV = NVEC == 1 ? 0 : NVEC == 2 ? 1 : NVEC == 4 ? 2 : 3;
            if ((lane & ((1 << (6 - LOGV)) - 1)) == 0 && (which == 0 || two))
                ws[wid][kk + which][lane >> (6 - LOGV)] = t;
        }
        if (kk + 2 < kc) {
#pragma unroll
            for (int j = 0; j < kProjPer; ++j) {
                a0[j] = n0[j];
                a1[j] = n1[j];
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kc * NVEC; t += blockDim.x) {
        const int k = t / NVEC, w = t % NVEC;
        part[((size_t)blockIdx.x * NVEC + w) * r + k0 + k] =
            ((ws[0][k][w] + ws[1][k][w]) + ws[2][k][w]) + ws[3][k][w];
    }
}

// Every system's reduced Newton correction d = -(H_s)^-1 g with H_s =
// (H + H^T)/2: one workgroup per system, Cholesky factorization of H_s in LDS
// (packed lower triangle, FP64), forward and back substitution.  A pivot that
// is not above 1e-13 times the largest diagonal entry, or a non-finite
// result, gives d = 0 (no correction) for that system.
//
// Two barriers per factorization step (the column, then the trailing
// update) and one per substitution step: every thread forms the step's
// unknown itself from values no thread writes in that step.
constexpr int kSolveMaxR = 128;
constexpr int kSolveBlock = 256;

__global__ __launch_bounds__(kSolveBlock) void reduced_solve_kernel(
    int r, const double *__restrict__ H, const double *__restrict__ g, float *__restrict__ d) {
    __shared__ double L[kSolveMaxR * (kSolveMaxR + 1) / 2];
    __shared__ double y[kSolveMaxR], z[kSolveMaxR], c[kSolveMaxR], cs[kSolveMaxR];
    __shared__ int fail;
    const int tid = threadIdx.x, sys = blockIdx.x;
    const double *Hs = H + (size_t)sys * r * r;
    auto at = [](int i, int j) { return i * (i + 1) / 2 + j; };     // j <= i
    for (int t = tid; t < r * r; t += kSolveBlock) {
        const int i = t / r, j = t % r;
        if (j <= i) L[at(i, j)] = 0.5 * (Hs[(size_t)i * r + j] + Hs[(size_t)j * r + i]);
    }
    for (int i = tid; i < r; i += kSolveBlock) y[i] = -g[(size_t)sys * r + i];
    if (tid == 0) fail = 0;
    __syncthreads();
    double dmax = 0.0;
    for (int i = 0; i < r; ++i) dmax = fmax(dmax, fabs(L[at(i, i)]));
    const bool bad_scale = !(dmax > 0.0) || !(dmax < INFINITY);
    // factorization, right-looking: column k is copied to c (and c / A_kk to
    // cs), then the trailing triangle A_ij -= c_i cs_j, 16 x 16 threads over
    // its rows and columns
    const int tr = tid / 16, tc = tid % 16;
    for (int k = 0; k < r && !bad_scale; ++k) {
        const double dk = L[at(k, k)];
        if (!(dk > 1e-13 * dmax) || !(dk < INFINITY)) {
            if (tid == 0) fail = 1;
            break;                                   // uniform: every thread read the same dk
        }
        const double lk = sqrt(dk);
        for (int i = k + 1 + tid; i < r; i += kSolveBlock) {
            const double a = L[at(i, k)];
            c[i] = a;
            cs[i] = a / dk;
            L[at(i, k)] = a / lk;                    // the factor's column
        }
        __syncthreads();
        if (tid == 0) L[at(k, k)] = lk;
        for (int i = k + 1 + tr; i < r; i += 16) {
            const double ci = c[i];
            for (int j = k + 1 + tc; j <= i; j += 16) L[at(i, j)] -= ci * cs[j];
        }
        __syncthreads();
    }
    __syncthreads();
    const bool ok = !bad_scale && !fail;
    if (ok) {
        for (int k = 0; k < r; ++k) {                // L z = y
            const double zk = y[k] / L[at(k, k)];
            for (int i = k + 1 + tid; i < r; i += kSolveBlock) y[i] -= L[at(i, k)] * zk;
            if (tid == 0) z[k] = zk;
            __syncthreads();
        }
        for (int k = r - 1; k >= 0; --k) {           // L^T x = z
            const double xk = z[k] / L[at(k, k)];
            for (int i = tid; i < k; i += kSolveBlock) z[i] -= L[at(k, i)] * xk;
            if (tid == 0) y[k] = xk;
            __syncthreads();
        }
    }
    bool finite = ok;
    if (ok)
        for (int i = 0; i < r; ++i) finite = finite && isfinite(y[i]);
    for (int i = tid; i < r; i += kSolveBlock)
        d[(size_t)sys * r + i] = finite ? (float)y[i] : 0.0f;
}

}  // namespace sn

using namespace sn;

extern "C" {

static int project_many_to(int n, int r, int nvec, const float *U, const double *v, double *g,
                           hipStream_t s) {
    SN_REQUIRE(n > 0 && r > 0 && U && v && g, "sn_pod_project_many: bad argument");
    SN_REQUIRE(nvec == 1 || nvec == 2 || nvec == 4 || nvec == 8,
               "sn_pod_project_many: nvec must be 1, 2, 4 or 8 (%d)", nvec);
    const int G = (int)((n + 256LL * kProjPer - 1) / (256LL * kProjPer));
    const int m = nvec * r;
    double *part = nullptr;
    SN_HIP(hipMallocAsync((void **)&part, (size_t)G * m * sizeof(double), s));
    const dim3 grid(G, (r + kProjModes - 1) / kProjModes);
    switch (nvec) {
        case 1: hipLaunchKernelGGL(pod_project_many_kernel<1>, grid, dim3(256), 0, s, n, r, U, v, part); break;
        case 2: hipLaunchKernelGGL(pod_project_many_kernel<2>, grid, dim3(256), 0, s, n, r, U, v, part); break;
        case 4: hipLaunchKernelGGL(pod_project_many_kernel<4>, grid, dim3(256), 0, s, n, r, U, v, part); break;
        default: hipLaunchKernelGGL(pod_project_many_kernel<8>, grid, dim3(256), 0, s, n, r, U, v, part); break;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sum_partials_kernel<double>, dim3((m + 63) / 64), dim3(256), 0, s, G,
                           m, part, g);
        e = hipGetLastError();
    }
    hipError_t f = hipFreeAsync(part, s);
    if (e == hipSuccess) e = f;
    if (e != hipSuccess) {
        set_error("sn_pod_project_many: %s", hipGetErrorString(e));
        return SN_ERR_HIP;
    }
    return SN_OK;
}

int sn_pod_project_many(int n, int r, int nvec, const float *U, const double *v,
                        double *g_host, void *stream) {
    SN_REQUIRE(g_host && r > 0 && nvec > 0, "sn_pod_project_many: bad argument");
    hipStream_t s = as_stream(stream);
    const int m = nvec * r;
    double *g = nullptr;
    SN_HIP(hipMallocAsync((void **)&g, (size_t)m * sizeof(double), s));
    int rc = project_many_to(n, r, nvec, U, v, g, s);
    hipError_t e = hipSuccess;
    if (rc == SN_OK)
        e = hipMemcpyAsync(g_host, g, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, s);
    hipError_t f = hipFreeAsync(g, s);
    if (e == hipSuccess) e = f;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (rc != SN_OK) return rc;
    if (e != hipSuccess) {
        set_error("sn_pod_project_many: %s", hipGetErrorString(e));
        return SN_ERR_HIP;
    }
    return SN_OK;
}

int sn_pod_project_many_dev(int n, int r, int nvec, const float *U, const double *v,
                            double *g_dev, void *stream) {
    return project_many_to(n, r, nvec, U, v, g_dev, as_stream(stream));
}

int sn_pod_element_modes(int num_el, int n, int r, const float *U, const int *gids,
                         const int *orient, float *W, void *stream) {
    SN_REQUIRE(num_el > 0 && n > 0 && r > 0 && U && gids && orient && W,
               "sn_pod_element_modes: bad argument");
    const long long tot = (long long)num_el * r;
    hipLaunchKernelGGL(pod_element_modes_kernel, dim3(grid_for(tot, 256, 1 << 30)), dim3(256), 0,
                       as_stream(stream), num_el, n, r, U, gids, orient, W);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_pod_reduced_jacobian(int num_el, int r, const float *W, const float *Mr, double *H_host,
                            void *stream) {
    return sn_pod_reduced_jacobian_range(num_el, 0, num_el, r, W, Mr, H_host, stream);
}

static int reduced_jacobian_to(int num_el, int e0, int e1, int r, const float *W,
                               const float *Mr, double *H_dev, hipStream_t s) {
    SN_REQUIRE(num_el > 0 && W && Mr && H_dev, "sn_pod_reduced_jacobian: bad argument");
    SN_REQUIRE(0 <= e0 && e0 < e1 && e1 <= num_el,
               "sn_pod_reduced_jacobian: element range [%d, %d) outside [0, %d)", e0, e1, num_el);
    SN_REQUIRE(r == 32 || r == 64 || r == 128,
               "sn_pod_reduced_jacobian: r must be 32, 64 or 128 (r=%d)", r);
    const int ne = e1 - e0;
    const float *Wr = W + (size_t)e0 * 3 * r;
    const float *Mrr = Mr + e0;
    const int nb = (ne + kGalEB - 1) / kGalEB;      // kGalMfmaEB is the same chunk
    const int G = nb < 512 ? nb : 512;
    const int m = r * r;
    float *part = nullptr;
    SN_HIP(hipMallocAsync((void **)&part, (size_t)G * m * sizeof(float), s));
    if (r == 32)
        hipLaunchKernelGGL(pod_reduced_jacobian_kernel<32>, dim3(G), dim3(kGalBlock), 0, s,
                           ne, num_el, Wr, Mrr, part);
    else if (r == 64)
        hipLaunchKernelGGL(pod_reduced_jacobian_kernel<64>, dim3(G), dim3(kGalBlock), 0, s,
                           ne, num_el, Wr, Mrr, part);
    else
        hipLaunchKernelGGL(pod_reduced_jacobian_mfma_kernel, dim3(G), dim3(kGalMfmaBlock), 0, s,
                           ne, num_el, Wr, Mrr, part);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sum_partials_kernel<float>, dim3((m + 63) / 64), dim3(256), 0, s, G,
                           m, part, H_dev);
        e = hipGetLastError();
    }
    hipError_t f = hipFreeAsync(part, s);
    if (e == hipSuccess) e = f;
    if (e != hipSuccess) {
        set_error("sn_pod_reduced_jacobian: %s", hipGetErrorString(e));
        return SN_ERR_HIP;
    }
    return SN_OK;
}

int sn_pod_reduced_jacobian_range(int num_el, int e0, int e1, int r, const float *W,
                                  const float *Mr, double *H_host, void *stream) {
    SN_REQUIRE(H_host && r > 0, "sn_pod_reduced_jacobian: bad argument");
    hipStream_t s = as_stream(stream);
    const int m = r * r;
    double *H = nullptr;
    SN_HIP(hipMallocAsync((void **)&H, (size_t)m * sizeof(double), s));
    int rc = reduced_jacobian_to(num_el, e0, e1, r, W, Mr, H, s);
    hipError_t e = hipSuccess;
    if (rc == SN_OK)
        e = hipMemcpyAsync(H_host, H, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, s);
    hipError_t f = hipFreeAsync(H, s);
    if (e == hipSuccess) e = f;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (rc != SN_OK) return rc;
    if (e != hipSuccess) {
        set_error("sn_pod_reduced_jacobian: %s", hipGetErrorString(e));
        return SN_ERR_HIP;
    }
    return SN_OK;
}

int sn_pod_reduced_jacobian_range_dev(int num_el, int e0, int e1, int r, const float *W,
                                      const float *Mr, double *H_dev, void *stream) {
    return reduced_jacobian_to(num_el, e0, e1, r, W, Mr, H_dev, as_stream(stream));
}

int sn_pod_reduced_solve_many(int nsys, int r, const double *H_dev, const double *g_dev,
                              float *d_dev, void *stream) {
    SN_REQUIRE(nsys >= 1 && r >= 1 && r <= kSolveMaxR && H_dev && g_dev && d_dev,
               "sn_pod_reduced_solve_many: bad argument (nsys=%d, r=%d, r <= %d)", nsys, r,
               kSolveMaxR);
    hipLaunchKernelGGL(reduced_solve_kernel, dim3(nsys), dim3(kSolveBlock), 0, as_stream(stream),
                       r, H_dev, g_dev, d_dev);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_pod_project(int n, int r, const float *U, const double *v, double *g_host, void *stream) {
    SN_REQUIRE(n > 0 && r > 0 && U && v && g_host, "sn_pod_project: bad argument");
    hipStream_t s = as_stream(stream);
    const int G = (int)((n + 256LL * kProjPer - 1) / (256LL * kProjPer));
    double *part = nullptr, *g = nullptr;
    SN_HIP(hipMallocAsync((void **)&part, (size_t)G * r * sizeof(double), s));
    SN_HIP(hipMallocAsync((void **)&g, (size_t)r * sizeof(double), s));
    hipLaunchKernelGGL(pod_project_kernel, dim3(G, (r + kProjModes - 1) / kProjModes), dim3(256),
                       0, s, n, r, U, v, part);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sum_partials_kernel<double>, dim3((r + 63) / 64), dim3(256), 0, s, G,
                           r, part, g);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(g_host, g, (size_t)r * sizeof(double),
                                            hipMemcpyDeviceToHost, s);
    hipError_t f1 = hipFreeAsync(part, s), f2 = hipFreeAsync(g, s);
    if (e == hipSuccess) e = f1 != hipSuccess ? f1 : f2;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        set_error("sn_pod_project: %s", hipGetErrorString(e));
        return SN_ERR_HIP;
    }
    return SN_OK;
}

}  // extern "C"
