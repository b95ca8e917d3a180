// Auxiliary-space preconditioner for the curl-curl Jacobian (no counterpart
// in the reference, which preconditions its CG with the Jacobi diagonal
// only, snoopy/red_mvp_solver.py:198-199).
//
// The Jacobian J of the reduced-vector-potential problem is singular: every
// discrete gradient is in its kernel (curl grad = 0), so point smoothers and
// nodal AMG alone cannot reach its smooth error.  The Hiptmair-Xu
// preconditioner splits the edge space into high-frequency edge functions
// (handled by the Jacobi diagonal) and interpolated vector nodal fields
// (handled by a multigrid cycle of the nodal vector Laplacian):
//
//   M r = D^-1 r + w * Pi B_L Pi^T r
//
//   Pi      edge DOF of the Whitney interpolant of a P1 vector field:
//           (Pi y)_e = (x_hi - x_lo)/2 . (y_lo + y_hi)
//   B_L     one V(1,1)-cycle of unsmoothed-aggregation AMG applied to the
//           three components of y at once; L is the P1 Laplacian with the
//           element reluctivity, sum_e nu_e vol_e grad(l_a).grad(l_b), and
//           Dirichlet nodes (nodes of Dirichlet edges) removed.  Since
//           grad-fields lie in the kernel of J, no gradient-space correction
//           is needed.
//
// Set-up: the aggregation (host, once per mesh) orders every level so that
// the rows of each aggregate are contiguous; restriction is then a sum over
// a contiguous range and prolongation a broadcast.  All matrices are formed
// on the device by fixed-order gathers (element values -> level 0, level l ->
// level l+1), so a refresh per Newton step costs a few bandwidth-bound
// passes; the coarsest matrix (<= a few hundred rows) is inverted by one
// block with Gauss-Jordan elimination (SPD, no pivoting).
//
// Apply (all float32, 3 right-hand sides packed in float4):
//   edge_restrict    b0 = Pi^T r                    (thread per free node)
//   down (per level) x = w D^-1 b, b' = P^T (b - A x) (wave per aggregate:
//                    the aggregate's rows and their nonzeros are contiguous)
//   coarse           y = C^-1 b                     (wave per row, dense)
//   up (per level)   x' = x + P e, y = x' + w D^-1 (b - A x')
//   edge_prolong     z = D_J^-1 r + w Pi y0, partial sums of r.z
// Pre- and post-smoothing are the same damped Jacobi step, so the cycle is a
// symmetric operator and M is symmetric positive definite on the range of J.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "sn_ams.hpp"
#include "sn_common.hpp"

namespace sn {

constexpr int kAmsBlock = 256;
constexpr int kAmsEdgeBlock = 1024;
constexpr int kAmsMaxCoarse = 1024;

__device__ __forceinline__ float4 f4_zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
__device__ __forceinline__ void f4_fma(float4 &a, float s, const float4 &b) {
    a.x += s * b.x;
    a.y += s * b.y;
    a.z += s * b.z;
}
__device__ __forceinline__ float4 f4_add(const float4 &a, const float4 &b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, 0.f);
}
template <int W>
__device__ __forceinline__ void f4_reduce(float4 &a) {
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) {
        a.x += __shfl_xor(a.x, off, kWave);
        a.y += __shfl_xor(a.y, off, kWave);
        a.z += __shfl_xor(a.z, off, kWave);
    }
}

// Work index of block b of nb.  xcd: workgroups are dealt round-robin to the
// 8 XCDs; renumber them so that each XCD takes one contiguous eighth of the
// grid (a multiple of 8 blocks), whose gathers of one mesh neighbourhood then
// meet in its L2.
__device__ __forceinline__ long long ams_block(int xcd, int b, int nb) {
    if (xcd && (nb & 7) == 0) return (long long)(b & 7) * (nb >> 3) + (b >> 3);
    return b;
}

inline int ams_grid(long long threads, int block, int xcd) {
    const int g = grid_for(threads, block);
    return xcd ? (g + 7) / 8 * 8 : g;
}

// ------------------------------------------------------------ numeric set-up
// P1 stiffness of each element with its reluctivity: 10 pair values
// (a <= b, row-major upper triangle) nu_e vol_e grad(l_a).grad(l_b).
__global__ __launch_bounds__(kAmsBlock) void ams_lap_local_kernel(
    int n_el, const double *__restrict__ nodes, const int *__restrict__ cells,
    const float *__restrict__ nu_el, float *__restrict__ loc) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    double X[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int nd = cells[4 * e + a];
#pragma unroll
        for (int k = 0; k < 3; ++k) X[a][k] = nodes[3 * nd + k];
    }
    double E[3][3];  // E[c] = X[c+1] - X[0]
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k) E[c][k] = X[c + 1][k] - X[0][k];
    // rows of E^-1 (E with columns e1, e2, e3): (e2 x e3, e3 x e1, e1 x e2) / det
    double g[4][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double *u = E[(c + 1) % 3], *v = E[(c + 2) % 3];
        g[c + 1][0] = u[1] * v[2] - u[2] * v[1];
        g[c + 1][1] = u[2] * v[0] - u[0] * v[2];
        g[c + 1][2] = u[0] * v[1] - u[1] * v[0];
    }
    const double det = E[0][0] * g[1][0] + E[0][1] * g[1][1] + E[0][2] * g[1][2];
    const double inv = 1.0 / det;
#pragma unroll
    for (int a = 1; a < 4; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) g[a][k] *= inv;
#pragma unroll
    for (int k = 0; k < 3; ++k) g[0][k] = -(g[1][k] + g[2][k] + g[3][k]);
    const double s = (double)nu_el[e] * fabs(det) / 6.0;
    int p = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = a; b < 4; ++b) {
            const double v = s * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
            loc[(long long)p * n_el + e] = (float)v;
            ++p;
        }
}

// node coordinates in float, in the preconditioner's node order; the edge
// tangents (x_hi - x_lo)/2 are formed from them where they are used
// (a 16-byte gather from a few MB that stays in cache, instead of a
// 16-byte tangent streamed per edge and per incidence)
__global__ void ams_coords_kernel(int n_nodes, int n_free, const double *__restrict__ nodes,
                                  const int *__restrict__ node_mesh, float4 *__restrict__ xyz,
                                  float4 *__restrict__ yx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    const int m = node_mesh[i];
    const float4 c = make_float4((float)nodes[3 * m], (float)nodes[3 * m + 1],
                                 (float)nodes[3 * m + 2], 0.f);
    xyz[i] = c;
    yx[2 * i + 1] = c;
    if (i >= n_free) yx[2 * i] = f4_zero();   // fixed nodes: y stays 0
}

// half the difference of two coordinates: both the restriction and the
// prolongation form a tangent as half_diff(x_hi, x_lo), so the two use the
// same rounded values and stay exact transposes
__device__ __forceinline__ float4 half_diff(const float4 &hi, const float4 &lo) {
    return make_float4(0.5f * (hi.x - lo.x), 0.5f * (hi.y - lo.y), 0.5f * (hi.z - lo.z), 0.f);
}

// (Pi y)_e = t_e . (y_lo + y_hi) from the {y, coordinates} pairs of the
// edge's end points {lo, hi}; y is 0 at fixed nodes
__device__ __forceinline__ float edge_aux(const float4 *__restrict__ yx, int2 en) {
    const float4 y_lo = yx[2 * (long long)en.x], c_lo = yx[2 * (long long)en.x + 1];
    const float4 y_hi = yx[2 * (long long)en.y], c_hi = yx[2 * (long long)en.y + 1];
    const float4 s = f4_add(y_lo, y_hi);
    const float4 t = half_diff(c_hi, c_lo);
    return t.x * s.x + t.y * s.y + t.z * s.z;
}

// R = P^T (I - omega A_f D_f^-1) of a coarse level: entry e (aggregate I,
// fine row j) = [j in I] - omega dinv_f[j] sum_{i in I} a_ij, the sum over the
// plan's fine nonzeros (fixed order)
__global__ void ams_rt_kernel(int n_rt, const int *__restrict__ cols, const int *__restrict__ gptr,
                              const int *__restrict__ glist, const float *__restrict__ vals_f,
                              const float *__restrict__ dinv_f, float omega,
                              float *__restrict__ rt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_rt) return;
    const int c = cols[e];
    const int j = c & 0x7fffffff;
    float a = 0.f;
    for (int k = gptr[e]; k < gptr[e + 1]; ++k) a += vals_f[glist[k]];
    rt[e] = (c < 0 ? 1.f : 0.f) - omega * dinv_f[j] * a;
}

__global__ void ams_gather_kernel(int n_out, const int *__restrict__ gptr,
                                  const int *__restrict__ glist, const float *__restrict__ in,
                                  float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    // four list entries, then their four values, in flight at once; the sum
    // keeps the list order
    double s = 0.0;
    int k = gptr[i];
    const int k1 = gptr[i + 1];
    for (; k + 4 <= k1; k += 4) {
        const int g0 = glist[k], g1 = glist[k + 1], g2 = glist[k + 2], g3 = glist[k + 3];
        const float v0 = in[g0], v1 = in[g1], v2 = in[g2], v3 = in[g3];
        s += (double)v0;
        s += (double)v1;
        s += (double)v2;
        s += (double)v3;
    }
    for (; k < k1; ++k) s += (double)in[glist[k]];
    out[i] = (float)s;
}

__global__ void ams_dinv_kernel(int n, const int *__restrict__ diagpos,
                                const float *__restrict__ vals, float *__restrict__ dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dinv[i] = 1.0f / vals[diagpos[i]];
}

// One block: dense inverse of the coarsest matrix by Gauss-Jordan elimination
// in place (symmetric positive definite, so no pivoting is needed).
// A coarsest matrix of up to kAmsLdsCoarse rows is eliminated in LDS (one
// pivot step costs a few LDS round trips instead of L2 ones); larger ones
// in the global scratch `a_glob`.  Same arithmetic either way.
// Dynamic LDS: the matrix (when it fits, 80^2 doubles = 50 KB) followed by the
// staged pivot column and row (2 nc doubles), within the 64 KB a workgroup
// may allocate
constexpr int kAmsLdsCoarse = 80;

// One pivot step on entry (i, j), from the old pivot row and column staged
// in `row` and `col` and pk = 1 / a_kk: the pivot row is scaled, the pivot
// column becomes -a_ik pk, every other entry loses a_ik (a_kj pk).
__device__ __forceinline__ double gj_entry(double a_ij, int i, int j, int k, double pk,
                                           const double *col, const double *row) {
    if (i == k) return (j == k) ? pk : row[j] * pk;
    if (j == k) return -col[i] * pk;
    return a_ij - col[i] * (row[j] * pk);
}

// launched with 1024 threads: entries per thread of an LDS-resident matrix
constexpr int kAmsInvThreads = 1024;
constexpr int kAmsLdsEntries = (kAmsLdsCoarse * kAmsLdsCoarse + kAmsInvThreads - 1) / kAmsInvThreads;

__global__ __launch_bounds__(kAmsInvThreads) void ams_coarse_invert_kernel(
    int nc, const int *__restrict__ indptr, const int *__restrict__ indices,
    const float *__restrict__ vals, double *__restrict__ a_glob, float *__restrict__ cinv) {
    extern __shared__ double dyn[];
    const bool in_lds = nc <= kAmsLdsCoarse;
    double *a = in_lds ? dyn : a_glob;
    double *col = dyn + (in_lds ? nc * nc : 0);
    double *row = col + nc;
    const int t = threadIdx.x, T = blockDim.x;
    for (int k = t; k < nc * nc; k += T) a[k] = 0.0;
    __syncthreads();
    for (int i = t; i < nc; i += T)
        for (int k = indptr[i]; k < indptr[i + 1]; ++k) a[(size_t)i * nc + indices[k]] = vals[k];
    __syncthreads();
    // a thread's entries of an LDS-resident matrix: (row, column) formed once
    int ei[kAmsLdsEntries], ej[kAmsLdsEntries];
#pragma unroll
    for (int m = 0; m < kAmsLdsEntries; ++m) {
        const int idx = t + m * T;
        ei[m] = (in_lds && idx < nc * nc) ? idx / nc : -1;
        ej[m] = (in_lds && idx < nc * nc) ? idx % nc : 0;
    }
    // two barriers per pivot: stage the old pivot row and column, then
    // update every entry from them
    for (int k = 0; k < nc; ++k) {
        for (int i = t; i < nc; i += T) {
            col[i] = a[(size_t)i * nc + k];
            row[i] = a[(size_t)k * nc + i];
        }
        __syncthreads();
        const double pk = 1.0 / row[k];
        if (in_lds) {
#pragma unroll
            for (int m = 0; m < kAmsLdsEntries; ++m) {
                if (ei[m] < 0) continue;
                const int idx = t + m * T;
                a[idx] = gj_entry(a[idx], ei[m], ej[m], k, pk, col, row);
            }
        } else {
            for (int idx = t; idx < nc * nc; idx += T)
                a[idx] = gj_entry(a[idx], idx / nc, idx % nc, k, pk, col, row);
        }
        __syncthreads();
    }
    for (int k = t; k < nc * nc; k += T) cinv[k] = (float)a[k];
}

// ------------------------------------------------------------------- apply
// LPR lanes per free node a: b0 = Pi^T r over the incidences of a (edge id
// and other end point streamed in order) and the pre-smoothed x0 = w D^-1 b0.
template <typename V, int LPR>
__global__ __launch_bounds__(kAmsBlock) void ams_edge_restrict_kernel(
    int n_free, const int *__restrict__ inc_ptr, const int2 *__restrict__ inc,
    const float4 *__restrict__ xyz, const V *__restrict__ r,
    const float *__restrict__ dinv0, float omega, float4 *__restrict__ b0,
    float4 *__restrict__ x0, const int *__restrict__ done, int xcd) {
    const int a =
        (int)((ams_block(xcd, blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) / LPR);
    const int lane = threadIdx.x % LPR;
    // the node's incidence range and coordinates and the stop flag:
    // independent loads, issued together
    const bool in = a < n_free;
    const int k0 = in ? inc_ptr[a] : 0, k1 = in ? inc_ptr[a + 1] : 0;
    const float4 xa = in ? xyz[a] : f4_zero();
    if (done && *done) return;
    float4 acc = f4_zero();
    // two incidences per trip (k and k + LPR, the order of a one-at-a-time
    // loop), both chains inc -> {coordinates, r} in flight at once; a
    // missing second one reloads the first and is dropped
    for (int k = k0 + lane; k < k1; k += 2 * LPR) {
        const bool h2 = k + LPR < k1;
        const int2 ie = inc[k], ie2 = inc[h2 ? k + LPR : k];   // {edge, other end | sign}
        const float4 xb = xyz[ie.y & 0x7fffffff], xb2 = xyz[ie2.y & 0x7fffffff];
        const float rv = (float)r[ie.x], rv2 = (float)r[ie2.x];
        f4_fma(acc, rv, ie.y < 0 ? half_diff(xa, xb) : half_diff(xb, xa));
        if (h2) f4_fma(acc, rv2, ie2.y < 0 ? half_diff(xa, xb2) : half_diff(xb2, xa));
    }
    f4_reduce<LPR>(acc);
    if (a < n_free && lane == 0) {
        b0[a] = acc;
        const float d = omega * dinv0[a];
        x0[a] = make_float4(d * acc.x, d * acc.y, d * acc.z, 0.f);
    }
}

// wave per aggregate of the next level: b_c = sum_rows b - sum_nnz a_ij x_j
// (x = w D^-1 b already stored for this level), and x_c = w D_c^-1 b_c unless
// the next level is the coarsest
__global__ __launch_bounds__(kAmsBlock) void ams_down_kernel(
    int n_agg, const int *__restrict__ agg_ptr, const int *__restrict__ agg_kp,
    const int *__restrict__ indices, const float *__restrict__ vals, const float4 *__restrict__ b,
    const float4 *__restrict__ x, const float *__restrict__ dinv_c, float omega,
    float4 *__restrict__ bc, float4 *__restrict__ xc, const int *__restrict__ done, int xcd) {
    const int wid =
        (int)((ams_block(xcd, blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    if (wid >= n_agg) return;  // uniform over the wave
    // the aggregate's rows and nonzero range (agg_kp = indptr at agg_ptr) and
    // the stop flag are independent loads, issued together
    const int r0 = agg_ptr[wid], r1 = agg_ptr[wid + 1];
    const int k0 = agg_kp[wid], k1 = agg_kp[wid + 1];
    if (done && *done) return;
    float4 s = f4_zero();
    for (int i = r0 + lane; i < r1; i += kWave) s = f4_add(s, b[i]);
    // two nonzeros per trip (k and k + kWave, the order of a one-at-a-time
    // loop), both index -> x chains in flight at once; a missing second one
    // reloads the first and is dropped
    for (int k = k0 + lane; k < k1; k += 2 * kWave) {
        const bool h2 = k + kWave < k1;
        const int k2 = h2 ? k + kWave : k;
        const float v = vals[k], v2 = vals[k2];
        const float4 xc = x[indices[k]], xc2 = x[indices[k2]];
        f4_fma(s, -v, xc);
        if (h2) f4_fma(s, -v2, xc2);
    }
    f4_reduce<kWave>(s);
    if (lane == 0) {
        bc[wid] = s;
        if (xc) {
            const float d = omega * dinv_c[wid];
            xc[wid] = make_float4(d * s.x, d * s.y, d * s.z, 0.f);
        }
    }
}

// The same step through the coarse level's combined restriction
// R = P^T (I - omega A_f D_f^-1): b_c = R b_f, one gather of b_f per distinct
// fine row reached from the aggregate (on level 0 about half the aggregate's
// fine nonzeros), the same operator up to float rounding
__global__ __launch_bounds__(kAmsBlock) void ams_down_rt_kernel(
    int n_agg, const int *__restrict__ rt_ptr, const int *__restrict__ rt_cols,
    const float *__restrict__ rt_vals, const float4 *__restrict__ b,
    const float *__restrict__ dinv_c, float omega, float4 *__restrict__ bc,
    float4 *__restrict__ xc, const int *__restrict__ done, int xcd) {
    const int wid =
        (int)((ams_block(xcd, blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    if (wid >= n_agg) return;  // uniform over the wave
    const int k0 = rt_ptr[wid], k1 = rt_ptr[wid + 1];
    if (done && *done) return;
    float4 s = f4_zero();
    for (int k = k0 + lane; k < k1; k += 2 * kWave) {
        const bool h2 = k + kWave < k1;
        const int k2 = h2 ? k + kWave : k;
        const float v = rt_vals[k], v2 = rt_vals[k2];
        const float4 b1 = b[rt_cols[k] & 0x7fffffff], b2 = b[rt_cols[k2] & 0x7fffffff];
        f4_fma(s, v, b1);
        if (h2) f4_fma(s, v2, b2);
    }
    f4_reduce<kWave>(s);
    if (lane == 0) {
        bc[wid] = s;
        if (xc) {
            const float d = omega * dinv_c[wid];
            xc[wid] = make_float4(d * s.x, d * s.y, d * s.z, 0.f);
        }
    }
}

// wave per row I of the coarsest level: e_I = (C^-1 b)_I, then prolongation
// into the finer level's pre-smoothed iterate, x_f[i] += e_I for the rows i
// of aggregate I (or, with a single level, y = e, stored as y[2I] of the
// per-node {y, coordinates} pairs)
__global__ __launch_bounds__(kAmsBlock) void ams_coarse_kernel(
    int nc, const float *__restrict__ cinv, const float4 *__restrict__ b,
    const int *__restrict__ agg_ptr, float4 *__restrict__ x_f, float4 *__restrict__ y,
    const int *__restrict__ done) {
    if (done && *done) return;
    const int wid = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (wid >= nc) return;
    float4 acc = f4_zero();
    for (int j = lane; j < nc; j += kWave) f4_fma(acc, cinv[(size_t)wid * nc + j], b[j]);
    f4_reduce<kWave>(acc);
    if (x_f) {
        for (int i = agg_ptr[wid] + lane; i < agg_ptr[wid + 1]; i += kWave)
            x_f[i] = f4_add(x_f[i], acc);
    } else if (lane == 0) {
        y[2 * wid] = acc;
    }
}

// LPR lanes per row of level l, whose x already holds x + P e_(l+1):
// y = x + w D^-1 (b - A x); then y is prolongated into the finer level
// (x_f[i] += y for the rows i of this row's aggregate), or stored at level 0
// as y[2 row] of the per-node {y, coordinates} pairs
template <int LPR>
__global__ __launch_bounds__(kAmsBlock) void ams_up_kernel(
    int n, const int *__restrict__ indptr, const int *__restrict__ indices,
    const float *__restrict__ vals, const float *__restrict__ dinv, const float4 *__restrict__ b,
    const float4 *__restrict__ x, float omega, const int *__restrict__ agg_ptr,
    float4 *__restrict__ x_f, float4 *__restrict__ y, const int *__restrict__ done, int xcd) {
    const int row =
        (int)((ams_block(xcd, blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) / LPR);
    const int lane = threadIdx.x % LPR;
    // the row's range, its own values and the stop flag: independent loads
    const bool in = row < n;
    const int k0 = in ? indptr[row] : 0, k1 = in ? indptr[row + 1] : 0;
    const float4 xi = in ? x[row] : f4_zero(), bi = in ? b[row] : f4_zero();
    const float di = in ? dinv[row] : 0.f;
    if (done && *done) return;
    float4 acc = f4_zero();
    // two nonzeros per trip, as in ams_down_kernel
    for (int k = k0 + lane; k < k1; k += 2 * LPR) {
        const bool h2 = k + LPR < k1;
        const int k2 = h2 ? k + LPR : k;
        const float v = vals[k], v2 = vals[k2];
        const float4 xc = x[indices[k]], xc2 = x[indices[k2]];
        f4_fma(acc, v, xc);
        if (h2) f4_fma(acc, v2, xc2);
    }
    f4_reduce<LPR>(acc);
    if (!in) return;
    const float d = omega * di;
    const float4 yi = make_float4(xi.x + d * (bi.x - acc.x), xi.y + d * (bi.y - acc.y),
                                  xi.z + d * (bi.z - acc.z), 0.f);
    if (x_f) {
        for (int i = agg_ptr[row] + lane; i < agg_ptr[row + 1]; i += LPR)
            x_f[i] = f4_add(x_f[i], yi);
    } else if (lane == 0) {
        y[2 * row] = yi;
    }
}

// The coarse solve and the first up step in one kernel: level f is the one
// above the coarsest level c (nc rows; b_c formed by ams_down_kernel).  Every
// block stages C^-1, b_c and x_f in LDS, forms e = C^-1 b_c (thread per
// coarse row) and x_f' = x_f + P e for all nf rows of f (thread per row;
// with C^-1 read from global memory and a thread per aggregate the kernel
// took 9.1 us at 1M tets, 8.7 us now), then smooths
// its own rows of f as ams_up_kernel does, from x_f' in LDS:
// y = x_f' + w D^-1 (b - A x_f'), prolongated into the next finer level's x
// (or stored as y[2 row] when f is level 0).  The redundant part is nc^2
// multiply-adds and nf loads per block, all independent of each other; it
// replaces a launch that costs its latency (about 4.5 us) more than its
// work.  (Forming b_c here too, wave per aggregate as the down kernel does,
// made every block walk all aggregates one after another: 39 instead of 32 ms
// per 1M-tet solve.)
constexpr int kAmsTailRows = 2048;

template <int LPR>
__global__ __launch_bounds__(kAmsBlock) void ams_tail_kernel(
    int nf, int nc, const int *__restrict__ indptr, const int *__restrict__ indices,
    const float *__restrict__ vals, const float *__restrict__ dinv, const float4 *__restrict__ b,
    const float4 *__restrict__ x, float omega, const float4 *__restrict__ bc,
    const int *__restrict__ c_agg_ptr, const float *__restrict__ cinv,
    const int *__restrict__ agg_ptr, float4 *__restrict__ x_f, float4 *__restrict__ y,
    const int *__restrict__ done, int xcd) {
    __shared__ float4 s_bc[kAmsLdsCoarse];
    __shared__ float4 s_e[kAmsLdsCoarse];
    __shared__ float4 s_x[kAmsTailRows];
    __shared__ float s_ci[kAmsLdsCoarse * kAmsLdsCoarse];
    __shared__ int s_cap[kAmsLdsCoarse + 1];
    if (done && *done) return;   // the same flag for every thread
    const int t = threadIdx.x;
    // every global input of the redundant part in one coalesced round,
    // through registers (fixed trip counts, fully unrolled: all loads are
    // issued before the first LDS store): C^-1, x_f, b_c and the aggregate
    // bounds of c
    constexpr int kCi = (kAmsLdsCoarse * kAmsLdsCoarse + kAmsBlock - 1) / kAmsBlock;
    constexpr int kXr = kAmsTailRows / kAmsBlock;
    static_assert(kAmsLdsCoarse + 1 <= kAmsBlock, "one thread per b_c entry and bound");
    float ci[kCi];
    float4 xr[kXr];
#pragma unroll
    for (int m = 0; m < kCi; ++m) {
        const int k = t + m * kAmsBlock;
        ci[m] = k < nc * nc ? cinv[k] : 0.f;
    }
#pragma unroll
    for (int m = 0; m < kXr; ++m) {
        const int i = t + m * kAmsBlock;
        xr[m] = i < nf ? x[i] : f4_zero();
    }
    const float4 bct = t < nc ? bc[t] : f4_zero();
    const int capt = t <= nc ? c_agg_ptr[t] : 0;
#pragma unroll
    for (int m = 0; m < kCi; ++m) {
        const int k = t + m * kAmsBlock;
        if (k < nc * nc) s_ci[k] = ci[m];
    }
#pragma unroll
    for (int m = 0; m < kXr; ++m) {
        const int i = t + m * kAmsBlock;
        if (i < nf) s_x[i] = xr[m];
    }
    if (t < nc) s_bc[t] = bct;
    if (t <= nc) s_cap[t] = capt;
    __syncthreads();
    for (int I = t; I < nc; I += blockDim.x) {
        float4 acc = f4_zero();
        for (int j = 0; j < nc; ++j) f4_fma(acc, s_ci[I * nc + j], s_bc[j]);
        s_e[I] = acc;
    }
    __syncthreads();
    // x_f' = x_f + P e, thread per row of f; its aggregate I by bisection,
    // s_cap[I] <= i < s_cap[I + 1]
    for (int i = t; i < nf; i += blockDim.x) {
        int lo = 0, hi = nc;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_cap[mid] <= i) lo = mid;
            else hi = mid;
        }
        s_x[i] = f4_add(s_x[i], s_e[lo]);
    }
    __syncthreads();
    const int row =
        (int)((ams_block(xcd, blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) / LPR);
    const int lane = threadIdx.x % LPR;
    const bool in = row < nf;
    const int k0 = in ? indptr[row] : 0, k1 = in ? indptr[row + 1] : 0;
    const float4 xi = in ? s_x[row] : f4_zero(), bi = in ? b[row] : f4_zero();
    const float di = in ? dinv[row] : 0.f;
    float4 acc = f4_zero();
    for (int k = k0 + lane; k < k1; k += LPR) f4_fma(acc, vals[k], s_x[indices[k]]);
    f4_reduce<LPR>(acc);
    if (!in) return;
    const float d = omega * di;
    const float4 yi = make_float4(xi.x + d * (bi.x - acc.x), xi.y + d * (bi.y - acc.y),
                                  xi.z + d * (bi.z - acc.z), 0.f);
    if (x_f) {
        for (int i = agg_ptr[row] + lane; i < agg_ptr[row + 1]; i += LPR)
            x_f[i] = f4_add(x_f[i], yi);
    } else if (lane == 0) {
        y[2 * row] = yi;
    }
}

// Blocks of a batched solve serve one row segment each (seg_of_blk, blk0,
// row0 as in the CG's vector kernel), so every partial sum of r.z belongs to
// one segment; without segments the grid strides over all edges.
template <typename V>
__global__ __launch_bounds__(kAmsEdgeBlock) void ams_edge_prolong_kernel(
    int n_edges, int n_free, const unsigned char *__restrict__ mask, const V *__restrict__ r,
    const V *__restrict__ dinv, const int2 *__restrict__ edge_nodes,
    const float4 *__restrict__ yx, float weight,
    V *__restrict__ z, int zs, const V *__restrict__ pin, double *__restrict__ partials,
    int stride,
    const int *__restrict__ row0,
    const int *__restrict__ blk0, const int *__restrict__ seg_of_blk,
    const int *__restrict__ done, int xcd) {
    __shared__ double lds[kAmsEdgeBlock / kWave];
    if (done && *done) return;
    long long e0 = 0, e1 = n_edges;
    int lb = blockIdx.x, nbl = gridDim.x;
    if (seg_of_blk) {
        const int seg = seg_of_blk[blockIdx.x];
        e0 = row0[seg];
        e1 = row0[seg + 1];
        lb = blockIdx.x - blk0[seg];
        nbl = blk0[seg + 1] - blk0[seg];
    }
    lb = (int)ams_block(xcd, lb, nbl);
    using V2 = typename std::conditional<sizeof(V) == 4, float2, double2>::type;
    double rz = 0.0;
    for (long long i = e0 + (long long)lb * blockDim.x + threadIdx.x; i < e1;
         i += (long long)nbl * blockDim.x) {
        const bool f = mask ? mask[i] != 0 : true;
        V zi = 0;
        if (f) {
            const V ri = r[i];
            zi = dinv[i] * ri + (V)(weight * edge_aux(yx, edge_nodes[i]));
            rz += (double)ri * (double)zi;
        }
        if (pin) {
            // the whole pair {z, p}: no half-written sectors
            reinterpret_cast<V2 *>(z)[i] = V2{zi, pin[i]};
        } else {
            z[i * zs] = zi;
        }
    }
    const double sum = block_sum(rz, lds);
    if (threadIdx.x == 0 && partials) partials[(long long)stride * blockIdx.x] = sum;
}

template <typename V, int LPR>
static void launch_apply_t(sn_ams *A, const V *r, const V *dinv, const unsigned char *mask, V *z,
                           int zs, const V *pin, double *partials, int stride, int grid,
                           const AmsSegments *seg, const int *done, int xcd, hipStream_t s) {
    const int L = (int)A->lv.size();
    AmsLevel &l0 = A->lv[0];
    hipLaunchKernelGGL((ams_edge_restrict_kernel<V, LPR>),
                       dim3(ams_grid((long long)A->n_free * LPR, kAmsBlock, xcd)),
                       dim3(kAmsBlock), 0, s, A->n_free, A->inc_ptr, A->inc, A->xyz, r, l0.dinv,
                       A->omega, l0.b, l0.x, done, xcd);
    const int waves_per_block = kAmsBlock / kWave;
    AmsLevel &lc = A->lv[L - 1];
    // the coarse solve and the first up step in one kernel when the two
    // lowest levels fit its LDS
    const bool tail = L >= 2 && lc.n <= kAmsLdsCoarse && A->lv[L - 2].n <= kAmsTailRows;
    for (int l = 0; l + 1 < L; ++l) {
        AmsLevel &f = A->lv[l], &c = A->lv[l + 1];
        if (A->rt && c.rt_vals)
            hipLaunchKernelGGL(ams_down_rt_kernel, dim3(ams_grid(c.n, waves_per_block, xcd)),
                               dim3(kAmsBlock), 0, s, c.n, c.rt_ptr, c.rt_cols, c.rt_vals, f.b,
                               c.dinv, A->omega, c.b, (l + 2 < L) ? c.x : (float4 *)nullptr,
                               done, xcd);
        else
            hipLaunchKernelGGL(ams_down_kernel, dim3(ams_grid(c.n, waves_per_block, xcd)),
                               dim3(kAmsBlock), 0, s, c.n, c.agg_ptr, c.agg_kp, f.indices,
                               f.vals, f.b, f.x, c.dinv, A->omega, c.b,
                               (l + 2 < L) ? c.x : (float4 *)nullptr, done, xcd);
    }
    if (!tail)
        hipLaunchKernelGGL(ams_coarse_kernel, dim3(grid_for(lc.n, waves_per_block)),
                           dim3(kAmsBlock), 0, s, lc.n, A->cinv, lc.b, lc.agg_ptr,
                           L > 1 ? A->lv[L - 2].x : (float4 *)nullptr, A->yx, done);
    for (int l = L - 2; l >= 0; --l) {
        AmsLevel &f = A->lv[l];
        float4 *x_next = l > 0 ? A->lv[l - 1].x : (float4 *)nullptr;
        if (tail && l == L - 2) {
            hipLaunchKernelGGL((ams_tail_kernel<LPR>),
                               dim3(ams_grid((long long)f.n * LPR, kAmsBlock, xcd)),
                               dim3(kAmsBlock), 0, s, f.n, lc.n, f.indptr, f.indices, f.vals,
                               f.dinv, f.b, f.x, A->omega, lc.b, lc.agg_ptr, A->cinv,
                               f.agg_ptr, x_next, A->yx, done, xcd);
            continue;
        }
        hipLaunchKernelGGL((ams_up_kernel<LPR>), dim3(ams_grid((long long)f.n * LPR, kAmsBlock, xcd)),
                           dim3(kAmsBlock), 0, s, f.n, f.indptr, f.indices, f.vals, f.dinv, f.b,
                           f.x, A->omega, f.agg_ptr, x_next, A->yx, done, xcd);
    }
    hipLaunchKernelGGL((ams_edge_prolong_kernel<V>), dim3(grid), dim3(kAmsEdgeBlock), 0, s,
                       A->n_edges, A->n_free, mask, r, dinv, A->edge_nodes, A->yx,
                       A->weight, z, zs, pin, partials, stride, seg ? seg->row0 : nullptr,
                       seg ? seg->blk0 : nullptr, seg ? seg->seg_of_blk : nullptr, done, xcd);
}

int ams_launch_apply(sn_ams *A, bool v64, const void *r, const void *dinv,
                     const unsigned char *mask, void *z, int zstride, const void *pin,
                     double *partials, int stride, int grid, const AmsSegments *seg,
                     const int *done, int sched, hipStream_t s) {
    SN_REQUIRE(A && A->ready, "AMS preconditioner used before sn_ams_update");
    SN_REQUIRE(zstride >= 1, "ams_launch_apply: zstride must be >= 1");
    SN_REQUIRE(!pin || zstride == 2, "ams_launch_apply: pairs need zstride 2");
    const int xcd = (sched & 2) ? 1 : 0;
#define SN_AMS_APPLY(LPR)                                                                      \
    do {                                                                                       \
        if (v64)                                                                               \
            launch_apply_t<double, LPR>(A, (const double *)r, (const double *)dinv, mask,      \
                                        (double *)z, zstride, (const double *)pin, partials,   \
                                        stride, grid, seg, done, xcd, s);                      \
        else                                                                                   \
            launch_apply_t<float, LPR>(A, (const float *)r, (const float *)dinv, mask,         \
                                       (float *)z, zstride, (const float *)pin, partials,      \
                                       stride, grid, seg, done, xcd, s);                       \
    } while (0)
    if (A->lanes == 4) SN_AMS_APPLY(4);
    else if (A->lanes == 16) SN_AMS_APPLY(16);
    else SN_AMS_APPLY(8);
#undef SN_AMS_APPLY
    SN_CHECK_LAUNCH();
    return SN_OK;
}

template <typename T>
static int upload(T **dst, const T *src, size_t count, hipStream_t s) {
    SN_HIP(hipMalloc(dst, (count > 0 ? count : 1) * sizeof(T)));
    if (count > 0) SN_HIP(hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    return SN_OK;
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_ams_create(sn_ams **out, int n_edges, int n_free_nodes, int n_el, const int *edge_nodes,
                  const int *edge_mesh_nodes, const int *inc_ptr, const int *inc_edge, float omega,
                  float weight, void *stream) {
    SN_REQUIRE(out && edge_nodes && edge_mesh_nodes && inc_ptr && inc_edge,
               "sn_ams_create: null argument");
    SN_REQUIRE(n_edges > 0 && n_free_nodes > 0 && n_el > 0, "sn_ams_create: empty mesh");
    SN_REQUIRE((long long)n_el * 10 < (1LL << 31), "sn_ams_create: too many elements");
    hipStream_t s = as_stream(stream);
    sn_ams *A = new sn_ams();
    A->n_edges = n_edges;
    A->n_free = n_free_nodes;
    A->n_el = n_el;
    A->omega = omega;
    A->weight = weight;
    for (int e = 0; e < 2 * n_edges; ++e) {
        SN_REQUIRE(edge_nodes[e] >= 0, "sn_ams_create: negative node id at edge %d", e / 2);
        SN_REQUIRE(edge_mesh_nodes[e] >= 0, "sn_ams_create: negative mesh node at edge %d", e / 2);
    }
    for (int a = 0; a < n_free_nodes; ++a)
        SN_REQUIRE(inc_ptr[a] <= inc_ptr[a + 1], "sn_ams_create: incidence pointers not ascending");
    const int ninc = inc_ptr[n_free_nodes];
    for (int k = 0; k < ninc; ++k)
        SN_REQUIRE(inc_edge[k] >= 0 && inc_edge[k] < n_edges, "sn_ams_create: edge id out of range");
    // node ids: the free nodes keep theirs; every fixed mesh node (ids >=
    // n_free, possibly one shared id) gets its own, in mesh order, so the
    // coordinates of both end points of every edge are at hand
    std::vector<int> node_mesh(n_free_nodes, -1), fixed_mesh;
    for (int e = 0; e < 2 * n_edges; ++e) {
        const int v = edge_nodes[e], m = edge_mesh_nodes[e];
        if (v < n_free_nodes) {
            SN_REQUIRE(node_mesh[v] < 0 || node_mesh[v] == m,
                       "sn_ams_create: node %d stands for two mesh nodes", v);
            node_mesh[v] = m;
        } else {
            fixed_mesh.push_back(m);
        }
    }
    for (int v = 0; v < n_free_nodes; ++v)
        SN_REQUIRE(node_mesh[v] >= 0, "sn_ams_create: free node %d on no edge", v);
    std::sort(fixed_mesh.begin(), fixed_mesh.end());
    fixed_mesh.erase(std::unique(fixed_mesh.begin(), fixed_mesh.end()), fixed_mesh.end());
    std::vector<int2> en(n_edges);
    for (int e = 0; e < n_edges; ++e) {
        int v[2];
        for (int j = 0; j < 2; ++j) {
            v[j] = edge_nodes[2 * e + j];
            if (v[j] >= n_free_nodes)
                v[j] = n_free_nodes + (int)(std::lower_bound(fixed_mesh.begin(), fixed_mesh.end(),
                                                             edge_mesh_nodes[2 * e + j]) -
                                            fixed_mesh.begin());
        }
        en[e] = make_int2(v[0], v[1]);
    }
    node_mesh.insert(node_mesh.end(), fixed_mesh.begin(), fixed_mesh.end());
    A->n_nodes = (int)node_mesh.size();
    // incidences: {edge, other end point with the sign bit set when this node is hi}
    std::vector<int2> inc(ninc > 0 ? ninc : 1, make_int2(0, 0));
    for (int a = 0; a < n_free_nodes; ++a)
        for (int k = inc_ptr[a]; k < inc_ptr[a + 1]; ++k) {
            const int2 v = en[inc_edge[k]];
            SN_REQUIRE(v.x == a || v.y == a, "sn_ams_create: incidence %d: edge %d misses node %d",
                       k, inc_edge[k], a);
            inc[k] = make_int2(inc_edge[k],
                               (v.x == a) ? v.y : (int)((unsigned)v.x | 0x80000000u));
        }
    int rc;
    if ((rc = upload(&A->edge_nodes, en.data(), (size_t)n_edges, s))) return rc;
    if ((rc = upload(&A->node_mesh, node_mesh.data(), node_mesh.size(), s))) return rc;
    SN_HIP(hipMalloc(&A->xyz, node_mesh.size() * sizeof(float4)));
    SN_HIP(hipMalloc(&A->yx, 2 * node_mesh.size() * sizeof(float4)));
    if ((rc = upload(&A->inc_ptr, inc_ptr, (size_t)n_free_nodes + 1, s))) return rc;
    if ((rc = upload(&A->inc, inc.data(), (size_t)ninc, s))) return rc;
    A->n_inc = ninc;
    SN_HIP(hipMalloc(&A->loc, (size_t)10 * n_el * sizeof(float)));
    SN_HIP(hipStreamSynchronize(s));
    *out = A;
    return SN_OK;
}

int sn_ams_add_level(sn_ams *A, int n, int nnz, const int *indptr, const int *indices,
                     const int *gather_ptr, const int *gather_list, const int *agg_ptr,
                     void *stream) {
    SN_REQUIRE(A && indptr && indices && gather_ptr && gather_list, "sn_ams_add_level: null argument");
    SN_REQUIRE(n > 0 && nnz >= n, "sn_ams_add_level: bad sizes (n=%d, nnz=%d)", n, nnz);
    const int L = (int)A->lv.size();
    SN_REQUIRE(L > 0 || n == A->n_free, "sn_ams_add_level: level 0 must have one row per free node");
    SN_REQUIRE(L == 0 || agg_ptr, "sn_ams_add_level: coarse level without aggregates");
    SN_REQUIRE(indptr[0] == 0 && indptr[n] == nnz, "sn_ams_add_level: indptr does not span nnz");
    const long long src_n = (L == 0) ? 10LL * A->n_el : (long long)A->lv[L - 1].nnz;
    std::vector<int> diag(n, -1);
    for (int i = 0; i < n; ++i) {
        SN_REQUIRE(indptr[i] <= indptr[i + 1], "sn_ams_add_level: indptr not ascending");
        for (int k = indptr[i]; k < indptr[i + 1]; ++k) {
            SN_REQUIRE(indices[k] >= 0 && indices[k] < n, "sn_ams_add_level: column out of range");
            if (indices[k] == i) diag[i] = k;
        }
        SN_REQUIRE(diag[i] >= 0, "sn_ams_add_level: row %d has no diagonal", i);
    }
    SN_REQUIRE(gather_ptr[0] == 0, "sn_ams_add_level: gather plan must start at 0");
    for (int k = 0; k < nnz; ++k)
        SN_REQUIRE(gather_ptr[k] <= gather_ptr[k + 1], "sn_ams_add_level: gather plan not ascending");
    const int ng = gather_ptr[nnz];
    for (int k = 0; k < ng; ++k)
        SN_REQUIRE(gather_list[k] >= 0 && gather_list[k] < src_n,
                   "sn_ams_add_level: gather entry %d out of range", k);
    if (L > 0) {
        const int nf = A->lv[L - 1].n;
        SN_REQUIRE(agg_ptr[0] == 0 && agg_ptr[n] == nf,
                   "sn_ams_add_level: aggregates must cover the previous level's rows");
        for (int I = 0; I < n; ++I)
            SN_REQUIRE(agg_ptr[I] < agg_ptr[I + 1], "sn_ams_add_level: empty aggregate %d", I);
    }
    hipStream_t s = as_stream(stream);
    AmsLevel lv;
    lv.n = n;
    lv.nnz = nnz;
    lv.indptr_h.assign(indptr, indptr + n + 1);
    lv.indices_h.assign(indices, indices + nnz);
    std::vector<int> agg_kp;
    if (L > 0) {
        const std::vector<int> &fp = A->lv[L - 1].indptr_h;
        agg_kp.resize((size_t)n + 1);
        for (int I = 0; I <= n; ++I) agg_kp[I] = fp[agg_ptr[I]];
    }
    int rc;
    // host sources of the uploads below, alive until the stream is synchronised
    std::vector<int> rptr, rcols, gptr_h, glist_h;
    if ((rc = upload(&lv.indptr, indptr, (size_t)n + 1, s))) return rc;
    if ((rc = upload(&lv.indices, indices, (size_t)nnz, s))) return rc;
    if ((rc = upload(&lv.diagpos, diag.data(), (size_t)n, s))) return rc;
    if ((rc = upload(&lv.gptr, gather_ptr, (size_t)nnz + 1, s))) return rc;
    if ((rc = upload(&lv.glist, gather_list, (size_t)ng, s))) return rc;
    if (L > 0) {
        if ((rc = upload(&lv.agg_ptr, agg_ptr, (size_t)n + 1, s))) return rc;
        if ((rc = upload(&lv.agg_kp, agg_kp.data(), (size_t)n + 1, s))) return rc;
        // the pattern of R = P^T (I - omega A_f D_f^-1): per aggregate the
        // sorted distinct fine rows reached by its rows' nonzeros, and for
        // each entry the fine nonzeros that sum to it, in nonzero order
        const AmsLevel &f = A->lv[L - 1];
        rptr.assign((size_t)n + 1, 0);
        glist_h.resize((size_t)f.nnz);
        std::vector<int> cnt;
        gptr_h.push_back(0);
        for (int I = 0; I < n; ++I) {
            const int k0 = f.indptr_h[agg_ptr[I]], k1 = f.indptr_h[agg_ptr[I + 1]];
            std::vector<int> cols(f.indices_h.begin() + k0, f.indices_h.begin() + k1);
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            cnt.assign(cols.size(), 0);
            for (int k = k0; k < k1; ++k)
                ++cnt[std::lower_bound(cols.begin(), cols.end(), f.indices_h[k]) - cols.begin()];
            std::vector<int> pos(cols.size());
            for (size_t q = 0; q < cols.size(); ++q) {
                const int j = cols[q];
                const bool own = j >= agg_ptr[I] && j < agg_ptr[I + 1];
                rcols.push_back(own ? (int)(j | 0x80000000u) : j);
                pos[q] = gptr_h.back();
                gptr_h.push_back(gptr_h.back() + cnt[q]);
            }
            for (int k = k0; k < k1; ++k) {
                const size_t q = std::lower_bound(cols.begin(), cols.end(), f.indices_h[k]) -
                                 cols.begin();
                glist_h[pos[q]++] = k;
            }
            rptr[I + 1] = (int)rcols.size();
        }
        SN_REQUIRE(gptr_h.back() == f.nnz, "sn_ams_add_level: restriction plan does not cover");
        lv.rt_nnz = (int)rcols.size();
        if ((rc = upload(&lv.rt_ptr, rptr.data(), rptr.size(), s))) return rc;
        if ((rc = upload(&lv.rt_cols, rcols.data(), rcols.size(), s))) return rc;
        if ((rc = upload(&lv.rt_gptr, gptr_h.data(), gptr_h.size(), s))) return rc;
        if ((rc = upload(&lv.rt_glist, glist_h.data(), glist_h.size(), s))) return rc;
        SN_HIP(hipMalloc(&lv.rt_vals, (size_t)std::max(1, lv.rt_nnz) * sizeof(float)));
    }
    SN_HIP(hipMalloc(&lv.vals, (size_t)nnz * sizeof(float)));
    SN_HIP(hipMalloc(&lv.dinv, (size_t)n * sizeof(float)));
    SN_HIP(hipMalloc(&lv.b, (size_t)n * sizeof(float4)));
    SN_HIP(hipMalloc(&lv.x, (size_t)n * sizeof(float4)));
    SN_HIP(hipStreamSynchronize(s));
    A->lv.push_back(lv);
    A->ready = false;
    return SN_OK;
}

int sn_ams_update(sn_ams *A, const double *nodes, const int *cells, const float *nu_el,
                  void *stream) {
    SN_REQUIRE(A && nodes && cells && nu_el, "sn_ams_update: null argument");
    SN_REQUIRE(!A->lv.empty(), "sn_ams_update: no levels");
    const int L = (int)A->lv.size();
    const int nc = A->lv[L - 1].n;
    SN_REQUIRE(nc <= kAmsMaxCoarse, "sn_ams_update: coarsest level has %d rows (max %d)", nc,
               kAmsMaxCoarse);
    hipStream_t s = as_stream(stream);
    if (!A->cinv) {
        SN_HIP(hipMalloc(&A->dense, (size_t)nc * nc * sizeof(double)));
        SN_HIP(hipMalloc(&A->cinv, (size_t)nc * nc * sizeof(float)));
    }
    hipLaunchKernelGGL(ams_coords_kernel, dim3(grid_for(A->n_nodes, 256)), dim3(256), 0, s,
                       A->n_nodes, A->n_free, nodes, A->node_mesh, A->xyz, A->yx);
    hipLaunchKernelGGL(ams_lap_local_kernel, dim3(grid_for(A->n_el, kAmsBlock)), dim3(kAmsBlock),
                       0, s, A->n_el, nodes, cells, nu_el, A->loc);
    for (int l = 0; l < L; ++l) {
        AmsLevel &v = A->lv[l];
        const float *src = (l == 0) ? A->loc : A->lv[l - 1].vals;
        hipLaunchKernelGGL(ams_gather_kernel, dim3(grid_for(v.nnz, 256)), dim3(256), 0, s, v.nnz,
                           v.gptr, v.glist, src, v.vals);
        hipLaunchKernelGGL(ams_dinv_kernel, dim3(grid_for(v.n, 256)), dim3(256), 0, s, v.n,
                           v.diagpos, v.vals, v.dinv);
    }
    for (int l = 1; l < L; ++l) {
        AmsLevel &v = A->lv[l], &f = A->lv[l - 1];
        if (v.rt_vals)
            hipLaunchKernelGGL(ams_rt_kernel, dim3(grid_for(v.rt_nnz, 256)), dim3(256), 0, s,
                               v.rt_nnz, v.rt_cols, v.rt_gptr, v.rt_glist, f.vals, f.dinv,
                               A->omega, v.rt_vals);
    }
    AmsLevel &c = A->lv[L - 1];
    const size_t lds = ((nc <= kAmsLdsCoarse) ? (size_t)nc * nc + 2 * nc : 2 * (size_t)nc) *
                       sizeof(double);
    hipLaunchKernelGGL(ams_coarse_invert_kernel, dim3(1), dim3(kAmsInvThreads), lds, s, nc,
                       c.indptr,
                       c.indices, c.vals, A->dense, A->cinv);
    SN_CHECK_LAUNCH();
    A->ready = true;
    return SN_OK;
}

int sn_ams_apply(sn_ams *A, const unsigned char *mask, const double *r, const double *dinv,
                 double *z, void *stream) {
    SN_REQUIRE(A && r && dinv && z, "sn_ams_apply: null argument");
    return ams_launch_apply(A, true, r, dinv, mask, z, 1, nullptr, nullptr, 0,
                            grid_for(A->n_edges, kAmsEdgeBlock, 4 * num_cus()), nullptr, nullptr,
                            0, as_stream(stream));
}

int sn_ams_level_info(const sn_ams *A, int level, int *n, int *nnz) {
    SN_REQUIRE(A && level >= 0 && level < (int)A->lv.size(), "sn_ams_level_info: bad level");
    if (n) *n = A->lv[level].n;
    if (nnz) *nnz = A->lv[level].nnz;
    return SN_OK;
}

int sn_ams_level_values(const sn_ams *A, int level, float *vals, void *stream) {
    SN_REQUIRE(A && vals && level >= 0 && level < (int)A->lv.size(), "sn_ams_level_values: bad argument");
    SN_HIP(hipMemcpyAsync(vals, A->lv[level].vals, (size_t)A->lv[level].nnz * sizeof(float),
                          hipMemcpyDeviceToDevice, as_stream(stream)));
    return SN_OK;
}

int sn_ams_num_levels(const sn_ams *A) { return A ? (int)A->lv.size() : 0; }

int sn_ams_set_restriction(sn_ams *A, int combined) {
    SN_REQUIRE(A, "sn_ams_set_restriction: null handle");
    A->rt = combined != 0;
    return SN_OK;
}

int sn_ams_set_lanes(sn_ams *A, int lanes) {
    SN_REQUIRE(A, "sn_ams_set_lanes: null handle");
    SN_REQUIRE(lanes == 4 || lanes == 8 || lanes == 16,
               "sn_ams_set_lanes: lanes must be 4, 8 or 16 (%d)", lanes);
    A->lanes = lanes;
    return SN_OK;
}

// Greedy aggregation on the strength graph (host): j is a strong neighbour
// of i when |a_ij| >= theta sqrt(|a_ii a_jj|).  Pass 1 makes an aggregate of
// every node whose strong neighbourhood is still unassigned; pass 2 attaches
// each remaining node to the aggregate of its strongest assigned neighbour
// (assignments of pass 1 only); pass 3 groups what is left with its
// unassigned strong neighbours.  Nodes are visited in index order, so the
// result is deterministic.
int sn_amg_aggregate(int n, const int *indptr, const int *indices, const double *vals,
                     double theta, int *agg, int *n_agg) {
    SN_REQUIRE(n > 0 && indptr && indices && vals && agg && n_agg, "sn_amg_aggregate: bad argument");
    std::vector<double> d(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = indptr[i]; k < indptr[i + 1]; ++k) {
            SN_REQUIRE(indices[k] >= 0 && indices[k] < n, "sn_amg_aggregate: column out of range");
            if (indices[k] == i) d[i] = std::fabs(vals[k]);
        }
    auto strong = [&](int i, int k) {
        const int j = indices[k];
        return j != i && std::fabs(vals[k]) >= theta * std::sqrt(d[i] * d[j]);
    };
    for (int i = 0; i < n; ++i) agg[i] = -1;
    int na = 0;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        bool free_nb = true;
        for (int k = indptr[i]; k < indptr[i + 1] && free_nb; ++k)
            if (strong(i, k) && agg[indices[k]] >= 0) free_nb = false;
        if (!free_nb) continue;
        agg[i] = na;
        for (int k = indptr[i]; k < indptr[i + 1]; ++k)
            if (strong(i, k)) agg[indices[k]] = na;
        ++na;
    }
    std::vector<int> pass1(agg, agg + n);
    for (int i = 0; i < n; ++i) {
        if (pass1[i] >= 0) continue;
        double best = -1.0;
        for (int k = indptr[i]; k < indptr[i + 1]; ++k) {
            const int j = indices[k];
            if (!strong(i, k) || pass1[j] < 0) continue;
            const double w = std::fabs(vals[k]);
            if (w > best) {
                best = w;
                agg[i] = pass1[j];
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        agg[i] = na;
        for (int k = indptr[i]; k < indptr[i + 1]; ++k)
            if (strong(i, k) && agg[indices[k]] < 0) agg[indices[k]] = na;
        ++na;
    }
    *n_agg = na;
    return SN_OK;
}

void sn_ams_destroy(sn_ams *A) {
    if (!A) return;
    for (AmsLevel &v : A->lv) {
        void *bufs[] = {v.indptr, v.indices, v.diagpos, v.gptr, v.glist, v.agg_ptr, v.agg_kp,
                        v.vals, v.dinv, v.b, v.x, v.rt_ptr, v.rt_cols, v.rt_gptr, v.rt_glist,
                        v.rt_vals};
        for (void *p : bufs)
            if (p) (void)hipFree(p);
    }
    void *bufs[] = {A->edge_nodes, A->node_mesh, A->xyz, A->yx, A->inc_ptr, A->inc,
                    A->loc, A->dense, A->cinv};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    delete A;
}

}  // extern "C"
