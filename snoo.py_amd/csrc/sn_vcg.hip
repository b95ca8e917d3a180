// Variant-minor CG on the element operator: V systems that share one mesh
// topology (DOF numbering, element-to-DOF lists, Dirichlet mask) and differ
// in their element matrices, right-hand sides and stop tests, solved by the
// same kernels.  No counterpart in the reference, which solves one magnet per
// call (snoopy/red_mvp_solver.py:121-224; scipy cg at :202-205).
//
// Layout: every vector is stored variant-minor, x[i*V + v], and the element
// matrices and products likewise, Mr[(k*N + e)*V + v], so
//   - an index (element DOF ids, DOF contribution lists, incidences, columns,
//     aggregate bounds) is read once for the V systems;
//   - a gather at a DOF or node brings the V variants' values in one
//     contiguous piece (V*8 bytes of {z, p} pairs instead of 8 bytes of a
//     64-byte line).
// The element pass runs V consecutive lanes per element (one per variant);
// the DOF pass, the start and the preconditioner run one thread (or lane
// group) per DOF, node or row that carries all V variants in registers and
// loads them with 16-byte loads, which keeps the dependent index -> value
// chains as short per variant as in the single CG.
// The auxiliary-space preconditioner (sn_ams: hierarchy, matrix values,
// coordinates) is shared by the V systems; only its vectors are per variant.
// It is the preconditioner of one of the systems (the caller's choice) and a
// fixed symmetric positive definite operator for all of them, so each system
// runs a valid preconditioned CG.
//
// Per variant the arithmetic is that of the single element-operator CG
// (sn_cg.hip, cg_elem_kernel / cg_edge_kernel and the AMS apply): scipy's cg
// rule, float32 vectors and products, float64 inner products.  Reductions are
// per variant and fixed-order (per-block partials re-reduced by every block of
// the consuming kernel), so V identical systems give V bit-identical
// results.  Each variant stops on its own test; the element and DOF passes
// then leave it alone.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "sn_ams.hpp"
#include "sn_common.hpp"
#include "sn_pattern.hpp"

namespace sn {

struct VcgState {
    double rho[2];   // rho of even / odd iterations
    double pq, rr, tol, bnorm;
    double beta;     // this iteration's beta, formed by the element pass
    int it, maxiter, done, pad;  // done: 0 running, 1 converged, 2 iteration cap, 3 breakdown
};

constexpr int kVElBlock = 256;    // element pass
constexpr int kVVecBlock = 1024;  // DOF pass, init, prolongation
constexpr int kVAmsBlock = 256;   // restriction, down, coarse, up

namespace {

__device__ __forceinline__ float4 v4_zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
__device__ __forceinline__ float4 half_diff4(const float4 &hi, const float4 &lo) {
    return make_float4(0.5f * (hi.x - lo.x), 0.5f * (hi.y - lo.y), 0.5f * (hi.z - lo.z), 0.f);
}

// blocks dealt round-robin to the 8 XCDs, renumbered so each XCD takes one
// contiguous eighth of the grid (when the grid is a multiple of 8)
__device__ __forceinline__ long long v_block(int xcd) {
    const int b = blockIdx.x, nb = gridDim.x;
    if (xcd && (nb & 7) == 0) return (long long)(b & 7) * (nb >> 3) + (b >> 3);
    return b;
}

// Class sums over a block: thread t contributes to class t % C (C a power of
// two <= 64); the C sums land in out[0..C), visible to every thread on return.
// Fixed order: wave shuffles, then the waves in order.  `lds` holds
// C * blockDim/64 doubles.
template <int C>
__device__ __forceinline__ void class_sums(double s, double *lds, double *out) {
#pragma unroll
    for (int off = kWave / 2; off >= C; off >>= 1) s += __shfl_xor(s, off, kWave);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    __syncthreads();
    if (lane < C) lds[wid * C + lane] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        double t = 0.0;
        for (int w = 0; w < nw; ++w) t += lds[w * C + threadIdx.x];
        out[threadIdx.x] = t;
    }
    __syncthreads();
}

// The C class sums of per-block partials part[b*C + c], b < nb, in the same
// order in every block of a grid of equal blocks.
template <int C>
__device__ __forceinline__ void reduce_classes(const double *__restrict__ part, int nb,
                                               double *lds, double *out) {
    const int c = threadIdx.x % C;
    double s = 0.0;
    for (int b = threadIdx.x / C; b < nb; b += blockDim.x / C) s += part[(long long)b * C + c];
    class_sums<C>(s, lds, out);
}

// the V values of one index: V contiguous floats (doubles), 16-byte loads
template <int NV>
__device__ __forceinline__ void ldv(const float *__restrict__ p, float (&o)[NV]) {
    if constexpr (NV % 4 == 0) {
#pragma unroll
        for (int q = 0; q < NV / 4; ++q) {
            const float4 t = reinterpret_cast<const float4 *>(p)[q];
            o[4 * q] = t.x;
            o[4 * q + 1] = t.y;
            o[4 * q + 2] = t.z;
            o[4 * q + 3] = t.w;
        }
    } else if constexpr (NV == 2) {
        const float2 t = *reinterpret_cast<const float2 *>(p);
        o[0] = t.x;
        o[1] = t.y;
    } else {
        o[0] = p[0];
    }
}
template <int NV>
__device__ __forceinline__ void stv(float *__restrict__ p, const float (&o)[NV]) {
    if constexpr (NV % 4 == 0) {
#pragma unroll
        for (int q = 0; q < NV / 4; ++q)
            reinterpret_cast<float4 *>(p)[q] =
                make_float4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else if constexpr (NV == 2) {
        *reinterpret_cast<float2 *>(p) = make_float2(o[0], o[1]);
    } else {
        p[0] = o[0];
    }
}
template <int NV>
__device__ __forceinline__ void ldv_d(const double *__restrict__ p, double (&o)[NV]) {
    if constexpr (NV % 2 == 0) {
#pragma unroll
        for (int q = 0; q < NV / 2; ++q) {
            const double2 t = reinterpret_cast<const double2 *>(p)[q];
            o[2 * q] = t.x;
            o[2 * q + 1] = t.y;
        }
    } else {
        o[0] = p[0];
    }
}
// the record of index i in a pair array: z of the V variants, then p (2V
// contiguous floats); each half is one contiguous run of 4V bytes, so a kernel
// that writes only z or only p stores whole 16-byte pieces
template <int NV>
__device__ __forceinline__ void ld_pairs(const float *w, long long i, float (&z)[NV],
                                         float (&p)[NV]) {
    ldv<NV>(w + i * 2 * NV, z);
    ldv<NV>(w + i * 2 * NV + NV, p);
}
template <int NV>
__device__ __forceinline__ void st_pairs(float *w, long long i, const float (&z)[NV],
                                         const float (&p)[NV]) {
    stv<NV>(w + i * 2 * NV, z);
    stv<NV>(w + i * 2 * NV + NV, p);
}

// Per-variant block sums of the threads' values vals[v]: the V sums in
// lds[0..V) on return, visible to every thread (fixed order: wave shuffles,
// then the waves in order).  `lds` holds V * blockDim/64 doubles.
template <int NV>
__device__ __forceinline__ void block_sums(const double (&vals)[NV], double *lds) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    double s[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = wave_sum(vals[v]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) lds[wid * NV + v] = s[v];
    }
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < NV)
        for (int w = 0; w < nw; ++w) t += lds[w * NV + threadIdx.x];
    __syncthreads();
    if (threadIdx.x < NV) lds[threadIdx.x] = t;
    __syncthreads();
}

// c_j . Mr c_j for local edge j from the upper triangle 00 01 02 11 12 22 of
// Mr: a + b - 2c with the components {a, b, c} below for j < 3 (edges 0-2),
// the single component a for j >= 3
__constant__ int kDiagComp[6][3] = {{3, 5, 4}, {0, 5, 2}, {0, 3, 1}, {5, 0, 0}, {3, 0, 0},
                                    {0, 0, 0}};

}  // namespace

// signed DOF ids, slot j*n_el + e: id | bit 31 when the edge's orientation is
// flipped in element e
__global__ void vcg_el_dofs_kernel(int n_el, const int *__restrict__ gids,
                                   const int *__restrict__ orient, int *__restrict__ el_dofs) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    const int o = orient[e];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int g = gids[6 * e + j];
        el_dofs[j * (long long)n_el + e] = ((o >> j) & 1) ? (int)(g | 0x80000000u) : g;
    }
}

// r = b, the diagonal of every variant's operator from its element matrices,
// dinv, z = dinv r, p = 0 in both pair arrays, xs = 0 (thread per DOF, all
// variants); per-block partials of (r.z, r.r) and b.b per variant
template <int NV>
__global__ __launch_bounds__(kVVecBlock) void vcg_init_kernel(
    int n, int n_el, const int *__restrict__ dof_ptr, const int *__restrict__ dof_contrib,
    const float *__restrict__ Mr, const unsigned char *__restrict__ mask,
    const double *__restrict__ b, float *__restrict__ r, float *__restrict__ w,
    float *__restrict__ w_other, float *__restrict__ dinv, float *__restrict__ xs,
    double *__restrict__ partB, double *__restrict__ partbb) {
    __shared__ double lds[NV * kVVecBlock / kWave];
    double bb[NV], rr[NV], rz[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) bb[v] = rr[v] = rz[v] = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const bool free_row = mask ? mask[i] != 0 : true;
        float ri[NV], di[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) ri[v] = di[v] = 0.f;
        if (free_row) {
            double dg[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) dg[v] = 0.0;
            for (int k = dof_ptr[i]; k < dof_ptr[i + 1]; ++k) {
                const int slot = dof_contrib[k];
                const int j = slot / n_el;
                const long long e = slot - (long long)j * n_el;
                // c_j . Mr c_j needs three components for j < 3, one else
                const int ca = kDiagComp[j][0], cb = kDiagComp[j][1], cc = kDiagComp[j][2];
                float ma[NV], mb[NV], mc[NV];
                ldv<NV>(Mr + (ca * (long long)n_el + e) * NV, ma);
                if (j < 3) {
                    ldv<NV>(Mr + (cb * (long long)n_el + e) * NV, mb);
                    ldv<NV>(Mr + (cc * (long long)n_el + e) * NV, mc);
#pragma unroll
                    for (int v = 0; v < NV; ++v)
                        dg[v] += (double)ma[v] + (double)mb[v] - 2.0 * (double)mc[v];
                } else {
#pragma unroll
                    for (int v = 0; v < NV; ++v) dg[v] += (double)ma[v];
                }
            }
            double bi[NV];
            ldv_d<NV>(b + i * NV, bi);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                bb[v] += bi[v] * bi[v];
                ri[v] = (float)bi[v];
                di[v] = dg[v] != 0.0 ? (float)(1.0 / dg[v]) : 0.f;
            }
        }
        float zi[NV], zero[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            zi[v] = di[v] * ri[v];
            zero[v] = 0.f;
            rr[v] += (double)ri[v] * (double)ri[v];
            rz[v] += (double)ri[v] * (double)zi[v];
        }
        stv<NV>(r + i * NV, ri);
        stv<NV>(dinv + i * NV, di);
        stv<NV>(xs + i * NV, zero);
        st_pairs<NV>(w, i, zi, zero);
        st_pairs<NV>(w_other, i, zero, zero);
    }
    block_sums<NV>(bb, lds);
    if (threadIdx.x < NV) partbb[(long long)blockIdx.x * NV + threadIdx.x] = lds[threadIdx.x];
    block_sums<NV>(rz, lds);
    if (threadIdx.x < NV) partB[((long long)blockIdx.x * NV + threadIdx.x) * 2] = lds[threadIdx.x];
    block_sums<NV>(rr, lds);
    if (threadIdx.x < NV)
        partB[((long long)blockIdx.x * NV + threadIdx.x) * 2 + 1] = lds[threadIdx.x];
}

// one block: every variant's tolerance, ||b|| and iteration-0 state
template <int NV>
__global__ __launch_bounds__(kVVecBlock) void vcg_init_state_kernel(
    const double *__restrict__ partbb, int nb, VcgState *__restrict__ st,
    const double *__restrict__ rtol, double atol, int maxiter) {
    __shared__ double lds[NV * kVVecBlock / kWave];
    __shared__ double sums[NV];
    reduce_classes<NV>(partbb, nb, lds, sums);
    if (threadIdx.x < NV) {
        VcgState *S = st + threadIdx.x;
        const double bnorm = sqrt(sums[threadIdx.x]);
        const double rt = rtol[threadIdx.x] * bnorm;
        S->bnorm = bnorm;
        S->tol = atol > rt ? atol : rt;
        S->rho[0] = S->rho[1] = 0.0;
        S->pq = 0.0;
        S->rr = sums[threadIdx.x];   // ||r_0||^2 = ||b||^2 (x_0 = 0), for the host's rate estimate
        S->beta = 0.0;
        S->it = 0;
        S->maxiter = maxiter;
        S->done = (bnorm == 0.0) ? 1 : 0;
    }
}

// Element pass (thread per element and variant): every variant's stop test
// from the previous kernel's partials; p = z + beta p at the element's 6 DOFs,
// face circulations u, m = Mr u, the 6 element products, p.q partials.
// parity = iteration & 1: reads rho[parity ^ 1], writes rho[parity].
// DM: the products go to their places in the DOF lists (ypos[j*N + e]), so
// the DOF pass reads each DOF's products as one contiguous run; otherwise to
// slot j*N + e, which the DOF pass gathers through the lists.
template <int NV, bool DM>
__global__ __launch_bounds__(kVElBlock) void vcg_elem_kernel(
    int n_el, const int *__restrict__ el_dofs, const int *__restrict__ ypos,
    const float *__restrict__ Mr, const float *__restrict__ zp, float *__restrict__ yloc,
    const double *__restrict__ partB, int nbB, double *__restrict__ partE,
    VcgState *__restrict__ st, int parity, int xcd) {
    __shared__ double lds[2 * NV * kVElBlock / kWave];
    __shared__ double sums[2 * NV];
    const int v = threadIdx.x % NV;
    const int was_done = st[v].done;
    const int it = st[v].it, maxit = st[v].maxiter;
    const double tol = st[v].tol, rho_prev = st[v].rho[parity ^ 1];
    reduce_classes<2 * NV>(partB, nbB, lds, sums);
    const double rho = sums[2 * v], rr = sums[2 * v + 1];
    const int stop = was_done ? 0 : ((it >= maxit) ? 2 : ((sqrt(rr) < tol) ? 1 : 0));
    const bool active = !was_done && !stop;
    const double beta_d = it > 0 ? rho / rho_prev : 0.0;
    if (blockIdx.x == 0 && threadIdx.x < NV && !was_done) {
        VcgState *S = st + v;
        S->rr = rr;
        if (stop) {
            S->done = stop;
        } else {
            S->rho[parity] = rho;
            S->beta = beta_d;
        }
    }
    const float beta = (float)beta_d;
    double pq = 0.0;
    if (active) {
        const long long step = (long long)gridDim.x * blockDim.x / NV;
        for (long long e = (v_block(xcd) * blockDim.x + threadIdx.x) / NV; e < n_el; e += step) {
            int d[6], pos[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                d[j] = el_dofs[j * (long long)n_el + e];
                pos[j] = DM ? ypos[j * (long long)n_el + e] : 0;
            }
            float a[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) a[k] = Mr[(k * (long long)n_el + e) * NV + v];
            float xh[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                const float *w = zp + (long long)(d[j] & 0x7fffffff) * 2 * NV;
                const float pj = w[v] + beta * w[NV + v];
                xh[j] = d[j] < 0 ? -pj : pj;
            }
            const float u0 = xh[1] - xh[2] + xh[5];
            const float u1 = xh[2] - xh[0] - xh[4];
            const float u2 = xh[0] - xh[1] + xh[3];
            const float m0 = a[0] * u0 + a[1] * u1 + a[2] * u2;
            const float m1 = a[1] * u0 + a[3] * u1 + a[4] * u2;
            const float m2 = a[2] * u0 + a[4] * u1 + a[5] * u2;
            pq += (double)u0 * (double)m0 + (double)u1 * (double)m1 + (double)u2 * (double)m2;
            const float yh[6] = {m2 - m1, m0 - m2, m1 - m0, m2, -m1, m0};
#pragma unroll
            for (int j = 0; j < 6; ++j)
                yloc[(DM ? (long long)pos[j] : j * (long long)n_el + e) * NV + v] =
                    d[j] < 0 ? -yh[j] : yh[j];
        }
    }
    class_sums<NV>(pq, lds, sums);
    if (threadIdx.x < NV) partE[(long long)blockIdx.x * NV + threadIdx.x] = sums[threadIdx.x];
}

// DOF pass (thread per DOF, all variants): every variant's alpha from the
// element pass's partials; q_i = fixed-order sum of the element products;
// x += alpha p, r -= alpha q, r.r partials; JAC: z = dinv r and the {z, p}
// record of wnew, r.z partials; otherwise p into the p half of the record it
// has read from zp (the preconditioner then writes the z half).  A variant
// that is done keeps x, r as they are.  DM: the DOF's
// products are the run [dof_ptr[i], dof_ptr[i+1]) of yloc (vcg_elem_kernel
// with DM), else gathered through the DOF lists; the same values are summed
// in the same order either way.
template <int NV, bool JAC, bool DM>
__global__ __launch_bounds__(kVVecBlock) void vcg_edge_kernel(
    int n, const int *__restrict__ dof_ptr, const int *__restrict__ dof_contrib,
    const float *__restrict__ yloc, const unsigned char *__restrict__ mask,
    const float *__restrict__ dinv, float *__restrict__ zp, float *__restrict__ wnew,
    float *__restrict__ r, float *__restrict__ x,
    const double *__restrict__ partE, int nbE, double *__restrict__ partB,
    VcgState *__restrict__ st, int parity, int xcd) {
    __shared__ double lds[NV * kVVecBlock / kWave];
    __shared__ double sums[NV];
    bool active[NV];
    float alpha[NV], beta[NV];
    double rho[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        active[v] = st[v].done == 0;   // set by the element pass when v stops there
        rho[v] = st[v].rho[parity];
        beta[v] = (float)st[v].beta;
    }
    reduce_classes<NV>(partE, nbE, lds, sums);
    bool any = false;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double pq = sums[v];
        if (!(pq == pq) || pq == 0.0 || isinf(pq)) active[v] = false;
        alpha[v] = active[v] ? (float)(rho[v] / pq) : 0.f;
        any = any || active[v];
    }
    double rz[NV], rr[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) rz[v] = rr[v] = 0.0;
    if (any) {
        // the first kHead list entries and their products loaded at once (a
        // missing entry adds 0); each DOF's sum keeps the list order
        constexpr int kHead = NV >= 8 ? 6 : 8;
        for (long long i = v_block(xcd) * blockDim.x + threadIdx.x; i < n;
             i += (long long)gridDim.x * blockDim.x) {
            const bool fr = mask ? mask[i] != 0 : true;
            const int k0 = dof_ptr[i], k1 = dof_ptr[i + 1];
            int c[kHead];
#pragma unroll
            for (int m = 0; m < kHead; ++m)
                c[m] = (fr && k0 + m < k1) ? (DM ? k0 + m : dof_contrib[k0 + m]) : -1;
            float zz[NV], pp[NV];
            ld_pairs<NV>(zp, i, zz, pp);
            float q[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) q[v] = 0.f;
#pragma unroll
            for (int m = 0; m < kHead; ++m) {
                float y[NV];
                if (c[m] >= 0) {
                    ldv<NV>(yloc + (long long)c[m] * NV, y);
                } else {
#pragma unroll
                    for (int v = 0; v < NV; ++v) y[v] = 0.f;
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) q[v] += y[v];
            }
            float zero[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) zero[v] = 0.f;
            if (!fr) {
                if (JAC) st_pairs<NV>(wnew, i, zero, zero);
                else stv<NV>(zp + i * 2 * NV + NV, zero);
                continue;
            }
            for (int k = k0 + kHead; k < k1; ++k) {
                float y[NV];
                ldv<NV>(yloc + (long long)(DM ? k : dof_contrib[k]) * NV, y);
#pragma unroll
                for (int v = 0; v < NV; ++v) q[v] += y[v];
            }
            float xi[NV], ri[NV], pi[NV], zi[NV];
            ldv<NV>(x + i * NV, xi);
            ldv<NV>(r + i * NV, ri);
            float di[NV];
            if (JAC) ldv<NV>(dinv + i * NV, di);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                pi[v] = zz[v] + beta[v] * pp[v];
                if (active[v]) {
                    xi[v] += alpha[v] * pi[v];
                    ri[v] = ri[v] - alpha[v] * q[v];
                    rr[v] += (double)ri[v] * (double)ri[v];
                }
                if (JAC) {
                    zi[v] = di[v] * ri[v];
                    if (active[v]) rz[v] += (double)ri[v] * (double)zi[v];
                }
            }
            stv<NV>(x + i * NV, xi);
            stv<NV>(r + i * NV, ri);
            if (JAC) st_pairs<NV>(wnew, i, zi, pi);
            else stv<NV>(zp + i * 2 * NV + NV, pi);
        }
    }
    if (JAC) {
        block_sums<NV>(rz, lds);
        if (threadIdx.x < NV) partB[((long long)blockIdx.x * NV + threadIdx.x) * 2] = lds[threadIdx.x];
    }
    block_sums<NV>(rr, lds);
    if (threadIdx.x < NV) partB[((long long)blockIdx.x * NV + threadIdx.x) * 2 + 1] = lds[threadIdx.x];
    if (blockIdx.x == 0 && threadIdx.x < NV) {
        VcgState *S = st + threadIdx.x;
        const double pq = sums[threadIdx.x];
        if (S->done == 0) {
            S->pq = pq;
            if (!(pq == pq) || pq == 0.0 || isinf(pq)) S->done = 3;
            else S->it = S->it + 1;   // no other block writes it, and no block reads it here
        }
    }
}

// ---------------------------------------------------- the preconditioner, V-wide
// Thread (or lane group) per node, row or edge handling all V variants.  The
// preconditioner's vectors hold three components per node and variant,
// packed [node][component][variant] (12 V bytes per node, no padding), so
// each gathered index brings its V variants' values in 3 V contiguous
// floats.  The matrices are shared, so one variant's values never mix with
// another's; a variant that is done is computed along (its vectors are
// stale and finite) and its results are not used.

template <int NV>
__device__ __forceinline__ void ld3(const float *__restrict__ p, float (&o)[3][NV]) {
    ldv<NV>(p, o[0]);
    ldv<NV>(p + NV, o[1]);
    ldv<NV>(p + 2 * NV, o[2]);
}
template <int NV>
__device__ __forceinline__ void st3(float *__restrict__ p, const float (&o)[3][NV]) {
    stv<NV>(p, o[0]);
    stv<NV>(p + NV, o[1]);
    stv<NV>(p + 2 * NV, o[2]);
}
template <int NV>
__device__ __forceinline__ void zero3(float (&a)[3][NV]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int v = 0; v < NV; ++v) a[c][v] = 0.f;
}
// a += s * b
template <int NV>
__device__ __forceinline__ void fma3(float (&a)[3][NV], float s, const float (&b)[3][NV]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int v = 0; v < NV; ++v) a[c][v] += s * b[c][v];
}
// sum over the lanes that differ in bits [0, log2 G) of the lane id (G lanes
// of one node, row or aggregate)
template <int NV, int G>
__device__ __forceinline__ void reduce3(float (&a)[3][NV]) {
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int v = 0; v < NV; ++v) a[c][v] += __shfl_xor(a[c][v], off, kWave);
}

// all variants done: every kernel of the apply returns at once
template <int NV>
__device__ __forceinline__ bool all_done(const VcgState *__restrict__ st) {
    bool d = true;
#pragma unroll
    for (int v = 0; v < NV; ++v) d = d && st[v].done != 0;
    return d;
}

// LPR lanes per free node: b0 = Pi^T r, x0 = w D^-1 b0
template <int NV, int LPR>
__global__ __launch_bounds__(kVAmsBlock) void vams_restrict_kernel(
    int n_free, const int *__restrict__ inc_ptr, const int2 *__restrict__ inc,
    const float4 *__restrict__ xyz, const float *__restrict__ r, const float *__restrict__ dinv0,
    float omega, float *__restrict__ b0, float *__restrict__ x0,
    const VcgState *__restrict__ st, int xcd) {
    if (all_done<NV>(st)) return;
    const long long a = (v_block(xcd) * blockDim.x + threadIdx.x) / LPR;
    const int lane = threadIdx.x % LPR;
    const bool in = a < n_free;
    const int k0 = in ? inc_ptr[a] : 0, k1 = in ? inc_ptr[a + 1] : 0;
    const float4 xa = in ? xyz[a] : v4_zero();
    float acc[3][NV];
    zero3<NV>(acc);
    for (int k = k0 + lane; k < k1; k += LPR) {
        const int2 ie = inc[k];   // {edge, other end | sign}
        const float4 xb = xyz[ie.y & 0x7fffffff];
        float rv[NV];
        ldv<NV>(r + (long long)ie.x * NV, rv);
        const float4 tg = ie.y < 0 ? half_diff4(xa, xb) : half_diff4(xb, xa);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            acc[0][v] += rv[v] * tg.x;
            acc[1][v] += rv[v] * tg.y;
            acc[2][v] += rv[v] * tg.z;
        }
    }
    reduce3<NV, LPR>(acc);
    if (!in || lane != 0) return;
    const float d = omega * dinv0[a];
    st3<NV>(b0 + a * 3 * NV, acc);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[c][v] *= d;
    st3<NV>(x0 + a * 3 * NV, acc);
}

// SL lanes per aggregate of the next level: b_c = sum_rows b - sum_nnz a x;
// x_c = w D_c^-1 b_c unless the next level is the coarsest.  TWO: two
// nonzeros per trip (small levels, bound by the dependent chains; on level 0
// the extra loads in flight cost more than they hide)
template <int NV, int SL, bool TWO>
__global__ __launch_bounds__(kVAmsBlock) void vams_down_kernel(
    int n_agg, const int *__restrict__ agg_ptr, const int *__restrict__ agg_kp,
    const int *__restrict__ indices, const float *__restrict__ vals, const float *__restrict__ b,
    const float *__restrict__ x, const float *__restrict__ dinv_c, float omega,
    float *__restrict__ bc, float *__restrict__ xc, const VcgState *__restrict__ st, int xcd) {
    if (all_done<NV>(st)) return;
    const long long ag = (v_block(xcd) * blockDim.x + threadIdx.x) / SL;
    const int slot = threadIdx.x % SL;
    const bool in = ag < n_agg;
    const int r0 = in ? agg_ptr[ag] : 0, r1 = in ? agg_ptr[ag + 1] : 0;
    const int k0 = in ? agg_kp[ag] : 0, k1 = in ? agg_kp[ag + 1] : 0;
    float s[3][NV];
    zero3<NV>(s);
    for (int i = r0 + slot; i < r1; i += SL) {
        float bi[3][NV];
        ld3<NV>(b + (long long)i * 3 * NV, bi);
        fma3<NV>(s, 1.f, bi);
    }
    // two nonzeros per trip (k and k + SL, the order of a one-at-a-time
    // loop), both index -> x chains in flight at once
    for (int k = k0 + slot; k < k1; k += (TWO ? 2 : 1) * SL) {
        const bool h2 = TWO && k + SL < k1;
        const int k2 = h2 ? k + SL : k;
        const float a = vals[k];
        float xa[3][NV];
        ld3<NV>(x + (long long)indices[k] * 3 * NV, xa);
        if (TWO) {
            const float a2 = h2 ? vals[k2] : 0.f;
            float xb[3][NV];
            ld3<NV>(x + (long long)indices[k2] * 3 * NV, xb);
            fma3<NV>(s, -a, xa);
            fma3<NV>(s, -a2, xb);
        } else {
            fma3<NV>(s, -a, xa);
        }
    }
    reduce3<NV, SL>(s);
    if (!in || slot != 0) return;
    st3<NV>(bc + ag * 3 * NV, s);
    if (xc) {
        const float d = omega * dinv_c[ag];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int v = 0; v < NV; ++v) s[c][v] *= d;
        st3<NV>(xc + ag * 3 * NV, s);
    }
}

// The same step through the combined restriction R = P^T (I - w A_f D_f^-1)
// of the coarse level: b_c = R b_f, a gather of b_f per entry of R (its
// distinct fine rows, about half the fine nonzeros of the aggregate on level
// 0), instead of b summed over the aggregate's rows minus A_f applied to the
// pre-smoothed x_f = w D_f^-1 b_f; x_c = w D_c^-1 b_c unless the next level
// is the coarsest.  The values differ from that step in float rounding only.
template <int NV, int SL, bool TWO>
__global__ __launch_bounds__(kVAmsBlock) void vams_down_rt_kernel(
    int n_agg, const int *__restrict__ rt_ptr, const int *__restrict__ rt_cols,
    const float *__restrict__ rt_vals, const float *__restrict__ b,
    const float *__restrict__ dinv_c, float omega, float *__restrict__ bc,
    float *__restrict__ xc, const VcgState *__restrict__ st, int xcd) {
    if (all_done<NV>(st)) return;
    const long long ag = (v_block(xcd) * blockDim.x + threadIdx.x) / SL;
    const int slot = threadIdx.x % SL;
    const bool in = ag < n_agg;
    const int k0 = in ? rt_ptr[ag] : 0, k1 = in ? rt_ptr[ag + 1] : 0;
    float s[3][NV];
    zero3<NV>(s);
    for (int k = k0 + slot; k < k1; k += (TWO ? 2 : 1) * SL) {
        const bool h2 = TWO && k + SL < k1;
        const int k2 = h2 ? k + SL : k;
        const float a = rt_vals[k];
        float ba[3][NV];
        ld3<NV>(b + (long long)(rt_cols[k] & 0x7fffffff) * 3 * NV, ba);
        if (TWO) {
            const float a2 = h2 ? rt_vals[k2] : 0.f;
            float bb[3][NV];
            ld3<NV>(b + (long long)(rt_cols[k2] & 0x7fffffff) * 3 * NV, bb);
            fma3<NV>(s, a, ba);
            fma3<NV>(s, a2, bb);
        } else {
            fma3<NV>(s, a, ba);
        }
    }
    reduce3<NV, SL>(s);
    if (!in || slot != 0) return;
    st3<NV>(bc + ag * 3 * NV, s);
    if (xc) {
        const float d = omega * dinv_c[ag];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int v = 0; v < NV; ++v) s[c][v] *= d;
        st3<NV>(xc + ag * 3 * NV, s);
    }
}

// wave per coarsest row I: e_I = (C^-1 b)_I, prolongated into the finer
// level's x (x_f[i] += e_I for the rows i of aggregate I), or stored as the
// level-0 result y when there is one level
template <int NV>
__global__ __launch_bounds__(kVAmsBlock) void vams_coarse_kernel(
    int nc, const float *__restrict__ cinv, const float *__restrict__ b,
    const int *__restrict__ agg_ptr, float *__restrict__ x_f, float *__restrict__ y,
    const VcgState *__restrict__ st) {
    if (all_done<NV>(st)) return;
    const long long I = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (I >= nc) return;   // uniform over the wave
    float acc[3][NV];
    zero3<NV>(acc);
    for (int j = lane; j < nc; j += kWave) {
        float bj[3][NV];
        ld3<NV>(b + (long long)j * 3 * NV, bj);
        fma3<NV>(acc, cinv[I * nc + j], bj);
    }
    reduce3<NV, kWave>(acc);
    if (x_f) {
        for (int i = agg_ptr[I] + lane; i < agg_ptr[I + 1]; i += kWave) {
            float xi[3][NV];
            ld3<NV>(x_f + (long long)i * 3 * NV, xi);
            fma3<NV>(xi, 1.f, acc);
            st3<NV>(x_f + (long long)i * 3 * NV, xi);
        }
    } else if (lane == 0) {
        st3<NV>(y + I * 3 * NV, acc);
    }
}

// LPR lanes per row of level l, x already holding x + P e:
// y = x + w D^-1 (b - A x), prolongated into the finer level (x_f[i] += y
// for the rows i of this row's aggregate) or stored as the level-0 result.
// TWO as in vams_down_kernel.
template <int NV, int LPR, bool TWO>
__global__ __launch_bounds__(kVAmsBlock) void vams_up_kernel(
    int n, const int *__restrict__ indptr, const int *__restrict__ indices,
    const float *__restrict__ vals, const float *__restrict__ dinv, const float *__restrict__ b,
    const float *__restrict__ x, float omega, const int *__restrict__ agg_ptr,
    float *__restrict__ x_f, float *__restrict__ y, const VcgState *__restrict__ st, int xcd) {
    if (all_done<NV>(st)) return;
    const long long row = (v_block(xcd) * blockDim.x + threadIdx.x) / LPR;
    const int lane = threadIdx.x % LPR;
    const bool in = row < n;
    const int k0 = in ? indptr[row] : 0, k1 = in ? indptr[row + 1] : 0;
    float acc[3][NV];
    zero3<NV>(acc);
    // two nonzeros per trip, as in vams_down_kernel
    for (int k = k0 + lane; k < k1; k += (TWO ? 2 : 1) * LPR) {
        const bool h2 = TWO && k + LPR < k1;
        const int k2 = h2 ? k + LPR : k;
        const float a = vals[k];
        float xa[3][NV];
        ld3<NV>(x + (long long)indices[k] * 3 * NV, xa);
        if (TWO) {
            const float a2 = h2 ? vals[k2] : 0.f;
            float xb[3][NV];
            ld3<NV>(x + (long long)indices[k2] * 3 * NV, xb);
            fma3<NV>(acc, a, xa);
            fma3<NV>(acc, a2, xb);
        } else {
            fma3<NV>(acc, a, xa);
        }
    }
    reduce3<NV, LPR>(acc);
    if (!in) return;
    const float d = omega * dinv[row];
    float yi[3][NV], bi[3][NV];
    ld3<NV>(x + row * 3 * NV, yi);
    ld3<NV>(b + row * 3 * NV, bi);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int v = 0; v < NV; ++v) yi[c][v] += d * (bi[c][v] - acc[c][v]);
    if (x_f) {
        for (int i = agg_ptr[row] + lane; i < agg_ptr[row + 1]; i += LPR) {
            float xi[3][NV];
            ld3<NV>(x_f + (long long)i * 3 * NV, xi);
            fma3<NV>(xi, 1.f, yi);
            st3<NV>(x_f + (long long)i * 3 * NV, xi);
        }
    } else if (lane == 0) {
        st3<NV>(y + row * 3 * NV, yi);
    }
}

// thread per edge, all variants: z = D_J^-1 r + w Pi y0 on the free edges (0
// elsewhere), stored in the z half of the edge's record in w (the p half holds
// the CG's p, written by the DOF pass, or 0 at the start); per-block partials
// of r.z
template <int NV>
__global__ __launch_bounds__(kVVecBlock) void vams_prolong_kernel(
    int n_edges, const unsigned char *__restrict__ mask, const float *__restrict__ r,
    const float *__restrict__ dinv, const int2 *__restrict__ edge_nodes,
    const float4 *__restrict__ xyz, const float *__restrict__ y, float weight,
    float *__restrict__ w, double *__restrict__ partB,
    const VcgState *__restrict__ st, int xcd) {
    __shared__ double lds[NV * kVVecBlock / kWave];
    double rz[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) rz[v] = 0.0;
    if (!all_done<NV>(st)) {
        for (long long i = v_block(xcd) * blockDim.x + threadIdx.x; i < n_edges;
             i += (long long)gridDim.x * blockDim.x) {
            const bool f = mask ? mask[i] != 0 : true;
            float zi[NV];
            if (f) {
                float ri[NV], di[NV], ylo[3][NV], yhi[3][NV];
                ldv<NV>(r + i * NV, ri);
                ldv<NV>(dinv + i * NV, di);
                const int2 en = edge_nodes[i];
                ld3<NV>(y + (long long)en.x * 3 * NV, ylo);
                ld3<NV>(y + (long long)en.y * 3 * NV, yhi);
                const float4 t = half_diff4(xyz[en.y], xyz[en.x]);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const float dot = t.x * (ylo[0][v] + yhi[0][v]) +
                                      t.y * (ylo[1][v] + yhi[1][v]) +
                                      t.z * (ylo[2][v] + yhi[2][v]);
                    zi[v] = di[v] * ri[v] + weight * dot;
                    rz[v] += (double)ri[v] * (double)zi[v];
                }
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) zi[v] = 0.f;
            }
            stv<NV>(w + i * 2 * NV, zi);
        }
    }
    block_sums<NV>(rz, lds);
    if (threadIdx.x < NV) partB[((long long)blockIdx.x * NV + threadIdx.x) * 2] = lds[threadIdx.x];
}

// place of every element product in the DOF lists: ypos[dof_contrib[k]] = k
// (every slot j*N + e appears in exactly one list)
__global__ void vcg_slot_pos_kernel(int n_slots, const int *__restrict__ dof_contrib,
                                    int *__restrict__ ypos) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_slots) return;
    ypos[dof_contrib[k]] = k;
}

// timing state: partials that reduce to rho = ||r||^2 = 1 for every variant,
// z = 1 on the free rows and p = 0 in the records read by parity 1
__global__ void vcg_fill_timing_kernel(long long n_nv, int nv, int nb,
                                       const unsigned char *__restrict__ mask,
                                       double *__restrict__ partB, float *__restrict__ w1) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 2LL * nb * nv) partB[k] = (k < 2LL * nv) ? 1.0 : 0.0;
    if (k < n_nv) {
        const long long i = k / nv, v = k - i * nv;
        w1[i * 2 * nv + v] = (mask && !mask[i]) ? 0.f : 1.f;
        w1[i * 2 * nv + nv + v] = 0.f;
    }
}

// Variant-minor operands from per-variant arrays: Mr[(j*N + e)*V + v] =
// src_v[j*N + e] and b[i*V + v] = -R_v[i] (0 where R_v is null).  A thread
// per slot or DOF reads the V sources (each coalesced across the threads) and
// writes its V values as one contiguous run.
struct VcgPackSrc {
    const float *mr[8];
    const double *r[8];
};

template <int NV>
__global__ __launch_bounds__(256) void vcg_pack_kernel(long long n_slots, long long n,
                                                       VcgPackSrc src, float *__restrict__ Mr,
                                                       double *__restrict__ b) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_slots) {
        float m[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) m[v] = src.mr[v][k];
#pragma unroll
        for (int v = 0; v < NV; ++v) Mr[k * NV + v] = m[v];
    }
    if (k < n) {
        double r[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) r[v] = src.r[v] ? -src.r[v][k] : 0.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) b[k * NV + v] = r[v];
    }
}

// x64[i*V + v] += xs (the float32 solution) on free rows
__global__ void vcg_finalize_kernel(long long n_nv, int nv, const unsigned char *__restrict__ mask,
                                    const float *__restrict__ xs, double *__restrict__ x) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_nv) return;
    if (mask && !mask[k / nv]) return;
    x[k] += (double)xs[k];
}

}  // namespace sn

struct sn_vcg {
    const sn_pattern *pat = nullptr;
    int nv = 0, n = 0, n_el = 0;
    int chunk = 32, chunk_short = 8, chunk_tail = 2;
    int grid_el = 0, grid_vec = 0;
    int xcd = 1;
    int *el_dofs = nullptr;            // [6*n_el] signed DOF ids (shared)
    int *ypos = nullptr;               // [6*n_el] place of slot j*n_el + e in the DOF lists
    // element products stored in DOF-list order (read as one run per DOF)
    // instead of by element slot: the DOF pass took 128 -> 119 us at 1M tets x
    // 8 variants, the element pass's scattered stores 103 -> 161 us, so off
    bool dm = false;
    // the down steps through the coarse levels' combined restrictions R
    // (default; off: the fine residual formed from the pre-smoothed x)
    bool rt = true;
    float *yloc = nullptr;             // [6*n_el*nv]
    float *r = nullptr, *dinv = nullptr, *xs = nullptr;   // [n*nv]
    // [n][2*nv] records {z of the variants, p of the variants}; with the
    // preconditioner every iteration reads and rewrites w0 in place (the DOF
    // pass its p half, the preconditioner its z half), without it (Jacobi)
    // the iterations alternate between w0 and w1
    float *w0 = nullptr, *w1 = nullptr;
    double *partE = nullptr, *partB = nullptr, *partbb = nullptr, *rtol_dev = nullptr;
    sn::VcgState *st = nullptr, *st_host = nullptr;
    sn_ams *ams = nullptr;
    int ams_lanes = 8;
    std::vector<float *> vb, vx;       // per level [n_l*3*nv], [row][component][variant]
    float *vy = nullptr;               // [n_nodes*3*nv] level-0 result, 0 at fixed nodes
    hipGraphExec_t graph = nullptr, graph_short = nullptr, graph_tail = nullptr,
                   graph_quarter = nullptr;
    const float *g_Mr = nullptr;
    const unsigned char *g_mask = nullptr;
    hipStream_t g_stream = nullptr;
};

namespace sn {

static hipError_t vcg_drop_graphs(sn_vcg *C) {
    hipError_t e = hipSuccess;
    for (hipGraphExec_t *g : {&C->graph, &C->graph_short, &C->graph_tail, &C->graph_quarter}) {
        if (*g) {
            const hipError_t f = hipGraphExecDestroy(*g);
            if (e == hipSuccess) e = f;
            *g = nullptr;
        }
    }
    return e;
}

static void vcg_free_ams_vectors(sn_vcg *C) {
    for (float *p : C->vb)
        if (p) (void)hipFree(p);
    for (float *p : C->vx)
        if (p) (void)hipFree(p);
    if (C->vy) (void)hipFree(C->vy);
    C->vb.clear();
    C->vx.clear();
    C->vy = nullptr;
}

inline int vams_grid(long long threads, int block, int xcd) {
    const int g = grid_for(threads, block);
    return xcd ? (g + 7) / 8 * 8 : g;
}

// Lanes per aggregate of the down kernels and per row of the up kernels:
// levels of a few thousand rows or fewer are bound by the dependent
// index -> value chains, so they take more lanes per item
constexpr int kVSmallLevel = 8192;

template <int NV, int LPR>
static void vcg_apply_t(sn_vcg *C, const unsigned char *mask, float *w, hipStream_t s) {
    sn_ams *A = C->ams;
    const int L = (int)A->lv.size();
    const int xcd = C->xcd;
    AmsLevel &l0 = A->lv[0];
    hipLaunchKernelGGL((vams_restrict_kernel<NV, LPR>),
                       dim3(vams_grid((long long)A->n_free * LPR, kVAmsBlock, xcd)),
                       dim3(kVAmsBlock), 0, s, A->n_free, A->inc_ptr, A->inc, A->xyz, C->r,
                       l0.dinv, A->omega, C->vb[0], C->vx[0], C->st, xcd);
    for (int l = 0; l + 1 < L; ++l) {
        AmsLevel &f = A->lv[l], &c = A->lv[l + 1];
        float *xc = (l + 2 < L) ? C->vx[l + 1] : (float *)nullptr;
        if (C->rt && c.rt_vals) {
            if (c.n <= kVSmallLevel)
                hipLaunchKernelGGL((vams_down_rt_kernel<NV, kWave, true>),
                                   dim3(vams_grid((long long)c.n * kWave, kVAmsBlock, xcd)),
                                   dim3(kVAmsBlock), 0, s, c.n, c.rt_ptr, c.rt_cols, c.rt_vals,
                                   C->vb[l], c.dinv, A->omega, C->vb[l + 1], xc, C->st, xcd);
            else
                hipLaunchKernelGGL((vams_down_rt_kernel<NV, 16, false>),
                                   dim3(vams_grid((long long)c.n * 16, kVAmsBlock, xcd)),
                                   dim3(kVAmsBlock), 0, s, c.n, c.rt_ptr, c.rt_cols, c.rt_vals,
                                   C->vb[l], c.dinv, A->omega, C->vb[l + 1], xc, C->st, xcd);
        } else if (c.n <= kVSmallLevel)
            hipLaunchKernelGGL((vams_down_kernel<NV, kWave, true>),
                               dim3(vams_grid((long long)c.n * kWave, kVAmsBlock, xcd)),
                               dim3(kVAmsBlock), 0, s, c.n, c.agg_ptr, c.agg_kp, f.indices,
                               f.vals, C->vb[l], C->vx[l], c.dinv, A->omega, C->vb[l + 1], xc,
                               C->st, xcd);
        else
            hipLaunchKernelGGL((vams_down_kernel<NV, 16, false>),
                               dim3(vams_grid((long long)c.n * 16, kVAmsBlock, xcd)),
                               dim3(kVAmsBlock), 0, s, c.n, c.agg_ptr, c.agg_kp, f.indices,
                               f.vals, C->vb[l], C->vx[l], c.dinv, A->omega, C->vb[l + 1], xc,
                               C->st, xcd);
    }
    AmsLevel &lc = A->lv[L - 1];
    hipLaunchKernelGGL((vams_coarse_kernel<NV>), dim3(grid_for(lc.n, kVAmsBlock / kWave)),
                       dim3(kVAmsBlock), 0, s, lc.n, A->cinv, C->vb[L - 1], lc.agg_ptr,
                       L > 1 ? C->vx[L - 2] : (float *)nullptr, C->vy, C->st);
    for (int l = L - 2; l >= 0; --l) {
        AmsLevel &f = A->lv[l];
        float *x_f = l > 0 ? C->vx[l - 1] : (float *)nullptr;
        if (f.n <= kVSmallLevel)
            hipLaunchKernelGGL((vams_up_kernel<NV, 16, true>),
                               dim3(vams_grid((long long)f.n * 16, kVAmsBlock, xcd)),
                               dim3(kVAmsBlock), 0, s, f.n, f.indptr, f.indices, f.vals, f.dinv,
                               C->vb[l], C->vx[l], A->omega, f.agg_ptr, x_f, C->vy, C->st, xcd);
        else
            hipLaunchKernelGGL((vams_up_kernel<NV, LPR, false>),
                               dim3(vams_grid((long long)f.n * LPR, kVAmsBlock, xcd)),
                               dim3(kVAmsBlock), 0, s, f.n, f.indptr, f.indices, f.vals, f.dinv,
                               C->vb[l], C->vx[l], A->omega, f.agg_ptr, x_f, C->vy, C->st, xcd);
    }
    hipLaunchKernelGGL((vams_prolong_kernel<NV>), dim3(C->grid_vec), dim3(kVVecBlock), 0, s,
                       A->n_edges, mask, C->r, C->dinv, A->edge_nodes, A->xyz, C->vy, A->weight,
                       w, C->partB, C->st, xcd);
}

template <int NV>
static void vcg_apply(sn_vcg *C, const unsigned char *mask, float *w, hipStream_t s) {
    if (C->ams_lanes >= 16) vcg_apply_t<NV, 16>(C, mask, w, s);
    else if (C->ams_lanes >= 8) vcg_apply_t<NV, 8>(C, mask, w, s);
    else vcg_apply_t<NV, 4>(C, mask, w, s);
}

template <int NV, bool DM>
static void vcg_iteration_t(sn_vcg *C, const float *Mr, const unsigned char *mask, int parity,
                            hipStream_t s) {
    const sn_pattern *P = C->pat;
    // with the preconditioner the records of w0 are rewritten in place
    float *wold = (parity && !C->ams) ? C->w1 : C->w0;
    float *wnew = (parity || C->ams) ? C->w0 : C->w1;
    hipLaunchKernelGGL((vcg_elem_kernel<NV, DM>), dim3(C->grid_el), dim3(kVElBlock), 0, s,
                       C->n_el, C->el_dofs, C->ypos, Mr, wold, C->yloc, C->partB, C->grid_vec,
                       C->partE, C->st, parity, C->xcd);
    if (C->ams) {
        hipLaunchKernelGGL((vcg_edge_kernel<NV, false, DM>), dim3(C->grid_vec), dim3(kVVecBlock),
                           0, s, P->n, P->dof_ptr, P->dof_contrib, C->yloc, mask, C->dinv, wold,
                           nullptr, C->r, C->xs, C->partE, C->grid_el, C->partB, C->st, parity,
                           C->xcd);
        vcg_apply<NV>(C, mask, wnew, s);
    } else {
        hipLaunchKernelGGL((vcg_edge_kernel<NV, true, DM>), dim3(C->grid_vec), dim3(kVVecBlock),
                           0, s, P->n, P->dof_ptr, P->dof_contrib, C->yloc, mask, C->dinv, wold,
                           wnew, C->r, C->xs, C->partE, C->grid_el, C->partB, C->st, parity,
                           C->xcd);
    }
}

template <int NV>
static void vcg_iteration(sn_vcg *C, const float *Mr, const unsigned char *mask, int parity,
                          hipStream_t s) {
    if (C->dm) vcg_iteration_t<NV, true>(C, Mr, mask, parity, s);
    else vcg_iteration_t<NV, false>(C, Mr, mask, parity, s);
}

template <int NV>
static void vcg_init(sn_vcg *C, const float *Mr, const unsigned char *mask, const double *b,
                     double atol, int maxiter, hipStream_t s) {
    const sn_pattern *P = C->pat;
    hipLaunchKernelGGL((vcg_init_kernel<NV>), dim3(C->grid_vec), dim3(kVVecBlock), 0, s, P->n,
                       C->n_el, P->dof_ptr, P->dof_contrib, Mr, mask, b, C->r, C->w0, C->w1,
                       C->dinv, C->xs, C->partB, C->partbb);
    hipLaunchKernelGGL((vcg_init_state_kernel<NV>), dim3(1), dim3(kVVecBlock), 0, s, C->partbb,
                       C->grid_vec, C->st, C->rtol_dev, atol, maxiter);
    if (C->ams) vcg_apply<NV>(C, mask, C->w0, s);
}

template <int NV>
static void vcg_chunk(sn_vcg *C, const float *Mr, const unsigned char *mask, int count,
                      hipStream_t s) {
    // an even number of iterations keeps the pair buffers and rho slots in step
    for (int k = 0; k < count; ++k) vcg_iteration<NV>(C, Mr, mask, k & 1, s);
}

#define SN_VCG_DISPATCH(C, FN, ...)                      \
    do {                                                 \
        switch ((C)->nv) {                               \
            case 1: FN<1>(C, __VA_ARGS__); break;        \
            case 2: FN<2>(C, __VA_ARGS__); break;        \
            case 4: FN<4>(C, __VA_ARGS__); break;        \
            default: FN<8>(C, __VA_ARGS__); break;       \
        }                                                \
    } while (0)

static int vcg_upload_ints(int **dst, const int *h, size_t count, hipStream_t s) {
    SN_HIP(hipMalloc(dst, (count > 0 ? count : 1) * sizeof(int)));
    if (count > 0) SN_HIP(hipMemcpyAsync(*dst, h, count * sizeof(int), hipMemcpyHostToDevice, s));
    return SN_OK;
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_vcg_create(sn_vcg **out, const sn_pattern *P, const int *gids, const int *orient, int nvar,
                  int chunk, void *stream) {
    SN_REQUIRE(out && P && gids && orient, "sn_vcg_create: null argument");
    SN_REQUIRE(nvar == 1 || nvar == 2 || nvar == 4 || nvar == 8,
               "sn_vcg_create: nvar must be 1, 2, 4 or 8 (%d)", nvar);
    SN_REQUIRE(P->dof_ptr && P->dof_contrib && P->num_el > 0,
               "sn_vcg_create: pattern without element contribution lists");
    SN_REQUIRE((long long)P->num_el * 6 * nvar < (1LL << 31) && (long long)P->n * nvar < (1LL << 31),
               "sn_vcg_create: %d elements x %d variants exceed the index range", P->num_el, nvar);
    hipStream_t s = as_stream(stream);
    if (chunk <= 0) chunk = 32;
    if (chunk & 1) chunk += 1;
    sn_vcg *C = new sn_vcg();
    C->pat = P;
    C->nv = nvar;
    C->n = P->n;
    C->n_el = P->num_el;
    C->chunk = chunk;
    C->chunk_short = std::max(2, (chunk / 4) & ~1);
    const int cus = num_cus();
    // element pass: a multiple of the 8 XCDs, at most 8 blocks per CU
    int ge = grid_for((long long)C->n_el * nvar, kVElBlock, 8 * cus);
    if (ge >= 8) ge = ge / 8 * 8;
    C->grid_el = ge;
    int gv = grid_for(C->n, kVVecBlock, cus);
    if (gv >= 8) gv = gv / 8 * 8;
    C->grid_vec = gv;
    const size_t nn = (size_t)C->n * nvar;
    SN_HIP(hipMalloc(&C->el_dofs, (size_t)6 * C->n_el * sizeof(int)));
    SN_HIP(hipMalloc(&C->yloc, (size_t)6 * C->n_el * nvar * sizeof(float)));
    SN_HIP(hipMalloc(&C->r, nn * sizeof(float)));
    SN_HIP(hipMalloc(&C->dinv, nn * sizeof(float)));
    SN_HIP(hipMalloc(&C->xs, nn * sizeof(float)));
    SN_HIP(hipMalloc(&C->w0, 2 * nn * sizeof(float)));
    SN_HIP(hipMalloc(&C->w1, 2 * nn * sizeof(float)));
    SN_HIP(hipMalloc(&C->partE, (size_t)C->grid_el * nvar * sizeof(double)));
    SN_HIP(hipMalloc(&C->partB, (size_t)2 * C->grid_vec * nvar * sizeof(double)));
    SN_HIP(hipMalloc(&C->partbb, (size_t)C->grid_vec * nvar * sizeof(double)));
    SN_HIP(hipMalloc(&C->rtol_dev, (size_t)nvar * sizeof(double)));
    SN_HIP(hipMalloc(&C->st, (size_t)nvar * sizeof(VcgState)));
    SN_HIP(hipHostMalloc((void **)&C->st_host, (size_t)nvar * sizeof(VcgState),
                         hipHostMallocDefault));
    std::memset(C->st_host, 0, (size_t)nvar * sizeof(VcgState));
    SN_HIP(hipMemcpyAsync(C->st, C->st_host, (size_t)nvar * sizeof(VcgState),
                          hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(vcg_el_dofs_kernel, dim3(grid_for(C->n_el, 256)), dim3(256), 0, s, C->n_el,
                       gids, orient, C->el_dofs);
    SN_HIP(hipMalloc(&C->ypos, (size_t)6 * C->n_el * sizeof(int)));
    hipLaunchKernelGGL(vcg_slot_pos_kernel, dim3(grid_for(6LL * C->n_el, 256)), dim3(256), 0, s,
                       6 * C->n_el, P->dof_contrib, C->ypos);
    SN_CHECK_LAUNCH();
    SN_HIP(hipStreamSynchronize(s));
    *out = C;
    return SN_OK;
}

int sn_vcg_set_ams(sn_vcg *C, sn_ams *A, void *stream) {
    SN_REQUIRE(C, "sn_vcg_set_ams: null workspace");
    SN_REQUIRE(!A || A->n_edges == C->n, "sn_vcg_set_ams: %d edges, the system has %d rows",
               A ? A->n_edges : 0, C->n);
    SN_REQUIRE(!A || !A->lv.empty(), "sn_vcg_set_ams: preconditioner without levels");
    hipStream_t s = as_stream(stream);
    SN_HIP(vcg_drop_graphs(C));
    SN_HIP(hipStreamSynchronize(s));
    vcg_free_ams_vectors(C);
    C->ams = A;
    if (!A) return SN_OK;
    C->ams_lanes = A->lanes;
    const size_t nv = (size_t)C->nv;
    for (const AmsLevel &l : A->lv) {
        float *b = nullptr, *x = nullptr;
        SN_HIP(hipMalloc(&b, (size_t)l.n * 3 * nv * sizeof(float)));
        SN_HIP(hipMalloc(&x, (size_t)l.n * 3 * nv * sizeof(float)));
        C->vb.push_back(b);
        C->vx.push_back(x);
    }
    // the level-0 result is read at both end points of every edge; fixed
    // nodes (ids >= n_free) keep y = 0 for good
    SN_HIP(hipMalloc(&C->vy, (size_t)A->n_nodes * 3 * nv * sizeof(float)));
    SN_HIP(hipMemsetAsync(C->vy, 0, (size_t)A->n_nodes * 3 * nv * sizeof(float), s));
    SN_HIP(hipStreamSynchronize(s));
    return SN_OK;
}

int sn_vcg_set_schedule(sn_vcg *C, int flags) {
    SN_REQUIRE(C && flags >= 0 && flags <= 7, "sn_vcg_set_schedule: flags in [0, 7]");
    SN_HIP(vcg_drop_graphs(C));
    C->xcd = (flags & 1) ? 1 : 0;
    C->dm = (flags & 2) == 0;
    C->rt = (flags & 4) == 0;
    return SN_OK;
}

int sn_vcg_solve(sn_vcg *C, const float *Mr, const unsigned char *mask, const double *b,
                 double *x, const double *rtol_host, double atol, int maxiter, int *iters_host,
                 double *rnorm_host, int *status_host, void *stream) {
    SN_REQUIRE(C && Mr && b && x && rtol_host, "sn_vcg_solve: null argument");
    SN_REQUIRE(!C->ams || C->ams->ready, "sn_vcg_solve: preconditioner has no values yet");
    SN_REQUIRE(!C->ams || (int)C->vb.size() == (int)C->ams->lv.size(),
               "sn_vcg_solve: preconditioner changed its levels since sn_vcg_set_ams");
    hipStream_t s = as_stream(stream);
    const int nv = C->nv;
    if (maxiter < 0) maxiter = (int)std::min<long long>(10LL * C->n, INT_MAX);
    SN_HIP(hipMemcpyAsync(C->rtol_dev, rtol_host, nv * sizeof(double), hipMemcpyHostToDevice, s));
    SN_VCG_DISPATCH(C, vcg_init, Mr, mask, b, atol, maxiter, s);
    SN_CHECK_LAUNCH();
    if (!C->graph || C->g_Mr != Mr || C->g_mask != mask || C->g_stream != s) {
        SN_HIP(vcg_drop_graphs(C));
        hipStream_t cap = s;
        bool own = false;
        if (cap == nullptr) {
            SN_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
            own = true;
        }
        hipGraphExec_t *dst[4] = {&C->graph, &C->graph_short, &C->graph_quarter, &C->graph_tail};
        const int len[4] = {C->chunk, C->chunk_short, 4, C->chunk_tail};
        for (int which = 0; which < 4; ++which) {
            hipGraph_t g;
            SN_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
            SN_VCG_DISPATCH(C, vcg_chunk, Mr, mask, len[which], cap);
            SN_HIP(hipStreamEndCapture(cap, &g));
            SN_HIP(hipGraphInstantiate(dst[which], g, nullptr, nullptr, 0));
            SN_HIP(hipGraphDestroy(g));
        }
        if (own) SN_HIP(hipStreamDestroy(cap));
        C->g_Mr = Mr;
        C->g_mask = mask;
        C->g_stream = s;
    }
    const size_t sb = nv * sizeof(VcgState);
    SN_HIP(hipMemcpyAsync(C->st_host, C->st, sb, hipMemcpyDeviceToHost, s));
    SN_HIP(hipStreamSynchronize(s));
    std::vector<double> r_start(nv), r_last(nv);
    std::vector<int> it_last(nv, 0);
    for (int v = 0; v < nv; ++v) r_start[v] = r_last[v] = sqrt(C->st_host[v].rr);
    // Replays: the first a short one; then a full chunk only while the
    // running systems' estimate of the iterations still needed (at their
    // rates over the last replay) is at least two full chunks, short chunks (8) down to
    // 10 still needed, 4 down to 5, two-iteration ones after that.  An
    // iteration replayed after every system stopped still launches its
    // kernels (they exit at entry, about 60 us at 1M tets x 8 systems), a
    // poll costs about 40 us.
    for (;;) {
        bool running = false, first = false;
        double need_max = 0.0;
        for (int v = 0; v < nv; ++v) {
            const VcgState &h = C->st_host[v];
            if (h.done != 0) continue;
            running = true;
            if (h.it <= 0) {
                first = true;
                continue;
            }
            // the rate over the last replay, else the mean rate since the start
            const double r = sqrt(h.rr), r0 = r_start[v], rl = r_last[v];
            double rate = 0.0;
            if (h.it > it_last[v] && r > 0.0 && rl > 0.0 && r < rl)
                rate = log(r / rl) / (h.it - it_last[v]);
            else if (r > 0.0 && r0 > 0.0 && r < r0)
                rate = log(r / r0) / h.it;
            r_last[v] = r;
            it_last[v] = h.it;
            // no rate (the residual has not gone below its start): a long
            // solve once it has run two full chunks, else poll soon
            const double need = rate < 0.0 ? log(h.tol / r) / rate
                                           : (h.it >= 2 * C->chunk ? 1e30 : C->chunk_short);
            need_max = std::max(need_max, need);
        }
        if (!running) break;
        hipGraphExec_t g = C->graph_short;
        if (!first) {
            if (need_max >= 2.0 * C->chunk) g = C->graph;
            else if (need_max < 5.0) g = C->graph_tail;
            else if (need_max < 10.0) g = C->graph_quarter;
        }
        SN_HIP(hipGraphLaunch(g, s));
        SN_HIP(hipMemcpyAsync(C->st_host, C->st, sb, hipMemcpyDeviceToHost, s));
        SN_HIP(hipStreamSynchronize(s));
    }
    const long long n_nv = (long long)C->n * nv;
    hipLaunchKernelGGL(vcg_finalize_kernel, dim3(grid_for(n_nv, 256)), dim3(256), 0, s, n_nv, nv,
                       mask, C->xs, x);
    SN_CHECK_LAUNCH();
    int worst = SN_OK;
    for (int v = 0; v < nv; ++v) {
        const VcgState &h = C->st_host[v];
        if (iters_host) iters_host[v] = h.it;
        if (rnorm_host) rnorm_host[v] = sqrt(h.rr);
        const int code = h.done == 2 ? SN_ERR_NOT_CONVERGED
                                     : (h.done == 3 ? SN_ERR_BREAKDOWN : SN_OK);
        if (status_host) status_host[v] = code;
        if (code != SN_OK && worst == SN_OK) {
            worst = code;
            set_error("sn_vcg_solve: variant %d %s at iteration %d (||r|| = %.3e, tol = %.3e)", v,
                      code == SN_ERR_BREAKDOWN ? "broke down" : "reached its iteration cap",
                      h.it, sqrt(h.rr), h.tol);
        }
    }
    return worst;
}

int sn_vcg_pack(sn_vcg *C, const float *const *Mr_src, const double *const *R_src,
                float *Mr, double *b, void *stream) {
    SN_REQUIRE(C && Mr_src && R_src && Mr && b, "sn_vcg_pack: null argument");
    VcgPackSrc src{};
    for (int v = 0; v < C->nv; ++v) {
        SN_REQUIRE(Mr_src[v], "sn_vcg_pack: every system needs its element operators");
        src.mr[v] = Mr_src[v];
        src.r[v] = R_src[v];
    }
    const long long n_slots = 6LL * C->n_el;
    const long long cover = std::max<long long>(n_slots, C->n);
    const dim3 g((unsigned)((cover + 255) / 256)), blk(256);
    hipStream_t s = as_stream(stream);
    switch (C->nv) {
    case 1: hipLaunchKernelGGL((vcg_pack_kernel<1>), g, blk, 0, s, n_slots, (long long)C->n, src, Mr, b); break;
    case 2: hipLaunchKernelGGL((vcg_pack_kernel<2>), g, blk, 0, s, n_slots, (long long)C->n, src, Mr, b); break;
    case 4: hipLaunchKernelGGL((vcg_pack_kernel<4>), g, blk, 0, s, n_slots, (long long)C->n, src, Mr, b); break;
    default: hipLaunchKernelGGL((vcg_pack_kernel<8>), g, blk, 0, s, n_slots, (long long)C->n, src, Mr, b); break;
    }
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_vcg_time_pair(sn_vcg *C, const float *Mr, const unsigned char *mask, int reps,
                     double *avg_ms, void *stream) {
    SN_REQUIRE(C && Mr && avg_ms && reps > 0, "sn_vcg_time_pair: bad argument");
    hipStream_t s = as_stream(stream);
    const int nv = C->nv;
    // every variant running: rho = 1 now and before (beta = 1), ||r||^2 = 1
    // above a zero tolerance, z = 1 on the free rows, p_old = 0
    std::vector<VcgState> h(nv);
    for (auto &e : h) {
        std::memset(&e, 0, sizeof(e));
        e.it = 1;
        e.rho[0] = e.rho[1] = 1.0;
        e.tol = 0.0;
        e.maxiter = INT_MAX;
    }
    SN_HIP(hipMemcpyAsync(C->st, h.data(), nv * sizeof(VcgState), hipMemcpyHostToDevice, s));
    const long long n_nv = (long long)C->n * nv;
    const long long cover = std::max(n_nv, 2LL * C->grid_vec * nv);
    hipLaunchKernelGGL(vcg_fill_timing_kernel, dim3(grid_for(cover, 256)), dim3(256), 0, s, n_nv,
                       nv, C->grid_vec, mask, C->partB, C->w1);
    SN_CHECK_LAUNCH();
    hipEvent_t e0, e1;
    SN_HIP(hipEventCreate(&e0));
    SN_HIP(hipEventCreate(&e1));
    SN_HIP(hipEventRecord(e0, s));
    // parity 1 reads the records of w1 (z = 1, p = 0 at first) and rho[0],
    // writes rho[1] and p in place: beta = 1, so p grows by z per repetition
    // (values only; the kernels' work is the same in every repetition)
    const sn_pattern *P = C->pat;
    for (int r = 0; r < reps; ++r) {
#define SN_VCG_PAIR(NV)                                                                        \
    do {                                                                                       \
        if (C->dm) {                                                                            \
            hipLaunchKernelGGL((vcg_elem_kernel<NV, true>), dim3(C->grid_el), dim3(kVElBlock), 0, \
                               s, C->n_el, C->el_dofs, C->ypos, Mr, C->w1, C->yloc, C->partB,  \
                               C->grid_vec, C->partE, C->st, 1, C->xcd);                       \
            hipLaunchKernelGGL((vcg_edge_kernel<NV, false, true>), dim3(C->grid_vec),          \
                               dim3(kVVecBlock), 0, s, P->n, P->dof_ptr, P->dof_contrib,       \
                               C->yloc, mask, C->dinv, C->w1, nullptr, C->r, C->xs,            \
                               C->partE, C->grid_el, C->partB, C->st, 1, C->xcd);              \
        } else {                                                                               \
            hipLaunchKernelGGL((vcg_elem_kernel<NV, false>), dim3(C->grid_el), dim3(kVElBlock),  \
                               0, s, C->n_el, C->el_dofs, C->ypos, Mr, C->w1, C->yloc,         \
                               C->partB, C->grid_vec, C->partE, C->st, 1, C->xcd);             \
            hipLaunchKernelGGL((vcg_edge_kernel<NV, false, false>), dim3(C->grid_vec),         \
                               dim3(kVVecBlock), 0, s, P->n, P->dof_ptr, P->dof_contrib,       \
                               C->yloc, mask, C->dinv, C->w1, nullptr, C->r, C->xs,            \
                               C->partE, C->grid_el, C->partB, C->st, 1, C->xcd);              \
        }                                                                                      \
    } while (0)
        switch (nv) {
            case 1: SN_VCG_PAIR(1); break;
            case 2: SN_VCG_PAIR(2); break;
            case 4: SN_VCG_PAIR(4); break;
            default: SN_VCG_PAIR(8); break;
        }
#undef SN_VCG_PAIR
    }
    SN_HIP(hipEventRecord(e1, s));
    SN_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    SN_HIP(hipEventElapsedTime(&ms, e0, e1));
    SN_HIP(hipEventDestroy(e0));
    SN_HIP(hipEventDestroy(e1));
    SN_HIP(hipMemcpyAsync(h.data(), C->st, nv * sizeof(VcgState), hipMemcpyDeviceToHost, s));
    SN_HIP(hipStreamSynchronize(s));
    for (const auto &e : h)
        SN_REQUIRE(e.done == 0, "sn_vcg_time_pair: the timed launches stopped early (done=%d)",
                   e.done);
    *avg_ms = (double)ms / reps;
    return SN_OK;
}

int sn_vcg_set_grids(sn_vcg *C, int grid_el, int vec_per_cu, void *stream) {
    SN_REQUIRE(C, "sn_vcg_set_grids: null workspace");
    SN_REQUIRE(grid_el >= 0 && vec_per_cu >= 1 && vec_per_cu <= 8,
               "sn_vcg_set_grids: grid_el >= 0 (0: default), vec_per_cu in [1, 8]");
    hipStream_t s = as_stream(stream);
    SN_HIP(vcg_drop_graphs(C));
    // no kernel of this workspace may still read the partials replaced here
    SN_HIP(hipStreamSynchronize(s));
    const int cus = num_cus();
    int ge = grid_el > 0 ? grid_el : grid_for((long long)C->n_el * C->nv, kVElBlock, 8 * cus);
    if (ge >= 8) ge = ge / 8 * 8;
    int gv = grid_for(C->n, kVVecBlock, vec_per_cu * cus);
    if (gv >= 8) gv = gv / 8 * 8;
    for (void *p : {(void *)C->partE, (void *)C->partB, (void *)C->partbb})
        if (p) SN_HIP(hipFree(p));
    C->partE = C->partB = C->partbb = nullptr;
    C->grid_el = ge;
    C->grid_vec = gv;
    SN_HIP(hipMalloc(&C->partE, (size_t)C->grid_el * C->nv * sizeof(double)));
    SN_HIP(hipMalloc(&C->partB, (size_t)2 * C->grid_vec * C->nv * sizeof(double)));
    SN_HIP(hipMalloc(&C->partbb, (size_t)C->grid_vec * C->nv * sizeof(double)));
    return SN_OK;
}

int sn_vcg_info(const sn_vcg *C, int *nvar, int *grid_el, int *grid_vec) {
    SN_REQUIRE(C, "sn_vcg_info: null workspace");
    if (nvar) *nvar = C->nv;
    if (grid_el) *grid_el = C->grid_el;
    if (grid_vec) *grid_vec = C->grid_vec;
    return SN_OK;
}

void sn_vcg_destroy(sn_vcg *C) {
    if (!C) return;
    (void)vcg_drop_graphs(C);
    vcg_free_ams_vectors(C);
    void *bufs[] = {C->el_dofs, C->ypos, C->yloc, C->r, C->dinv, C->xs, C->w0, C->w1,
                    C->partE, C->partB, C->partbb, C->rtol_dev, C->st};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (C->st_host) (void)hipHostFree(C->st_host);
    delete C;
}

}  // extern "C"
