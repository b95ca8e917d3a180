// Jacobi-preconditioned CG on the CSR Jacobian, scipy 1.15.3 `cg` semantics
// (scipy/sparse/linalg/_isolve/iterative.py: r = b - A x0; loop: stop when
// ||r|| < max(atol, rtol ||b||); z = M r; rho = r.z; p = z + (rho/rho_prev) p;
// q = A p; alpha = rho/(p.q); x += alpha p; r -= alpha q), as driven by
// snoopy/red_mvp_solver.py:195-205 with Dirichlet rows/columns removed by a
// mask instead of a fancy-indexed copy of the matrix.
//
// Two kernels per iteration:
//   A (row-parallel): reduce B's partials -> rho, ||r||^2; stop test;
//      p_new = z + beta p_old formed on the fly for every column read;
//      q = A p_new; per-block partial of p.q
//   B (vector): reduce A's partials -> p.q; alpha = rho / p.q;
//      x += alpha p_new, r -= alpha q, z = dinv r; per-block partials of r.z, r.r
// Every block of a kernel re-reduces the previous kernel's per-block partials
// in the same fixed order, so all blocks hold bit-identical scalars without
// any cross-block synchronisation inside a kernel (no atomics, no device-scope
// fences: on a multi-XCD part those would write back L2 per block).  Kernel
// boundaries order the partial writes before their reads.  The host polls the
// stop flag once per chunk of iterations; each chunk is captured once as a
// hipGraph and replayed.  Two chunk graphs are kept, the full chunk and a
// quarter of it: a solve starts with the short one, and later replays take
// the short one whenever the residual's geometric rate so far puts the stop
// within three short chunks, so a solve that stops early in a chunk does not
// replay most of a full one as launches that exit at entry.
#include <algorithm>
#include <climits>
#include <initializer_list>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <cstring>
#include <vector>

#include <cstddef>

#include "sn_ams.hpp"
#include "sn_common.hpp"
#include "sn_pattern.hpp"

namespace sn {

struct CgState {
    double rho[2];   // rho of even / odd iterations (written by the segment's first A block)
    double pq, rr, tol, bnorm;
    double beta;     // element operator: this iteration's beta, written by A1 for A2
    int it, maxiter, done, pad;  // done: 0 running, 1 converged, 2 iteration cap, 3 breakdown
};

// Row segments of a batched solve (one segment per independent system of a
// block-diagonal matrix) and the blocks of each grid that serve them.
struct SegMap {
    const int *row0;   // [nseg+1] first row of each segment
    const int *blkA0;  // [nseg+1] first block of each segment in kernel A's grid
    const int *blkB0;  // [nseg+1] first block of each segment in kernel B's / init's grid
    const int *segA;   // [grid A] segment of each block
    const int *segB;   // [grid B] segment of each block
};

constexpr int kRowBlock = 256;   // kernel A (SpMV) block
constexpr int kVecBlock = 1024;  // kernel B and init block

// Fixed-order block sum of partials[k*stride + off], k in [b0, b1).
template <int BLOCK>
__device__ __forceinline__ double reduce_range(const double *__restrict__ partials, int b0, int b1,
                                               int stride, int off, double *lds) {
    double s = 0.0;
    for (int k = b0 + threadIdx.x; k < b1; k += BLOCK) s += partials[k * stride + off];
    return block_sum(s, lds);
}

// z and the search direction p live interleaved, w[2i] = z_i, w[2i+1] = p_i,
// so the SpMV fetches both at a column with one 8- or 16-byte load.  Two
// such arrays alternate: the SpMV of iteration k reads {z, p_old} from one
// and writes p_new into the other, where the update (or the preconditioner)
// then puts the next z.
template <typename V> struct Pair;
template <> struct Pair<float> { using T = float2; };
template <> struct Pair<double> { using T = double2; };

// ---------------------------------------------------------------- init
// r = b - A_c x0 (x0 masked; the f32 solve starts from 0), dinv, z, p_old = 0,
// per-block partials: (r.z, r.r) in the layout kernel B uses, b.b separately.
template <typename V, typename M, int LPR>
__global__ __launch_bounds__(kVecBlock) void cg_init_kernel(
    SegMap sm, const int *__restrict__ indptr, const int *__restrict__ indices,
    const int *__restrict__ diag, const M *__restrict__ vals, const unsigned char *__restrict__ mask,
    const double *__restrict__ b, const double *__restrict__ x0, V *__restrict__ r,
    V *__restrict__ w, V *__restrict__ w_other, V *__restrict__ dinv, float *__restrict__ xs,
    double *__restrict__ partB, double *__restrict__ partbb) {
    __shared__ double lds[kVecBlock / kWave];
    const int seg = sm.segB[blockIdx.x];
    const int lb = blockIdx.x - sm.blkB0[seg], nbl = sm.blkB0[seg + 1] - sm.blkB0[seg];
    const int r0 = sm.row0[seg], r1 = sm.row0[seg + 1];
    const int lane = threadIdx.x % LPR;
    const int rows_per_block = kVecBlock / LPR;
    double bb = 0.0, rr = 0.0, rz = 0.0;
    for (long long row = r0 + (long long)lb * rows_per_block + threadIdx.x / LPR; row < r1;
         row += (long long)nbl * rows_per_block) {
        const int i = (int)row;
        const bool free_row = mask ? mask[i] != 0 : true;
        double acc = 0.0;
        if (free_row && x0) {
            for (int k = indptr[i] + lane; k < indptr[i + 1]; k += LPR) {
                const int j = indices[k];
                const bool fj = mask ? mask[j] != 0 : true;
                if (fj) acc += (double)vals[k] * x0[j];
            }
        }
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
        if (lane == 0) {
            V ri = 0, di = 0;
            if (free_row) {
                const double bi = b[i];
                bb += bi * bi;
                ri = (V)(bi - acc);
                const int d = diag[i];
                di = (d >= 0) ? (V)(1.0 / (double)vals[d]) : (V)0;
            }
            const V zi = di * ri;
            r[i] = ri;
            w[2 * (long long)i] = zi;
            w[2 * (long long)i + 1] = 0;
            // the other array's z of fixed rows stays 0 for the whole solve
            // (only the free rows' z is written by the iterations)
            w_other[2 * (long long)i] = 0;
            w_other[2 * (long long)i + 1] = 0;
            dinv[i] = di;
            if (xs) xs[i] = 0;
            rr += (double)ri * (double)ri;
            rz += (double)ri * (double)zi;
        }
    }
    const double sbb = block_sum(bb, lds);
    const double srz = block_sum(rz, lds);
    const double srr = block_sum(rr, lds);
    if (threadIdx.x == 0) {
        partB[2 * blockIdx.x] = srz;
        partB[2 * blockIdx.x + 1] = srr;
        partbb[blockIdx.x] = sbb;
    }
}

// one block per segment: tol, ||b||, state for iteration 0
__global__ __launch_bounds__(kVecBlock) void cg_init_state_kernel(
    SegMap sm, const double *__restrict__ partbb, const double *__restrict__ partB,
    CgState *__restrict__ st, const double *__restrict__ rtol, double atol, int maxiter) {
    __shared__ double lds[kVecBlock / kWave];
    const int seg = blockIdx.x;
    const double bb = reduce_range<kVecBlock>(partbb, sm.blkB0[seg], sm.blkB0[seg + 1], 1, 0, lds);
    // ||r_0||^2 from the init pass's partials (the host's rate estimate)
    const double rr0 = reduce_range<kVecBlock>(partB, sm.blkB0[seg], sm.blkB0[seg + 1], 2, 1, lds);
    if (threadIdx.x == 0) {
        CgState *S = st + seg;
        const double bnorm = sqrt(bb);
        const double rt = rtol[seg] * bnorm;
        S->bnorm = bnorm;
        S->tol = atol > rt ? atol : rt;  // _get_atol_rtol
        S->rho[0] = S->rho[1] = 0.0;
        S->pq = 0.0;
        S->rr = rr0;
        S->it = 0;
        S->maxiter = maxiter;
        S->done = (bnorm == 0.0) ? 1 : 0;
    }
}

// ---------------------------------------------------------------- kernel A
// Column of nonzero k in row i: int32 indices, or 16-bit offsets j - i when
// every row's columns lie within 32767 of the row (edge numberings of
// structured meshes up to a few million elements), which cuts the index
// stream of the SpMV in half.
__device__ __forceinline__ int col_at(const int *__restrict__ c, int k, int) { return c[k]; }
__device__ __forceinline__ int col_at(const short *__restrict__ c, int k, int i) {
    return i + (int)c[k];
}

// parity = iteration & 1: reads rho[parity ^ 1] (previous), writes rho[parity].
template <typename V, typename M, int LPR, typename I>
__global__ __launch_bounds__(kRowBlock) void cg_spmv_kernel(
    SegMap sm, const int *__restrict__ indptr, const I *__restrict__ indices,
    const M *__restrict__ vals, const unsigned char *__restrict__ mask,
    const V *__restrict__ wold, V *__restrict__ wnew, V *__restrict__ q,
    const double *__restrict__ partB, double *__restrict__ partA, CgState *__restrict__ st,
    int parity, int sched) {
    using V2 = typename Pair<V>::T;
    const V2 *__restrict__ zp = reinterpret_cast<const V2 *>(wold);   // {z, p_old}
    __shared__ double lds[kRowBlock / kWave];
    const int seg = sm.segA[blockIdx.x];
    CgState *S = st + seg;
    // the stop flag is read together with the partials (independent loads)
    // and tested after them
    const int was_done = S->done;
    const int nbl = sm.blkA0[seg + 1] - sm.blkA0[seg];
    int lb = blockIdx.x - sm.blkA0[seg];
    // schedule bit 1: workgroups are dealt round-robin to the 8 XCDs;
    // renumber them so that each XCD takes one contiguous eighth of the
    // blocks (needs a multiple of 8 blocks)
    if ((sched & 2) && (nbl & 7) == 0) lb = (lb & 7) * (nbl >> 3) + (lb >> 3);
    const int b0 = sm.blkB0[seg], b1 = sm.blkB0[seg + 1];
    const double rho = reduce_range<kRowBlock>(partB, b0, b1, 2, 0, lds);
    const double rr = reduce_range<kRowBlock>(partB, b0, b1, 2, 1, lds);
    if (was_done) return;
    const int it = S->it;
    // scipy: the iteration cap ends the loop before the residual test
    const int stop = (it >= S->maxiter) ? 2 : ((sqrt(rr) < S->tol) ? 1 : 0);
    if (stop) {
        if (lb == 0 && threadIdx.x == 0) {
            S->rr = rr;
            S->done = stop;
        }
        return;
    }
    const V beta = (V)(it > 0 ? rho / S->rho[parity ^ 1] : 0.0);
    const int r0 = sm.row0[seg], r1 = sm.row0[seg + 1];
    const int lane = threadIdx.x % LPR;
    const int rows_per_block = kRowBlock / LPR;
    // schedule bit 0: each block takes one contiguous span of rows;
    // otherwise blocks stride over the rows in steps of the grid
    const long long per = (long long)nbl * rows_per_block;
    long long c0, c1, step;
    if (sched & 1) {
        const long long span = (r1 - r0 + per - 1) / per * rows_per_block;
        c0 = r0 + (long long)lb * span;
        c1 = c0 + span < r1 ? c0 + span : r1;
        step = rows_per_block;
    } else {
        c0 = r0 + (long long)lb * rows_per_block;
        c1 = r1;
        step = per;
    }
    double pq = 0.0;
    for (long long row = c0 + threadIdx.x / LPR; row < c1; row += step) {
        const int i = (int)row;
        V acc = 0;
        const int k1 = indptr[i + 1];
        for (int k = indptr[i] + lane; k < k1; k += LPR) {
            const V2 w = zp[col_at(indices, k, i)];
            acc += (V)vals[k] * (w.x + beta * w.y);
        }
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
        if (lane == 0) {
            const bool free_row = mask ? mask[i] != 0 : true;
            const V2 w = zp[i];
            const V pi = free_row ? w.x + beta * w.y : (V)0;
            const V qi = free_row ? acc : (V)0;
            wnew[2 * row + 1] = pi;
            q[i] = qi;
            pq += (double)pi * (double)qi;
        }
    }
    const double s = block_sum(pq, lds);
    if (threadIdx.x == 0) {
        partA[blockIdx.x] = s;
        if (lb == 0) {
            S->rho[parity] = rho;
            S->rr = rr;
        }
    }
}

// ------------------------------------------------- element operator (A1, A2)
// The Jacobian as a sum of element products, A = sum_e P_e^T S_e c^T Mr_e c S_e P_e
// (P_e the element's 6 DOFs, S_e their orientation signs, c the 3x6 integer
// reference curls, Mr_e the symmetric 3x3 of sn_element_newton mode 2).  With
// float32 values this keeps A's gradient null space exact, which the rounded
// CSR values of the assembled matrix do not: on meshes of ~1M elements the
// float32 CG on the rounded CSR matrix stalls near a relative residual of
// 1e-2, this one converges like the float64 CG.
//
// Two kernels per iteration do the work of kernels A and B:
// A1 (thread per element): kernel A's stop test; p = z + beta p_old at the
//    6 DOFs; u = c (S p), the circulations of the three faces at local node 0;
//    m = Mr u; the element products y_j = s_j c_j . m at slot j*n_el + e; and
//    per-block partials of p.q = p^T A p = sum_e u . m (the element energies:
//    no pass over q is needed for it).
// A2 (thread per DOF, kernel B's grid): alpha = rho / p.q from A1's
//    partials; q_i = sum of DOF i's element products in ascending element
//    order (the pattern's DOF contribution lists); p_i = z_i + beta p_old_i;
//    x_i += alpha p_i; r_i -= alpha q_i; per-block partials of r.r (and r.z)
//    in kernel B's layout.  With Jacobi, z_i = dinv_i r_i and the pair
//    {z_i, p_i} goes to the other {z, p} array; with the auxiliary-space
//    preconditioner p_i goes to a plain array and the preconditioner's last
//    kernel writes the pair {z_i, p_i}.  q itself is never stored.
// Measured on the 1M-tet dipole: with the products in planes of slots
// j*n_el + e the DOF pass took 19 us (as a product, before the update moved
// into it); storing an element's 6 products side by side (6e + j) made it
// 35 us, storing m (16 bytes per element) and forming c_j . m there 25 us.
// Work index of this block.  Schedule bit 1: workgroups are dealt
// round-robin to the 8 XCDs, so renumber them to give each XCD one
// contiguous eighth of the grid (needs a multiple of 8 blocks).
__device__ __forceinline__ long long xcd_block(int sched) {
    const int b = blockIdx.x, nb = gridDim.x;
    if ((sched & 2) && (nb & 7) == 0) return (long long)(b & 7) * (nb >> 3) + (b >> 3);
    return b;
}

template <typename V>
__global__ __launch_bounds__(kRowBlock) void cg_elem_kernel(
    int n_el, const int *__restrict__ el_dofs, const float *__restrict__ Mr,
    const V *__restrict__ wold, V *__restrict__ yloc, const double *__restrict__ partB,
    int nbB, double *__restrict__ partE, CgState *__restrict__ st, int parity, int sched) {
    using V2 = typename Pair<V>::T;
    const V2 *__restrict__ zp = reinterpret_cast<const V2 *>(wold);
    __shared__ double lds[kRowBlock / kWave];
    CgState *S = st;
    const int was_done = S->done;
    const long long lb = xcd_block(sched);
    const double rho = reduce_range<kRowBlock>(partB, 0, nbB, 2, 0, lds);
    const double rr = reduce_range<kRowBlock>(partB, 0, nbB, 2, 1, lds);
    if (was_done) return;
    const int it = S->it;
    const int stop = (it >= S->maxiter) ? 2 : ((sqrt(rr) < S->tol) ? 1 : 0);
    if (stop) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            S->rr = rr;
            S->done = stop;
        }
        return;
    }
    const double beta_d = it > 0 ? rho / S->rho[parity ^ 1] : 0.0;
    const V beta = (V)beta_d;
    double pq = 0.0;
    for (long long e = lb * blockDim.x + threadIdx.x; e < n_el;
         e += (long long)gridDim.x * blockDim.x) {
        int d[6];
        V xh[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) d[j] = el_dofs[j * (long long)n_el + e];
        V a[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) a[k] = (V)Mr[k * (long long)n_el + e];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const V2 w = zp[d[j] & 0x7fffffff];
            const V pj = w.x + beta * w.y;
            xh[j] = d[j] < 0 ? -pj : pj;
        }
        const V u0 = xh[1] - xh[2] + xh[5];
        const V u1 = xh[2] - xh[0] - xh[4];
        const V u2 = xh[0] - xh[1] + xh[3];
        const V m0 = a[0] * u0 + a[1] * u1 + a[2] * u2;
        const V m1 = a[1] * u0 + a[3] * u1 + a[4] * u2;
        const V m2 = a[2] * u0 + a[4] * u1 + a[5] * u2;
        pq += (double)u0 * (double)m0 + (double)u1 * (double)m1 + (double)u2 * (double)m2;
        const V yh[6] = {m2 - m1, m0 - m2, m1 - m0, m2, -m1, m0};
#pragma unroll
        for (int j = 0; j < 6; ++j) yloc[j * (long long)n_el + e] = d[j] < 0 ? -yh[j] : yh[j];
    }
    const double s = block_sum(pq, lds);
    if (threadIdx.x == 0) {
        partE[blockIdx.x] = s;
        if (blockIdx.x == 0) {
            S->rho[parity] = rho;
            S->rr = rr;
            S->beta = beta_d;
        }
    }
}

template <typename V, typename X, bool JAC>
__global__ __launch_bounds__(kVecBlock) void cg_edge_kernel(
    int n, const int *__restrict__ dof_ptr, const int *__restrict__ dof_contrib,
    const V *__restrict__ yloc, const unsigned char *__restrict__ mask,
    const V *__restrict__ dinv, const V *__restrict__ wold, V *__restrict__ wnew,
    V *__restrict__ pout, V *__restrict__ r, X *__restrict__ x,
    const double *__restrict__ partE, int nbE, double *__restrict__ partB,
    CgState *__restrict__ st, int parity, int sched) {
    using V2 = typename Pair<V>::T;
    const V2 *__restrict__ zp = reinterpret_cast<const V2 *>(wold);
    __shared__ double lds[kVecBlock / kWave];
    CgState *S = st;
    const int was_done = S->done;   // set by A1 when the solve stops here
    const long long lb = xcd_block(sched);
    const double pq = reduce_range<kVecBlock>(partE, 0, nbE, 1, 0, lds);
    if (was_done) return;
    if (!(pq == pq) || pq == 0.0 || isinf(pq)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            S->pq = pq;
            S->done = 3;
        }
        return;
    }
    // rho and beta as A1 formed them.  Not S->it: block 0 advances it at its
    // end, and a block that starts later (other streams can hold the CUs)
    // would read the advanced count
    const double rho = S->rho[parity];
    const V beta = (V)S->beta;
    const V alpha = (V)(rho / pq);
    double rz = 0.0, rr = 0.0;
    // Two DOFs per trip (i and i + step, the grid-stride order), and the
    // first kHead list entries of both, then their products, loaded at once
    // (predicated; a missing entry adds 0, which leaves the sum unchanged):
    // the DOF range -> list -> product chain is this kernel's latency.  The
    // sum of each DOF keeps the list order.
    constexpr int kHead = 8;
    const long long step = (long long)gridDim.x * blockDim.x;
    for (long long i0 = lb * blockDim.x + threadIdx.x; i0 < n; i0 += 2 * step) {
        const bool two = i0 + step < n;
        const long long id[2] = {i0, two ? i0 + step : i0};
        bool fr[2];
        int k0[2], k1[2];
        V2 w[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            fr[u] = (u == 0 || two) && (mask ? mask[id[u]] != 0 : true);
            k0[u] = dof_ptr[id[u]];
            k1[u] = dof_ptr[id[u] + 1];
            w[u] = zp[id[u]];
        }
        int c[2][kHead];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int m = 0; m < kHead; ++m)
                c[u][m] = (fr[u] && k0[u] + m < k1[u]) ? dof_contrib[k0[u] + m] : -1;
        V q[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            V y[kHead];
#pragma unroll
            for (int m = 0; m < kHead; ++m) y[m] = c[u][m] >= 0 ? yloc[c[u][m]] : (V)0;
            q[u] = 0;
#pragma unroll
            for (int m = 0; m < kHead; ++m) q[u] += y[m];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (u == 1 && !two) break;
            const long long i = id[u];
            if (!fr[u]) {
                // p (and, with Jacobi, z) of a fixed row stays 0
                if (JAC) reinterpret_cast<V2 *>(wnew)[i] = V2{(V)0, (V)0};
                else pout[i] = (V)0;
                continue;
            }
            V qi = q[u];
            for (int k = k0[u] + kHead; k < k1[u]; ++k) qi += yloc[dof_contrib[k]];
            const V pi = w[u].x + beta * w[u].y;
            x[i] += (X)(alpha * pi);
            const V ri = r[i] - alpha * qi;
            r[i] = ri;
            rr += (double)ri * (double)ri;
            if (JAC) {
                const V zi = dinv[i] * ri;
                reinterpret_cast<V2 *>(wnew)[i] = V2{zi, pi};
                rz += (double)ri * (double)zi;
            } else {
                // a plain array: the preconditioner writes the whole {z, p}
                // pair of the next iteration from it (a store to every other
                // float writes whole sectors: this kernel wrote 18.8 MB for
                // 14.3 MB of data at 1M tets, now 14.1 MB)
                pout[i] = pi;
            }
        }
    }
    const double s1 = JAC ? block_sum(rz, lds) : 0.0;
    const double s2 = block_sum(rr, lds);
    if (threadIdx.x == 0) {
        if (JAC) partB[2 * blockIdx.x] = s1;
        partB[2 * blockIdx.x + 1] = s2;
        if (blockIdx.x == 0) {
            S->pq = pq;
            S->it = S->it + 1;   // no other block writes it
        }
    }
}

// c_j . Mr c_j for local edge j (Mr upper triangle 00 01 02 11 12 22)
__device__ __forceinline__ double elem_diag(const float *__restrict__ Mr, long long n_el,
                                            long long e, int j) {
    const auto m = [&](int k) { return (double)Mr[k * n_el + e]; };
    switch (j) {
        case 0: return m(3) + m(5) - 2.0 * m(4);
        case 1: return m(0) + m(5) - 2.0 * m(2);
        case 2: return m(0) + m(3) - 2.0 * m(1);
        case 3: return m(5);
        case 4: return m(3);
        default: return m(0);
    }
}

// init of the element operator's solve (one segment, start from 0): r = b,
// the diagonal from the element matrices, dinv, z = dinv r, p_old = 0, and
// the partials of cg_init_kernel
template <typename V>
__global__ __launch_bounds__(kVecBlock) void cg_init_elem_kernel(
    int n, int n_el, const int *__restrict__ dof_ptr, const int *__restrict__ dof_contrib,
    const float *__restrict__ Mr, const unsigned char *__restrict__ mask,
    const double *__restrict__ b, V *__restrict__ r, V *__restrict__ w, V *__restrict__ w_other,
    V *__restrict__ dinv, float *__restrict__ xs, double *__restrict__ partB,
    double *__restrict__ partbb) {
    __shared__ double lds[kVecBlock / kWave];
    double bb = 0.0, rr = 0.0, rz = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const bool free_row = mask ? mask[i] != 0 : true;
        V ri = 0, di = 0;
        if (free_row) {
            double dg = 0.0;
            for (int k = dof_ptr[i]; k < dof_ptr[i + 1]; ++k) {
                const int slot = dof_contrib[k];
                const int j = slot / n_el;
                dg += elem_diag(Mr, n_el, slot - (long long)j * n_el, j);
            }
            const double bi = b[i];
            bb += bi * bi;
            ri = (V)bi;
            di = dg != 0.0 ? (V)(1.0 / dg) : (V)0;
        }
        const V zi = di * ri;
        r[i] = ri;
        w[2 * i] = zi;
        w[2 * i + 1] = 0;
        w_other[2 * i] = 0;
        w_other[2 * i + 1] = 0;
        dinv[i] = di;
        if (xs) xs[i] = 0;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    const double sbb = block_sum(bb, lds);
    const double srz = block_sum(rz, lds);
    const double srr = block_sum(rr, lds);
    if (threadIdx.x == 0) {
        partB[2 * blockIdx.x] = srz;
        partB[2 * blockIdx.x + 1] = srr;
        partbb[blockIdx.x] = sbb;
    }
}

// float64 product with the element operator (the arithmetic of A1 on x,
// masked: fixed DOFs carry 0), element products into work[j*n_el + e]
__global__ void elem_apply_f64_kernel(int n_el, const int *__restrict__ gids,
                                      const int *__restrict__ orient,
                                      const double *__restrict__ Mr,
                                      const unsigned char *__restrict__ mask,
                                      const double *__restrict__ x, double *__restrict__ work) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    const int o = orient[e];
    double xh[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        const int g = gids[6 * e + j];
        const double v = (mask && !mask[g]) ? 0.0 : x[g];
        xh[j] = ((o >> j) & 1) ? -v : v;
    }
    double a[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) a[k] = Mr[k * (long long)n_el + e];
    const double u0 = xh[1] - xh[2] + xh[5];
    const double u1 = xh[2] - xh[0] - xh[4];
    const double u2 = xh[0] - xh[1] + xh[3];
    const double m0 = a[0] * u0 + a[1] * u1 + a[2] * u2;
    const double m1 = a[1] * u0 + a[3] * u1 + a[4] * u2;
    const double m2 = a[2] * u0 + a[4] * u1 + a[5] * u2;
    const double yh[6] = {m2 - m1, m0 - m2, m1 - m0, m2, -m1, m0};
#pragma unroll
    for (int j = 0; j < 6; ++j) work[j * (long long)n_el + e] = ((o >> j) & 1) ? -yh[j] : yh[j];
}

__global__ void elem_gather_f64_kernel(int n, const int *__restrict__ dof_ptr,
                                       const int *__restrict__ dof_contrib,
                                       const unsigned char *__restrict__ mask,
                                       const double *__restrict__ work, double *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    if (!mask || mask[i])
        for (int k = dof_ptr[i]; k < dof_ptr[i + 1]; ++k) s += work[dof_contrib[k]];
    y[i] = s;
}


// signed DOF ids of the element operator, slot j*n_el + e: id | bit 31 when
// the edge's orientation is flipped in element e
__global__ void el_dofs_kernel(int n_el, const int *__restrict__ gids,
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

// ---------------------------------------------------------------- kernel B
// JAC: z = dinv r here (Jacobi); otherwise z and the r.z partials come from
// the auxiliary-space apply that follows, and only r.r is written.
template <typename V, typename X, bool JAC>
__global__ __launch_bounds__(kVecBlock) void cg_update_kernel(
    SegMap sm, const unsigned char *__restrict__ mask, V *__restrict__ w,
    const V *__restrict__ q, const V *__restrict__ dinv, V *__restrict__ r,
    X *__restrict__ x, const double *__restrict__ partA, double *__restrict__ partB,
    CgState *__restrict__ st, int parity) {
    __shared__ double lds[kVecBlock / kWave];
    const int seg = sm.segB[blockIdx.x];
    CgState *S = st + seg;
    const int was_done = S->done;   // read with the partials, tested after them
    const int lb = blockIdx.x - sm.blkB0[seg], nbl = sm.blkB0[seg + 1] - sm.blkB0[seg];
    const double pq = reduce_range<kVecBlock>(partA, sm.blkA0[seg], sm.blkA0[seg + 1], 1, 0, lds);
    if (was_done) return;
    if (!(pq == pq) || pq == 0.0 || isinf(pq)) {
        if (lb == 0 && threadIdx.x == 0) {
            S->pq = pq;
            S->done = 3;
        }
        return;
    }
    const double alpha_d = S->rho[parity] / pq;
    const V alpha = (V)alpha_d;
    const int r0 = sm.row0[seg], r1 = sm.row0[seg + 1];
    double rz = 0.0, rr = 0.0;
    for (long long i = r0 + (long long)lb * blockDim.x + threadIdx.x; i < r1;
         i += (long long)nbl * blockDim.x) {
        const bool free_row = mask ? mask[i] != 0 : true;
        if (!free_row) continue;
        x[i] += (X)(alpha * w[2 * i + 1]);
        const V ri = r[i] - alpha * q[i];
        r[i] = ri;
        if (JAC) {
            const V zi = dinv[i] * ri;
            w[2 * i] = zi;
            rz += (double)ri * (double)zi;
        }
        rr += (double)ri * (double)ri;
    }
    const double s1 = JAC ? block_sum(rz, lds) : 0.0;
    const double s2 = block_sum(rr, lds);
    if (threadIdx.x == 0) {
        if (JAC) partB[2 * blockIdx.x] = s1;
        partB[2 * blockIdx.x + 1] = s2;
        if (lb == 0) {
            S->pq = pq;
            S->it = S->it + 1;
        }
    }
}

// 16-bit column offsets j - i (thread per row) and, per wave, the largest
// |j - i| seen, so the host can tell whether the offsets are exact
__global__ __launch_bounds__(256) void col_offset_kernel(int n, const int *__restrict__ indptr,
                                                         const int *__restrict__ indices,
                                                         short *__restrict__ dcol,
                                                         int *__restrict__ wave_max_out) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = (int)gid;
    int m = 0;
    if (gid < n) {
        for (int k = indptr[i]; k < indptr[i + 1]; ++k) {
            const int d = indices[k] - i;
            const int ad = d < 0 ? -d : d;
            m = ad > m ? ad : m;
            dcol[k] = (short)(ad <= 32767 ? d : 0);
        }
    }
    m = wave_max(m);
    if ((threadIdx.x & (kWave - 1)) == 0) wave_max_out[gid / kWave] = m;
}

// x64 += xs (f32 correction) on free rows
__global__ void cg_finalize_kernel(int n, const unsigned char *__restrict__ mask,
                                   const float *__restrict__ xs, double *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mask && !mask[i]) return;
    x[i] += (double)xs[i];
}

__global__ void zero_free_kernel(int n, const unsigned char *__restrict__ mask,
                                 double *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mask && !mask[i]) return;
    x[i] = 0.0;
}

template <typename V>
__global__ void fill_free_kernel(int n, const unsigned char *__restrict__ mask, V *__restrict__ w,
                                 V v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool f = mask ? mask[i] != 0 : true;
    w[2 * (long long)i] = f ? v : (V)0;      // z
    w[2 * (long long)i + 1] = (V)0;          // p
}

// partials that reduce to (rho = 1, ||r||^2 = 1) in every segment, for the
// SpMV timing helper
__global__ void fill_partials_kernel(SegMap sm, int nb, double *__restrict__ partB) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const bool first = sm.blkB0[sm.segB[k]] == k;
    partB[2 * k] = first ? 1.0 : 0.0;
    partB[2 * k + 1] = first ? 1.0 : 0.0;
}

// one block per segment: out[v] = sum over rows of segment v (mask != 0) of a*b
__global__ __launch_bounds__(kVecBlock) void seg_dot_kernel(const int *__restrict__ row0,
                                                            const double *__restrict__ a,
                                                            const double *__restrict__ b,
                                                            const unsigned char *__restrict__ mask,
                                                            double *__restrict__ out) {
    __shared__ double lds[kVecBlock / kWave];
    const int v = blockIdx.x;
    double s = 0.0;
    for (int i = row0[v] + threadIdx.x; i < row0[v + 1]; i += kVecBlock)
        if (!mask || mask[i]) s += a[i] * b[i];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) out[v] = s;
}

// ---------------------------------------------------------------- helpers
template <int LPR>
__global__ __launch_bounds__(kRowBlock) void spmv_f64_kernel(
    int n, const int *__restrict__ indptr, const int *__restrict__ indices,
    const double *__restrict__ vals, const unsigned char *__restrict__ mask,
    const double *__restrict__ x, double *__restrict__ y) {
    const int lane = threadIdx.x % LPR;
    const int rows_per_block = kRowBlock / LPR;
    for (long long row = (long long)blockIdx.x * rows_per_block + threadIdx.x / LPR; row < n;
         row += (long long)gridDim.x * rows_per_block) {
        const int i = (int)row;
        const bool fi = mask ? mask[i] != 0 : true;
        double acc = 0.0;
        if (fi) {
            for (int k = indptr[i] + lane; k < indptr[i + 1]; k += LPR) {
                const int j = indices[k];
                if (!mask || mask[j]) acc += vals[k] * x[j];
            }
        }
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
        if (lane == 0) y[i] = acc;
    }
}

__global__ __launch_bounds__(kRowBlock) void dot_partials_kernel(
    int n, const double *__restrict__ a, const double *__restrict__ b,
    const unsigned char *__restrict__ mask, int op, double *__restrict__ partials) {
    __shared__ double lds[kRowBlock / kWave];
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        if (mask && !mask[i]) continue;
        if (op == 0) s += a[i] * b[i];
        else s = fmax(s, fabs(a[i]));
    }
    if (op == 0) {
        s = block_sum(s, lds);
    } else {
        s = wave_max(s);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x / 64] = s;
        __syncthreads();
        s = 0.0;
        for (int w = 0; w < (int)blockDim.x / 64; ++w) s = fmax(s, lds[w]);
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// The partials (at most 1024) summed in index order by one thread: the same
// additions, in the same order, as the host sum of reduce_host, so the same
// bits.  The workgroup stages them in LDS first, so the serial sum waits on
// LDS rather than on one global load after another.
constexpr int kSumInOrderMax = 1024;

__global__ __launch_bounds__(kSumInOrderMax) void sum_in_order_kernel(
    const double *__restrict__ partials, int G, double *__restrict__ out) {
    __shared__ double buf[kSumInOrderMax];
    for (int k = threadIdx.x; k < G; k += blockDim.x) buf[k] = partials[k];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double r = 0.0;
    for (int k = 0; k < G; ++k) r = r + buf[k];
    *out = r;
}

}  // namespace sn

using namespace sn;

struct sn_cg {
    const sn_pattern *pat = nullptr;
    int precision = 64;  // vector precision
    bool mat32 = false;  // float32 matrix values (with float64 vectors when precision = 64)
    int lpr = 8;
    int sched = 0;   // SpMV row schedule: bit 0 contiguous spans, bit 1 XCD-grouped blocks
    int chunk = 32;
    int nseg = 1;
    int grid_rows = 0;   // blocks of kernel A
    int grid_vec = 0;    // blocks of kernel B and init
    std::vector<int> row0_h;            // [nseg+1]
    int *row0 = nullptr, *blkA0 = nullptr, *blkB0 = nullptr, *segA = nullptr, *segB = nullptr;
    double *rtol_dev = nullptr;         // [nseg]
    // q: [n] the product of kernel A (CSR); with the element operator, A2's p
    void *r = nullptr, *q = nullptr, *dinv = nullptr;
    void *w0 = nullptr, *w1 = nullptr;   // interleaved {z, p}: even iterations read w0
    float *xs = nullptr;
    double *partA = nullptr, *partB = nullptr, *partbb = nullptr;
    short *dcol = nullptr;       // [nnz] 16-bit column offsets j - i, when they are exact
    bool use_dcol = true;        // SpMV reads dcol (when present) instead of the indices
    sn_ams *ams = nullptr;       // auxiliary-space preconditioner (null: Jacobi)
    // element operator (sn_cg_set_element_operator): solves take the element
    // matrices Mr instead of CSR values
    bool el_mode = false;
    int n_el = 0;
    int grid_el = 0;
    int *el_dofs = nullptr;      // [6*n_el] signed DOF ids, slot j*n_el + e
    float *yloc = nullptr;       // [6*n_el] element products, slot j*n_el + e
    CgState *st = nullptr;       // [nseg]
    CgState *st_host = nullptr;  // [nseg], pinned
    // cached graphs of one chunk and of a short chunk (chunk_short
    // iterations), keyed by their pointers
    hipGraphExec_t graph = nullptr;
    hipGraphExec_t graph_short = nullptr;
    hipGraphExec_t graph_tail = nullptr;      // two iterations
    hipGraphExec_t graph_quarter = nullptr;   // four iterations
    int chunk_short = 8;
    const void *g_vals = nullptr;
    const unsigned char *g_mask = nullptr;
    void *g_x = nullptr;
    hipStream_t g_stream = nullptr;

    SegMap map() const { return SegMap{row0, blkA0, blkB0, segA, segB}; }
    const int *done_flag() const {
        return reinterpret_cast<const int *>(reinterpret_cast<const char *>(st) +
                                             offsetof(CgState, done));
    }
};

namespace sn {

template <typename V, typename M, int LPR>
static void launch_spmv(sn_cg *C, const M *vals, const unsigned char *mask, int parity,
                        hipStream_t s) {
    const sn_pattern *P = C->pat;
    const V *wold = (const V *)(parity ? C->w1 : C->w0);
    V *wnew = (V *)(parity ? C->w0 : C->w1);
    if (C->dcol && C->use_dcol)
        hipLaunchKernelGGL((cg_spmv_kernel<V, M, LPR, short>), dim3(C->grid_rows),
                           dim3(kRowBlock), 0, s, C->map(), P->indptr, C->dcol, vals, mask,
                           wold, wnew, (V *)C->q, C->partB, C->partA, C->st, parity, C->sched);
    else
        hipLaunchKernelGGL((cg_spmv_kernel<V, M, LPR, int>), dim3(C->grid_rows),
                           dim3(kRowBlock), 0, s, C->map(), P->indptr, P->indices, vals, mask,
                           wold, wnew, (V *)C->q, C->partB, C->partA, C->st, parity, C->sched);
}

// 16-bit column offsets for the SpMV when every |j - i| fits (kept otherwise
// on int32 indices)
static int build_col_offsets(sn_cg *C, hipStream_t s) {
    const sn_pattern *P = C->pat;
    if (P->nnz == 0 || P->n == 0) return SN_OK;
    const int grid = grid_for(P->n, 256);
    const size_t nw = (size_t)grid * (256 / kWave);
    int *wmax = nullptr;
    SN_HIP(hipMalloc(&C->dcol, (size_t)P->nnz * sizeof(short)));
    SN_HIP(hipMalloc(&wmax, nw * sizeof(int)));
    hipLaunchKernelGGL(col_offset_kernel, dim3(grid), dim3(256), 0, s, P->n, P->indptr,
                       P->indices, C->dcol, wmax);
    std::vector<int> h(nw);
    SN_HIP(hipMemcpyAsync(h.data(), wmax, nw * sizeof(int), hipMemcpyDeviceToHost, s));
    SN_HIP(hipStreamSynchronize(s));
    (void)hipFree(wmax);
    int m = 0;
    for (int v : h) m = v > m ? v : m;
    if (m > 32767) {
        (void)hipFree(C->dcol);
        C->dcol = nullptr;
    }
    return SN_OK;
}

template <typename V, typename M>
static void launch_spmv_lpr(sn_cg *C, const M *vals, const unsigned char *mask, int parity,
                            hipStream_t s) {
    switch (C->lpr) {
        case 4: launch_spmv<V, M, 4>(C, vals, mask, parity, s); break;
        case 16: launch_spmv<V, M, 16>(C, vals, mask, parity, s); break;
        default: launch_spmv<V, M, 8>(C, vals, mask, parity, s); break;
    }
}

// kernel B and the preconditioner, after the product of an iteration
template <typename V>
static void launch_update_precond(sn_cg *C, const unsigned char *mask, void *xdst, int parity,
                                  hipStream_t s) {
    V *wnew = (V *)(parity ? C->w0 : C->w1);
#define SN_LAUNCH_UPD(X, JAC)                                                                 \
    hipLaunchKernelGGL((cg_update_kernel<V, X, JAC>), dim3(C->grid_vec), dim3(kVecBlock), 0, s, \
                       C->map(), mask, wnew, (const V *)C->q, (const V *)C->dinv, (V *)C->r,  \
                       (X *)xdst, C->partA, C->partB, C->st, parity)
    if (C->ams) {
        if (C->precision == 64) SN_LAUNCH_UPD(double, false);
        else SN_LAUNCH_UPD(float, false);
        const AmsSegments seg{C->row0, C->blkB0, C->segB};
        (void)ams_launch_apply(C->ams, C->precision == 64, C->r, C->dinv, mask, wnew, 2,
                               nullptr, C->partB, 2, C->grid_vec, &seg,
                               C->nseg == 1 ? C->done_flag() : nullptr, C->sched, s);
    } else {
        if (C->precision == 64) SN_LAUNCH_UPD(double, true);
        else SN_LAUNCH_UPD(float, true);
    }
#undef SN_LAUNCH_UPD
}

template <typename V, typename M>
static void launch_iteration(sn_cg *C, const M *vals, const unsigned char *mask, void *xdst,
                             int parity, hipStream_t s) {
    launch_spmv_lpr<V, M>(C, vals, mask, parity, s);
    launch_update_precond<V>(C, mask, xdst, parity, s);
}

// A1 and A2 of one iteration on the element operator (float32 workspaces):
// the product with the CG update (and the Jacobi z when there is no
// auxiliary-space preconditioner).  A1's p.q partials go to partA
// (grid_el <= grid_rows blocks).
static void launch_elem_product(sn_cg *C, const float *Mr, const unsigned char *mask, void *xdst,
                                int parity, hipStream_t s) {
    const sn_pattern *P = C->pat;
    const float *wold = (const float *)(parity ? C->w1 : C->w0);
    float *wnew = (float *)(parity ? C->w0 : C->w1);
    hipLaunchKernelGGL(cg_elem_kernel<float>, dim3(C->grid_el), dim3(kRowBlock), 0, s, C->n_el,
                       C->el_dofs, Mr, wold, C->yloc, C->partB, C->grid_vec, C->partA, C->st,
                       parity, C->sched);
#define SN_LAUNCH_EDGE(JAC)                                                                    \
    hipLaunchKernelGGL((cg_edge_kernel<float, float, JAC>), dim3(C->grid_vec), dim3(kVecBlock), \
                       0, s, P->n, P->dof_ptr, P->dof_contrib, C->yloc, mask,                  \
                       (const float *)C->dinv, wold, wnew, (float *)C->q, (float *)C->r,       \
                       (float *)xdst, C->partA, C->grid_el, C->partB, C->st, parity, C->sched)
    if (C->ams) SN_LAUNCH_EDGE(false);
    else SN_LAUNCH_EDGE(true);
#undef SN_LAUNCH_EDGE
}

// one iteration on the element operator: A1, A2, then the auxiliary-space
// preconditioner (z and the r.z partials)
static void launch_elem_iteration(sn_cg *C, const float *Mr, const unsigned char *mask,
                                  void *xdst, int parity, hipStream_t s) {
    launch_elem_product(C, Mr, mask, xdst, parity, s);
    if (C->ams) {
        // the pairs {z, p} of the next iteration: z from the preconditioner,
        // p from A2's plain array
        float *wnew = (float *)(parity ? C->w0 : C->w1);
        const AmsSegments seg{C->row0, C->blkB0, C->segB};
        (void)ams_launch_apply(C->ams, false, C->r, C->dinv, mask, wnew, 2, C->q, C->partB, 2,
                               C->grid_vec, &seg, C->done_flag(), C->sched, s);
    }
}

template <typename V, typename M, int LPR>
static void launch_init(sn_cg *C, const M *vals, const unsigned char *mask, const double *b,
                        const double *x0, hipStream_t s) {
    const sn_pattern *P = C->pat;
    hipLaunchKernelGGL((cg_init_kernel<V, M, LPR>), dim3(C->grid_vec), dim3(kVecBlock), 0, s,
                       C->map(), P->indptr, P->indices, P->diag, vals, mask, b, x0, (V *)C->r,
                       (V *)C->w0, (V *)C->w1, (V *)C->dinv,
                       C->precision == 32 ? C->xs : (float *)nullptr, C->partB, C->partbb);
}

template <typename V, typename M>
static void launch_init_lpr(sn_cg *C, const M *vals, const unsigned char *mask, const double *b,
                            const double *x0, double atol, int maxiter, hipStream_t s) {
    switch (C->lpr) {
        case 4: launch_init<V, M, 4>(C, vals, mask, b, x0, s); break;
        case 16: launch_init<V, M, 16>(C, vals, mask, b, x0, s); break;
        default: launch_init<V, M, 8>(C, vals, mask, b, x0, s); break;
    }
    hipLaunchKernelGGL(cg_init_state_kernel, dim3(C->nseg), dim3(kVecBlock), 0, s, C->map(),
                       C->partbb, C->partB, C->st, C->rtol_dev, atol, maxiter);
    if (C->ams) {
        const AmsSegments seg{C->row0, C->blkB0, C->segB};
        (void)ams_launch_apply(C->ams, C->precision == 64, C->r, C->dinv, mask, C->w0, 2,
                               nullptr, C->partB, 2, C->grid_vec, &seg, nullptr, C->sched, s);
    }
}

static void launch_init_elem(sn_cg *C, const float *Mr, const unsigned char *mask,
                             const double *b, double atol, int maxiter, hipStream_t s) {
    const sn_pattern *P = C->pat;
    hipLaunchKernelGGL(cg_init_elem_kernel<float>, dim3(C->grid_vec), dim3(kVecBlock), 0, s, P->n,
                       C->n_el, P->dof_ptr, P->dof_contrib, Mr, mask, b, (float *)C->r,
                       (float *)C->w0, (float *)C->w1, (float *)C->dinv, C->xs, C->partB,
                       C->partbb);
    hipLaunchKernelGGL(cg_init_state_kernel, dim3(1), dim3(kVecBlock), 0, s, C->map(), C->partbb,
                       C->partB, C->st, C->rtol_dev, atol, maxiter);
    if (C->ams) {
        const AmsSegments seg{C->row0, C->blkB0, C->segB};
        (void)ams_launch_apply(C->ams, false, C->r, C->dinv, mask, C->w0, 2, nullptr, C->partB,
                               2, C->grid_vec, &seg, nullptr, C->sched, s);
    }
}

static hipError_t drop_graphs(sn_cg *C) {
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

static void launch_chunk(sn_cg *C, const void *vals, const unsigned char *mask, void *xdst,
                         int count, hipStream_t s) {
    // an even number of iterations keeps the p buffers and rho slots in step
    for (int k = 0; k < count; ++k) {
        if (C->el_mode)
            launch_elem_iteration(C, (const float *)vals, mask, xdst, k & 1, s);
        else if (C->precision == 64 && !C->mat32)
            launch_iteration<double, double>(C, (const double *)vals, mask, xdst, k & 1, s);
        else if (C->precision == 64)
            launch_iteration<double, float>(C, (const float *)vals, mask, xdst, k & 1, s);
        else
            launch_iteration<float, float>(C, (const float *)vals, mask, xdst, k & 1, s);
    }
}

// Blocks of a grid of about `target` blocks shared out over the segments in
// proportion to their rows (at least one, at most one per `rows_per_block`).
static void share_blocks(const std::vector<int> &row0, int target, int rows_per_block,
                         std::vector<int> &blk0, std::vector<int> &seg_of) {
    const int nseg = (int)row0.size() - 1;
    const long long n = row0[nseg] - row0[0];
    blk0.assign(nseg + 1, 0);
    seg_of.clear();
    for (int v = 0; v < nseg; ++v) {
        const long long nv = row0[v + 1] - row0[v];
        long long b = (nv * target + n - 1) / (n > 0 ? n : 1);
        const long long cap = (nv + rows_per_block - 1) / rows_per_block;
        if (b > cap) b = cap;
        if (b < 1) b = 1;
        blk0[v + 1] = blk0[v] + (int)b;
        for (long long k = 0; k < b; ++k) seg_of.push_back(v);
    }
}

// Copies go through the caller's stream (never the legacy stream, which
// would have to wait for streams another thread may be capturing).
static int upload_ints(int **dst, const std::vector<int> &h, hipStream_t s) {
    SN_HIP(hipMalloc(dst, h.size() * sizeof(int)));
    SN_HIP(hipMemcpyAsync(*dst, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, s));
    SN_HIP(hipStreamSynchronize(s));
    return SN_OK;
}

static int cg_run(sn_cg *C, const void *vals, const unsigned char *mask, const double *b,
                  double *x, const double *rtol_host, double atol, int maxiter, hipStream_t s) {
    const sn_pattern *P = C->pat;
    if (maxiter < 0) {
        long long nfree = P->n;
        if (mask) {
            std::vector<unsigned char> h(P->n);
            SN_HIP(hipMemcpyAsync(h.data(), mask, P->n, hipMemcpyDeviceToHost, s));
            SN_HIP(hipStreamSynchronize(s));
            nfree = 0;
            for (int i = 0; i < P->n; ++i) nfree += h[i] != 0;
        }
        maxiter = (int)(nfree * 10 > INT_MAX ? INT_MAX : nfree * 10);
    }
    SN_REQUIRE(!C->ams || C->ams->ready, "CG: auxiliary-space preconditioner has no values yet");
    SN_HIP(hipMemcpyAsync(C->rtol_dev, rtol_host, C->nseg * sizeof(double),
                          hipMemcpyHostToDevice, s));
    void *xdst = (C->precision == 64) ? (void *)x : (void *)C->xs;
    const double *x0 = (C->precision == 64) ? x : nullptr;
    if (C->el_mode)
        launch_init_elem(C, (const float *)vals, mask, b, atol, maxiter, s);
    else if (C->precision == 64 && !C->mat32)
        launch_init_lpr<double, double>(C, (const double *)vals, mask, b, x0, atol, maxiter, s);
    else if (C->precision == 64)
        launch_init_lpr<double, float>(C, (const float *)vals, mask, b, x0, atol, maxiter, s);
    else
        launch_init_lpr<float, float>(C, (const float *)vals, mask, b, x0, atol, maxiter, s);
    SN_CHECK_LAUNCH();

    // (re)capture the chunk graph when the pointers or stream changed
    if (!C->graph || C->g_vals != vals || C->g_mask != mask || C->g_x != xdst || C->g_stream != s) {
        SN_HIP(drop_graphs(C));
        hipStream_t cap = s;
        bool own = false;
        if (cap == nullptr) {
            SN_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
            own = true;
        }
        hipGraphExec_t *dst[4] = {&C->graph, &C->graph_short, &C->graph_quarter, &C->graph_tail};
        const int len[4] = {C->chunk, C->chunk_short, 4, 2};
        for (int which = 0; which < 4; ++which) {
            hipGraph_t g;
            SN_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
            launch_chunk(C, vals, mask, xdst, len[which], cap);
            SN_HIP(hipStreamEndCapture(cap, &g));
            SN_HIP(hipGraphInstantiate(dst[which], g, nullptr, nullptr, 0));
            SN_HIP(hipGraphDestroy(g));
        }
        if (own) SN_HIP(hipStreamDestroy(cap));
        C->g_vals = vals;
        C->g_mask = mask;
        C->g_x = xdst;
        C->g_stream = s;
    }
    const size_t sb = C->nseg * sizeof(CgState);
    SN_HIP(hipMemcpyAsync(C->st_host, C->st, sb, hipMemcpyDeviceToHost, s));
    SN_HIP(hipStreamSynchronize(s));
    // ||r|| of every segment at iteration 0, for the rate estimate
    std::vector<double> r_start(C->nseg), r_last(C->nseg);
    std::vector<int> it_last(C->nseg, 0);
    for (int v = 0; v < C->nseg; ++v) r_start[v] = r_last[v] = sqrt(C->st_host[v].rr);
    // Replays: the first a short one; then a full chunk only while some
    // running segment's estimate of the iterations still needed (at its rate
    // over the last replay) is at least two full chunks, short chunks (8) down to 10
    // still needed, 4 down to 5, two-iteration ones after that (iterations
    // replayed after every segment stopped still launch kernels that exit at
    // entry).
    for (;;) {
        bool running = false, first = false;
        double need_max = 0.0;
        for (int v = 0; v < C->nseg; ++v) {
            const CgState &h = C->st_host[v];
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
            // no rate: a long solve once it has run two full chunks
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
    if (C->precision == 64) {
        // scipy returns b itself (all zeros) when ||b|| == 0
        for (int v = 0; v < C->nseg; ++v) {
            if (C->st_host[v].bnorm != 0.0) continue;
            const int r0 = C->row0_h[v], nv = C->row0_h[v + 1] - r0;
            if (nv > 0)
                hipLaunchKernelGGL(zero_free_kernel, dim3(grid_for(nv, 256)), dim3(256), 0, s, nv,
                                   mask ? mask + r0 : nullptr, x + r0);
        }
    } else {
        hipLaunchKernelGGL(cg_finalize_kernel, dim3(grid_for(P->n, 256)), dim3(256), 0, s, P->n,
                           mask, C->xs, x);
    }
    SN_CHECK_LAUNCH();
    return SN_OK;
}

}  // namespace sn

extern "C" {

int sn_cg_create_batched(sn_cg **out, const sn_pattern *P, int precision, int lpr, int chunk,
                         int nseg, const int *seg_row0_host, void *stream) {
    SN_REQUIRE(out && P, "sn_cg_create: null argument");
    SN_REQUIRE(precision == 64 || precision == 32 || precision == 3264,
               "sn_cg_create: precision must be 64, 32 or 3264");
    if (lpr == 0) lpr = 8;
    SN_REQUIRE(lpr == 4 || lpr == 8 || lpr == 16, "sn_cg_create: lanes_per_row in {4,8,16}");
    SN_REQUIRE(nseg >= 1, "sn_cg_create: nseg must be >= 1");
    const bool mat32 = precision != 64;
    if (precision == 3264) precision = 64;
    std::vector<int> row0(nseg + 1);
    if (seg_row0_host) {
        for (int v = 0; v <= nseg; ++v) row0[v] = seg_row0_host[v];
    } else {
        SN_REQUIRE(nseg == 1, "sn_cg_create: segment rows missing");
        row0[0] = 0;
        row0[1] = P->n;
    }
    SN_REQUIRE(row0[0] == 0 && row0[nseg] == P->n, "sn_cg_create: segments must cover [0, n)");
    for (int v = 0; v < nseg; ++v)
        SN_REQUIRE(row0[v + 1] >= row0[v], "sn_cg_create: segment rows must be ascending");
    if (chunk <= 0) chunk = 32;
    if (chunk & 1) chunk += 1;
    sn_cg *C = new sn_cg();
    C->pat = P;
    C->precision = precision;
    C->mat32 = mat32;
    C->lpr = lpr;
    C->chunk = chunk;
    // a quarter chunk, even, at least 2
    C->chunk_short = std::max(2, (chunk / 4) & ~1);
    C->nseg = nseg;
    C->row0_h = row0;
    const int cus = num_cus();
    std::vector<int> blkA0, segA, blkB0, segB;
    share_blocks(row0, 8 * cus, kRowBlock / lpr, blkA0, segA);
    share_blocks(row0, cus, kVecBlock, blkB0, segB);
    C->grid_rows = blkA0[nseg];
    C->grid_vec = blkB0[nseg];
    int rc;
    if ((rc = upload_ints(&C->row0, row0, as_stream(stream)))) return rc;
    if ((rc = upload_ints(&C->blkA0, blkA0, as_stream(stream)))) return rc;
    if ((rc = upload_ints(&C->blkB0, blkB0, as_stream(stream)))) return rc;
    if ((rc = upload_ints(&C->segA, segA, as_stream(stream)))) return rc;
    if ((rc = upload_ints(&C->segB, segB, as_stream(stream)))) return rc;
    const size_t vb = (size_t)P->n * (precision == 64 ? 8 : 4);
    SN_HIP(hipMalloc(&C->r, vb));
    SN_HIP(hipMalloc(&C->w0, 2 * vb));
    SN_HIP(hipMalloc(&C->w1, 2 * vb));
    SN_HIP(hipMalloc(&C->q, vb));
    SN_HIP(hipMalloc(&C->dinv, vb));
    if (precision == 32) SN_HIP(hipMalloc(&C->xs, (size_t)P->n * 4));
    SN_HIP(hipMalloc(&C->partA, (size_t)C->grid_rows * sizeof(double)));
    SN_HIP(hipMalloc(&C->partB, (size_t)2 * C->grid_vec * sizeof(double)));
    SN_HIP(hipMalloc(&C->partbb, (size_t)C->grid_vec * sizeof(double)));
    SN_HIP(hipMalloc(&C->rtol_dev, (size_t)nseg * sizeof(double)));
    SN_HIP(hipMalloc(&C->st, nseg * sizeof(CgState)));
    SN_HIP(hipHostMalloc((void **)&C->st_host, nseg * sizeof(CgState), hipHostMallocDefault));
    std::memset(C->st_host, 0, nseg * sizeof(CgState));
    SN_HIP(hipMemcpyAsync(C->st, C->st_host, nseg * sizeof(CgState), hipMemcpyHostToDevice,
                          as_stream(stream)));
    if ((rc = build_col_offsets(C, as_stream(stream)))) return rc;
    SN_HIP(hipStreamSynchronize(as_stream(stream)));
    *out = C;
    return SN_OK;
}

int sn_cg_create(sn_cg **out, const sn_pattern *P, int precision, int lpr, int chunk,
                 void *stream) {
    return sn_cg_create_batched(out, P, precision, lpr, chunk, 1, nullptr, stream);
}

int sn_cg_set_schedule(sn_cg *C, int sched) {
    SN_REQUIRE(C && sched >= 0 && sched <= 3, "sn_cg_set_schedule: schedule in [0, 3]");
    SN_HIP(drop_graphs(C));
    C->sched = sched;
    return SN_OK;
}

int sn_cg_set_col_offsets(sn_cg *C, int enable) {
    SN_REQUIRE(C, "sn_cg_set_col_offsets: null workspace");
    SN_REQUIRE(!enable || C->dcol, "sn_cg_set_col_offsets: some |column - row| exceeds 32767");
    SN_HIP(drop_graphs(C));
    C->use_dcol = enable != 0;
    return SN_OK;
}

int sn_cg_index_bytes(const sn_cg *C) { return (C && C->dcol && C->use_dcol) ? 2 : 4; }

int sn_cg_set_ams(sn_cg *C, sn_ams *A) {
    SN_REQUIRE(C, "sn_cg_set_ams: null workspace");
    SN_REQUIRE(!A || A->n_edges == C->pat->n, "sn_cg_set_ams: %d edges, matrix has %d rows",
               A ? A->n_edges : 0, C->pat->n);
    SN_HIP(drop_graphs(C));
    C->ams = A;
    return SN_OK;
}

int sn_cg_set_element_operator(sn_cg *C, const int *gids, const int *orient, int enable,
                               void *stream) {
    SN_REQUIRE(C, "sn_cg_set_element_operator: null workspace");
    if (enable) {
        SN_REQUIRE(gids && orient, "sn_cg_set_element_operator: null element data");
        SN_REQUIRE(C->precision == 32 && C->mat32,
                   "sn_cg_set_element_operator: needs a float32 workspace (precision 32)");
        SN_REQUIRE(C->nseg == 1, "sn_cg_set_element_operator: one segment only");
        SN_REQUIRE(C->pat->dof_ptr && C->pat->dof_contrib && C->pat->num_el > 0,
                   "sn_cg_set_element_operator: pattern without element contribution lists");
        SN_REQUIRE((long long)C->pat->num_el * 6 < (1LL << 31),
                   "sn_cg_set_element_operator: too many elements");
    }
    SN_HIP(drop_graphs(C));
    if (!enable) {
        C->el_mode = false;
        return SN_OK;
    }
    hipStream_t s = as_stream(stream);
    const int ne = C->pat->num_el;
    if (C->n_el != ne || !C->el_dofs) {
        if (C->el_dofs) (void)hipFree(C->el_dofs);
        if (C->yloc) (void)hipFree(C->yloc);
        C->el_dofs = nullptr;
        C->yloc = nullptr;
        SN_HIP(hipMalloc(&C->el_dofs, (size_t)6 * ne * sizeof(int)));
        SN_HIP(hipMalloc(&C->yloc, (size_t)6 * ne * sizeof(float)));
        C->n_el = ne;
    }
    hipLaunchKernelGGL(el_dofs_kernel, dim3(grid_for(ne, 256)), dim3(256), 0, s, ne, gids, orient,
                       C->el_dofs);
    SN_CHECK_LAUNCH();
    SN_HIP(hipStreamSynchronize(s));
    // A1 strides over the elements with at most grid_rows blocks (its p.q
    // partials go to partA), a multiple of the 8 XCDs
    C->grid_el = grid_for(ne, kRowBlock);
    if (C->grid_el > C->grid_rows) C->grid_el = C->grid_rows;
    if (C->grid_el >= 8) C->grid_el = C->grid_el / 8 * 8;
    // XCD-grouped blocks: the element and DOF ranges of one XCD are
    // neighbours in the mesh, so their shared gathers meet in its L2
    // (1M-tet dipole: 43.3 -> 41.0 us per product, back to back)
    C->sched = 2;
    C->el_mode = true;
    return SN_OK;
}

int sn_cg_set_grids(sn_cg *C, int grid_el, int vec_per_cu, void *stream) {
    SN_REQUIRE(C, "sn_cg_set_grids: null workspace");
    SN_REQUIRE(grid_el >= 0, "sn_cg_set_grids: grid_el must be >= 0 (0: default)");
    SN_REQUIRE(vec_per_cu >= 1 && vec_per_cu <= 8, "sn_cg_set_grids: vec_per_cu in [1, 8]");
    hipStream_t s = as_stream(stream);
    SN_HIP(drop_graphs(C));
    // no kernel of this workspace may still read the buffers replaced here
    SN_HIP(hipStreamSynchronize(s));
    std::vector<int> blkB0, segB;
    share_blocks(C->row0_h, vec_per_cu * num_cus(), kVecBlock, blkB0, segB);
    for (void *p : {(void *)C->blkB0, (void *)C->segB, (void *)C->partB, (void *)C->partbb})
        if (p) SN_HIP(hipFree(p));
    C->blkB0 = C->segB = nullptr;
    C->partB = C->partbb = nullptr;
    int rc;
    if ((rc = upload_ints(&C->blkB0, blkB0, s))) return rc;
    if ((rc = upload_ints(&C->segB, segB, s))) return rc;
    C->grid_vec = blkB0[C->nseg];
    SN_HIP(hipMalloc(&C->partB, (size_t)2 * C->grid_vec * sizeof(double)));
    SN_HIP(hipMalloc(&C->partbb, (size_t)C->grid_vec * sizeof(double)));
    if (C->el_mode) {
        int g = grid_el > 0 ? grid_el : grid_for(C->n_el, kRowBlock);
        if (g > C->grid_rows) g = C->grid_rows;
        if (g >= 8) g = g / 8 * 8;
        C->grid_el = g;
    }
    return SN_OK;
}

int sn_cg_grids(const sn_cg *C, int *grid_el, int *grid_vec) {
    SN_REQUIRE(C, "sn_cg_grids: null workspace");
    if (grid_el) *grid_el = C->grid_el;
    if (grid_vec) *grid_vec = C->grid_vec;
    return SN_OK;
}

int sn_cg_element_operator(const sn_cg *C) { return (C && C->el_mode) ? 1 : 0; }

int sn_element_apply_f64(const sn_pattern *P, const int *gids, const int *orient,
                         const double *Mr, const unsigned char *mask, const double *x,
                         double *work, double *y, void *stream) {
    SN_REQUIRE(P && gids && orient && Mr && x && work && y, "sn_element_apply_f64: null argument");
    SN_REQUIRE(P->dof_ptr && P->dof_contrib, "sn_element_apply_f64: pattern without DOF lists");
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(elem_apply_f64_kernel, dim3(grid_for(P->num_el, 256)), dim3(256), 0, s,
                       P->num_el, gids, orient, Mr, mask, x, work);
    hipLaunchKernelGGL(elem_gather_f64_kernel, dim3(grid_for(P->n, 256)), dim3(256), 0, s, P->n,
                       P->dof_ptr, P->dof_contrib, mask, work, y);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

void sn_cg_destroy(sn_cg *C) {
    if (!C) return;
    (void)drop_graphs(C);
    void *bufs[] = {C->r,    C->w0,       C->w1, C->q,     C->dinv,  C->xs,
                    C->partA, C->partB,   C->partbb, C->rtol_dev, C->st, C->row0, C->blkA0,
                    C->blkB0, C->segA,    C->segB, C->dcol, C->el_dofs, C->yloc};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    (void)hipHostFree(C->st_host);
    delete C;
}

int sn_cg_solve(sn_cg *C, const void *vals, const unsigned char *mask, const double *b, double *x,
                double rtol, double atol, int maxiter, int *iters_out, double *rnorm_out,
                void *stream) {
    SN_REQUIRE(C && vals && b && x, "sn_cg_solve: null argument");
    SN_REQUIRE(C->nseg == 1, "sn_cg_solve: batched workspace, use sn_cg_solve_batched");
    int rc = cg_run(C, vals, mask, b, x, &rtol, atol, maxiter, as_stream(stream));
    if (rc) return rc;
    const CgState h = C->st_host[0];
    if (iters_out) *iters_out = h.it;
    if (rnorm_out) *rnorm_out = sqrt(h.rr);
    if (h.done == 2) {
        set_error("sn_cg_solve: iteration cap %d reached (||r|| = %.3e, tol = %.3e)", h.maxiter,
                  sqrt(h.rr), h.tol);
        return SN_ERR_NOT_CONVERGED;
    }
    if (h.done == 3) {
        set_error("sn_cg_solve: breakdown at iteration %d (p.Ap = %.3e)", h.it, h.pq);
        return SN_ERR_BREAKDOWN;
    }
    return SN_OK;
}

int sn_cg_solve_batched(sn_cg *C, const void *vals, const unsigned char *mask, const double *b,
                        double *x, const double *rtol_host, double atol, int maxiter,
                        int *iters_host, double *rnorm_host, int *status_host, void *stream) {
    SN_REQUIRE(C && vals && b && x && rtol_host, "sn_cg_solve_batched: null argument");
    int rc = cg_run(C, vals, mask, b, x, rtol_host, atol, maxiter, as_stream(stream));
    if (rc) return rc;
    int worst = SN_OK;
    for (int v = 0; v < C->nseg; ++v) {
        const CgState &h = C->st_host[v];
        if (iters_host) iters_host[v] = h.it;
        if (rnorm_host) rnorm_host[v] = sqrt(h.rr);
        const int st = h.done == 2 ? SN_ERR_NOT_CONVERGED : (h.done == 3 ? SN_ERR_BREAKDOWN : SN_OK);
        if (status_host) status_host[v] = st;
        if (st == SN_ERR_BREAKDOWN || (st == SN_ERR_NOT_CONVERGED && worst == SN_OK)) worst = st;
    }
    if (worst != SN_OK) set_error("sn_cg_solve_batched: a segment did not converge");
    return worst;
}

int sn_cg_time_spmv(sn_cg *C, const void *vals, const unsigned char *mask, int reps,
                    double *avg_ms, void *stream) {
    SN_REQUIRE(C && vals && avg_ms && reps > 0, "sn_cg_time_spmv: bad argument");
    const sn_pattern *P = C->pat;
    hipStream_t s = as_stream(stream);
    // a running state in every segment: rho = 1 now and before (beta = 1),
    // ||r||^2 = 1 above a zero tolerance, z = 1 on the free rows, p_old = 0
    std::vector<CgState> h(C->nseg);
    for (auto &e : h) {
        std::memset(&e, 0, sizeof(e));
        e.it = 1;
        e.rho[0] = e.rho[1] = 1.0;
        e.tol = 0.0;
        e.maxiter = INT_MAX;
    }
    SN_HIP(hipMemcpyAsync(C->st, h.data(), C->nseg * sizeof(CgState), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(fill_partials_kernel, dim3(grid_for(C->grid_vec, 256)), dim3(256), 0, s,
                       C->map(), C->grid_vec, C->partB);
    if (C->precision == 64)
        hipLaunchKernelGGL(fill_free_kernel<double>, dim3(grid_for(P->n, 256)), dim3(256), 0, s,
                           P->n, mask, (double *)C->w1, 1.0);
    else
        hipLaunchKernelGGL(fill_free_kernel<float>, dim3(grid_for(P->n, 256)), dim3(256), 0, s,
                           P->n, mask, (float *)C->w1, 1.0f);
    SN_CHECK_LAUNCH();
    hipEvent_t e0, e1;
    SN_HIP(hipEventCreate(&e0));
    SN_HIP(hipEventCreate(&e1));
    SN_HIP(hipEventRecord(e0, s));
    // parity 1 reads {z, p} of w1 (p zero) and rho[0], writes p into w0 and rho[1]: the state stays
    // the same for every repetition
    for (int r = 0; r < reps; ++r) {
        if (C->el_mode)   // A1 and A2, whose update moves x and r each time
            launch_elem_product(C, (const float *)vals, mask, C->xs, 1, s);
        else if (C->precision == 64 && !C->mat32)
            launch_spmv_lpr<double, double>(C, (const double *)vals, mask, 1, s);
        else if (C->precision == 64)
            launch_spmv_lpr<double, float>(C, (const float *)vals, mask, 1, s);
        else
            launch_spmv_lpr<float, float>(C, (const float *)vals, mask, 1, s);
    }
    SN_HIP(hipEventRecord(e1, s));
    SN_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    SN_HIP(hipEventElapsedTime(&ms, e0, e1));
    SN_HIP(hipEventDestroy(e0));
    SN_HIP(hipEventDestroy(e1));
    SN_HIP(hipMemcpyAsync(h.data(), C->st, C->nseg * sizeof(CgState), hipMemcpyDeviceToHost, s));
    SN_HIP(hipStreamSynchronize(s));
    for (const auto &e : h)
        SN_REQUIRE(e.done == 0, "sn_cg_time_spmv: the timed launches stopped early (done=%d)",
                   e.done);
    *avg_ms = (double)ms / reps;
    return SN_OK;
}

int sn_spmv_f64(const sn_pattern *P, const double *vals, const unsigned char *mask,
                const double *x, double *y, void *stream) {
    SN_REQUIRE(P && vals && x && y, "sn_spmv_f64: null argument");
    hipLaunchKernelGGL(spmv_f64_kernel<8>, dim3(grid_for(P->n, kRowBlock / 8, 8 * num_cus())),
                       dim3(kRowBlock), 0, as_stream(stream), P->n, P->indptr, P->indices, vals,
                       mask, x, y);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

// Partial-sum buffer of the host reductions, one per (device, stream, host
// thread), made on the first reduction of that combination and kept for the
// life of the process: a call reads its partials back before it returns, and
// the calls of one host thread on one stream cannot overlap, so the buffer is
// never shared by two calls in flight.
static std::mutex g_part_mu;
static std::map<std::tuple<int, hipStream_t, std::thread::id>, double *> g_part;

static int reduce_partials(hipStream_t s, double **out) {
    int dev = 0;
    SN_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_part_mu);
    double *&p = g_part[std::make_tuple(dev, s, std::this_thread::get_id())];
    if (!p) SN_HIP(hipMalloc((void **)&p, 1024 * sizeof(double)));
    *out = p;
    return SN_OK;
}

static int reduce_host(int n, const double *a, const double *b, const unsigned char *mask, int op,
                       double *out, void *stream) {
    SN_REQUIRE(a && out && (op != 0 || b), "reduction: null argument");
    hipStream_t s = as_stream(stream);
    const int G = grid_for(n, kRowBlock, 1024);
    double *part = nullptr;
    int rc = reduce_partials(s, &part);
    if (rc) return rc;
    hipLaunchKernelGGL(dot_partials_kernel, dim3(G), dim3(kRowBlock), 0, s, n, a, b, mask, op,
                       part);
    SN_CHECK_LAUNCH();
    std::vector<double> h(G);
    SN_HIP(hipMemcpyAsync(h.data(), part, G * sizeof(double), hipMemcpyDeviceToHost, s));
    SN_HIP(hipStreamSynchronize(s));
    double r = 0.0;
    for (int k = 0; k < G; ++k) r = (op == 0) ? r + h[k] : (h[k] > r ? h[k] : r);
    *out = r;
    return SN_OK;
}

int sn_seg_dot_f64(int nseg, const int *row0_dev, const double *a, const double *b,
                   const unsigned char *mask, double *out_host, void *stream) {
    SN_REQUIRE(nseg >= 1 && row0_dev && a && b && out_host, "sn_seg_dot_f64: bad argument");
    hipStream_t s = as_stream(stream);
    double *dev = nullptr;
    SN_HIP(hipMallocAsync(&dev, nseg * sizeof(double), s));
    hipLaunchKernelGGL(seg_dot_kernel, dim3(nseg), dim3(kVecBlock), 0, s, row0_dev, a, b, mask, dev);
    SN_CHECK_LAUNCH();
    SN_HIP(hipMemcpyAsync(out_host, dev, nseg * sizeof(double), hipMemcpyDeviceToHost, s));
    SN_HIP(hipFreeAsync(dev, s));
    SN_HIP(hipStreamSynchronize(s));
    return SN_OK;
}

// Allocation-free segmented dot: the per-segment sums go to a file-scope
// device array instead of a buffer allocated per call.  The array's host-side
// name is not a device address: its address on each device is looked up once
// per device (hipGetSymbolAddress, with that device current) and only that
// address is passed to kernels and copies.
constexpr int kSegDotScratchMax = 4096;
static __device__ __align__(256) double scratch[kSegDotScratchMax];

static std::mutex g_scratch_mu;
static double *g_scratch_dev[64] = {nullptr};

static int seg_dot_scratch(double **out) {
    int dev = 0;
    SN_HIP(hipGetDevice(&dev));
    SN_REQUIRE(dev >= 0 && dev < 64, "sn_seg_dot_f64_noalloc: device %d beyond 64", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        SN_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        SN_REQUIRE(ptr, "sn_seg_dot_f64_noalloc: no device address for the scratch array");
        g_scratch_dev[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch_dev[dev];
    return SN_OK;
}

int sn_seg_dot_f64_noalloc(int nseg, const int *row0_dev, const double *a, const double *b,
                           const unsigned char *mask, double *out_host, void *stream) {
    SN_REQUIRE(nseg >= 1 && row0_dev && a && b && out_host,
               "sn_seg_dot_f64_noalloc: bad argument");
    SN_REQUIRE(nseg <= kSegDotScratchMax,
               "sn_seg_dot_f64_noalloc: %d segments exceed the scratch maximum of %d", nseg,
               kSegDotScratchMax);
    double *sc = nullptr;
    int rc = seg_dot_scratch(&sc);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(seg_dot_kernel, dim3(nseg), dim3(kVecBlock), 0, s, row0_dev, a, b, mask, sc);
    SN_CHECK_LAUNCH();
    SN_HIP(hipMemcpyAsync(out_host, sc, nseg * sizeof(double), hipMemcpyDeviceToHost, s));
    SN_HIP(hipStreamSynchronize(s));
    return SN_OK;
}

int sn_dot_f64(int n, const double *a, const double *b, const unsigned char *mask, double *out,
               void *stream) {
    return reduce_host(n, a, b, mask, 0, out, stream);
}

int sn_dot_f64_dev(int n, const double *a, const double *b, const unsigned char *mask,
                   double *out_dev, void *stream) {
    SN_REQUIRE(n >= 0 && a && b && out_dev, "sn_dot_f64_dev: null argument");
    hipStream_t s = as_stream(stream);
    const int G = grid_for(n, kRowBlock, 1024);
    double *part = nullptr;
    int rc = reduce_partials(s, &part);
    if (rc) return rc;
    hipLaunchKernelGGL(dot_partials_kernel, dim3(G), dim3(kRowBlock), 0, s, n, a, b, mask, 0,
                       part);
    SN_CHECK_LAUNCH();
    SN_REQUIRE(G <= kSumInOrderMax, "sn_dot_f64_dev: too many partials");
    hipLaunchKernelGGL(sum_in_order_kernel, dim3(1), dim3(kSumInOrderMax), 0, s, part, G,
                       out_dev);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_absmax_f64(int n, const double *a, const unsigned char *mask, double *out, void *stream) {
    return reduce_host(n, a, nullptr, mask, 1, out, stream);
}

}  // extern "C"
