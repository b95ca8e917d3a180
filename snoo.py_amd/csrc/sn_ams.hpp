// Auxiliary-space (Hiptmair-Xu) preconditioner state, shared by sn_ams.hip
// (set-up, numeric refresh, kernels) and sn_cg.hip (which launches the apply
// inside its captured CG chunk).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

struct AmsLevel {
    int n = 0;              // unknowns (level 0: free mesh nodes)
    int nnz = 0;
    int *indptr = nullptr;  // [n+1]
    int *indices = nullptr; // [nnz]
    int *diagpos = nullptr; // [n]
    int *gptr = nullptr;    // [nnz+1] gather plan of vals
    int *glist = nullptr;   // level 0: slot*n_el + e of the element values; else nnz of level-1
    int *agg_ptr = nullptr; // [n+1] rows of level-1 forming each row here (null at level 0)
    int *agg_kp = nullptr;  // [n+1] level-1 indptr at agg_ptr: each aggregate's nonzeros
    std::vector<int> indptr_h;  // host copy of indptr (forms the next level's agg_kp)
    std::vector<int> indices_h; // host copy of indices (forms the next level's rt pattern)
    // coarse levels: the restriction of the previous level's pre-smoothed
    // residual, R = P^T (I - omega A_f D_f^-1), one row per aggregate, so that
    // b_c = R b_f (the columns: rows of the previous level reached from the
    // aggregate; sign bit set for the aggregate's own rows, where P^T adds 1)
    int rt_nnz = 0;
    int *rt_ptr = nullptr;   // [n+1]
    int *rt_cols = nullptr;  // [rt_nnz]
    int *rt_gptr = nullptr;  // [rt_nnz+1] plan over the previous level's nonzeros
    int *rt_glist = nullptr; // [previous nnz]
    float *rt_vals = nullptr;// [rt_nnz]
    float *vals = nullptr;  // [nnz]
    float *dinv = nullptr;  // [n]
    float4 *b = nullptr;    // right-hand side of the cycle at this level
    float4 *x = nullptr;    // pre-smoothed iterate, then + prolongated coarse correction
};

struct sn_ams {
    int n_edges = 0, n_free = 0, n_el = 0;
    int n_nodes = 0;             // free nodes [0, n_free), then fixed nodes
    int2 *edge_nodes = nullptr;  // [n_edges] node ids of (lo, hi); >= n_free: fixed node
    int *node_mesh = nullptr;    // [n_nodes] mesh node id of each node
    float4 *xyz = nullptr;       // [n_nodes] coordinates (float), refreshed by sn_ams_update
    float4 *yx = nullptr;        // [2*n_nodes] per node: the cycle's result y (0 at fixed
                                 // nodes), then the coordinates: what the prolongation
                                 // reads at an end point, in one 32-byte piece
    int *inc_ptr = nullptr;      // [n_free+1] free edges at each free node
    int2 *inc = nullptr;         // [n_inc] {edge, the edge's other node | sign bit when this
                                 // node is the edge's hi end}
    int n_inc = 0;
    float *loc = nullptr;        // [10*n_el] element Laplacian pair values
    std::vector<AmsLevel> lv;
    double *dense = nullptr;     // [nc*nc] scratch of the coarsest inversion
    float *cinv = nullptr;       // [nc*nc] inverse of the coarsest matrix
    float omega = 0.6f;          // damped-Jacobi weight of the nodal smoother
    float weight = 1.0f;         // weight of the auxiliary correction
    bool ready = false;          // numeric values present
    int lanes = 8;               // lanes per row of the restriction, up and tail kernels
    // down steps of the single-system apply through the combined restrictions
    // R: off by default (one variant's level-0 step is latency-bound, 9.4 us
    // with R against 9.7 us without, at 1M tets); the batched CG uses R
    bool rt = false;
};

namespace sn {

// Row segments of a batched CG and the blocks serving each (device arrays).
struct AmsSegments {
    const int *row0;        // [nseg+1]
    const int *blk0;        // [nseg+1]
    const int *seg_of_blk;  // [grid]
};

// z = dinv .* r + weight * Pi V(Pi^T r) on the free edges (0 elsewhere),
// stored at z[zstride * i], with per-block partial sums of r.z written to
// partials[stride*b], b < grid (blocks of 1024 threads striding over the
// edges, or over the rows of their segment when `seg` is given).  With `pin`
// (zstride 2 only) the pair {z_i, pin[i]} is stored at z[2i], whole lines
// instead of every other value.  `done` (may be null) points to a device
// flag that turns every kernel into a no-op when set.  sched bit 1:
// XCD-grouped workgroups (as the CG's schedule knob).
int ams_launch_apply(sn_ams *A, bool v64, const void *r, const void *dinv,
                     const unsigned char *mask, void *z, int zstride, const void *pin,
                     double *partials, int stride, int grid, const AmsSegments *seg,
                     const int *done, int sched, hipStream_t s);

}  // namespace sn
