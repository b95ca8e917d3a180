// Element kernels that take the reference's own tables: the curls of the
// edge basis for every orientation the mesher enumerates, indexed by a
// per-element orientation number, and the P1 gradients at every quadrature
// point.  These are the arguments the reference's FFI receives
// (fem_c/interface_c.pyx:162-374 for the stiffness and Jacobian,
// :555-696 for the B field) and the tables its kernels index
// (fem_c/ctools.c:1195-1388, 960-1193, 1500-1563):
//   J_q[k][l]  = sum_a X[n_a][k] dphi_a/du_l (xi_q)     d_phi [Q][4][3]
//   c_{j,q}    = curls[((o Q + q) 6 + j) 3 + :]          curls [n_orient][Q][6][3]
//   g_{j,q}    = J_q c_{j,q},  det_q = det J_q
//   B_q        = Bs_q + sum_k g_{k,q} x[g_k] / det_q
//   K_jk       = sum_q nu_q (g_j.g_k) w_q / det_q
//   dK_jk      = sum_q [|B_q| >= 1e-14] nu'_q (g_j.B_q)(g_k.B_q) w_q / (det_q |B_q|)
//   rhs_j      = sum_q (nu0 - nu_q)(g_j.Bs_q) w_q        (source domains only)
// Nothing is assumed about the tables: the Jacobian and the curls are read
// per quadrature point, so any element map and any orientation enumeration
// the tables describe are evaluated as given.  The hot Newton path uses the
// affine-element kernels of sn_assemble.hip; these kernels serve callers that
// hand over the reference's arrays unchanged.  An orientation outside
// [0, n_orient) is counted, its element is skipped, and the entry point
// reports the count as an error.
#include "sn_common.hpp"

namespace sn {

constexpr int kRtBlock = 128;
constexpr double kNu0Rt = 1.0 / 4.0 / 3.141592653589793238462643383279502984 * 1e7;

// nu(|B|), dnu/d|B| on the reference's piecewise-cubic tables
// (fem_c/ctools.c:130-204): interval i holds x_i <= B < x_{i+1}; at or beyond
// the last knot the last cubic is extrapolated; below the first knot 0;
// n = -1 or type 0 is a constant.
__device__ __forceinline__ void rt_nu(int type, int n, const double *__restrict__ c,
                                      const double *__restrict__ x, double B, double &nu,
                                      double &dnu) {
    if (type == SN_NU_CONST || n == -1) {
        nu = c[0];
        dnu = 0.0;
        return;
    }
    const double t = (type == SN_NU_BSQ) ? B * B : B;
    if (t < x[0]) {
        nu = dnu = 0.0;
        return;
    }
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (x[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const double d = t - x[lo];
    const double c0 = c[lo], c1 = c[n + lo], c2 = c[2 * n + lo], c3 = c[3 * n + lo];
    nu = ((c0 * d + c1) * d + c2) * d + c3;
    const double dd = (3.0 * c0 * d + 2.0 * c1) * d + c2;
    dnu = (type == SN_NU_BSQ) ? 2.0 * B * dd : dd;
}

// J_q from the element's nodes and the gradient table of point q.
__device__ __forceinline__ double rt_jacobian(const double (&X)[4][3],
                                              const double *__restrict__ dphi_q,
                                              double (&Jq)[3][3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < 4; ++a) s += X[a][k] * dphi_q[3 * a + l];
            Jq[k][l] = s;
        }
    return Jq[0][0] * (Jq[1][1] * Jq[2][2] - Jq[2][1] * Jq[1][2]) -
           Jq[0][1] * (Jq[1][0] * Jq[2][2] - Jq[1][2] * Jq[2][0]) +
           Jq[0][2] * (Jq[1][0] * Jq[2][1] - Jq[1][1] * Jq[2][0]);
}

__global__ __launch_bounds__(kRtBlock) void element_table_kernel(
    int num_el, const double *__restrict__ nodes, const int *__restrict__ cells,
    const int *__restrict__ gids, const int *__restrict__ orient, int n_orient,
    const double *__restrict__ curls, const double *__restrict__ d_phi,
    const double *__restrict__ w, int Q, int type_spec, int n_int,
    const double *__restrict__ coeffs, const double *__restrict__ knots,
    const double *__restrict__ x, const double *__restrict__ Bs, double *__restrict__ locK,
    double *__restrict__ locdK, double *__restrict__ loc6, int *__restrict__ n_bad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= num_el) return;
    const int o = orient[e];
    if (o < 0 || o >= n_orient) {
        atomicAdd(n_bad, 1);
        for (int p = 0; p < 21; ++p) {
            locK[(long long)p * num_el + e] = 0.0;
            locdK[(long long)p * num_el + e] = 0.0;
        }
        for (int j = 0; j < 6; ++j) loc6[(long long)j * num_el + e] = 0.0;
        return;
    }
    double X[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int nd = cells[4 * e + a];
#pragma unroll
        for (int k = 0; k < 3; ++k) X[a][k] = nodes[3 * nd + k];
    }
    double xv[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) xv[j] = x[gids[6 * e + j]];

    double K[21], dK[21], rhs[6];
#pragma unroll
    for (int p = 0; p < 21; ++p) K[p] = dK[p] = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) rhs[j] = 0.0;

    const double *ctab = curls + (size_t)o * Q * 18;
    for (int q = 0; q < Q; ++q) {
        double Jq[3][3];
        const double det = rt_jacobian(X, d_phi + 12 * q, Jq);
        const double inv_det = 1.0 / det;
        double G[6][3];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const double *c = ctab + 18 * q + 3 * j;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                G[j][k] = Jq[k][0] * c[0] + Jq[k][1] * c[1] + Jq[k][2] * c[2];
        }
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        if (Bs) {
            const double *b = Bs + ((size_t)e * Q + q) * 3;
            s0 = b[0];
            s1 = b[1];
            s2 = b[2];
        }
        double b0 = s0, b1 = s1, b2 = s2;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            b0 += G[j][0] * xv[j] * inv_det;
            b1 += G[j][1] * xv[j] * inv_det;
            b2 += G[j][2] * xv[j] * inv_det;
        }
        const double bm = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
        double nu, dnu;
        rt_nu(type_spec, n_int, coeffs, knots, bm, nu, dnu);
        const double wq = w[q];
        const double kf = nu * wq * inv_det;
        double P[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) P[j] = G[j][0] * b0 + G[j][1] * b1 + G[j][2] * b2;
        const double df = (bm >= 1e-14) ? dnu * wq * inv_det / bm : 0.0;
        int p = 0;
#pragma unroll
        for (int j = 0; j < 6; ++j)
#pragma unroll
            for (int k = j; k < 6; ++k) {
                K[p] += kf * (G[j][0] * G[k][0] + G[j][1] * G[k][1] + G[j][2] * G[k][2]);
                dK[p] += df * P[j] * P[k];
                ++p;
            }
        if (Bs) {
            const double f = (kNu0Rt - nu) * wq;
#pragma unroll
            for (int j = 0; j < 6; ++j) rhs[j] += f * (G[j][0] * s0 + G[j][1] * s1 + G[j][2] * s2);
        }
    }
#pragma unroll
    for (int p = 0; p < 21; ++p) {
        locK[(long long)p * num_el + e] = K[p];
        locdK[(long long)p * num_el + e] = dK[p];
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) loc6[(long long)j * num_el + e] = rhs[j];
}

// Quadrature points and curl A per element and point (fem_c/ctools.c:1500-1563):
// points = sum_a phi_a(xi_q) X_a, B = sum_k g_{k,q} x[g_k] / det_q.
__global__ __launch_bounds__(kRtBlock) void b_hcurl_table_kernel(
    int num_el, const double *__restrict__ nodes, const int *__restrict__ cells,
    const int *__restrict__ gids, const int *__restrict__ orient, int n_orient,
    const double *__restrict__ curls, const double *__restrict__ phi,
    const double *__restrict__ d_phi, int Q, const double *__restrict__ x,
    double *__restrict__ pts, double *__restrict__ B, int *__restrict__ n_bad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= num_el) return;
    const int o = orient[e];
    const bool ok = o >= 0 && o < n_orient;
    if (!ok) atomicAdd(n_bad, 1);
    double X[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int nd = cells[4 * e + a];
#pragma unroll
        for (int k = 0; k < 3; ++k) X[a][k] = nodes[3 * nd + k];
    }
    double xv[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) xv[j] = x[gids[6 * e + j]];
    for (int q = 0; q < Q; ++q) {
        const size_t out = ((size_t)e * Q + q) * 3;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < 4; ++a) s += phi[4 * q + a] * X[a][k];
            pts[out + k] = s;
        }
        if (!ok) {
            B[out] = B[out + 1] = B[out + 2] = 0.0;
            continue;
        }
        double Jq[3][3];
        const double inv_det = 1.0 / rt_jacobian(X, d_phi + 12 * q, Jq);
        const double *ctab = curls + ((size_t)o * Q + q) * 18;
        double b[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const double *c = ctab + 3 * j;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                b[k] += (Jq[k][0] * c[0] + Jq[k][1] * c[1] + Jq[k][2] * c[2]) * xv[j] * inv_det;
        }
        B[out] = b[0];
        B[out + 1] = b[1];
        B[out + 2] = b[2];
    }
}

// Launch with a device counter of invalid orientations; synchronises the
// stream to read it back (these entry points are synchronous, like the
// reference FFI calls they stand for).
template <typename Launch>
static int run_counted(hipStream_t s, const char *name, Launch launch) {
    int *bad = nullptr;
    SN_HIP(hipMallocAsync((void **)&bad, sizeof(int), s));
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(int), s);
    if (e == hipSuccess) {
        launch(bad);
        e = hipGetLastError();
    }
    int nbad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&nbad, bad, sizeof(int), hipMemcpyDeviceToHost, s);
    hipError_t ef = hipFreeAsync(bad, s);
    if (e == hipSuccess) e = ef;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        set_error("%s: %s", name, hipGetErrorString(e));
        return SN_ERR_HIP;
    }
    SN_REQUIRE(nbad == 0, "%s: %d element(s) with an orientation outside [0, n_orient)", name,
               nbad);
    return SN_OK;
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_element_table_f64(int num_el, const double *nodes, const int *cells, const int *gids,
                         const int *orient, int n_orient, const double *curls,
                         const double *d_phi, const double *w, int Q, int type_spec, int n_int,
                         const double *coeffs, const double *knots, const double *x,
                         const double *Bs, double *locK, double *locdK, double *loc6,
                         void *stream) {
    SN_REQUIRE(num_el >= 0 && Q >= 1 && n_orient >= 1,
               "sn_element_table_f64: bad size (num_el=%d, Q=%d, n_orient=%d)", num_el, Q,
               n_orient);
    SN_REQUIRE((long long)num_el * 21 < (1LL << 31), "sn_element_table_f64: too many elements");
    SN_REQUIRE(type_spec >= 0 && type_spec <= 2, "unknown reluctivity type %d", type_spec);
    SN_REQUIRE(type_spec == SN_NU_CONST || n_int == -1 || (n_int >= 1 && knots),
               "reluctivity type %d needs knots and at least one interval (n_int=%d)",
               type_spec, n_int);
    if (num_el == 0) return SN_OK;
    SN_REQUIRE(nodes && cells && gids && orient && curls && d_phi && w && coeffs && x && locK &&
                   locdK && loc6,
               "sn_element_table_f64: null pointer");
    hipStream_t s = as_stream(stream);
    return run_counted(s, "sn_element_table_f64", [&](int *bad) {
        hipLaunchKernelGGL(element_table_kernel, dim3((num_el + kRtBlock - 1) / kRtBlock),
                           dim3(kRtBlock), 0, s, num_el, nodes, cells, gids, orient, n_orient,
                           curls, d_phi, w, Q, type_spec, n_int, coeffs, knots, x, Bs, locK,
                           locdK, loc6, bad);
    });
}

int sn_B_Hcurl_table_f64(int num_el, const double *nodes, const int *cells, const int *gids,
                         const int *orient, int n_orient, const double *curls,
                         const double *phi, const double *d_phi, int Q, const double *x,
                         double *pts, double *B, void *stream) {
    SN_REQUIRE(num_el >= 0 && Q >= 1 && n_orient >= 1,
               "sn_B_Hcurl_table_f64: bad size (num_el=%d, Q=%d, n_orient=%d)", num_el, Q,
               n_orient);
    if (num_el == 0) return SN_OK;
    SN_REQUIRE(nodes && cells && gids && orient && curls && phi && d_phi && x && pts && B,
               "sn_B_Hcurl_table_f64: null pointer");
    hipStream_t s = as_stream(stream);
    return run_counted(s, "sn_B_Hcurl_table_f64", [&](int *bad) {
        hipLaunchKernelGGL(b_hcurl_table_kernel, dim3((num_el + kRtBlock - 1) / kRtBlock),
                           dim3(kRtBlock), 0, s, num_el, nodes, cells, gids, orient, n_orient,
                           curls, phi, d_phi, Q, x, pts, B, bad);
    });
}

}  // extern "C"
