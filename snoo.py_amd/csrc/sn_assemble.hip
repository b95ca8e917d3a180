// Element kernels of the curl-curl stiffness and Newton Jacobian, and the
// deterministic scatter (gather through the pattern's contribution lists).
//
// Per element (fem_c/ctools.c:1195-1388 for iron, :960-1193 for air):
//   g_j = J c_j           (J = dx/du, c_j the reference curl of local edge j
//                          with the element's orientation sign)
//   B_q = Bs_q + sum_k g_k x[g_k] / det
//   K_jk  = sum_q nu_q (g_j.g_k) w_q / det
//   dK_jk = sum_q [|B_q| >= 1e-14] nu'_q (g_j.B_q)(g_k.B_q) w_q / (det |B_q|)
//   rhs_j = sum_q (nu0 - nu_q)(g_j.Bs_q) w_q          (iron domains only)
// On an affine tetrahedron g_j and det do not depend on q, so they are formed
// once per element; nu, nu' and B_q vary with q through the source field.
// One thread per element; geometry and DOF data are streamed once, the curl
// signs and reluctivity tables are staged in LDS, and the 21 unique pair
// values are written structure-of-arrays so the writes coalesce.
#include "sn_common.hpp"
#include "sn_pattern.hpp"

namespace sn {

constexpr int kElBlock = 256;
constexpr int kMaxLdsIntervals = 512;
constexpr double kNu0 = 1.0 / 4.0 / 3.141592653589793238462643383279502984 * 1e7;

// MODE_NEWTON_C: the residual of MODE_NEWTON_R, and instead of the 21 pair
// values the Jacobian in reference-curl form.  With c_j = 2 c^_j the curl of
// local edge j on the reference element (c^_j has entries 0 and +-1), s_j the
// edge's orientation sign and Jm the element map's Jacobian, g_j = s_j Jm c_j,
// so J_jk = s_j s_k c^_j . Mr c^_k with the symmetric 3x3
//   Mr = 4 Jm^T (kscale I + sum_q a_q B_q B_q^T) Jm.
// A product of this form keeps the gradient null space of J exact whatever
// the rounding of Mr: c^ applied to a gradient's edge values sums them
// around faces, which cancels exactly.
enum { MODE_LOCAL = 0, MODE_NEWTON_B = 1, MODE_NEWTON_R = 2, MODE_NEWTON_C = 3 };

// Diluted steel (fem_c/ctools.c:124-128, 1565-1759): at a quadrature point
// with physical z, s = (c_num + z m_num) / (c_den + z m_den) scales the
// reluctivity as nu = s nu(s |B|), dnu/d|B| = s^2 nu'(s |B|).
struct Dilution {
    int on;
    double c_num, m_num, c_den, m_den;
};

struct NuModel {
    int type;
    int n;
    const double *c;  // [4*n] (LDS or global)
    const double *x;  // [n+1]
};

// nu(|B|), dnu/d|B| (fem_c/ctools.c:130-204 semantics; type 2 in s = B^2).
__device__ __forceinline__ void nu_eval(const NuModel &m, double B, double &nu, double &dnu) {
    if (m.type == SN_NU_CONST || m.n == -1) {
        nu = m.c[0];
        dnu = 0.0;
        return;
    }
    const double t = (m.type == SN_NU_BSQ) ? B * B : B;
    if (t < m.x[0]) {  // below the first knot the reference returns 0
        nu = 0.0;
        dnu = 0.0;
        return;
    }
    // largest i with x[i] <= t, at most n-1 (extrapolates the last cubic)
    int lo = 0, hi = m.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (m.x[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const double dt = t - m.x[lo];
    const double c0 = m.c[lo], c1 = m.c[m.n + lo], c2 = m.c[2 * m.n + lo], c3 = m.c[3 * m.n + lo];
    nu = ((c0 * dt + c1) * dt + c2) * dt + c3;
    const double d = (3.0 * c0 * dt + 2.0 * c1) * dt + c2;
    dnu = (m.type == SN_NU_BSQ) ? 2.0 * B * d : d;
}

template <typename T>
__device__ __forceinline__ void store_as(void *base, long long idx, double v) {
    reinterpret_cast<T *>(base)[idx] = (T)v;
}

template <int MODE, typename TJ>
__global__ __launch_bounds__(kElBlock) void element_kernel(
    int num_el_total, int el_begin, int el_count, const double *__restrict__ nodes,
    const int *__restrict__ cells, const int *__restrict__ gids, const int *__restrict__ orient,
    const double *__restrict__ curls, const double *__restrict__ w, int Q, int type_spec,
    int n_int, const double *__restrict__ coeffs, const double *__restrict__ knots,
    const double *__restrict__ x, const double *__restrict__ Bs, void *__restrict__ out_a,
    double *__restrict__ out_b, double *__restrict__ out6, float *__restrict__ nu_el,
    const double *__restrict__ phi, Dilution dil) {
    __shared__ double s_curl[64 * 18];
    __shared__ double s_c[4 * kMaxLdsIntervals];
    __shared__ double s_x[kMaxLdsIntervals + 1];
    for (int k = threadIdx.x; k < 64 * 18; k += blockDim.x)
        s_curl[k] = curls[(size_t)(k / 18) * Q * 18 + (k % 18)];
    NuModel m;
    m.type = type_spec;
    m.n = n_int;
    if (type_spec != SN_NU_CONST && n_int > 0 && n_int <= kMaxLdsIntervals) {
        for (int k = threadIdx.x; k < 4 * n_int; k += blockDim.x) s_c[k] = coeffs[k];
        for (int k = threadIdx.x; k <= n_int; k += blockDim.x) s_x[k] = knots[k];
        m.c = s_c;
        m.x = s_x;
    } else {
        m.c = coeffs;
        m.x = knots;
    }
    __syncthreads();

    const int le = blockIdx.x * blockDim.x + threadIdx.x;
    if (le >= el_count) return;
    const int e = el_begin + le;

    // element geometry: J[k][l] = X_{l+1,k} - X_{0,k} (P1 gradients, ctools.c:373-413)
    int nd[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) nd[a] = cells[4 * e + a];
    double X0[3], J[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) X0[k] = nodes[3 * nd[0] + k];
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int k = 0; k < 3; ++k) J[k][l] = nodes[3 * nd[l + 1] + k] - X0[k];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double inv_det = 1.0 / det;

    const int o = orient[e];
    int g[6];
    double xv[6], G[6][3];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        g[j] = gids[6 * e + j];
        xv[j] = x[g[j]];
        const double *c = &s_curl[o * 18 + 3 * j];
#pragma unroll
        for (int k = 0; k < 3; ++k) G[j][k] = J[k][0] * c[0] + J[k][1] * c[1] + J[k][2] * c[2];
    }
    // curl A part of B (constant on the element), ctools.c:922-958
    double Bc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) Bc[k] += G[j][k] * xv[j] * inv_det;

    double sum_nu_w = 0.0, sum_w = 0.0;
    double dK[21];
    double rhs[6];
    double dM[6];   // MODE_NEWTON_C: sum_q a_q t t^T, t = Jm^T B_q (upper triangle)
#pragma unroll
    for (int p = 0; p < 21; ++p) dK[p] = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) rhs[j] = dM[j] = 0.0;

    const double *bs = Bs ? Bs + (size_t)le * 3 * Q : nullptr;
    for (int q = 0; q < Q; ++q) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        if (bs) {
            s0 = bs[3 * q];
            s1 = bs[3 * q + 1];
            s2 = bs[3 * q + 2];
        }
        const double b0 = s0 + Bc[0], b1 = s1 + Bc[1], b2 = s2 + Bc[2];
        const double bm = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
        double nu, dnu;
        if (dil.on) {
            double z = 0.0;
#pragma unroll
            for (int a = 0; a < 4; ++a) z += phi[4 * q + a] * nodes[3 * nd[a] + 2];
            const double sf = (dil.c_num + z * dil.m_num) / (dil.c_den + z * dil.m_den);
            nu_eval(m, sf * bm, nu, dnu);
            nu *= sf;
            dnu *= sf * sf;
        } else {
            nu_eval(m, bm, nu, dnu);
        }
        const double wq = w[q];
        sum_nu_w += nu * wq;
        sum_w += wq;
        if (MODE == MODE_NEWTON_C && bm >= 1e-14) {
            const double a = dnu * wq * inv_det / bm;
            double t[3];
#pragma unroll
            for (int l = 0; l < 3; ++l) t[l] = J[0][l] * b0 + J[1][l] * b1 + J[2][l] * b2;
            dM[0] += a * t[0] * t[0];
            dM[1] += a * t[0] * t[1];
            dM[2] += a * t[0] * t[2];
            dM[3] += a * t[1] * t[1];
            dM[4] += a * t[1] * t[2];
            dM[5] += a * t[2] * t[2];
        } else if (MODE != MODE_NEWTON_C && bm >= 1e-14) {
            const double a = dnu * wq * inv_det / bm;
            double P[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) P[j] = G[j][0] * b0 + G[j][1] * b1 + G[j][2] * b2;
            int p = 0;
#pragma unroll
            for (int j = 0; j < 6; ++j)
#pragma unroll
                for (int k = j; k < 6; ++k) dK[p++] += a * P[j] * P[k];
        }
        if (bs) {
            const double f = (kNu0 - nu) * wq;
#pragma unroll
            for (int j = 0; j < 6; ++j) rhs[j] += f * (G[j][0] * s0 + G[j][1] * s1 + G[j][2] * s2);
        }
    }

    if (nu_el) nu_el[e] = (float)(sum_nu_w / sum_w);  // mean reluctivity (nodal preconditioner)
    const double kscale = sum_nu_w * inv_det;
    double Kx[6], dKx[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) Kx[j] = dKx[j] = 0.0;
    int p = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
#pragma unroll
        for (int k = j; k < 6; ++k) {
            const double kv = kscale * (G[j][0] * G[k][0] + G[j][1] * G[k][1] + G[j][2] * G[k][2]);
            const double dv = dK[p];
            const long long slot = (long long)p * num_el_total + e;
            if (MODE == MODE_LOCAL) {
                reinterpret_cast<double *>(out_a)[slot] = kv;
                out_b[slot] = dv;
            } else {
                // J = dK + K (curl_curl_assembler.py:932)
                if (MODE != MODE_NEWTON_C) store_as<TJ>(out_a, slot, dv + kv);
                Kx[j] += kv * xv[k];
                dKx[j] += dv * xv[k];
                if (k != j) {
                    Kx[k] += kv * xv[j];
                    dKx[k] += dv * xv[j];
                }
            }
            ++p;
        }
    }
    if (MODE == MODE_NEWTON_C) {
        // Mr = 4 (kscale Jm^T Jm + dM), upper triangle (00, 01, 02, 11, 12, 22)
        int u = 0;
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int m2 = l; m2 < 3; ++m2) {
                const double jtj = J[0][l] * J[0][m2] + J[1][l] * J[1][m2] + J[2][l] * J[2][m2];
                store_as<TJ>(out_a, (long long)u * num_el_total + e, 4.0 * (kscale * jtj + dM[u]));
                ++u;
            }
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double v;
        if (MODE == MODE_LOCAL) v = rhs[j];
        else if (MODE == MODE_NEWTON_B) v = dKx[j] + rhs[j];
        else v = Kx[j] - rhs[j];
        out6[(long long)j * num_el_total + e] = v;
    }
}

// ---------------------------------------------------------------------------
// Deterministic scatter: every nonzero sums its contributing local slots in
// ascending element order.
// ---------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ void csr_gather_one_kernel(long long nnz, const int *__restrict__ nz_ptr,
                                      const int *__restrict__ contrib, const TI *__restrict__ loc,
                                      TO *__restrict__ out) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    TI s = 0;
    for (int k = nz_ptr[p]; k < nz_ptr[p + 1]; ++k) s += loc[contrib[k]];
    out[p] = (TO)s;
}

__global__ void csr_gather_kdk_kernel(long long nnz, const int *__restrict__ nz_ptr,
                                      const int *__restrict__ contrib,
                                      const double *__restrict__ locK,
                                      const double *__restrict__ locdK, double *__restrict__ K,
                                      double *__restrict__ dK, double *__restrict__ Jv) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    double sk = 0.0, sd = 0.0;
    for (int k = nz_ptr[p]; k < nz_ptr[p + 1]; ++k) {
        const int c = contrib[k];
        sk += locK[c];
        sd += locdK[c];
    }
    if (K) K[p] = sk;
    if (dK) dK[p] = sd;
    if (Jv) Jv[p] = sd + sk;
}

__global__ void dof_gather_kernel(int n, const int *__restrict__ dof_ptr,
                                  const int *__restrict__ dof_contrib,
                                  const double *__restrict__ loc6, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // four list entries, then their four values, in flight at once; the sum
    // keeps the list order
    double s = 0.0;
    int k = dof_ptr[i];
    const int k1 = dof_ptr[i + 1];
    for (; k + 4 <= k1; k += 4) {
        const int c0 = dof_contrib[k], c1 = dof_contrib[k + 1];
        const int c2 = dof_contrib[k + 2], c3 = dof_contrib[k + 3];
        const double v0 = loc6[c0], v1 = loc6[c1], v2 = loc6[c2], v3 = loc6[c3];
        s += v0;
        s += v1;
        s += v2;
        s += v3;
    }
    for (; k < k1; ++k) s += loc6[dof_contrib[k]];
    out[i] = s;
}

static int check_elem_args(int num_el_total, int el_begin, int el_count, int Q, int type_spec,
                           int n_int, const void *nodes, const void *cells, const void *gids,
                           const void *orient, const void *curls, const void *w,
                           const void *coeffs, const void *knots, const void *x) {
    SN_REQUIRE(el_begin >= 0 && el_count >= 0 && el_begin + el_count <= num_el_total,
               "element range [%d, %d) outside [0, %d)", el_begin, el_begin + el_count,
               num_el_total);
    SN_REQUIRE(Q >= 1, "Q must be >= 1 (Q=%d)", Q);
    SN_REQUIRE(type_spec >= 0 && type_spec <= 2, "unknown reluctivity type %d", type_spec);
    SN_REQUIRE(type_spec == SN_NU_CONST || n_int >= 1,
               "reluctivity type %d needs at least one interval (n_int=%d)", type_spec, n_int);
    SN_REQUIRE((long long)num_el_total * 21 < (1LL << 31), "too many elements");
    SN_REQUIRE(nodes && cells && gids && orient && curls && w && coeffs && x,
               "null pointer argument");
    SN_REQUIRE(type_spec == SN_NU_CONST || knots, "null knots");
    return SN_OK;
}

}  // namespace sn

using namespace sn;

extern "C" {

static int make_dilution(const double *dil_host, const double *phi, const double *Bs,
                         Dilution &d) {
    d.on = 0;
    d.c_num = d.m_num = d.c_den = d.m_den = 0.0;
    // as the reference, the dilution applies to source (iron) domains only
    if (!dil_host || !Bs) return SN_OK;
    SN_REQUIRE(phi, "dilution needs the P1 basis values phi [Q*4]");
    d.on = 1;
    d.c_num = dil_host[0];
    d.m_num = dil_host[1];
    d.c_den = dil_host[2];
    d.m_den = dil_host[3];
    return SN_OK;
}

int sn_element_local_f64(int num_el_total, int el_begin, int el_count, const double *nodes,
                         const int *cells, const int *gids, const int *orient,
                         const double *curls, const double *d_phi, const double *w, int Q,
                         int type_spec, int n_int, const double *coeffs, const double *knots,
                         const double *x, const double *Bs, double *locK, double *locdK,
                         double *loc6, const double *phi, const double *dil_host, void *stream) {
    (void)d_phi;
    int rc = check_elem_args(num_el_total, el_begin, el_count, Q, type_spec, n_int, nodes, cells,
                             gids, orient, curls, w, coeffs, knots, x);
    if (rc) return rc;
    if (el_count == 0) return SN_OK;
    SN_REQUIRE(locK && locdK && loc6, "sn_element_local_f64: null output");
    Dilution dil;
    if ((rc = make_dilution(dil_host, phi, Bs, dil))) return rc;
    hipLaunchKernelGGL((element_kernel<MODE_LOCAL, double>),
                       dim3((el_count + kElBlock - 1) / kElBlock), dim3(kElBlock), 0,
                       as_stream(stream), num_el_total, el_begin, el_count, nodes, cells, gids,
                       orient, curls, w, Q, type_spec, n_int, coeffs, knots, x, Bs,
                       (void *)locK, locdK, loc6, (float *)nullptr, phi, dil);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_element_newton(int out_f32, int mode, int num_el_total, int el_begin, int el_count,
                      const double *nodes, const int *cells, const int *gids, const int *orient,
                      const double *curls, const double *d_phi, const double *w, int Q,
                      int type_spec, int n_int, const double *coeffs, const double *knots,
                      const double *x, const double *Bs, void *locJ, double *loc6,
                      float *nu_el, const double *phi, const double *dil_host, void *stream) {
    (void)d_phi;
    int rc = check_elem_args(num_el_total, el_begin, el_count, Q, type_spec, n_int, nodes, cells,
                             gids, orient, curls, w, coeffs, knots, x);
    if (rc) return rc;
    SN_REQUIRE(mode >= 0 && mode <= 2, "sn_element_newton: mode must be 0, 1 or 2");
    if (el_count == 0) return SN_OK;
    SN_REQUIRE(locJ && loc6, "sn_element_newton: null output");
    Dilution dil;
    if ((rc = make_dilution(dil_host, phi, Bs, dil))) return rc;
    const dim3 grid((el_count + kElBlock - 1) / kElBlock), block(kElBlock);
    hipStream_t s = as_stream(stream);
#define SN_LAUNCH_EL(M, T)                                                                     \
    hipLaunchKernelGGL((element_kernel<M, T>), grid, block, 0, s, num_el_total, el_begin,     \
                       el_count, nodes, cells, gids, orient, curls, w, Q, type_spec, n_int,    \
                       coeffs, knots, x, Bs, locJ, (double *)nullptr, loc6, nu_el, phi, dil)
    if (mode == 0) {
        if (out_f32) SN_LAUNCH_EL(MODE_NEWTON_B, float);
        else SN_LAUNCH_EL(MODE_NEWTON_B, double);
    } else if (mode == 1) {
        if (out_f32) SN_LAUNCH_EL(MODE_NEWTON_R, float);
        else SN_LAUNCH_EL(MODE_NEWTON_R, double);
    } else {
        if (out_f32) SN_LAUNCH_EL(MODE_NEWTON_C, float);
        else SN_LAUNCH_EL(MODE_NEWTON_C, double);
    }
#undef SN_LAUNCH_EL
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_csr_gather_f64(const sn_pattern *P, const double *locK, const double *locdK, double *K,
                      double *dK, double *Jv, void *stream) {
    SN_REQUIRE(P && locK && locdK, "sn_csr_gather_f64: null argument");
    hipLaunchKernelGGL(csr_gather_kdk_kernel, dim3(grid_for(P->nnz, 256, 1 << 30)), dim3(256), 0,
                       as_stream(stream), P->nnz, P->nz_ptr, P->contrib, locK, locdK, K, dK, Jv);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_csr_gather_one(const sn_pattern *P, int f32, const void *loc, void *vals, void *stream) {
    SN_REQUIRE(P && loc && vals, "sn_csr_gather_one: null argument");
    const dim3 grid(grid_for(P->nnz, 256, 1 << 30));
    if (f32)
        hipLaunchKernelGGL((csr_gather_one_kernel<float, float>), grid, dim3(256), 0,
                           as_stream(stream), P->nnz, P->nz_ptr, P->contrib,
                           (const float *)loc, (float *)vals);
    else
        hipLaunchKernelGGL((csr_gather_one_kernel<double, double>), grid, dim3(256), 0,
                           as_stream(stream), P->nnz, P->nz_ptr, P->contrib,
                           (const double *)loc, (double *)vals);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_dof_gather_f64(const sn_pattern *P, const double *loc6, double *out, void *stream) {
    SN_REQUIRE(P && loc6 && out, "sn_dof_gather_f64: null argument");
    hipLaunchKernelGGL(dof_gather_kernel, dim3(grid_for(P->n, 256)), dim3(256), 0,
                       as_stream(stream), P->n, P->dof_ptr, P->dof_contrib, loc6, out);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

}  // extern "C"
