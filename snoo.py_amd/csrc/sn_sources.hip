// Source fields and field evaluation: Biot-Savart, brick field, quadrature
// points and curl A at quadrature points.
#include "sn_common.hpp"

namespace sn {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

int num_cus() {
    static int cached = 0;
    if (cached == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cached = prop.multiProcessorCount;
        if (cached <= 0) cached = 256;
    }
    return cached;
}

// ---------------------------------------------------------------------------
// Biot-Savart (fem_c/ctools.c:1761-1923).
// One thread per target; segments staged through LDS in tiles.  Each target
// sums its segments in order, the reference's single-thread order, and the
// sum is scaled by current*1e-7 once at the end (ctools.c:1908-1912).
// ---------------------------------------------------------------------------
constexpr int kBsBlock = 256;
constexpr int kBsTile = 256;

// Per segment, staged in LDS once per tile: start and end point, direction
// n = end - start, 1/|n|^2, and a sphere around the segment's midpoint that
// holds every point within the core radius of it, with a margin.  A target
// outside that sphere is farther than the core radius from the segment and
// from both end points, so the core test of the reference is false there and
// is skipped; inside it the test runs as the reference writes it.
//
// Coils are polylines: a segment whose start point is bit-for-bit the end
// point of the segment before it takes that segment's end-point difference
// and distance instead of forming them again (the same values: same inputs,
// same operations), which saves one square root per pair.
//
// A workgroup's targets are neighbours (consecutive quadrature points): when
// the box around them lies outside a segment's sphere, with a relative margin
// of 1e-12 that covers the rounding of either distance, every target's own
// sphere test is false and is skipped for the whole workgroup.
constexpr int kBsSegVals = 14;

struct BsVertex {
    double dx, dy, dz, l;
};

__device__ __forceinline__ BsVertex bs_vertex(const double *p, double tx, double ty, double tz) {
    BsVertex v;
    v.dx = p[0] - tx;
    v.dy = p[1] - ty;
    v.dz = p[2] - tz;
    v.l = sqrt(v.dx * v.dx + v.dy * v.dy + v.dz * v.dz);
    return v;
}

__global__ __launch_bounds__(kBsBlock) void biot_savart_kernel(
    const double *__restrict__ src, int num_src, const double *__restrict__ tar,
    int num_tar, double scale, double rad, double *__restrict__ B, int accumulate) {
    __shared__ double seg[kBsTile * kBsSegVals];
    __shared__ int flags[kBsTile];              // bit 0: linked, bit 1: box near the sphere
    __shared__ double box[6][kBsBlock / kWave];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double tx = 0.0, ty = 0.0, tz = 0.0;
    if (i < num_tar) {
        tx = tar[3 * i];
        ty = tar[3 * i + 1];
        tz = tar[3 * i + 2];
    }
    {
        // the workgroup's bounding box: -max(-x) and max(x) per axis; a NaN
        // coordinate makes the box NaN and every segment "near"
        const bool has = i < num_tar;
        double b[6] = {has ? -tx : -INFINITY, has ? -ty : -INFINITY, has ? -tz : -INFINITY,
                       has ? tx : -INFINITY, has ? ty : -INFINITY, has ? tz : -INFINITY};
        const int w = threadIdx.x / kWave;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double v = b[k];
            for (int off = 32; off > 0; off >>= 1) {
                const double o = __shfl_xor(v, off, kWave);
                v = (o > v || o != o) ? o : v;
            }
            if (threadIdx.x % kWave == 0) box[k][w] = v;
        }
    }
    __syncthreads();
    double lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double a = box[k][0], c = box[k + 3][0];
        for (int w = 1; w < kBsBlock / kWave; ++w) {
            a = (box[k][w] > a || box[k][w] != box[k][w]) ? box[k][w] : a;
            c = (box[k + 3][w] > c || box[k + 3][w] != box[k + 3][w]) ? box[k + 3][w] : c;
        }
        lo[k] = -a;
        hi[k] = c;
    }
    double ax = 0.0, ay = 0.0, az = 0.0;
    const double r2 = rad * rad;
    BsVertex prev = {0.0, 0.0, 0.0, 0.0};       // end point of the previous segment
    for (int base = 0; base < num_src; base += kBsTile) {
        const int cnt = min(kBsTile, num_src - base);
        __syncthreads();
        for (int j = threadIdx.x; j < cnt; j += blockDim.x) {
            const long long g_ix = base + j;
            const double *g = src + 6 * g_ix;
            double *s = &seg[kBsSegVals * j];
            const double x0 = g[0], y0 = g[1], z0 = g[2], x1 = g[3], y1 = g[4], z1 = g[5];
            const double nx = x1 - x0, ny = y1 - y0, nz = z1 - z0;
            const double nn = nx * nx + ny * ny + nz * nz;
            const double half = 0.5 * sqrt(nn);
            const double R = (half + rad) * (1.0 + 1e-6) + 1e-12;
            s[0] = x0; s[1] = y0; s[2] = z0;
            s[3] = x1; s[4] = y1; s[5] = z1;
            s[6] = nx; s[7] = ny; s[8] = nz;
            s[9] = nn;
            s[10] = 0.5 * (x0 + x1); s[11] = 0.5 * (y0 + y1); s[12] = 0.5 * (z0 + z1);
            s[13] = R * R;
            // bitwise equality of the coordinates (not ==, which equates 0 and -0)
            const bool linked = g_ix > 0 &&
                                __double_as_longlong(g[-3]) == __double_as_longlong(x0) &&
                                __double_as_longlong(g[-2]) == __double_as_longlong(y0) &&
                                __double_as_longlong(g[-1]) == __double_as_longlong(z0);
            // distance from the midpoint to the box; far only when it is a number
            // above the sphere's radius with the margin
            const double gx = fmax(fmax(lo[0] - s[10], s[10] - hi[0]), 0.0);
            const double gy = fmax(fmax(lo[1] - s[11], s[11] - hi[1]), 0.0);
            const double gz = fmax(fmax(lo[2] - s[12], s[12] - hi[2]), 0.0);
            const double g2 = gx * gx + gy * gy + gz * gz;
            const bool far = g2 > s[13] * (1.0 + 1e-12) && lo[0] == lo[0] && lo[1] == lo[1] &&
                             lo[2] == lo[2] && hi[0] == hi[0] && hi[1] == hi[1] &&
                             hi[2] == hi[2];
            flags[j] = (linked ? 1 : 0) | (far ? 0 : 2);
        }
        __syncthreads();
        if (i < num_tar) {
            for (int j = 0; j < cnt; ++j) {
                const double *s = &seg[kBsSegVals * j];
                const double nx = s[6], ny = s[7], nz = s[8];
                // the same for every lane: a scalar branch
                const int fl = __builtin_amdgcn_readfirstlane(flags[j]);
                const BsVertex v1 = (fl & 1) ? prev : bs_vertex(s, tx, ty, tz);
                const BsVertex v2 = bs_vertex(s + 3, tx, ty, tz);
                prev = v2;
                const double d1x = v1.dx, d1y = v1.dy, d1z = v1.dz, l1 = v1.l;
                const double d2x = v2.dx, d2y = v2.dy, d2z = v2.dz, l2 = v2.l;
                bool core = false;
                double px = 0.0, py = 0.0, pz = 0.0, dd = 0.0;
                if (fl & 2) {
                    const double mx = tx - s[10], my = ty - s[11], mz = tz - s[12];
                    if (mx * mx + my * my + mz * mz <= s[13]) {
                        const double tp = -1.0 * (d1x * nx + d1y * ny + d1z * nz) / s[9];
                        px = tx - (s[0] + tp * nx);
                        py = ty - (s[1] + tp * ny);
                        pz = tz - (s[2] + tp * nz);
                        dd = px * px + py * py + pz * pz;
                        core = (dd < r2 && tp <= 1.0 && tp >= 0.0) || l1 < rad || l2 < rad;
                    }
                }
                if (core) {
                    // inside the filament core: mu0 I r/(2 pi R^2) along n x p
                    const double bx = ny * pz - nz * py;
                    const double by = nz * px - nx * pz;
                    const double bz = nx * py - ny * px;
                    const double nb = sqrt(bx * bx + by * by + bz * bz);
                    const double f = 2.0 * sqrt(dd) / rad / rad / nb;
                    ax += bx * f;
                    ay += by * f;
                    az += bz * f;
                } else {
                    // finite straight segment (Russenschuck 5.51)
                    const double d12 = d1x * d2x + d1y * d2y + d1z * d2z;
                    const double f = (l1 + l2) / ((l1 * l2 + d12) * l1 * l2);
                    ax += (d1y * d2z - d1z * d2y) * f;
                    ay += (d1z * d2x - d1x * d2z) * f;
                    az += (d1x * d2y - d1y * d2x) * f;
                }
            }
        }
    }
    if (i < num_tar) {
        double b0 = accumulate ? B[3 * i] : 0.0;
        double b1 = accumulate ? B[3 * i + 1] : 0.0;
        double b2 = accumulate ? B[3 * i + 2] : 0.0;
        B[3 * i] = b0 + ax * scale;
        B[3 * i + 1] = b1 + ay * scale;
        B[3 * i + 2] = b2 + az * scale;
    }
}

// Brick field (snoopy/brick_field.py:400-433): first brick whose six faces all
// have (p - r).n <= 0 sets the field.
struct BrickTable {
    double face[3 * 6 * 6];
    double field[9];
};

__global__ void brick_field_kernel(BrickTable tab, const double *__restrict__ tar,
                                   int num_tar, double *__restrict__ B, int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_tar) return;
    const double px = tar[3 * i], py = tar[3 * i + 1], pz = tar[3 * i + 2];
    double b[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; ++d) {
        bool inside = true;
        for (int f = 0; f < 6; ++f) {
            const double *F = &tab.face[(d * 6 + f) * 6];
            const double proj = (px - F[0]) * F[3] + (py - F[1]) * F[4] + (pz - F[2]) * F[5];
            if (proj > 0.0) inside = false;
        }
        if (inside) {
            b[0] = tab.field[3 * d];
            b[1] = tab.field[3 * d + 1];
            b[2] = tab.field[3 * d + 2];
            break;
        }
    }
    for (int c = 0; c < 3; ++c) B[3 * i + c] = (accumulate ? B[3 * i + c] : 0.0) + b[c];
}

// ---------------------------------------------------------------------------
// Quadrature points and curl A at quadrature points.
// ---------------------------------------------------------------------------
__global__ void quad_points_kernel(int num_el, const double *__restrict__ nodes,
                                   const int *__restrict__ cells, const double *__restrict__ phi,
                                   int Q, double *__restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)num_el * Q) return;
    const int e = (int)(t / Q), q = (int)(t % Q);
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    for (int a = 0; a < 4; ++a) {
        const int n = cells[4 * e + a];
        const double f = phi[4 * q + a];
        p0 += f * nodes[3 * n];
        p1 += f * nodes[3 * n + 1];
        p2 += f * nodes[3 * n + 2];
    }
    out[3 * t] = p0;
    out[3 * t + 1] = p1;
    out[3 * t + 2] = p2;
}

// curl A is constant on an affine tetrahedron: B = sum_k g_k x_k / det with
// g_k = J c_k (fem_c/ctools.c:922-958); the reference evaluates it per q.
__global__ __launch_bounds__(256) void b_hcurl_kernel(
    int num_el, const double *__restrict__ nodes, const int *__restrict__ cells,
    const int *__restrict__ gids, const int *__restrict__ orient, const double *__restrict__ curls,
    const double *__restrict__ phi, int Q, const double *__restrict__ x,
    double *__restrict__ pts, double *__restrict__ B) {
    // q = 0 slice of every orientation, curls[((o*Q)*6 + j)*3 + c], staged in LDS
    __shared__ double curl0[64 * 18];
    for (int k = threadIdx.x; k < 64 * 18; k += blockDim.x)
        curl0[k] = curls[(size_t)(k / 18) * Q * 18 + (k % 18)];
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= num_el) return;
    double X[4][3];
    for (int a = 0; a < 4; ++a) {
        const int n = cells[4 * e + a];
        X[a][0] = nodes[3 * n];
        X[a][1] = nodes[3 * n + 1];
        X[a][2] = nodes[3 * n + 2];
    }
    double J[3][3];
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) J[k][l] = X[l + 1][k] - X[0][k];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const int o = orient[e];
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
    for (int k = 0; k < 6; ++k) {
        const double *c = &curl0[(o * 6 + k) * 3];
        const double xv = x[gids[6 * e + k]];
        b0 += (J[0][0] * c[0] + J[0][1] * c[1] + J[0][2] * c[2]) * xv / det;
        b1 += (J[1][0] * c[0] + J[1][1] * c[1] + J[1][2] * c[2]) * xv / det;
        b2 += (J[2][0] * c[0] + J[2][1] * c[1] + J[2][2] * c[2]) * xv / det;
    }
    for (int q = 0; q < Q; ++q) {
        const long long t = (long long)e * Q + q;
        double p0 = 0.0, p1 = 0.0, p2 = 0.0;
        for (int a = 0; a < 4; ++a) {
            const double f = phi[4 * q + a];
            p0 += f * X[a][0];
            p1 += f * X[a][1];
            p2 += f * X[a][2];
        }
        pts[3 * t] = p0;
        pts[3 * t + 1] = p1;
        pts[3 * t + 2] = p2;
        B[3 * t] = b0;
        B[3 * t + 1] = b1;
        B[3 * t + 2] = b2;
    }
}


// ------------------------------------------------------------ probe fields
// Nodal fields by element averaging (snoopy/curl_curl_assembler.py:2089-2197):
// out[i] = mean of the element vectors listed for node i (0 when none); the
// lists are in ascending element order, so the sum is deterministic.
__global__ void nodal_average_kernel(int n_nodes, const int *__restrict__ node_ptr,
                                     const int *__restrict__ node_el,
                                     const double *__restrict__ Be, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    const int k0 = node_ptr[i], k1 = node_ptr[i + 1];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = k0; k < k1; ++k) {
        const int e = node_el[k];
        s0 += Be[3 * e];
        s1 += Be[3 * e + 1];
        s2 += Be[3 * e + 2];
    }
    const double c = (k1 > k0) ? (double)(k1 - k0) : 1.0;
    out[3 * i] = s0 / c;
    out[3 * i + 1] = s1 / c;
    out[3 * i + 2] = s2 / c;
}

// P1 interpolation of one of several nodal vector fields at points with a
// known containing element (snoopy/evaluators.py:297-395): barycentric
// coordinates from the element's nodes, field chosen per point; points
// without an element get 0.
__global__ void p1_interp_kernel(int n_pts, const double *__restrict__ pts,
                                 const int *__restrict__ el, const int *__restrict__ field,
                                 int n_nodes, const double *__restrict__ nodes,
                                 const int *__restrict__ cells, const double *__restrict__ nodal,
                                 double *__restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pts) return;
    const int e = el[p], f = field[p];
    double r[3] = {0.0, 0.0, 0.0};
    if (e >= 0 && f >= 0) {
        int nd[4];
        double X[4][3];
        for (int a = 0; a < 4; ++a) {
            nd[a] = cells[4 * e + a];
            for (int k = 0; k < 3; ++k) X[a][k] = nodes[3 * nd[a] + k];
        }
        // columns e1..e3 of the element map; lambda_1..3 = E^-1 (x - X0)
        double E[3][3], d[3];
        for (int k = 0; k < 3; ++k) {
            for (int c = 0; c < 3; ++c) E[k][c] = X[c + 1][k] - X[0][k];
            d[k] = pts[3 * p + k] - X[0][k];
        }
        const double det = E[0][0] * (E[1][1] * E[2][2] - E[1][2] * E[2][1]) -
                           E[0][1] * (E[1][0] * E[2][2] - E[1][2] * E[2][0]) +
                           E[0][2] * (E[1][0] * E[2][1] - E[1][1] * E[2][0]);
        double lam[4];
        // Cramer's rule for E lam = d
        for (int c = 0; c < 3; ++c) {
            double M[3][3];
            for (int k = 0; k < 3; ++k)
                for (int j = 0; j < 3; ++j) M[k][j] = (j == c) ? d[k] : E[k][j];
            lam[c + 1] = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                          M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                          M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])) / det;
        }
        lam[0] = 1.0 - lam[1] - lam[2] - lam[3];
        const double *F = nodal + (size_t)f * n_nodes * 3;
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 3; ++k) r[k] += lam[a] * F[3 * nd[a] + k];
    }
    for (int k = 0; k < 3; ++k) out[3 * p + k] = r[k];
}

}  // namespace sn

using namespace sn;

extern "C" {

const char *sn_last_error(void) { return g_err.c_str(); }

int sn_version(void) { return 1; }

int sn_biot_savart_f64(const double *src, int num_src, const double *tar, int num_tar,
                       double current, double radius, double *B, int accumulate,
                       void *stream) {
    SN_REQUIRE(num_src >= 0 && num_tar >= 0, "sn_biot_savart_f64: negative size");
    if (num_tar == 0) return SN_OK;
    SN_REQUIRE(tar && B && (num_src == 0 || src), "sn_biot_savart_f64: null pointer");
    hipStream_t s = as_stream(stream);
    const int grid = (num_tar + kBsBlock - 1) / kBsBlock;
    hipLaunchKernelGGL(biot_savart_kernel, dim3(grid), dim3(kBsBlock), 0, s, src, num_src,
                       tar, num_tar, current * 1e-7, radius, B, accumulate);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_brick_field_f64(const double *faces_host, const double *fields_host, const double *tar,
                       int num_tar, double *B, int accumulate, void *stream) {
    SN_REQUIRE(num_tar >= 0, "sn_brick_field_f64: negative size");
    if (num_tar == 0) return SN_OK;
    SN_REQUIRE(faces_host && fields_host && tar && B, "sn_brick_field_f64: null pointer");
    BrickTable tab;
    for (int k = 0; k < 108; ++k) tab.face[k] = faces_host[k];
    for (int k = 0; k < 9; ++k) tab.field[k] = fields_host[k];
    hipLaunchKernelGGL(brick_field_kernel, dim3((num_tar + 255) / 256), dim3(256), 0,
                       as_stream(stream), tab, tar, num_tar, B, accumulate);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_quad_points_f64(int num_el, const double *nodes, const int *cells,
                       const double *phi, int Q, double *out, void *stream) {
    SN_REQUIRE(num_el >= 0 && Q >= 1, "sn_quad_points_f64: bad size (Q=%d)", Q);
    if (num_el == 0) return SN_OK;
    SN_REQUIRE(nodes && cells && phi && out, "sn_quad_points_f64: null pointer");
    const long long n = (long long)num_el * Q;
    hipLaunchKernelGGL(quad_points_kernel, dim3(grid_for(n, 256, 1 << 30)), dim3(256), 0,
                       as_stream(stream), num_el, nodes, cells, phi, Q, out);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_B_Hcurl_f64(int num_el, const double *nodes, const int *cells, const int *gids,
                   const int *orient, const double *curls, const double *phi, int Q,
                   const double *x, double *pts, double *B, void *stream) {
    SN_REQUIRE(num_el >= 0 && Q >= 1, "sn_B_Hcurl_f64: bad size (Q=%d)", Q);
    if (num_el == 0) return SN_OK;
    SN_REQUIRE(nodes && cells && gids && orient && curls && phi && x && pts && B,
               "sn_B_Hcurl_f64: null pointer");
    hipLaunchKernelGGL(b_hcurl_kernel, dim3((num_el + 255) / 256), dim3(256), 0,
                       as_stream(stream), num_el, nodes, cells, gids, orient, curls, phi, Q, x,
                       pts, B);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_nodal_average_f64(int n_nodes, const int *node_ptr, const int *node_el, const double *Be,
                         double *out, void *stream) {
    SN_REQUIRE(n_nodes > 0 && node_ptr && node_el && Be && out,
               "sn_nodal_average_f64: bad argument");
    hipLaunchKernelGGL(nodal_average_kernel, dim3(grid_for(n_nodes, 256)), dim3(256), 0,
                       as_stream(stream), n_nodes, node_ptr, node_el, Be, out);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

int sn_p1_interp_f64(int n_pts, const double *pts, const int *el, const int *field, int n_nodes,
                     const double *nodes, const int *cells, const double *nodal, double *out,
                     void *stream) {
    SN_REQUIRE(n_pts >= 0 && n_nodes > 0, "sn_p1_interp_f64: bad sizes");
    if (n_pts == 0) return SN_OK;
    SN_REQUIRE(pts && el && field && nodes && cells && nodal && out,
               "sn_p1_interp_f64: null argument");
    hipLaunchKernelGGL(p1_interp_kernel, dim3(grid_for(n_pts, 256)), dim3(256), 0,
                       as_stream(stream), n_pts, pts, el, field, n_nodes, nodes, cells, nodal,
                       out);
    SN_CHECK_LAUNCH();
    return SN_OK;
}

}  // extern "C"
