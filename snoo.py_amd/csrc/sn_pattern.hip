// CSR pattern and element-to-nonzero contribution lists, built on the device.
//
// The reference forms K and dK per domain with scipy csr_array((v,(i,j)))
// from 36 triplets per element, which sums duplicate (i,j) entries
// (fem_c/interface_c.pyx:367-368), then adds the domains
// (snoopy/curl_curl_assembler.py:928-932).  Here the pattern of the union of
// all 6x6 element blocks is built once per mesh with a stable radix sort of
// the 36*N_el (row,col) keys; the sorted order gives, for every nonzero, the
// list of contributing local pair slots in ascending element order, and the
// per-Newton-step scatter becomes a fixed-order gather (deterministic, no
// atomics).
#include <hipcub/hipcub.hpp>

#include "sn_common.hpp"
#include "sn_pattern.hpp"

namespace sn {

__constant__ int c_pair_index[36] = {0,  1,  2,  3,  4,  5,  1,  6,  7,  8,  9,  10,
                                     2,  7,  11, 12, 13, 14, 3,  8,  12, 15, 16, 17,
                                     4,  9,  13, 16, 18, 19, 5,  10, 14, 17, 19, 20};

__global__ void make_pair_keys(int num_el, int n, const int *__restrict__ gids,
                               unsigned long long *__restrict__ keys, int *__restrict__ vals) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)num_el * 36) return;
    const int e = (int)(t / 36), s = (int)(t % 36);
    const int j = s / 6, k = s % 6;
    const unsigned long long gj = (unsigned)gids[6 * e + j], gk = (unsigned)gids[6 * e + k];
    keys[t] = gj * (unsigned long long)n + gk;
    vals[t] = c_pair_index[s] * num_el + e;
}

__global__ void head_flags64(long long N, const unsigned long long *__restrict__ keys,
                             int *__restrict__ flags) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    flags[t] = (t == 0 || keys[t] != keys[t - 1]) ? 1 : 0;
}

__global__ void emit_nonzeros(long long N, int n, const unsigned long long *__restrict__ keys,
                              const int *__restrict__ flags, const int *__restrict__ ids,
                              int *__restrict__ nz_ptr, int *__restrict__ indices,
                              int *__restrict__ row_count) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    if (!flags[t]) return;
    const int p = ids[t] - 1;
    const unsigned long long key = keys[t];
    nz_ptr[p] = (int)t;
    indices[p] = (int)(key % (unsigned long long)n);
    atomicAdd(&row_count[key / (unsigned long long)n], 1);
}

__global__ void find_diag(int n, const int *__restrict__ indptr, const int *__restrict__ indices,
                          int *__restrict__ diag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int d = -1;
    for (int k = indptr[i]; k < indptr[i + 1]; ++k)
        if (indices[k] == i) d = k;
    diag[i] = d;
}

__global__ void make_dof_keys(int num_el, const int *__restrict__ gids, int *__restrict__ keys,
                              int *__restrict__ vals, int *__restrict__ count) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)num_el * 6) return;
    const int e = (int)(t / 6), j = (int)(t % 6);
    const int g = gids[t];
    keys[t] = g;
    vals[t] = j * num_el + e;
    atomicAdd(&count[g], 1);
}

__global__ void set_last(int *p, int idx, int v) { p[idx] = v; }

static int exclusive_scan_int(const int *in, int *out, int n, hipStream_t s) {
    size_t bytes = 0;
    SN_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s));
    void *tmp = nullptr;
    SN_HIP(hipMallocAsync(&tmp, bytes, s));
    SN_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, s));
    SN_HIP(hipFreeAsync(tmp, s));
    return SN_OK;
}

static int bits_for(unsigned long long v) {
    int b = 1;
    while (b < 64 && (v >> b) != 0ULL) ++b;
    return b;
}

}  // namespace sn

using namespace sn;

extern "C" {

int sn_pattern_create(sn_pattern **out, int n, int num_el, const int *gids, void *stream) {
    SN_REQUIRE(out && gids, "sn_pattern_create: null pointer");
    SN_REQUIRE(n > 0 && num_el > 0, "sn_pattern_create: empty mesh (n=%d, num_el=%d)", n, num_el);
    SN_REQUIRE((long long)num_el * 36 < (1LL << 31), "sn_pattern_create: too many elements");
    hipStream_t s = as_stream(stream);
    sn_pattern *P = new sn_pattern();
    P->n = n;
    P->num_el = num_el;
    const long long N = (long long)num_el * 36;
    const int B = 256;

    unsigned long long *k_in = nullptr, *k_out = nullptr;
    int *v_in = nullptr, *flags = nullptr, *ids = nullptr, *row_count = nullptr;
    SN_HIP(hipMallocAsync(&k_in, N * 8, s));
    SN_HIP(hipMallocAsync(&k_out, N * 8, s));
    SN_HIP(hipMallocAsync(&v_in, N * 4, s));
    SN_HIP(hipMalloc(&P->contrib, N * 4));
    hipLaunchKernelGGL(make_pair_keys, dim3(grid_for(N, B, 1 << 30)), dim3(B), 0, s, num_el, n,
                       gids, k_in, v_in);
    SN_CHECK_LAUNCH();
    const int end_bit = bits_for((unsigned long long)n * (unsigned long long)n);
    size_t bytes = 0;
    SN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k_in, k_out, v_in, P->contrib,
                                              (int)N, 0, end_bit, s));
    void *tmp = nullptr;
    SN_HIP(hipMallocAsync(&tmp, bytes, s));
    SN_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, k_in, k_out, v_in, P->contrib, (int)N,
                                              0, end_bit, s));
    SN_HIP(hipFreeAsync(tmp, s));
    SN_HIP(hipFreeAsync(k_in, s));
    SN_HIP(hipFreeAsync(v_in, s));

    SN_HIP(hipMallocAsync(&flags, N * 4, s));
    SN_HIP(hipMallocAsync(&ids, N * 4, s));
    hipLaunchKernelGGL(head_flags64, dim3(grid_for(N, B, 1 << 30)), dim3(B), 0, s, N, k_out,
                       flags);
    SN_CHECK_LAUNCH();
    bytes = 0;
    SN_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, flags, ids, (int)N, s));
    SN_HIP(hipMallocAsync(&tmp, bytes, s));
    SN_HIP(hipcub::DeviceScan::InclusiveSum(tmp, bytes, flags, ids, (int)N, s));
    SN_HIP(hipFreeAsync(tmp, s));
    int nnz = 0;
    SN_HIP(hipMemcpyAsync(&nnz, ids + (N - 1), 4, hipMemcpyDeviceToHost, s));
    SN_HIP(hipStreamSynchronize(s));
    P->nnz = nnz;

    SN_HIP(hipMalloc(&P->nz_ptr, ((size_t)nnz + 1) * 4));
    SN_HIP(hipMalloc(&P->indices, (size_t)nnz * 4));
    SN_HIP(hipMalloc(&P->indptr, ((size_t)n + 1) * 4));
    SN_HIP(hipMalloc(&P->diag, (size_t)n * 4));
    SN_HIP(hipMallocAsync(&row_count, ((size_t)n + 1) * 4, s));
    SN_HIP(hipMemsetAsync(row_count, 0, ((size_t)n + 1) * 4, s));
    hipLaunchKernelGGL(emit_nonzeros, dim3(grid_for(N, B, 1 << 30)), dim3(B), 0, s, N, n, k_out,
                       flags, ids, P->nz_ptr, P->indices, row_count);
    SN_CHECK_LAUNCH();
    hipLaunchKernelGGL(set_last, dim3(1), dim3(1), 0, s, P->nz_ptr, nnz, (int)N);
    int rc = exclusive_scan_int(row_count, P->indptr, n + 1, s);
    if (rc) return rc;
    hipLaunchKernelGGL(find_diag, dim3(grid_for(n, B)), dim3(B), 0, s, n, P->indptr, P->indices,
                       P->diag);
    SN_CHECK_LAUNCH();
    SN_HIP(hipFreeAsync(k_out, s));
    SN_HIP(hipFreeAsync(flags, s));
    SN_HIP(hipFreeAsync(ids, s));
    SN_HIP(hipFreeAsync(row_count, s));

    // per-DOF lists of the 6 local DOF slots, ascending element order
    const long long N6 = (long long)num_el * 6;
    int *dk_in = nullptr, *dk_out = nullptr, *dv_in = nullptr, *cnt = nullptr;
    SN_HIP(hipMallocAsync(&dk_in, N6 * 4, s));
    SN_HIP(hipMallocAsync(&dk_out, N6 * 4, s));
    SN_HIP(hipMallocAsync(&dv_in, N6 * 4, s));
    SN_HIP(hipMallocAsync(&cnt, ((size_t)n + 1) * 4, s));
    SN_HIP(hipMemsetAsync(cnt, 0, ((size_t)n + 1) * 4, s));
    SN_HIP(hipMalloc(&P->dof_contrib, N6 * 4));
    SN_HIP(hipMalloc(&P->dof_ptr, ((size_t)n + 1) * 4));
    hipLaunchKernelGGL(make_dof_keys, dim3(grid_for(N6, B, 1 << 30)), dim3(B), 0, s, num_el, gids,
                       dk_in, dv_in, cnt);
    SN_CHECK_LAUNCH();
    bytes = 0;
    const int dbits = bits_for((unsigned long long)n);
    SN_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, dk_in, dk_out, dv_in,
                                              P->dof_contrib, (int)N6, 0, dbits, s));
    SN_HIP(hipMallocAsync(&tmp, bytes, s));
    SN_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, dk_in, dk_out, dv_in, P->dof_contrib,
                                              (int)N6, 0, dbits, s));
    SN_HIP(hipFreeAsync(tmp, s));
    rc = exclusive_scan_int(cnt, P->dof_ptr, n + 1, s);
    if (rc) return rc;
    SN_HIP(hipFreeAsync(dk_in, s));
    SN_HIP(hipFreeAsync(dk_out, s));
    SN_HIP(hipFreeAsync(dv_in, s));
    SN_HIP(hipFreeAsync(cnt, s));
    SN_HIP(hipStreamSynchronize(s));
    *out = P;
    return SN_OK;
}

void sn_pattern_destroy(sn_pattern *P) {
    if (!P) return;
    (void)hipFree(P->indptr);
    (void)hipFree(P->indices);
    (void)hipFree(P->diag);
    (void)hipFree(P->nz_ptr);
    (void)hipFree(P->contrib);
    (void)hipFree(P->dof_ptr);
    (void)hipFree(P->dof_contrib);
    delete P;
}

int sn_pattern_info(const sn_pattern *P, long long *nnz, const int **indptr, const int **indices,
                    const int **diag) {
    SN_REQUIRE(P, "sn_pattern_info: null pattern");
    if (nnz) *nnz = P->nnz;
    if (indptr) *indptr = P->indptr;
    if (indices) *indices = P->indices;
    if (diag) *diag = P->diag;
    return SN_OK;
}

int sn_pattern_copy(const sn_pattern *P, int *indptr, int *indices, int *diag, void *stream) {
    SN_REQUIRE(P, "sn_pattern_copy: null pattern");
    hipStream_t s = as_stream(stream);
    if (indptr)
        SN_HIP(hipMemcpyAsync(indptr, P->indptr, ((size_t)P->n + 1) * 4, hipMemcpyDeviceToDevice, s));
    if (indices)
        SN_HIP(hipMemcpyAsync(indices, P->indices, (size_t)P->nnz * 4, hipMemcpyDeviceToDevice, s));
    if (diag)
        SN_HIP(hipMemcpyAsync(diag, P->diag, (size_t)P->n * 4, hipMemcpyDeviceToDevice, s));
    return SN_OK;
}

}  // extern "C"
