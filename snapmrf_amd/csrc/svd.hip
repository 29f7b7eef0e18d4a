// svd.hip -- rank-K (SVD) compressed dictionary matching.
//
// The dictionary compression of McGivney et al., IEEE TMI 2014, as the
// reference's MATLAB path builds it (matlab/MRF_dict_B1.m:125-159): per B1
// slice, the left singular vectors U_r of the normalised atoms with the
// largest singular values; the rank is the number of singular values whose
// cumulative energy fraction is <= sfrac on the first slice; the compressed
// dictionary is U_r^H * atoms.  Matching then compresses each signal with its
// slice's U_r and runs the fused FP32 matcher over K instead of T
// timepoints (the reference's CUDA path has no compressed matching).
//
// The basis comes from the FP64 Gram matrix G = D D^H (rocBLAS zgemm over
// blocks of atoms widened to complex128) and its FP64 eigendecomposition
// (rocSOLVER zheevd, batched over the bases): the eigenvectors of G are the
// left singular vectors of D and its eigenvalues the squared singular values.  The compressions are
// plain complex GEMMs (rocBLAS cgemm).
#include "kernels.h"
#include "snapmrf.h"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <cstring>
#include <map>
#include <vector>

namespace mrf {
namespace {

__global__ void widen_kernel(const float2 *src, double2 *dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = make_double2((double)src[i].x, (double)src[i].y);
}

// basis row k = eigenvector column (T - 1 - k) of G (eigenvalues ascending)
__global__ void take_basis_kernel(const double2 *evec, int T, int K, float2 *basis)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)K * T) return;
    const int k = (int)(i / T), t = (int)(i % T);
    const double2 v = evec[(size_t)(T - 1 - k) * T + t];
    basis[i] = make_float2((float)v.x, (float)v.y);
}

// s'[v] = S'_{slice(v)}[v]: the compression by the voxel's own slice basis
__global__ void pick_slice_kernel(const float2 *all, const int32_t *vox_slice, size_t V, int K,
                                  float2 *out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V * (size_t)K) return;
    const size_t v = i / K;
    out[i] = all[((size_t)vox_slice[v] * V + v) * K + (i % K)];
}

// one rocBLAS handle per (host thread, device), kept for the process
rocblas_handle blas_handle(hipStream_t s, int *rc)
{
    thread_local std::map<int, rocblas_handle> handles;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        *rc = mrf_fail(MRF_E_GPU, "hipGetDevice failed");
        return nullptr;
    }
    rocblas_handle &h = handles[dev];
    if (!h && rocblas_create_handle(&h) != rocblas_status_success) {
        h = nullptr;
        *rc = mrf_fail(MRF_E_GPU, "rocblas_create_handle failed");
        return nullptr;
    }
    if (rocblas_set_stream(h, s) != rocblas_status_success ||
        rocblas_set_pointer_mode(h, rocblas_pointer_mode_host) != rocblas_status_success) {
        *rc = mrf_fail(MRF_E_GPU, "rocblas stream setup failed");
        return nullptr;
    }
    return h;
}

#define MRF_BLAS(call)                                                                  \
    do {                                                                                \
        rocblas_status st_ = (call);                                                    \
        if (st_ != rocblas_status_success)                                              \
            return mrf_fail(MRF_E_GPU, "%s: %s", #call, rocblas_status_to_string(st_)); \
    } while (0)

// U^H X for every slice b: X_b is [n][T] (rows), out_b [n][K]
int compress_rows(rocblas_handle h, const float2 *basis, int K, int T, int nb,
                  const float2 *x, size_t n_per, size_t x_stride_b, float2 *out,
                  size_t out_stride_b)
{
    const rocblas_float_complex one(1.0f, 0.0f), zero(0.0f, 0.0f);
    for (int b = 0; b < nb; b++) {
        MRF_BLAS(rocblas_cgemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none,
                               K, (rocblas_int)n_per, T, &one,
                               (const rocblas_float_complex *)(basis + (size_t)b * K * T), T,
                               (const rocblas_float_complex *)(x + (size_t)b * x_stride_b), T,
                               &zero, (rocblas_float_complex *)(out + (size_t)b * out_stride_b),
                               K));
    }
    return 0;
}

}  // namespace
}  // namespace mrf

using namespace mrf;

extern "C" int mrf_svd_basis(void *stream, const mrf_complex *d_atoms, size_t T, size_t NATOMS,
                             size_t natoms, int per_slice, double energy_frac, size_t K_max,
                             mrf_complex *d_basis, double *h_sv, size_t *K_out)
{
    mrf_clear_error();
    if (!d_atoms || !d_basis || !K_out) return mrf_fail(MRF_E_ARG, "null pointer argument");
    if (T == 0 || T > 0x7fffffff) return mrf_fail(MRF_E_ARG, "bad T");
    if (natoms == 0 || NATOMS == 0 || NATOMS % natoms)
        return mrf_fail(MRF_E_ARG, "NATOMS (%zu) must be a positive multiple of natoms (%zu)",
                        NATOMS, natoms);
    if (K_max == 0 || K_max > T) return mrf_fail(MRF_E_ARG, "need 0 < K_max <= T");
    const int nb = per_slice ? (int)(NATOMS / natoms) : 1;
    const size_t nat = per_slice ? natoms : NATOMS;
    if (nat > 0x7fffffff) return mrf_fail(MRF_E_ARG, "too many atoms per basis");
    hipStream_t s = (hipStream_t)stream;
    int rc = 0;
    rocblas_handle h = blas_handle(s, &rc);
    if (!h) return rc;
    const int Ti = (int)T;
    // atoms per Gram product: at most 1 GiB of complex128 rows
    size_t chunk = ((size_t)1 << 30) / (T * sizeof(double2));
    if (chunk < 1) chunk = 1;
    if (chunk > nat) chunk = nat;
    // the Gram matrices of up to 2 GiB worth of bases go through one batched
    // eigensolver call: its thousands of small launches are shared by the batch
    const size_t gmat = T * T;
    size_t group = ((size_t)2 << 30) / (gmat * sizeof(double2));
    if (group < 1) group = 1;
    if (group > (size_t)nb) group = (size_t)nb;
    double2 *G = nullptr, *buf = nullptr;
    double *W = nullptr, *E = nullptr;
    rocblas_int *info = nullptr;
    if (hipMalloc((void **)&G, group * gmat * sizeof(double2)) != hipSuccess ||
        hipMalloc((void **)&buf, chunk * T * sizeof(double2)) != hipSuccess ||
        hipMalloc((void **)&W, group * T * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&E, group * T * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&info, group * sizeof(rocblas_int)) != hipSuccess)
        rc = mrf_fail(MRF_E_NOMEM, "SVD basis: device allocation failed");
    std::vector<double> w(group * T);
    std::vector<rocblas_int> hinfo(group);
    size_t K = K_max;
    for (int b0 = 0; b0 < nb && !rc; b0 += (int)group) {
        const int gn = (nb - b0 < (int)group) ? nb - b0 : (int)group;
        for (int j = 0; j < gn && !rc; j++) {
            const float2 *src =
                reinterpret_cast<const float2 *>(d_atoms) + (size_t)(b0 + j) * nat * T;
            for (size_t c0 = 0; c0 < nat && !rc; c0 += chunk) {
                const size_t cn = (nat - c0 < chunk) ? nat - c0 : chunk;
                const size_t n = cn * T;
                widen_kernel<<<dim3((unsigned)((n + 255) / 256)), 256, 0, s>>>(src + c0 * T, buf,
                                                                               n);
                // G (+)= D D^H as a full complex GEMM: twice the flops of
                // zherk, several times faster in rocBLAS at these shapes
                const rocblas_double_complex one(1.0, 0.0), beta(c0 ? 1.0 : 0.0, 0.0);
                rocblas_status st = rocblas_zgemm(
                    h, rocblas_operation_none, rocblas_operation_conjugate_transpose, Ti, Ti,
                    (rocblas_int)cn, &one, (const rocblas_double_complex *)buf, Ti,
                    (const rocblas_double_complex *)buf, Ti, &beta,
                    (rocblas_double_complex *)(G + (size_t)j * gmat), Ti);
                if (st != rocblas_status_success)
                    rc = mrf_fail(MRF_E_GPU, "rocblas_zgemm: %s", rocblas_status_to_string(st));
            }
        }
        if (rc) break;
        rocblas_status st = rocsolver_zheevd_strided_batched(
            h, rocblas_evect_original, rocblas_fill_lower, Ti, (rocblas_double_complex *)G, Ti,
            (rocblas_stride)gmat, W, (rocblas_stride)T, E, (rocblas_stride)T, info, gn);
        if (st != rocblas_status_success) {
            rc = mrf_fail(MRF_E_GPU, "rocsolver_zheevd: %s", rocblas_status_to_string(st));
            break;
        }
        if (hipMemcpyAsync(w.data(), W, (size_t)gn * T * sizeof(double), hipMemcpyDeviceToHost,
                           s) != hipSuccess ||
            hipMemcpyAsync(hinfo.data(), info, (size_t)gn * sizeof(rocblas_int),
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = mrf_fail(MRF_E_GPU, "eigenvalue download failed");
            break;
        }
        for (int j = 0; j < gn && !rc; j++) {
            if (hinfo[j] != 0) {
                rc = mrf_fail(MRF_E_GPU, "zheevd did not converge (basis %d, info %d)", b0 + j,
                              (int)hinfo[j]);
                break;
            }
            const int b = b0 + j;
            const double *wj = w.data() + (size_t)j * T;   // ascending
            if (b == 0 && energy_frac > 0.0) {
                // nDictSpace = sum(cumsum(s.^2) / sum(s.^2) <= sfrac), s descending
                double tot = 0.0;
                for (size_t i = 0; i < T; i++) tot += wj[i] > 0.0 ? wj[i] : 0.0;
                double cum = 0.0;
                size_t k = 0;
                for (size_t i = 0; i < T; i++) {
                    const double l = wj[T - 1 - i];
                    cum += l > 0.0 ? l : 0.0;
                    if (tot > 0.0 && cum / tot <= energy_frac) k++;
                    else break;
                }
                K = k < 1 ? 1 : (k > K_max ? K_max : k);
            }
            if (h_sv)
                for (size_t k = 0; k < K; k++) {
                    const double l = wj[T - 1 - k];
                    h_sv[(size_t)b * K_max + k] = l > 0.0 ? std::sqrt(l) : 0.0;
                }
            const size_t nk = K * T;
            take_basis_kernel<<<dim3((unsigned)((nk + 255) / 256)), 256, 0, s>>>(
                G + (size_t)j * gmat, Ti, (int)K,
                reinterpret_cast<float2 *>(d_basis) + (size_t)b * K * T);
        }
    }
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = mrf_fail(MRF_E_GPU, "basis failed");
    for (void *p : {(void *)G, (void *)buf, (void *)W, (void *)E, (void *)info})
        if (p) (void)hipFree(p);
    if (!rc) *K_out = K;
    return rc;
}

extern "C" int mrf_svd_compress(void *stream, const mrf_complex *d_basis, size_t K, size_t T,
                                int per_slice, const mrf_complex *d_atoms, size_t NATOMS,
                                size_t natoms, mrf_complex *d_catoms)
{
    mrf_clear_error();
    if (!d_basis || !d_atoms || !d_catoms) return mrf_fail(MRF_E_ARG, "null pointer argument");
    if (K == 0 || K > T || T > 0x7fffffff) return mrf_fail(MRF_E_ARG, "need 0 < K <= T");
    if (natoms == 0 || NATOMS == 0 || NATOMS % natoms)
        return mrf_fail(MRF_E_ARG, "NATOMS (%zu) must be a positive multiple of natoms (%zu)",
                        NATOMS, natoms);
    const int nb = per_slice ? (int)(NATOMS / natoms) : 1;
    const size_t nat = per_slice ? natoms : NATOMS;
    if (nat > 0x7fffffff) return mrf_fail(MRF_E_ARG, "too many atoms per basis");
    hipStream_t s = (hipStream_t)stream;
    int rc = 0;
    rocblas_handle h = blas_handle(s, &rc);
    if (!h) return rc;
    return compress_rows(h, reinterpret_cast<const float2 *>(d_basis), (int)K, (int)T, nb,
                         reinterpret_cast<const float2 *>(d_atoms), nat, nat * T,
                         reinterpret_cast<float2 *>(d_catoms), nat * K);
}

extern "C" int mrf_svd_match_ex(void *stream, const mrf_complex *d_basis, size_t K, size_t T,
                                int per_slice, const mrf_complex *d_catoms, const float *d_params,
                                size_t NATOMS, size_t natoms, const mrf_complex *d_sig, size_t V,
                                const float *d_b1map, float *d_maps, int32_t *d_idx)
{
    mrf_clear_error();
    if (!d_basis || !d_catoms || !d_params || !d_maps || (V && !d_sig))
        return mrf_fail(MRF_E_ARG, "null pointer argument");
    if (K == 0 || K > T || T > 0x7fffffff) return mrf_fail(MRF_E_ARG, "need 0 < K <= T");
    if (natoms == 0 || NATOMS == 0 || NATOMS % natoms)
        return mrf_fail(MRF_E_ARG, "NATOMS (%zu) must be a positive multiple of natoms (%zu)",
                        NATOMS, natoms);
    const int n_slices = (int)(NATOMS / natoms);
    const int nb = per_slice ? n_slices : 1;
    if (nb > 1 && !d_b1map)
        return mrf_fail(MRF_E_ARG, "per-slice bases need a B1 map to pick each voxel's basis");
    if (V == 0) return MRF_OK;
    if (V > 0x7fffffff) return mrf_fail(MRF_E_ARG, "too many voxels");
    hipStream_t s = (hipStream_t)stream;
    int rc = 0;
    rocblas_handle h = blas_handle(s, &rc);
    if (!h) return rc;
    float2 *all = nullptr, *cs = nullptr;
    int32_t *vs = nullptr;
    if (hipMallocAsync((void **)&all, (size_t)nb * V * K * sizeof(float2), s) != hipSuccess)
        return mrf_fail(MRF_E_NOMEM, "SVD matching: device allocation failed");
    // every basis compresses every signal; the voxel's slice picks one
    for (int b = 0; b < nb && !rc; b++)
        rc = compress_rows(h, reinterpret_cast<const float2 *>(d_basis) + (size_t)b * K * T,
                           (int)K, (int)T, 1, reinterpret_cast<const float2 *>(d_sig), V, 0,
                           all + (size_t)b * V * K, 0);
    cs = all;
    if (!rc && nb > 1) {
        if (hipMallocAsync((void **)&cs, V * K * sizeof(float2), s) != hipSuccess ||
            hipMallocAsync((void **)&vs, V * sizeof(int32_t), s) != hipSuccess)
            rc = mrf_fail(MRF_E_NOMEM, "device allocation failed");
        if (!rc) rc = launch_b1_select(d_b1map, d_params, natoms, n_slices, V, vs, s);
        if (!rc) {
            const size_t n = V * K;
            pick_slice_kernel<<<dim3((unsigned)((n + 255) / 256)), 256, 0, s>>>(all, vs, V,
                                                                              (int)K, cs);
        }
    }
    if (!rc)
        rc = mrf_match_ex(stream, MRF_MATCH_FP32, d_catoms, nullptr, nullptr, d_params, K, NATOMS,
                          natoms, reinterpret_cast<const mrf_complex *>(cs), V, d_b1map, d_maps,
                          d_idx);
    if (cs && cs != all) (void)hipFreeAsync(cs, s);
    if (vs) (void)hipFreeAsync(vs, s);
    (void)hipFreeAsync(all, s);
    return rc;
}
