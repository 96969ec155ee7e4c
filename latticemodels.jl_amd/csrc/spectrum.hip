// Fermi projector initial state: densitymatrix(H, mu; T) for a one-particle
// system (src/spectrum.jl:226-232, 253-257, 317-333):
//   P0 = V * (f.(E) .* V'),  f(E) = 1 if E - mu == 0 at T == 0, else
//   1 / (exp((E - mu)/T) + 1)  (Fermi-Dirac; at T = 0 a step at mu).
// Steps on the device:
//   1. dense H from the device ELL Hamiltonian (never assembled on the host);
//   2. eigen-decomposition with rocSOLVER (ascending E, like LAPACK's zheevr
//      behind the reference's `eigen`): dsyevd on the real matrix when every
//      hopping is real (no magnetic field), eigenvectors widened to complex;
//      zheevd otherwise;
//   3. W = V[:, 1:M] * sqrt(f) for the M columns with f > 0 (a prefix, since f
//      is non-increasing in E);
//   4. P0 = W W^H with a hand-written FP64 MFMA kernel
//      (v_mfma_f64_16x16x4f64), complex product as four real products.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <vector>

#include "lmhip_internal.h"

namespace lmh {

namespace {

__global__ void __launch_bounds__(256) zero_kernel(double2 *__restrict__ A, int64_t count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count; t += stride)
        A[t] = make_double2(0.0, 0.0);
}

__global__ void __launch_bounds__(256) dense_from_ell_kernel(DevHam h, double2 *__restrict__ A, int64_t lda) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= h.n) return;
    A[r * lda + r] = make_double2(h.diag[r], 0.0);
    for (int s = 0; s < h.zmax; ++s) {
        const int64_t q = (int64_t)s * h.n + r;
        const int32_t c = h.col[q];
        if (c >= 0) A[(int64_t)c * lda + r] = h.val[q];   // H[r, c]
    }
}

__global__ void __launch_bounds__(256) scale_columns_kernel(double2 *__restrict__ V, int64_t lda, int64_t n,
                                                            const double *__restrict__ w, int64_t m) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n * m) return;
    const int64_t col = t / n, row = t - col * n;
    const double s = w[col];
    double2 v = V[col * lda + row];
    V[col * lda + row] = make_double2(v.x * s, v.y * s);
}

// P = W W^H, W: n x m (column-major, ldw), P: n x n (ldp).
// Workgroup 256 threads = 4 waves computes a 64x64 tile of P; each wave a
// 32x32 sub-tile as 2x2 MFMA 16x16 blocks.  MFMA f64 16x16x4 lane layout:
//   A (16x4): lane l holds A[l % 16][l / 16]
//   B (4x16): lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, element r holds D[l / 16 + 4 * r][l % 16]
// With B = W^H: B[k][j] = conj(W[j][k]).
//   Re P = Wr Wr^T + Wi Wi^T,  Im P = Wi Wr^T - Wr Wi^T.
typedef double v4d __attribute__((ext_vector_type(4)));

// sgn (may be null): column k of W enters with sign sgn[k] (+-1), so that
// P = V diag(f) V^H for weights f of either sign (W = V sqrt|f|, sgn = sign f).
__global__ void __launch_bounds__(256) herk_mfma_kernel(const double2 *__restrict__ W, int64_t ldw, int64_t n,
                                                        int64_t m, double2 *__restrict__ P, int64_t ldp,
                                                        const double *__restrict__ sgn) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * 64 + (wave & 1) * 32;
    const int64_t j0 = (int64_t)blockIdx.y * 64 + (wave >> 1) * 32;
    const int r16 = lane & 15;
    const int k4 = lane >> 4;
    v4d acc_re[2][2], acc_im[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            acc_re[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
            acc_im[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
        }
    for (int64_t k0 = 0; k0 < m; k0 += 4) {
        const int64_t k = k0 + k4;
        double2 av[2], bv[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int64_t ia = i0 + 16 * t + r16;
            const int64_t jb = j0 + 16 * t + r16;
            av[t] = (k < m && ia < n) ? W[k * ldw + ia] : make_double2(0.0, 0.0);
            bv[t] = (k < m && jb < n) ? W[k * ldw + jb] : make_double2(0.0, 0.0);
        }
        if (sgn && k < m) {
            const double sk = sgn[k];
            bv[0] = make_double2(sk * bv[0].x, sk * bv[0].y);
            bv[1] = make_double2(sk * bv[1].x, sk * bv[1].y);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                acc_re[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a].x, bv[b].x, acc_re[a][b], 0, 0, 0);
                acc_re[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a].y, bv[b].y, acc_re[a][b], 0, 0, 0);
                acc_im[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a].y, bv[b].x, acc_im[a][b], 0, 0, 0);
                acc_im[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[a].x, bv[b].y, acc_im[a][b], 0, 0, 0);
            }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = i0 + 16 * a + k4 + 4 * r;
                const int64_t col = j0 + 16 * b + r16;
                if (row < n && col < n) P[col * ldp + row] = make_double2(acc_re[a][b][r], acc_im[a][b][r]);
            }
}

// per-block max |A - B| over the n x n entries (grid-stride)
__global__ void __launch_bounds__(256) maxdiff_kernel(const double2 *__restrict__ A, const double2 *__restrict__ B,
                                                      int64_t n, int64_t ld, double *__restrict__ part) {
    __shared__ double red[256];
    double m = 0.0;
    const int64_t total = n * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t off = (e / n) * ld + e % n;
        const double2 a = A[off], b = B[off];
        m = fmax(m, hypot(a.x - b.x, a.y - b.y));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// flag[0] |= 1 if any of the n x n entries is not finite
__global__ void __launch_bounds__(256) nonfinite_kernel(const double2 *__restrict__ A, int64_t n, int64_t ld,
                                                        int *__restrict__ flag) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const double2 v = A[(e / n) * ld + e % n];
    if (!isfinite(v.x) || !isfinite(v.y)) atomicOr(flag, 1);
}

// flag[0] |= 1 if a stored off-diagonal entry of H has a nonzero imaginary
// part (the diagonal is real by construction)
__global__ void __launch_bounds__(256) imag_check_kernel(DevHam h, int *__restrict__ flag) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= (int64_t)h.zmax * h.n) return;
    if (h.col[q] >= 0 && h.val[q].y != 0.0) atomicOr(flag, 1);
}

// real dense H (column-major, lda) from an ELL Hamiltonian with real entries
__global__ void __launch_bounds__(256) dense_real_from_ell_kernel(DevHam h, double *__restrict__ A, int64_t lda) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= h.n) return;
    A[r * lda + r] = h.diag[r];
    for (int s = 0; s < h.zmax; ++s) {
        const int64_t q = (int64_t)s * h.n + r;
        const int32_t c = h.col[q];
        if (c >= 0) A[(int64_t)c * lda + r] = h.val[q].x;
    }
}

// A (complex, lda) = R (real, ldr), n x n
__global__ void __launch_bounds__(256) widen_kernel(const double *__restrict__ R, int64_t ldr, double2 *__restrict__ A,
                                                    int64_t lda, int64_t n) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const int64_t c = t / n, r = t - c * n;
    A[c * lda + r] = make_double2(R[c * ldr + r], 0.0);
}

}  // namespace

hipError_t launch_imag_check(const DevHam &h, int *flag, hipStream_t st) {
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), st);
    const int64_t total = (int64_t)h.zmax * h.n;
    if (e != hipSuccess || total == 0) return e;
    hipLaunchKernelGGL(imag_check_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, h, flag);
    return hipGetLastError();
}

hipError_t launch_dense_real_from_ell(const DevHam &h, double *A, int64_t lda, hipStream_t st) {
    hipError_t e = hipMemsetAsync(A, 0, sizeof(double) * lda * h.n, st);
    if (e != hipSuccess || h.n == 0) return e;
    hipLaunchKernelGGL(dense_real_from_ell_kernel, dim3((unsigned)((h.n + 255) / 256)), dim3(256), 0, st, h, A, lda);
    return hipGetLastError();
}

hipError_t launch_widen(const double *R, int64_t ldr, double2 *A, int64_t lda, int64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((n * n + 255) / 256)), dim3(256), 0, st, R, ldr, A, lda, n);
    return hipGetLastError();
}

hipError_t launch_nonfinite(const double2 *A, int64_t n, int64_t ld, int *flag, hipStream_t st) {
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), st);
    if (e != hipSuccess || n == 0) return e;
    hipLaunchKernelGGL(nonfinite_kernel, dim3((unsigned)((n * n + 255) / 256)), dim3(256), 0, st, A, n, ld, flag);
    return hipGetLastError();
}

hipError_t launch_zero(double2 *A, int64_t count, hipStream_t st) {
    if (count == 0) return hipSuccess;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0, st, A, count);
    return hipGetLastError();
}

hipError_t launch_dense_from_ell(const DevHam &h, double2 *A, int64_t lda, hipStream_t st) {
    hipError_t e = launch_zero(A, lda * h.n, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dense_from_ell_kernel, dim3((unsigned)((h.n + 255) / 256)), dim3(256), 0, st, h, A, lda);
    return hipGetLastError();
}

hipError_t launch_scale_columns(double2 *V, int64_t lda, int64_t n, const double *w, int64_t m, hipStream_t st) {
    if (n * m == 0) return hipSuccess;
    hipLaunchKernelGGL(scale_columns_kernel, dim3((unsigned)((n * m + 255) / 256)), dim3(256), 0, st, V, lda, n, w, m);
    return hipGetLastError();
}

hipError_t launch_herk(const double2 *W, int64_t ldw, int64_t n, int64_t m, double2 *P, int64_t ldp, hipStream_t st,
                       const double *sgn) {
    const int64_t nb = (n + 63) / 64;
    hipLaunchKernelGGL(herk_mfma_kernel, dim3((unsigned)nb, (unsigned)nb), dim3(256), 0, st, W, ldw, n, m, P, ldp, sgn);
    return hipGetLastError();
}

// max |P P - P| (dense product on the matrix cores through rocBLAS)
int idempotency(hipStream_t st, const double2 *P, int64_t n, int64_t ld, double2 *W, double *part, int nblocks,
                double *out) {
    rocblas_handle handle = nullptr;
    if (rocblas_create_handle(&handle) != rocblas_status_success) return -1;
    rocblas_set_stream(handle, st);
    const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
    const auto *p = reinterpret_cast<const rocblas_double_complex *>(P);
    rocblas_status s = rocblas_zgemm(handle, rocblas_operation_none, rocblas_operation_none, (rocblas_int)n,
                                     (rocblas_int)n, (rocblas_int)n, &one, p, (rocblas_int)ld, p, (rocblas_int)ld,
                                     &zero, reinterpret_cast<rocblas_double_complex *>(W), (rocblas_int)ld);
    rocblas_destroy_handle(handle);
    if (s != rocblas_status_success) return (int)s;
    hipLaunchKernelGGL(maxdiff_kernel, dim3((unsigned)nblocks), dim3(256), 0, st, W, P, n, ld, part);
    std::vector<double> h(nblocks);
    if (hipMemcpyAsync(h.data(), part, 8 * nblocks, hipMemcpyDeviceToHost, st) != hipSuccess) return -2;
    if (hipStreamSynchronize(st) != hipSuccess) return -2;
    double m = 0.0;
    for (double v : h) m = std::max(m, v);
    *out = m;
    return 0;
}

int eigh(hipStream_t st, int method, double2 *A, int64_t n, int64_t lda, double *E, double *work, int *info) {
    rocblas_handle handle = nullptr;
    if (rocblas_create_handle(&handle) != rocblas_status_success) return -1;
    rocblas_set_stream(handle, st);
    auto *a = reinterpret_cast<rocblas_double_complex *>(A);
    auto *inf = reinterpret_cast<rocblas_int *>(info);
    rocblas_status s;
    if (method == 1)
        s = rocsolver_zheev(handle, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, a, (rocblas_int)lda, E,
                            work, inf);
    else if (method == 2)
        s = rocsolver_zheevdj(handle, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, a, (rocblas_int)lda,
                              E, inf);
    else
        s = rocsolver_zheevd(handle, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, a, (rocblas_int)lda,
                             E, work, inf);
    (void)hipStreamSynchronize(st);
    rocblas_destroy_handle(handle);
    return s == rocblas_status_success ? 0 : (int)s;
}

int eigh_real(hipStream_t st, int method, double *A, int64_t n, int64_t lda, double *E, double *work, int *info) {
    rocblas_handle handle = nullptr;
    if (rocblas_create_handle(&handle) != rocblas_status_success) return -1;
    rocblas_set_stream(handle, st);
    auto *inf = reinterpret_cast<rocblas_int *>(info);
    rocblas_status s;
    if (method == 1)
        s = rocsolver_dsyev(handle, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, A, (rocblas_int)lda, E,
                            work, inf);
    else if (method == 2)
        s = rocsolver_dsyevdj(handle, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, A, (rocblas_int)lda,
                              E, inf);
    else
        s = rocsolver_dsyevd(handle, rocblas_evect_original, rocblas_fill_lower, (rocblas_int)n, A, (rocblas_int)lda,
                             E, work, inf);
    (void)hipStreamSynchronize(st);
    rocblas_destroy_handle(handle);
    return s == rocblas_status_success ? 0 : (int)s;
}

}  // namespace lmh
