// Lanczos kernels of the Krylov ket propagator (lmh_krylov_step): for every
// ket column c at once, with the basis v_j stored as (krylovdim + 1) blocks of
// ld x m complex entries and the unnormalised next vector w_j in one of two
// scratch blocks,
//     beta_j = |w_j|,  v_j = w_j / beta_j                (w_0 = psi)
//     w = H v_j - beta_j v_{j-1},  alpha_j = <v_j, w>
//     w_{j+1} = w - alpha_j v_j
// H is the context's ELL Hamiltonian (diagonal + zmax slots), applied row-wise
// with the neighbours gathered through L1/L2 (one ket column is N x 16 B; the
// matrix-free stencil is bound by the basis traffic, not by arithmetic).
// Reductions over rows go through per-block partial sums; the kernel after a
// reduction sums the partials of its column itself (nb <= a few hundred), so
// an iteration is two launches (the normalisation of w_j rides with the
// product by H) and no host round trip.  w_j is read by the neighbours of a
// row while v_j is written, hence two scratch blocks used alternately.
#include "lmhip_internal.h"

namespace lmh {

namespace {

constexpr int kLzThreads = 256;

__device__ __forceinline__ double block_sum(double v, double *red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int w = kLzThreads / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// sum of a column's nb partials, by the whole block (fixed order)
__device__ __forceinline__ double column_total(const double *part, int nb, double *red) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += kLzThreads) acc += part[b];
    return block_sum(acc, red);
}

// v_1 = psi / |psi| for every column: partial |psi|^2 per block
__global__ void __launch_bounds__(kLzThreads) lanczos_norm_kernel(const double2 *__restrict__ X, int64_t n, int64_t ld,
                                                                 double *__restrict__ part, int nb) {
    __shared__ double red[kLzThreads];
    const int64_t c = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * kLzThreads + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        const double2 v = X[c * ld + i];
        s = v.x * v.x + v.y * v.y;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[c * nb + blockIdx.x] = s;
}

// beta[c] = |w| from the column's partials of |w|^2 (one block per column)
__global__ void __launch_bounds__(kLzThreads) lanczos_beta_kernel(const double *__restrict__ part, int nb,
                                                                 double *__restrict__ beta_out) {
    __shared__ double red[kLzThreads];
    const int64_t c = blockIdx.x;
    const double b = sqrt(column_total(part + c * nb, nb, red));
    if (threadIdx.x == 0) beta_out[c] = b;
}

// beta_j = |W| (partials part_w); v_j = W / beta_j into Vj (0 stays 0), block 0
// recording beta_j; Wn = H v_j - beta_j v_{j-1}; partials of <v_j, Wn>
template <int ZR>
__global__ void __launch_bounds__(kLzThreads) lanczos_matvec_kernel(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag, int64_t n,
    int zmax, const double2 *__restrict__ W, const double *__restrict__ part_w, double2 *__restrict__ Vj,
    const double2 *__restrict__ Vprev, double2 *__restrict__ Wn, int64_t ld, double *__restrict__ beta_out,
    double *__restrict__ part_re, double *__restrict__ part_im, int nb) {
    __shared__ double red[kLzThreads];
    const int64_t c = blockIdx.y;
    const double b = sqrt(column_total(part_w + c * nb, nb, red));
    const double inv = b > 0.0 ? 1.0 / b : 0.0;
    const int64_t i = (int64_t)blockIdx.x * kLzThreads + threadIdx.x;
    double dre = 0.0, dim = 0.0;
    if (i < n) {
        const double2 *w = W + c * ld;
        const double2 wi = w[i];
        const double2 vi = make_double2(wi.x * inv, wi.y * inv);
        Vj[c * ld + i] = vi;
        const double d = diag[i];
        double2 hw = make_double2(d * wi.x, d * wi.y);
#pragma unroll
        for (int q = 0; q < ZR; ++q) {
            if (q < zmax) {
                const int32_t cq = col[(int64_t)q * n + i];
                if (cq >= 0) {
                    const double2 hv = val[(int64_t)q * n + i];
                    const double2 x = w[cq];
                    hw.x += hv.x * x.x - hv.y * x.y;
                    hw.y += hv.x * x.y + hv.y * x.x;
                }
            }
        }
        double2 r = make_double2(hw.x * inv, hw.y * inv);       // H v_j
        if (Vprev) {
            const double2 p = Vprev[c * ld + i];
            r.x -= b * p.x;
            r.y -= b * p.y;
        }
        Wn[c * ld + i] = r;
        // conj(v_i) r
        dre = vi.x * r.x + vi.y * r.y;
        dim = vi.x * r.y - vi.y * r.x;
    }
    dre = block_sum(dre, red);
    dim = block_sum(dim, red);
    if (threadIdx.x == 0) {
        part_re[c * nb + blockIdx.x] = dre;
        part_im[c * nb + blockIdx.x] = dim;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) beta_out[c] = b;
}

// alpha = <v_j, W> from the partials; W -= alpha v_j; partials of |W|^2; block
// 0 records Re alpha
__global__ void __launch_bounds__(kLzThreads) lanczos_orth_kernel(const double2 *__restrict__ Vj, double2 *__restrict__ W,
                                                                 int64_t n, int64_t ld, const double *__restrict__ part_re,
                                                                 const double *__restrict__ part_im, int nb,
                                                                 double *__restrict__ alpha_out,
                                                                 double *__restrict__ part_nrm) {
    __shared__ double red[kLzThreads];
    const int64_t c = blockIdx.y;
    const double are = column_total(part_re + c * nb, nb, red);
    const double aim = column_total(part_im + c * nb, nb, red);
    const int64_t i = (int64_t)blockIdx.x * kLzThreads + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        const double2 v = Vj[c * ld + i];
        double2 w = W[c * ld + i];
        w.x -= are * v.x - aim * v.y;
        w.y -= are * v.y + aim * v.x;
        W[c * ld + i] = w;
        s = w.x * w.x + w.y * w.y;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part_nrm[c * nb + blockIdx.x] = s;
    if (blockIdx.x == 0 && threadIdx.x == 0) alpha_out[c] = are;
}

// Y[:, c] = sum_{k < len[c]} coef[c][k] V_k[:, c]; len[c] < 0: a copy of X[:, c]
__global__ void __launch_bounds__(kLzThreads) krylov_combine_kernel(const double2 *__restrict__ V, int64_t vstride,
                                                                   const double2 *__restrict__ coef, int kmax,
                                                                   const int *__restrict__ len,
                                                                   const double2 *__restrict__ X, double2 *__restrict__ Y,
                                                                   int64_t n, int64_t ld) {
    const int64_t c = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * kLzThreads + threadIdx.x;
    if (i >= n) return;
    const int L = len[c];
    if (L < 0) {
        Y[c * ld + i] = X[c * ld + i];
        return;
    }
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k < L; ++k) {
        const double2 a = coef[c * kmax + k];
        const double2 v = V[k * vstride + c * ld + i];
        acc.x += a.x * v.x - a.y * v.y;
        acc.y += a.x * v.y + a.y * v.x;
    }
    Y[c * ld + i] = acc;
}

}  // namespace

int lanczos_blocks(int64_t n) { return (int)((n + kLzThreads - 1) / kLzThreads); }

hipError_t launch_lanczos_start(const double2 *X, int64_t n, int64_t ld, int64_t m, double *part, hipStream_t st) {
    const int nb = lanczos_blocks(n);
    hipLaunchKernelGGL(lanczos_norm_kernel, dim3((unsigned)nb, (unsigned)m), dim3(kLzThreads), 0, st, X, n, ld, part,
                       nb);
    return hipGetLastError();
}

hipError_t launch_lanczos_iteration(const DevHam &h, const double2 *W, const double *part_w, double2 *Vj,
                                    const double2 *Vprev, double2 *Wn, int64_t ld, int64_t m, double *part_re,
                                    double *part_im, double *part_wn, double *alpha_out, double *beta_out,
                                    hipStream_t st) {
    const int64_t n = h.n;
    const int nb = lanczos_blocks(n);
    const dim3 grid((unsigned)nb, (unsigned)m);
#define LMH_LZ(ZR)                                                                                                 \
    hipLaunchKernelGGL(lanczos_matvec_kernel<ZR>, grid, dim3(kLzThreads), 0, st, h.col, h.val, h.diag, n, h.zmax, W, \
                       part_w, Vj, Vprev, Wn, ld, beta_out, part_re, part_im, nb)
    if (h.zmax <= 4)
        LMH_LZ(4);
    else if (h.zmax <= 9)
        LMH_LZ(9);
    else if (h.zmax <= 16)
        LMH_LZ(16);
    else
        return hipErrorInvalidValue;
#undef LMH_LZ
    hipLaunchKernelGGL(lanczos_orth_kernel, grid, dim3(kLzThreads), 0, st, Vj, Wn, n, ld, part_re, part_im, nb,
                       alpha_out, part_wn);
    return hipGetLastError();
}

hipError_t launch_lanczos_beta(const double *part, int64_t n, int64_t m, double *beta_out, hipStream_t st) {
    hipLaunchKernelGGL(lanczos_beta_kernel, dim3((unsigned)m), dim3(kLzThreads), 0, st, part, lanczos_blocks(n),
                       beta_out);
    return hipGetLastError();
}

hipError_t launch_krylov_combine(const double2 *V, int64_t vstride, const double2 *coef, int kmax, const int *len,
                                 const double2 *X, double2 *Y, int64_t n, int64_t ld, int64_t m, hipStream_t st) {
    const dim3 grid((unsigned)lanczos_blocks(n), (unsigned)m);
    hipLaunchKernelGGL(krylov_combine_kernel, grid, dim3(kLzThreads), 0, st, V, vstride, coef, kmax, len, X, Y, n,
                       ld);
    return hipGetLastError();
}

}  // namespace lmh
