// Fused observables of one density matrix, one launch:
//   rho_i = real(P[i,i])                      src/operators/latticeutils.jl:41-45
//   J_b   = 2 Im(H[i_b, j_b] * P[j_b, i_b])   src/zoo/currents.jl:81-102
// with H the device-built Hamiltonian of the last update (the H that the
// reference's iterator yields together with P, src/evolution.jl:266-275).
#include "lmhip_internal.h"

namespace lmh {

namespace {

__global__ void __launch_bounds__(256) observables_kernel(DevHam h, const double2 *__restrict__ P, int64_t ld,
                                                          double *__restrict__ rho, double *__restrict__ J) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (rho && t < h.n) rho[t] = P[t * ld + t].x;
    if (J && t < h.nbonds) {
        const int64_t i = h.bi[t], j = h.bj[t];
        const double2 hv = h.val[h.bslot[t]];
        const double2 p = P[i * ld + j];   // P[j, i]
        J[t] = 2.0 * (hv.x * p.y + hv.y * p.x);
    }
}

__global__ void __launch_bounds__(256) gather_h_kernel(DevHam h, double *__restrict__ diag, double2 *__restrict__ hij) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (diag && t < h.n) diag[t] = h.diag[t];
    if (hij && t < h.nbonds) hij[t] = h.val[h.bslot[t]];
}

// Plain streaming copy, 16 B per thread per iteration (state copies; also the
// known-size reference kernel for HBM counter calibration).
__global__ void __launch_bounds__(256) copy_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                   int64_t count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count; t += stride) dst[t] = src[t];
}

// Observables of a ket psi (P = psi psi'): rho_i = |psi_i|^2 and
// J_b = 2 Im(H[i_b, j_b] psi_j conj(psi_i)) per bond.
__global__ void __launch_bounds__(256) ket_observables_kernel(DevHam h, const double2 *__restrict__ psi,
                                                              double *__restrict__ rho, double *__restrict__ J) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (rho && t < h.n) {
        const double2 v = psi[t];
        rho[t] = v.x * v.x + v.y * v.y;
    }
    if (J && t < h.nbonds) {
        const double2 hv = h.val[h.bslot[t]];
        const double2 a = psi[h.bj[t]], b = psi[h.bi[t]];
        // p = psi_j conj(psi_i)
        const double2 p = make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
        J[t] = 2.0 * (hv.x * p.y + hv.y * p.x);
    }
}

// P[rows[k], cols[k]] = vals[k] (distinct index pairs; P zeroed before)
__global__ void __launch_bounds__(256) scatter_kernel(double2 *__restrict__ P, int64_t ld, int64_t count,
                                                      const int64_t *__restrict__ rows,
                                                      const int64_t *__restrict__ cols,
                                                      const double2 *__restrict__ vals) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < count) P[cols[k] * ld + rows[k]] = vals[k];
}

// out[k] = P[rows[k], cols[k]]
__global__ void __launch_bounds__(256) gather_kernel(const double2 *__restrict__ P, int64_t ld, int64_t count,
                                                     const int64_t *__restrict__ rows,
                                                     const int64_t *__restrict__ cols,
                                                     double2 *__restrict__ out) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < count) out[k] = P[cols[k] * ld + rows[k]];
}

}  // namespace

hipError_t launch_scatter(double2 *P, int64_t ld, int64_t count, const int64_t *rows, const int64_t *cols,
                          const double2 *vals, hipStream_t st) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, P, ld, count, rows,
                       cols, vals);
    return hipGetLastError();
}

hipError_t launch_gather(const double2 *P, int64_t ld, int64_t count, const int64_t *rows, const int64_t *cols,
                         double2 *out, hipStream_t st) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, P, ld, count, rows,
                       cols, out);
    return hipGetLastError();
}

hipError_t launch_ket_observables(const DevHam &h, const double2 *psi, double *rho, double *J, hipStream_t st) {
    const int64_t m = h.n > h.nbonds ? h.n : h.nbonds;
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(ket_observables_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, h, psi, rho, J);
    return hipGetLastError();
}

hipError_t launch_copy(const double2 *src, double2 *dst, int64_t count, hipStream_t st) {
    if (count == 0) return hipSuccess;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, dst, count);
    return hipGetLastError();
}

hipError_t launch_observables(const DevHam &h, const double2 *P, int64_t ld, double *rho, double *J,
                              hipStream_t st) {
    const int64_t m = h.n > h.nbonds ? h.n : h.nbonds;
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(observables_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, h, P, ld, rho, J);
    return hipGetLastError();
}

hipError_t launch_gather_h(const DevHam &h, double *diag, double2 *hij, hipStream_t st) {
    const int64_t m = h.n > h.nbonds ? h.n : h.nbonds;
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_h_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, h, diag, hij);
    return hipGetLastError();
}

}  // namespace lmh
