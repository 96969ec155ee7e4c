// Fused observables of one density matrix, one launch:
//   rho_i = real(P[i,i])                      src/operators/latticeutils.jl:41-45
//   J_b   = 2 Im(H[i_b, j_b] * P[j_b, i_b])   src/zoo/currents.jl:81-102
// with H the device-built Hamiltonian of the last update (the H that the
// reference's iterator yields together with P, src/evolution.jl:266-275).
#include <atomic>
#include <mutex>

#include "lmhip_internal.h"

namespace lmh {

namespace {

// Scratch of the allocation-free region-current entry point: the per-block
// partial sums of region_currents_kernel.  Each device holds its own copy
// until the process ends; host code uses the address hipGetSymbolAddress
// gives for the current device, never the array's name.  Calls that overlap
// on one device share it and race.
constexpr int64_t kStaticParts = 1024;               // 1024 blocks of 256 sites: N <= 262144
static __device__ __align__(256) double scratch[kStaticParts];

// conj: the buffer holds P^H (a state after an odd number of fused steps),
// so P[j, i] = conj(buffer[i, j]); real(P[i,i]) is the same either way
__global__ void __launch_bounds__(256) observables_kernel(DevHam h, const double2 *__restrict__ P, int64_t ld,
                                                          double *__restrict__ rho, double *__restrict__ J, int conj) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (rho && t < h.n) rho[t] = P[t * ld + t].x;
    if (J && t < h.nbonds) {
        const int64_t i = h.bi[t], j = h.bj[t];
        const double2 hv = h.val[h.bslot[t]];
        double2 p = conj ? P[j * ld + i] : P[i * ld + j];   // P[j, i]
        if (conj) p.y = -p.y;
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

// Currents between regions (currentsfrom / currentsfromto on DensityCurrents,
// src/currents.jl:76-110, src/zoo/currents.jl:92-102): for site j in dst,
//   out[j] = sum over the stored entries H[j, i] with i in src of
//            J_ij = 2 Im(H[i, j] P[j, i]) = 2 Im(conj(H[j, i]) P[j, i]),
// each (i, j) pair once (ELL rows hold distinct columns); dst == nullptr
// means the complement of src.  Per-block partial sums of out go to part.
__global__ void __launch_bounds__(256) region_currents_kernel(DevHam h, const double2 *__restrict__ P, int64_t ld,
                                                              const uint8_t *__restrict__ src,
                                                              const uint8_t *__restrict__ dst,
                                                              double *__restrict__ out, double *__restrict__ part,
                                                              int conj) {
    __shared__ double red[256];
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (j < h.n && (dst ? dst[j] != 0 : src[j] == 0)) {
        for (int q = 0; q < h.zmax; ++q) {
            const int32_t i = h.col[(int64_t)q * h.n + j];
            if (i < 0 || !src[i]) continue;
            const double2 hji = h.val[(int64_t)q * h.n + j];
            double2 p = conj ? P[(int64_t)j * ld + i] : P[(int64_t)i * ld + j];   // P[j, i]
            if (conj) p.y = -p.y;
            acc += 2.0 * (hji.x * p.y - hji.y * p.x);
        }
    }
    if (out && j < h.n) out[j] = acc;
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dst[0] = sum of part[0 .. nparts) in a fixed order (one block)
__global__ void __launch_bounds__(256) sum_parts_kernel(const double *__restrict__ part, int64_t nparts,
                                                        double *__restrict__ dst) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t k = threadIdx.x; k < nparts; k += 256) acc += part[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) dst[0] = red[0];
}

// Time series along axis 0 (entry k at v[k * width]), one thread per column:
//   differentiate!  v[i-1] <- a_i v[i] + b_i v[i-1]       (a_i = 1/dt_i, b_i = -1/dt_i)
//   integrate!      trapezoids h_i v[i] + h_i v[i-1] (h_i = dt_i / 2), shifted by
//                   one entry behind a zero, then the running sum
// src/timesequence.jl:213-225, 255-268; products and sums rounded separately.
__global__ void __launch_bounds__(256) series_differentiate_kernel(double *__restrict__ v, int64_t width, int64_t len,
                                                                   const double *__restrict__ a,
                                                                   const double *__restrict__ b) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= width) return;
    double prev = v[c];
    for (int64_t i = 1; i < len; ++i) {
        const double cur = v[i * width + c];
        v[(i - 1) * width + c] = __dadd_rn(__dmul_rn(a[i - 1], cur), __dmul_rn(b[i - 1], prev));
        prev = cur;
    }
}

__global__ void __launch_bounds__(256) series_integrate_kernel(double *__restrict__ v, int64_t width, int64_t len,
                                                               const double *__restrict__ hh) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= width) return;
    double prev = v[c], run = 0.0;
    v[c] = 0.0;
    for (int64_t i = 1; i < len; ++i) {
        const double cur = v[i * width + c];
        const double w = __dadd_rn(__dmul_rn(hh[i - 1], cur), __dmul_rn(hh[i - 1], prev));
        run = __dadd_rn(run, w);
        v[i * width + c] = run;
        prev = cur;
    }
}

}  // namespace

hipError_t launch_region_currents(const DevHam &h, const double2 *P, int64_t ld, const uint8_t *src,
                                  const uint8_t *dst, double *out, double *part, double *total, hipStream_t st,
                                  int conj) {
    const int64_t nb = (h.n + 255) / 256;
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(region_currents_kernel, dim3((unsigned)nb), dim3(256), 0, st, h, P, ld, src, dst, out, part,
                       conj);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !total) return e;
    hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(256), 0, st, part, nb, total);
    return hipGetLastError();
}

int64_t region_currents_parts(const DevHam &h) { return (h.n + 255) / 256; }

int64_t region_currents_static_max_sites() { return kStaticParts * 256; }

// the device address of `scratch` on the current device, looked up once per
// device (with that device current)
static hipError_t static_scratch(double **out) {
    constexpr int kMaxDevices = 64;
    static std::atomic<double *> ptrs[kMaxDevices];
    static std::mutex mu;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    double *p = ptrs[dev].load();
    if (!p) {
        std::lock_guard<std::mutex> lock(mu);
        p = ptrs[dev].load();
        if (!p) {
            e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return e;
            if (!p) return hipErrorInvalidSymbol;
            ptrs[dev].store(p);
        }
    }
    *out = p;
    return hipSuccess;
}

hipError_t prepare_region_currents_static() {
    double *p = nullptr;
    return static_scratch(&p);
}

hipError_t launch_region_currents_static(const DevHam &h, const double2 *P, int64_t ld, const uint8_t *src,
                                         const uint8_t *dst, double *out, double *total, hipStream_t st, int conj) {
    const int64_t nb = region_currents_parts(h);
    if (nb > kStaticParts) return hipErrorInvalidValue;
    double *part = nullptr;
    hipError_t e = static_scratch(&part);
    if (e != hipSuccess) return e;
    // every block writes its own partial sum: nothing to clear
    return launch_region_currents(h, P, ld, src, dst, out, part, total, st, conj);
}

hipError_t launch_series_differentiate(double *v, int64_t width, int64_t len, const double *a, const double *b,
                                       hipStream_t st) {
    if (width == 0 || len < 2) return hipSuccess;
    hipLaunchKernelGGL(series_differentiate_kernel, dim3((unsigned)((width + 255) / 256)), dim3(256), 0, st, v, width,
                       len, a, b);
    return hipGetLastError();
}

hipError_t launch_series_integrate(double *v, int64_t width, int64_t len, const double *hh, hipStream_t st) {
    if (width == 0 || len < 2) return hipSuccess;
    hipLaunchKernelGGL(series_integrate_kernel, dim3((unsigned)((width + 255) / 256)), dim3(256), 0, st, v, width,
                       len, hh);
    return hipGetLastError();
}

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
                              hipStream_t st, int conj) {
    const int64_t m = h.n > h.nbonds ? h.n : h.nbonds;
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(observables_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, h, P, ld, rho, J,
                       conj);
    return hipGetLastError();
}

hipError_t launch_gather_h(const DevHam &h, double *diag, double2 *hij, hipStream_t st) {
    const int64_t m = h.n > h.nbonds ? h.n : h.nbonds;
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_h_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, h, diag, hij);
    return hipGetLastError();
}

}  // namespace lmh
