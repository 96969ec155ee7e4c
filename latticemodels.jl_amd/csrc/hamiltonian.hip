// In-kernel Hamiltonian construction: Peierls phases from the gauge field,
// hopping amplitudes, boundary twists and on-site energies, written to an
// ELL layout that the propagation kernels read.
//
// Follows LatticeModels.jl 1.0.7:
//   src/operators/builder.jl:282-285   H[i,j] = t * twist_j * exp(-2πi ∫A·dl) * conj(twist_i)
//                                      (the angle is fl(fl(-2π) * x))
//   src/operators/builder.jl:296-309   auto-Hermitian partner H[j,i] = conj(H[i,j]);
//                                      repeated writes to one entry accumulate in order
//   src/zoo/magneticfields.jl:11-15    Landau:    (x1 + x2)(y2 - y1) B / 2
//   src/zoo/magneticfields.jl:27-31    symmetric: (x1 y2 - x2 y1) / 2 B
//   src/zoo/magneticfields.jl:72-85    axial point flux: acos(p1·p2/(|p1||p2|)/(1+1e-11))
//                                      * sign(det[p1 p2]) * flux / 2π
//   src/zoo/magneticfields.jl:89-104   singular point flux
//   src/operators/magneticfield.jl:104 field sums add left to right
// Floating-point contraction is disabled in the phase code so the operation
// order is the reference's; the bar for the phase is a few ulp.
#include "peierls.h"

namespace lmh {

namespace {

using peierls::bond_value;

__device__ __forceinline__ double habs(double2 v) { return hypot(v.x, v.y); }

__global__ void __launch_bounds__(256) build_h_kernel(DevHam h) {
    __shared__ double red_mx[256], red_rs[256];
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double mx = 0.0, rowsum = 0.0;
    if (r < h.n) {
        const double d = h.onsite ? h.onsite[r] : 0.0;
        h.diag[r] = d;
        mx = fabs(d);
        rowsum = fabs(d);
        for (int s = 0; s < h.zmax; ++s) {
            const int64_t q = (int64_t)s * h.n + r;
            double2 acc = make_double2(0.0, 0.0);
            const int64_t c0 = h.cptr[q], c1 = h.cptr[q + 1];
            for (int64_t c = c0; c < c1; ++c) {
                const int64_t e = h.cbond[c];
                const bool partner = e < 0;
                const int64_t b = partner ? ~e : e;
                double2 v = bond_value(h, b, h.field, h.nfield);
                if (partner) v.y = -v.y;
                if (c == c0) {
                    acc = v;
                } else {
                    acc.x += v.x;
                    acc.y += v.y;
                }
            }
            h.val[q] = acc;
            const double a = habs(acc);
            mx = fmax(mx, a);
            rowsum += a;
        }
        if (h.rec) {
            unsigned long long cb = 0;
            for (int s = 0; s < h.zmax; ++s) {
                const int32_t c = h.col[(int64_t)s * h.n + r];
                cb |= (unsigned long long)(c < 0 ? 0 : c) << (16 * s);
            }
            h.rec[r] = make_double2(d, __longlong_as_double((long long)cb));
        }
    }
    // block maxima first: one device-scope atomic per block and quantity
    red_mx[threadIdx.x] = mx;
    red_rs[threadIdx.x] = rowsum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red_mx[threadIdx.x] = fmax(red_mx[threadIdx.x], red_mx[threadIdx.x + w]);
            red_rs[threadIdx.x] = fmax(red_rs[threadIdx.x], red_rs[threadIdx.x + w]);
        }
        __syncthreads();
    }
    // max over rows: nonnegative doubles order like their bit patterns
    if (threadIdx.x == 0) {
        atomicMax(&h.red[0], (unsigned long long)__double_as_longlong(red_mx[0]));
        atomicMax(&h.red[1], (unsigned long long)__double_as_longlong(red_rs[0]));
    }
}

}  // namespace

hipError_t launch_build_h(const DevHam &h, hipStream_t st) {
    if (h.n == 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(h.red, 0, 2 * sizeof(unsigned long long), st);
    if (e != hipSuccess) return e;
    const int bs = 256;
    const int64_t nb = (h.n + bs - 1) / bs;
    hipLaunchKernelGGL(build_h_kernel, dim3((unsigned)nb), dim3(bs), 0, st, h);
    return hipGetLastError();
}

}  // namespace lmh
