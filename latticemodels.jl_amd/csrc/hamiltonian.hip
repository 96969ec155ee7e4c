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

#include <algorithm>

namespace lmh {

namespace {

using peierls::bond_value;

__device__ __forceinline__ double habs(double2 v) { return hypot(v.x, v.y); }

// H's entry in ELL slot q for the field components f[0 .. nf): its bonds'
// values (conjugated for the Hermitian partner) added in order
__device__ __forceinline__ double2 slot_value(const DevHam &h, int64_t q, const FieldComp *f, int64_t nf) {
    double2 acc = make_double2(0.0, 0.0);
    const int64_t c0 = h.cptr[q], c1 = h.cptr[q + 1];
    for (int64_t c = c0; c < c1; ++c) {
        const int64_t e = h.cbond[c];
        const bool partner = e < 0;
        const int64_t b = partner ? ~e : e;
        double2 v = bond_value(h, b, f, nf);
        if (partner) v.y = -v.y;
        if (c == c0) {
            acc = v;
        } else {
            acc.x += v.x;
            acc.y += v.y;
        }
    }
    return acc;
}

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
            const double2 acc = slot_value(h, q, h.field, h.nfield);
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
        if (h.ndict) h.drec[r] = make_double2(d, __longlong_as_double((long long)h.dpack[r]));
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

// After build_h: the class table from each class's defining slot, and a
// check that every slot holds its class's value (conjugated where marked);
// any difference sets red[2], and the fused passes then read val instead.
__global__ void __launch_bounds__(256) dict_check_kernel(DevHam h) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < h.ndict) {
        const int32_t rp = h.drep[q];
        h.dval[q] = rp >= 0 ? h.val[rp] : make_double2(0.0, 0.0);
    }
    if (q >= (int64_t)h.zmax * h.n || h.col[q] < 0) return;
    const int32_t cc = h.dcls[q];
    const int32_t rp = h.drep[cc & 0x7ff];
    double2 v = rp >= 0 ? h.val[rp] : make_double2(0.0, 0.0);
    if (cc >> 15) v.y = -v.y;
    const double2 w = h.val[q];
    if (v.x != w.x || v.y != w.y) atomicOr(&h.red[2], 1ull);
}

__global__ void __launch_bounds__(256) slot_probe_kernel(DevHam h, const FieldComp *probe, int nprobe,
                                                         double2 *out) {
    const int64_t nslots = (int64_t)h.zmax * h.n;
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= nslots * nprobe) return;
    const int64_t p = e / nslots, q = e % nslots;
    out[e] = slot_value(h, q, probe + p * h.nfield, h.nfield);
}

// Hamiltonian given as the stored values of a fixed CSC pattern: bond b takes
// nz[kup[b]] (its entry H[i_b, j_b], i_b < j_b) as its amplitude, and the
// mirrored entry nz[klo[b]] must be its exact conjugate (flag bit 0); site i
// takes the real part of its diagonal entry nz[kd[i]] (0 when not stored), whose
// imaginary part must be zero (flag bit 1); non-finite values set bit 2.
__global__ void __launch_bounds__(256) csc_gather_kernel(int64_t nbonds, int64_t n, const int64_t *__restrict__ kup,
                                                         const int64_t *__restrict__ klo,
                                                         const int64_t *__restrict__ kd,
                                                         const double2 *__restrict__ nz, double2 *__restrict__ amp,
                                                         double *__restrict__ onsite, int *__restrict__ flag) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < nbonds) {
        const double2 a = nz[kup[t]], b = nz[klo[t]];
        amp[t] = a;
        if (!isfinite(a.x) || !isfinite(a.y)) atomicOr(flag, 4);
        else if (b.x != a.x || b.y != -a.y) atomicOr(flag, 1);
    } else if (t < nbonds + n) {
        const int64_t i = t - nbonds;
        const int64_t k = kd[i];
        double d = 0.0;
        if (k >= 0) {
            const double2 v = nz[k];
            d = v.x;
            if (!isfinite(v.x) || !isfinite(v.y)) atomicOr(flag, 4);
            else if (v.y != 0.0) atomicOr(flag, 2);
        }
        onsite[i] = d;
    }
}

}  // namespace

hipError_t launch_csc_gather(int64_t nbonds, int64_t n, const int64_t *kup, const int64_t *klo, const int64_t *kd,
                             const double2 *nz, double2 *amp, double *onsite, int *flag, hipStream_t st) {
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), st);
    const int64_t work = nbonds + n;
    if (e != hipSuccess || work == 0) return e;
    hipLaunchKernelGGL(csc_gather_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st, nbonds, n, kup, klo,
                       kd, nz, amp, onsite, flag);
    return hipGetLastError();
}

hipError_t launch_build_h(const DevHam &h, hipStream_t st) {
    if (h.n == 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(h.red, 0, 4 * sizeof(unsigned long long), st);
    if (e != hipSuccess) return e;
    const int bs = 256;
    const int64_t nb = (h.n + bs - 1) / bs;
    hipLaunchKernelGGL(build_h_kernel, dim3((unsigned)nb), dim3(bs), 0, st, h);
    if (h.ndict) {
        const int64_t work = std::max<int64_t>((int64_t)h.zmax * h.n, h.ndict);
        hipLaunchKernelGGL(dict_check_kernel, dim3((unsigned)((work + bs - 1) / bs)), dim3(bs), 0, st, h);
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_strip_check(h, st);
    return e;
}

hipError_t launch_slot_probe(const DevHam &h, const FieldComp *probe, int nprobe, double2 *out, hipStream_t st) {
    const int64_t work = (int64_t)h.zmax * h.n * nprobe;
    if (work == 0) return hipSuccess;
    hipLaunchKernelGGL(slot_probe_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st, h, probe, nprobe,
                       out);
    return hipGetLastError();
}

}  // namespace lmh
