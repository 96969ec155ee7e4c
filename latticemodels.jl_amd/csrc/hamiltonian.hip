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
#include "lmhip_internal.h"

#include "../../include/lmhip.h"

namespace lmh {

namespace {

__device__ __forceinline__ double sign_of(double x) { return (x > 0.0) - (x < 0.0); }

__device__ double line_integral(const FieldComp *f, int64_t nf, double x1, double y1, double x2, double y2) {
#pragma clang fp contract(off)
    double acc = 0.0;
    for (int64_t c = 0; c < nf; ++c) {
        const FieldComp fc = f[c];
        double v = 0.0;
        if (fc.kind == LMH_FIELD_LANDAU) {
            v = (x1 + x2) * (y2 - y1) * fc.a / 2.0;
        } else if (fc.kind == LMH_FIELD_SYMMETRIC) {
            v = (x1 * y2 - x2 * y1) / 2.0 * fc.a;
        } else if (fc.kind == LMH_FIELD_POINT_AXIAL) {
            const double ax = x1 - fc.px, ay = y1 - fc.py;
            const double bx = x2 - fc.px, by = y2 - fc.py;
            const double n1 = sqrt(ax * ax + ay * ay);
            const double n2 = sqrt(bx * bx + by * by);
            if (n1 < 1e-11 || n2 < 1e-11) {
                v = 0.0;
            } else {
                const double nnorm = n1 * n2;
                const double det = ax * by - bx * ay;
                const double cosang = (ax * bx + ay * by) / nnorm / (1.0 + 1e-11);
                const double ang = acos(cosang) * sign_of(det);
                v = ang * fc.a / (2.0 * M_PI);
            }
        } else if (fc.kind == LMH_FIELD_POINT_SINGULAR) {
            const double ax = x1 - fc.px, ay = y1 - fc.py;
            const double bx = x2 - fc.px, by = y2 - fc.py;
            const double sg = bx - ax;
            if (fabs(sg) < 1e-11) {
                v = 0.0;
            } else if (ax * bx > 0.0 || fmax(ax, bx) == 0.0) {
                v = 0.0;
            } else {
                const double yint = (-ay * bx + by * ax) / (ax - bx);
                v = yint > 0.0 ? 0.0 : fc.a * sign_of(sg);
            }
        }
        acc = (c == 0) ? v : acc + v;
    }
    return acc;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
#pragma clang fp contract(off)
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// amp * (twist * exp(-2πi x)), the value of H[i_b, j_b] from bond b alone.
__device__ double2 bond_value(const DevHam &h, int64_t b, const FieldComp *f, int64_t nf) {
#pragma clang fp contract(off)
    const double2 ri = h.ri[b], rj = h.rj[b];
    const double x = line_integral(f, nf, ri.x, ri.y, rj.x, rj.y);
    const double th = -(2.0 * M_PI) * x;
    double2 ff;
    if (th == 0.0) {
        ff = make_double2(1.0, 0.0);
    } else {
        double s, c;
        sincos(th, &s, &c);
        ff = make_double2(c, s);
    }
    const double2 total = cmul(h.twist[b], ff);
    return cmul(h.amp[b], total);
}

__device__ __forceinline__ double habs(double2 v) { return hypot(v.x, v.y); }

__global__ void __launch_bounds__(256) build_h_kernel(DevHam h) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= h.n) return;
    const double d = h.onsite ? h.onsite[r] : 0.0;
    h.diag[r] = d;
    double mx = fabs(d);
    double rowsum = fabs(d);
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
    // max over rows: nonnegative doubles order like their bit patterns
    atomicMax(&h.red[0], (unsigned long long)__double_as_longlong(mx));
    atomicMax(&h.red[1], (unsigned long long)__double_as_longlong(rowsum));
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
