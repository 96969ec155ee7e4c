// Peierls phases of bond values, shared by the Hamiltonian builder and the
// adjoint (d H / d field strength) kernels.
//
// Follows LatticeModels.jl 1.0.7:
//   src/operators/builder.jl:282-285   H[i,j] = t * twist_j * exp(-2πi ∫A·dl) * conj(twist_i)
//                                      (the angle is fl(fl(-2π) * x))
//   src/zoo/magneticfields.jl:11-15    Landau:    (x1 + x2)(y2 - y1) B / 2
//   src/zoo/magneticfields.jl:27-31    symmetric: (x1 y2 - x2 y1) / 2 B
//   src/zoo/magneticfields.jl:72-85    axial point flux: acos(p1·p2/(|p1||p2|)/(1+1e-11))
//                                      * sign(det[p1 p2]) * flux / 2π
//   src/zoo/magneticfields.jl:89-104   singular point flux
//   src/operators/magneticfield.jl:104 field sums add left to right
// Floating-point contraction is disabled so the operation order is the
// reference's.  Every field kind is linear in its strength `a`, so
// d x / d a_c = (line integral of component c at a = 1).
#pragma once

#include "lmhip_internal.h"

#include "../../include/lmhip.h"

namespace lmh {
namespace peierls {

__device__ __forceinline__ double sign_of(double x) { return (x > 0.0) - (x < 0.0); }

__device__ inline double line_integral(const FieldComp *f, int64_t nf, double x1, double y1, double x2, double y2) {
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
__device__ inline double2 bond_value(const DevHam &h, int64_t b, const FieldComp *f, int64_t nf) {
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

// d H[i,j] / d a_comp of one ELL slot (sum over the bonds it holds; the
// conjugated partner direction contributes the conjugate).
__device__ inline double2 slot_derivative(const DevHam &h, int64_t q, int64_t comp) {
#pragma clang fp contract(off)
    double2 acc = make_double2(0.0, 0.0);
    FieldComp unit = h.field[comp];
    unit.a = 1.0;
    for (int64_t c = h.cptr[q]; c < h.cptr[q + 1]; ++c) {
        const int64_t e = h.cbond[c];
        const bool partner = e < 0;
        const int64_t b = partner ? ~e : e;
        const double2 v = bond_value(h, b, h.field, h.nfield);
        const double2 ri = h.ri[b], rj = h.rj[b];
        const double phi = line_integral(&unit, 1, ri.x, ri.y, rj.x, rj.y);
        // v * (-2πi phi)
        const double w = -(2.0 * M_PI) * phi;
        double2 d = make_double2(-w * v.y, w * v.x);
        if (partner) d.y = -d.y;
        acc.x += d.x;
        acc.y += d.y;
    }
    return acc;
}

}  // namespace peierls
}  // namespace lmh
