// Discrete adjoint of the density-matrix step P' = U P U^H, U = p(-i dt H) the
// order-K Taylor polynomial the forward kernels apply (s = 1).
//
// A step is two sides S(X) = (U X)^H.  Going backwards through step k with
// the costate Lambda_{k+1} = dL/dP_{k+1} (Hermitian, dL = Re tr(Lambda^H dP)):
//
//   side kernel A:  x = U^H P_{k+1}  (an exact reverse unitary step, = the
//                   right side's input (U P_k)^H up to U's unitarity defect),
//                   written conjugate-transposed for kernel B;
//                   reverse Horner of U x with cotangent Lambda_{k+1}
//                   -> U^H Lambda_{k+1}, written conjugate-transposed;
//   side kernel B:  x = U^H (kernel A's x)^H = P_k, or the checkpoint P_k;
//                   reverse Horner with cotangent (U^H Lambda_{k+1})^H
//                   -> Lambda_k = U^H Lambda_{k+1} U.
//
// Reverse Horner for one column (y_{K+1} = x, y_m = x + c_m H y_{m+1},
// c_m = -i dt/m, output y_1):
//     g_1 = cotangent column;  for m = 1..K:
//         F_ij += c_m conj(g_m[i]) y_{m+1}[j]   on the pattern of H
//         xbar += g_m;  g_{m+1} = conj(c_m) H g_m
//     xbar += g_{K+1}
// and dL = Re sum_ij F_ij dH_ij over both sides.
//
// One workgroup owns one column at a time with x and the K-1 iterates
// y_K .. y_2 in LDS (K * N * 16 B <= 160 KB); the costate columns reuse the
// iterate buffers as they are consumed.  F is accumulated in registers over
// all columns a workgroup handles and written once per workgroup; a second
// kernel sums the workgroups' partials, adds Re F_ii to the on-site gradient
// and contracts F with dH/da for each field component.
#include "peierls.h"

#include <algorithm>

namespace lmh {

namespace {

constexpr int kAdjThreads = 1024;
constexpr int kAdjLdsBytes = 163840;

__device__ __forceinline__ double2 cmadd(double2 acc, double2 a, double2 b) {
    acc.x += a.x * b.x - a.y * b.y;
    acc.y += a.x * b.y + a.y * b.x;
    return acc;
}

// H row (ZM slots, empty ones as column 0 with value 0) and diagonal of a site
// held in registers for the whole launch; sites beyond n compute on site 0.
template <int ZM>
struct HRow {
    int32_t c[ZM];
    double2 v[ZM];
    double d;
};

template <int ZM>
__device__ __forceinline__ double2 apply_row(const HRow<ZM> &r, const double2 *__restrict__ y, int s) {
    const double2 y0 = y[s];
    double2 acc = make_double2(r.d * y0.x, r.d * y0.y);
#pragma unroll
    for (int q = 0; q < ZM; ++q) acc = cmadd(acc, r.v[q], y[r.c[q]]);
    return acc;
}

// row of site s read from the ELL arrays (no register cache); nn is an
// opaque copy of n so the loads stay at their use
template <int ZM>
__device__ __forceinline__ HRow<ZM> load_row(const int32_t *__restrict__ col, const double2 *__restrict__ val,
                                             const double *__restrict__ diag, int nn, int zmax, int s) {
    HRow<ZM> r;
    r.d = diag[s];
#pragma unroll
    for (int q = 0; q < ZM; ++q) {
        const int32_t cq = q < zmax ? col[(int64_t)q * nn + s] : -1;
        r.c[q] = cq < 0 ? 0 : cq;
        r.v[q] = cq < 0 ? make_double2(0.0, 0.0) : val[(int64_t)q * nn + s];
    }
    return r;
}

template <int SPT, int ZM, bool HREG>
__global__ void __launch_bounds__(kAdjThreads) adjoint_side_kernel(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag, int n,
    int zmax, const double2 *__restrict__ X, int x_given, double2 *__restrict__ Xout, int xout_t,
    const double2 *__restrict__ G, double2 *__restrict__ Gout, int gout_t, int64_t ld, double dt, int K,
    double2 *__restrict__ partial, int accumulate) {
    extern __shared__ double2 lds[];       // K buffers of n: 0 = x, b = y_{K+1-b}
    const int t = threadIdx.x;
    constexpr int HS = HREG ? SPT : 1;
    HRow<ZM> hr[HS];
    int sc[SPT];
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int s0 = t + m * kAdjThreads;
        sc[m] = s0 < n ? s0 : 0;
        if constexpr (HREG) hr[m] = load_row<ZM>(col, val, diag, n, zmax, sc[m]);
    }
    // the row of thread site m: cached, or re-read at its use
    auto row = [&](int m) -> HRow<ZM> {
        if constexpr (HREG) {
            return hr[m];
        } else {
            int nn = n;
            asm volatile("" : "+s"(nn));
            return load_row<ZM>(col, val, diag, nn, zmax, sc[m]);
        }
    };
    double2 accd[SPT];
    double2 acco[SPT][ZM];
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        accd[m] = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < ZM; ++q) acco[m][q] = make_double2(0.0, 0.0);
    }
    for (int64_t c = blockIdx.x; c < n; c += gridDim.x) {
        const double2 *xc = X + c * ld;
        double2 xr[SPT];                  // the input column, for every U^H term
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int s = t + m * kAdjThreads;
            xr[m] = s < n ? xc[s] : make_double2(0.0, 0.0);
            if (s < n) lds[s] = xr[m];
        }
        __syncthreads();
        if (!x_given) {
            // x = U^H X[:, c]: Horner with coefficients +i dt/k
            for (int k = K; k >= 1; --k) {
                const double cc = -dt / (double)k;
                double2 ynew[SPT];
#pragma unroll
                for (int m = 0; m < SPT; ++m) {
                    const double2 acc = apply_row(row(m), lds, sc[m]);
                    ynew[m] = make_double2(xr[m].x + cc * acc.y, xr[m].y - cc * acc.x);
                }
                __syncthreads();
#pragma unroll
                for (int m = 0; m < SPT; ++m) {
                    const int s = t + m * kAdjThreads;
                    if (s < n) lds[s] = ynew[m];
                }
                __syncthreads();
            }
        }
        double2 xv[SPT];                  // x (the side's input column)
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int s = t + m * kAdjThreads;
            xv[m] = s < n ? lds[s] : make_double2(0.0, 0.0);
            if (Xout && s < n) {
                if (xout_t)
                    Xout[(int64_t)s * ld + c] = make_double2(xv[m].x, -xv[m].y);
                else
                    Xout[c * ld + s] = xv[m];
            }
        }
        // forward iterates y_K .. y_2 of U x into buffers 1 .. K-1
        for (int b = 1; b < K; ++b) {
            const double cc = dt / (double)(K + 1 - b);
            const double2 *src = lds + (int64_t)(b - 1) * n;
            double2 *dst = lds + (int64_t)b * n;
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int s = t + m * kAdjThreads;
                const double2 acc = apply_row(row(m), src, sc[m]);
                if (s < n) dst[s] = make_double2(xv[m].x + cc * acc.y, xv[m].y - cc * acc.x);
            }
            __syncthreads();
        }
        // reverse sweep
        const double2 *gc = G + c * ld;
        double2 gown[SPT], xbar[SPT];
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int s = t + m * kAdjThreads;
            gown[m] = s < n ? gc[s] : make_double2(0.0, 0.0);
            xbar[m] = make_double2(0.0, 0.0);
        }
        for (int ms = 1; ms <= K; ++ms) {
            const double2 *yb = lds + (int64_t)(K - ms) * n;                       // y_{ms+1}
            const double2 *gs = ms == 1 ? gc : lds + (int64_t)(K - ms + 1) * n;    // g_ms
            const double cr = dt / (double)ms;                                      // c_ms = -i cr
            double2 gnew[SPT];
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int s = sc[m];
                const double2 gi = gown[m];
                // w = c_ms conj(g_ms[s]); zero for sites beyond n (gown = 0)
                const double2 w = make_double2(-cr * gi.y, -cr * gi.x);
                accd[m] = cmadd(accd[m], w, yb[s]);
                const HRow<ZM> rm = row(m);
                double2 hg = make_double2(rm.d * gi.x, rm.d * gi.y);
#pragma unroll
                for (int q = 0; q < ZM; ++q) {
                    const int32_t nb = rm.c[q];
                    acco[m][q] = cmadd(acco[m][q], w, yb[nb]);
                    hg = cmadd(hg, rm.v[q], gs[nb]);
                }
                xbar[m].x += gi.x;
                xbar[m].y += gi.y;
                // g_{ms+1} = conj(c_ms) H g_ms = +i cr H g_ms
                gnew[m] = make_double2(-cr * hg.y, cr * hg.x);
            }
            __syncthreads();
            if (ms < K) {
                double2 *gd = lds + (int64_t)(K - ms) * n;    // y_{ms+1} is consumed
#pragma unroll
                for (int m = 0; m < SPT; ++m) {
                    const int s = t + m * kAdjThreads;
                    if (s < n) gd[s] = gnew[m];
                }
            }
#pragma unroll
            for (int m = 0; m < SPT; ++m) gown[m] = gnew[m];
            __syncthreads();
        }
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int s = t + m * kAdjThreads;
            if (s < n) {
                const double2 v = make_double2(xbar[m].x + gown[m].x, xbar[m].y + gown[m].y);
                if (gout_t)
                    Gout[(int64_t)s * ld + c] = make_double2(v.x, -v.y);
                else
                    Gout[c * ld + s] = v;
            }
        }
    }
    // this workgroup's F sums (added to the other side's when accumulating)
    double2 *pb = partial + (int64_t)blockIdx.x * (zmax + 1) * n;
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int s = t + m * kAdjThreads;
        if (s < n) {
#pragma unroll
            for (int q = 0; q < ZM; ++q) {
                if (q < zmax) {
                    double2 v = acco[m][q];
                    if (accumulate) {
                        const double2 o = pb[(int64_t)q * n + s];
                        v.x += o.x;
                        v.y += o.y;
                    }
                    pb[(int64_t)q * n + s] = v;
                }
            }
            double2 v = accd[m];
            if (accumulate) {
                const double2 o = pb[(int64_t)zmax * n + s];
                v.x += o.x;
                v.y += o.y;
            }
            pb[(int64_t)zmax * n + s] = v;
        }
    }
}

// xbar = g_1 + ... + g_{K+1} accumulated in XB by the terms that `xbm` marks:
// kXbTouch: XB gets g_{m+1} = conj(c_m) H g_m (and is written); kXbRead: the
// sum so far is read (else it starts here); kXbAddG: g_m is added too.
constexpr int kXbTouch = 1, kXbRead = 2, kXbAddG = 4;

// One reverse Horner term over every column (any N): a thread owns row i,
// keeps H's row in registers and sweeps the columns of block blockIdx.y:
//     F_ij += c_m conj(g_m[i]) y_{m+1}[j],  g_{m+1} = conj(c_m) H g_m,
// xbar per `xbm` (last term: g_{K+1} not stored).  F partial sums per column
// block are added to `partial` (zeroed once per reverse step).
template <int ZR>
__global__ void __launch_bounds__(256) reverse_term_kernel(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag, int64_t n,
    int zmax, const double2 *__restrict__ G, const double2 *__restrict__ Y, double2 *__restrict__ XB,
    double2 *__restrict__ Gout, int last, int xbm, int64_t ld, double cr, int64_t cb_cols,
    double2 *__restrict__ partial) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t c0 = (int64_t)blockIdx.y * cb_cols;
    const int64_t c1 = c0 + cb_cols < n ? c0 + cb_cols : n;
    int32_t hc[ZR];
    double2 hv[ZR];
    const double hd = diag[i];
#pragma unroll
    for (int q = 0; q < ZR; ++q) {
        const int32_t cq = q < zmax ? col[(int64_t)q * n + i] : -1;
        hc[q] = cq < 0 ? 0 : cq;                       // empty slot: column 0, value 0
        hv[q] = cq < 0 ? make_double2(0.0, 0.0) : val[(int64_t)q * n + i];
    }
    double2 accd = make_double2(0.0, 0.0);
    double2 acco[ZR];
#pragma unroll
    for (int q = 0; q < ZR; ++q) acco[q] = make_double2(0.0, 0.0);
    for (int64_t c = c0; c < c1; ++c) {
        const double2 *gcol = G + c * ld;
        const double2 *ycol = Y + c * ld;
        const double2 gi = gcol[i];
        const double2 w = make_double2(-cr * gi.y, -cr * gi.x);     // c_m conj(g_m[i]), c_m = -i cr
        accd = cmadd(accd, w, ycol[i]);
        double2 hg = make_double2(hd * gi.x, hd * gi.y);
#pragma unroll
        for (int q = 0; q < ZR; ++q) {
            const int32_t nb = hc[q];
            acco[q] = cmadd(acco[q], w, ycol[nb]);
            hg = cmadd(hg, hv[q], gcol[nb]);
        }
        const double2 gn = make_double2(-cr * hg.y, cr * hg.x);      // conj(c_m) H g_m
        if (!last) Gout[c * ld + i] = gn;
        if (xbm & kXbTouch) {
            double2 xb = (xbm & kXbRead) ? XB[c * ld + i] : make_double2(0.0, 0.0);
            if (xbm & kXbAddG) {
                xb.x += gi.x;
                xb.y += gi.y;
            }
            xb.x += gn.x;
            xb.y += gn.y;
            XB[c * ld + i] = xb;
        }
    }
    double2 *pb = partial + (int64_t)blockIdx.y * (zmax + 1) * n;
#pragma unroll
    for (int q = 0; q < ZR; ++q) {
        if (q < zmax) {
            double2 v = pb[(int64_t)q * n + i];
            v.x += acco[q].x;
            v.y += acco[q].y;
            pb[(int64_t)q * n + i] = v;
        }
    }
    double2 v = pb[(int64_t)zmax * n + i];
    v.x += accd.x;
    v.y += accd.y;
    pb[(int64_t)zmax * n + i] = v;
}

// The same reverse term for a banded H (every neighbour within `band` rows):
// a block owns kRevRows rows and, per column, stages rows [r0 - band, r0 +
// kRevRows + band) of g_m and y_{m+1} in LDS (double-buffered, the next
// column's segments loaded into registers meanwhile), so the neighbour
// gathers read LDS instead of L1/L2.  Same partial-sum layout.
constexpr int kRevRows = 512;
constexpr int kRevStage = 3;         // staged elements per thread and array: band <= kRevRows

// logical block of workgroup pid: each XCD (pid % 8) takes a contiguous range,
// so row blocks that share halo rows run on one XCD and meet in its L2
__device__ __forceinline__ int64_t rev_xcd_logical(int64_t pid, int64_t nprog) {
    const int64_t per = nprog / kNumXCD;
    return (pid % kNumXCD) * per + pid / kNumXCD;
}

// CPI columns per barrier: the next CPI columns' loads are in flight behind
// CPI columns of compute (one column of compute does not cover HBM latency at
// the one workgroup per CU that the register count allows).
template <int ZR, int CPI>
__global__ void __launch_bounds__(kRevRows) reverse_term_band_kernel(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag, int64_t n,
    int zmax, int band, const double2 *__restrict__ G, const double2 *__restrict__ Y, double2 *__restrict__ XB,
    double2 *__restrict__ Gout, int last, int xbm, int64_t ld, double cr, int64_t cb_cols,
    double2 *__restrict__ partial, int64_t rblocks, int64_t nreal, int64_t nprog) {
    extern __shared__ double2 lds[];          // [2 buffers][CPI columns][g, y][span]
    const int64_t lb = rev_xcd_logical(blockIdx.x, nprog);
    if (lb >= nreal) return;                  // padding to whole XCD rounds: the whole block leaves
    const int64_t rblk = lb % rblocks, cpart = lb / rblocks;
    const int t = threadIdx.x;
    const int64_t r0 = rblk * kRevRows;
    const int64_t i = r0 + t;
    const bool live = i < n;
    const int span = kRevRows + 2 * band;
    const int64_t g0 = r0 - band;
    const int self = band + t;
    int nbs[ZR];
    double2 hv[ZR];
#pragma unroll
    for (int q = 0; q < ZR; ++q) {
        const int32_t cq = (live && q < zmax) ? col[(int64_t)q * n + i] : -1;
        nbs[q] = cq >= 0 ? (int)(cq - g0) : self;      // empty slot: own row, value 0 (its partial is ignored)
        hv[q] = cq >= 0 ? val[(int64_t)q * n + i] : make_double2(0.0, 0.0);
    }
    const double hd = live ? diag[i] : 0.0;
    const int64_t c0 = cpart * cb_cols;
    const int64_t c1 = c0 + cb_cols < n ? c0 + cb_cols : n;
    double2 accd = make_double2(0.0, 0.0);
    double2 acco[ZR];
#pragma unroll
    for (int q = 0; q < ZR; ++q) acco[q] = make_double2(0.0, 0.0);
    double2 pg[CPI][kRevStage], py[CPI][kRevStage], px[CPI];
    auto stage_load = [&](int64_t cb) {
#pragma unroll
        for (int u = 0; u < CPI; ++u) {
            const int64_t c = cb + u;
            px[u] = (live && c < c1 && (xbm & kXbRead)) ? XB[c * ld + i] : make_double2(0.0, 0.0);  // own row only
#pragma unroll
            for (int p = 0; p < kRevStage; ++p) {
                const int e = t + p * kRevRows;
                const int64_t g = g0 + e;
                const bool ok = c < c1 && e < span && g >= 0 && g < n;
                pg[u][p] = ok ? G[c * ld + g] : make_double2(0.0, 0.0);
                py[u][p] = ok ? Y[c * ld + g] : make_double2(0.0, 0.0);
            }
        }
    };
    if (c0 < c1) stage_load(c0);
    int buf = 0;
    for (int64_t cb = c0; cb < c1; cb += CPI) {
        double2 *base = lds + (size_t)buf * CPI * 2 * span;
#pragma unroll
        for (int u = 0; u < CPI; ++u) {
            double2 *gt = base + (size_t)u * 2 * span;
            double2 *yt = gt + span;
#pragma unroll
            for (int p = 0; p < kRevStage; ++p) {
                const int e = t + p * kRevRows;
                if (e < span) {
                    gt[e] = pg[u][p];
                    yt[e] = py[u][p];
                }
            }
        }
        // one barrier per CPI columns: two buffers (see horner_left_band)
        __syncthreads();
        double2 xcur[CPI];
#pragma unroll
        for (int u = 0; u < CPI; ++u) xcur[u] = px[u];
        if (cb + CPI < c1) stage_load(cb + CPI);
#pragma unroll
        for (int u = 0; u < CPI; ++u) {
            const int64_t c = cb + u;
            if (c >= c1) break;
            const double2 *gt = base + (size_t)u * 2 * span;
            const double2 *yt = gt + span;
            const double2 xb0 = xcur[u];      // loaded with the column's staging, a column block ahead
            const double2 gi = gt[self];
            const double2 w = make_double2(-cr * gi.y, -cr * gi.x);     // c_m conj(g_m[i]), c_m = -i cr
            accd = cmadd(accd, w, yt[self]);
            double2 hg = make_double2(hd * gi.x, hd * gi.y);
#pragma unroll
            for (int q = 0; q < ZR; ++q) {
                acco[q] = cmadd(acco[q], w, yt[nbs[q]]);
                hg = cmadd(hg, hv[q], gt[nbs[q]]);
            }
            const double2 gn = make_double2(-cr * hg.y, cr * hg.x);      // conj(c_m) H g_m
            double2 xb = xb0;                 // zero unless kXbRead
            if (xbm & kXbAddG) {
                xb.x += gi.x;
                xb.y += gi.y;
            }
            xb.x += gn.x;
            xb.y += gn.y;
            if (live) {
                if (!last) Gout[c * ld + i] = gn;
                if (xbm & kXbTouch) XB[c * ld + i] = xb;
            }
        }
        buf ^= 1;
    }
    if (!live) return;
    double2 *pb = partial + cpart * (zmax + 1) * n;
#pragma unroll
    for (int q = 0; q < ZR; ++q) {
        if (q < zmax) {
            double2 v = pb[(int64_t)q * n + i];
            v.x += acco[q].x;
            v.y += acco[q].y;
            pb[(int64_t)q * n + i] = v;
        }
    }
    double2 v = pb[(int64_t)zmax * n + i];
    v.x += accd.x;
    v.y += accd.y;
    pb[(int64_t)zmax * n + i] = v;
}

// K == 0 (U = I): no iterates; x and the cotangent pass through.
__global__ void __launch_bounds__(256) adjoint_identity_kernel(const double2 *__restrict__ X, int x_given,
                                                               double2 *__restrict__ Xout, int xout_t,
                                                               const double2 *__restrict__ G,
                                                               double2 *__restrict__ Gout, int gout_t, int64_t n,
                                                               int64_t ld) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int64_t j = e / n, i = e % n;
    (void)x_given;
    if (Xout) {
        const double2 v = X[j * ld + i];
        if (xout_t) Xout[i * ld + j] = make_double2(v.x, -v.y); else Xout[j * ld + i] = v;
    }
    const double2 g = G[j * ld + i];
    if (gout_t) Gout[i * ld + j] = make_double2(g.x, -g.y); else Gout[j * ld + i] = g;
}

// Sum the workgroups' partial F (times `scale`); on-site gradient += Re F_ii;
// per-component field gradient partial sums per block (contracted with
// dH/da).  A block owns 64 consecutive slots (coalesced across lanes); its 4
// waves each sum a quarter of the partials with 4 independent accumulators.
__global__ void __launch_bounds__(256) adjoint_reduce_kernel(DevHam h, const double2 *__restrict__ partial,
                                                             int nparts, double scale, double *__restrict__ gonsite,
                                                             double *__restrict__ gfield) {
    __shared__ double2 part[4][64];
    __shared__ double red[64];
    const int64_t n = h.n;
    const int64_t nslots = (int64_t)(h.zmax + 1) * n;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t q = blockIdx.x * (int64_t)64 + lane;
    double2 a[4] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    if (q < nslots) {
        const int p0 = (int)((int64_t)nparts * wv / 4), p1 = (int)((int64_t)nparts * (wv + 1) / 4);
        int p = p0;
        for (; p + 4 <= p1; p += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double2 v = partial[(int64_t)(p + u) * nslots + q];
                a[u].x += v.x;
                a[u].y += v.y;
            }
        }
        for (; p < p1; ++p) {
            const double2 v = partial[(int64_t)p * nslots + q];
            a[0].x += v.x;
            a[0].y += v.y;
        }
    }
    part[wv][lane] = make_double2((a[0].x + a[1].x) + (a[2].x + a[3].x), (a[0].y + a[1].y) + (a[2].y + a[3].y));
    __syncthreads();
    double2 F = make_double2(0.0, 0.0);
    if (wv == 0) {
        F = make_double2(scale * ((part[0][lane].x + part[1][lane].x) + (part[2][lane].x + part[3][lane].x)),
                         scale * ((part[0][lane].y + part[1][lane].y) + (part[2][lane].y + part[3][lane].y)));
        const bool offdiag = q < (int64_t)h.zmax * n;
        if (q < nslots && !offdiag) gonsite[q - (int64_t)h.zmax * n] += F.x;
    }
    const bool live = wv == 0 && q < nslots && q < (int64_t)h.zmax * n && h.col[q] >= 0;
    for (int64_t comp = 0; comp < h.nfield; ++comp) {
        if (wv == 0) {
            double g = 0.0;
            if (live) {
                const double2 d = peierls::slot_derivative(h, q, comp);
                g = F.x * d.x - F.y * d.y;          // Re(F dH)
            }
            for (int off = 32; off > 0; off >>= 1) g += __shfl_xor(g, off);
            if (lane == 0) gfield[(int64_t)blockIdx.x * h.nfield + comp] = g;
        }
    }
    (void)red;
}

// out[b] = d H[i_b, j_b] / d a_comp for every bond.
__global__ void __launch_bounds__(256) bond_derivative_kernel(DevHam h, int64_t comp, double2 *__restrict__ out) {
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b >= h.nbonds) return;
    out[b] = peierls::slot_derivative(h, h.bslot[b], comp);
}

// dst = (src + src^H) / 2, 32x32 tiles; dst must not alias src.
__global__ void __launch_bounds__(256) hermitian_part_kernel(const double2 *__restrict__ src,
                                                             double2 *__restrict__ dst, int64_t n, int64_t ld,
                                                             int64_t tiles) {
    __shared__ double2 tile[32][33];
    const int64_t tb = blockIdx.x;
    const int64_t ti = tb % tiles, tj = tb / tiles;      // tile (row block ti, column block tj)
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // load the mirrored tile (rows tj, columns ti) transposed into LDS
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = tj * 32 + tx, j = ti * 32 + r;   // src[i, j], column j
        if (i < n && j < n) tile[r][tx] = src[j * ld + i];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = ti * 32 + tx, j = tj * 32 + r;   // dst[i, j]
        if (i < n && j < n) {
            const double2 a = src[j * ld + i];
            const double2 b = tile[tx][r];                 // src[j, i]
            dst[j * ld + i] = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
        }
    }
}

template <int SPT, int ZM, bool HREG>
hipError_t launch_side_t(const DevHam &h, const AdjointSide &a, hipStream_t st) {
    static std::atomic<uint64_t> attr_done{0};
    hipError_t e = once_per_device(attr_done, [] {
        return hipFuncSetAttribute((const void *)adjoint_side_kernel<SPT, ZM, HREG>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, kAdjLdsBytes);
    });
    if (e != hipSuccess) return e;
    const size_t lds = (size_t)a.K * h.n * sizeof(double2);
    hipLaunchKernelGGL((adjoint_side_kernel<SPT, ZM, HREG>), dim3((unsigned)a.nparts), dim3(kAdjThreads), lds, st, h.col,
                       h.val, h.diag, (int)h.n, h.zmax, a.X, a.x_given, a.Xout, a.xout_t, a.G, a.Gout, a.gout_t,
                       a.ld, a.dt, a.K, a.partial, a.accumulate);
    return hipGetLastError();
}

}  // namespace

bool adjoint_supported(const DevHam &h, int K) {
    return h.n > 0 && h.n <= 2 * kAdjThreads && h.zmax <= 16 && (int64_t)K * h.n * 16 <= kAdjLdsBytes &&
           (h.n <= kAdjThreads || h.zmax <= 8);
}

int adjoint_parts(const DevHam &h) { return (int)std::min<int64_t>(h.n, 256); }

hipError_t launch_adjoint_side(const DevHam &h, const AdjointSide &a, hipStream_t st) {
    if (a.K == 0) {
        const int64_t total = h.n * h.n;
        hipLaunchKernelGGL(adjoint_identity_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a.X,
                           a.x_given, a.Xout, a.xout_t, a.G, a.Gout, a.gout_t, h.n, a.ld);
        return hipGetLastError();
    }
    const bool one = h.n <= kAdjThreads;
    if (h.zmax <= 4) return one ? launch_side_t<1, 4, true>(h, a, st) : launch_side_t<2, 4, false>(h, a, st);
    if (h.zmax <= 8) return one ? launch_side_t<1, 8, false>(h, a, st) : launch_side_t<2, 8, false>(h, a, st);
    if (h.zmax <= 12) return launch_side_t<1, 12, false>(h, a, st);
    return launch_side_t<1, 16, false>(h, a, st);
}

hipError_t launch_adjoint_reduce(const DevHam &h, const double2 *partial, int nparts, double scale,
                                 double *gonsite, double *gfield, hipStream_t st) {
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((unsigned)adjoint_reduce_blocks(h)), dim3(256), 0, st, h,
                       partial, nparts, scale, gonsite, gfield);
    return hipGetLastError();
}

int64_t adjoint_reduce_blocks(const DevHam &h) { return ((int64_t)(h.zmax + 1) * h.n + 63) / 64; }

int64_t reverse_term_block_cols(const DevHam &h) { return h.n >= 8192 ? 128 : 64; }

int64_t reverse_term_parts(const DevHam &h) {
    const int64_t cb = reverse_term_block_cols(h);
    return (h.n + cb - 1) / cb;
}

bool reverse_term_supported(const DevHam &h) { return h.zmax <= 16; }

bool reverse_term_band_fits(const DevHam &h) { return h.band >= 0 && h.band <= kRevRows && h.zmax <= 12; }

hipError_t prepare_reverse_term() {
    static std::atomic<uint64_t> done{0};
    return once_per_device(done, [] {
        const void *f[] = {(const void *)reverse_term_band_kernel<4, 1>, (const void *)reverse_term_band_kernel<8, 1>,
                           (const void *)reverse_term_band_kernel<9, 1>, (const void *)reverse_term_band_kernel<12, 1>,
                           (const void *)reverse_term_band_kernel<4, 2>, (const void *)reverse_term_band_kernel<8, 2>,
                           (const void *)reverse_term_band_kernel<9, 2>, (const void *)reverse_term_band_kernel<12, 2>};
        hipError_t e = hipSuccess;
        for (const void *k : f)
            if (e == hipSuccess) e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
        return e;
    });
}

hipError_t launch_reverse_term(const DevHam &h, const double2 *G, const double2 *Y, double2 *XB, double2 *Gout,
                               int m, int K, int64_t ld, double cr, double2 *partial, hipStream_t st, int variant) {
    const int64_t cb = reverse_term_block_cols(h);
    const bool last = m == K;
    // odd m: g_m + g_{m+1}; the last term: g_{K+1} (and g_K when K is odd)
    const int xbm = (m % 2 == 1 || last ? kXbTouch : 0) | (m > 1 ? kXbRead : 0) | (m % 2 == 1 ? kXbAddG : 0);
    if (variant != 1 && reverse_term_band_fits(h)) {
        hipError_t e = prepare_reverse_term();
        if (e != hipSuccess) return e;
        const int64_t rblocks = (h.n + kRevRows - 1) / kRevRows;
        const int64_t nreal = rblocks * reverse_term_parts(h);
        const int64_t nprog = ((nreal + kNumXCD - 1) / kNumXCD) * kNumXCD;
        // two columns per barrier when their double buffers fit the LDS
        const size_t lds1 = 2 * 2 * (size_t)(kRevRows + 2 * h.band) * sizeof(double2);
        const bool two = 2 * lds1 <= 163840;
        const size_t lds = two ? 2 * lds1 : lds1;
#define LMH_RTB(ZR, CPI)                                                                                          \
    hipLaunchKernelGGL((reverse_term_band_kernel<ZR, CPI>), dim3((unsigned)nprog), dim3(kRevRows), lds, st, h.col,  \
                       h.val, h.diag, h.n, h.zmax, (int)h.band, G, Y, XB, Gout, last ? 1 : 0, xbm, ld, cr, cb,        \
                       partial,                                                                               \
                       rblocks, nreal, nprog)
#define LMH_RTB_Z(ZR)        \
    if (two)                 \
        LMH_RTB(ZR, 2);      \
    else                     \
        LMH_RTB(ZR, 1);
        if (h.zmax <= 4) {
            LMH_RTB_Z(4)
        } else if (h.zmax <= 8) {
            LMH_RTB_Z(8)
        } else if (h.zmax <= 9) {
            LMH_RTB_Z(9)
        } else {
            LMH_RTB_Z(12)
        }
#undef LMH_RTB_Z
#undef LMH_RTB
        return hipGetLastError();
    }
    const dim3 grid((unsigned)((h.n + 255) / 256), (unsigned)reverse_term_parts(h));
#define LMH_RT(ZR)                                                                                                 \
    hipLaunchKernelGGL(reverse_term_kernel<ZR>, grid, dim3(256), 0, st, h.col, h.val, h.diag, h.n, h.zmax, G, Y, XB, \
                       Gout, last ? 1 : 0, xbm, ld, cr, cb, partial)
    if (h.zmax <= 4)
        LMH_RT(4);
    else if (h.zmax <= 8)
        LMH_RT(8);
    else if (h.zmax <= 12)
        LMH_RT(12);
    else
        LMH_RT(16);
#undef LMH_RT
    return hipGetLastError();
}

hipError_t launch_bond_derivative(const DevHam &h, int64_t comp, double2 *out, hipStream_t st) {
    if (h.nbonds == 0) return hipSuccess;
    hipLaunchKernelGGL(bond_derivative_kernel, dim3((unsigned)((h.nbonds + 255) / 256)), dim3(256), 0, st, h, comp,
                       out);
    return hipGetLastError();
}

hipError_t launch_hermitian_part(const double2 *src, double2 *dst, int64_t n, int64_t ld, hipStream_t st) {
    const int64_t tiles = (n + 31) / 32;
    hipLaunchKernelGGL(hermitian_part_kernel, dim3((unsigned)(tiles * tiles)), dim3(256), 0, st, src, dst, n, ld,
                       tiles);
    return hipGetLastError();
}

}  // namespace lmh
