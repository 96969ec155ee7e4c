// Fused Taylor pass for grid-stencil Hamiltonians: the "strip" kernel.
//
// The square lattices of LatticeModels.jl (SquareLattice, site i = x Ly + y,
// src/lattices/bravais/lattice.jl:101-111) give an H whose row i holds the
// diagonal and the four grid neighbours i +- 1 (y +- 1) and i +- Ly (x +- 1),
// across the edges for periodic boundaries.  Under a Landau gauge
// (src/zoo/magneticfields.jl:11-15) a y bond's phase depends on x only and an
// x bond carries none, so along a run of consecutive sites of one x column
// every bond of one direction has the same value.  This kernel uses both
// facts:
//
//   * a thread owns a strip of S consecutive sites (x, y0 .. y0+S-1) of a
//     column of P and keeps y and x of the Horner recursion in registers; the
//     y +- 1 neighbours are its own registers (the strip ends read one value
//     of the neighbouring strip), so per site and term only the two x +- 1
//     neighbours come from LDS, plus one write of the site's new value;
//   * the four bond values of a strip are four complex registers, read once
//     per launch; the diagonal (on-site disorder) is streamed from L2, six
//     sites ahead.
//
// Per Horner term  y <- x + (-i dt / (s k)) H y  the new values of a column
// are written to its LDS copy (one x column per ly-rounded-to-odd run, so the
// x +- 1 reads of a wave's consecutive lanes hit distinct banks), a barrier,
// then every thread reads its two neighbour strips and updates its sites.
// The column is loaded with gfx950 direct-to-LDS loads (global_load_lds
// 16 B per lane), which keeps 160 KB per CU in flight without VGPRs: one
// workgroup per CU loads at the HBM rate that a plain read kernel reaches.
// The workgroup is persistent over its share of the columns.
//
// Whether H qualifies is decided twice: the pattern on the host when the
// geometry is set (strip_detect), and the values after every H build on the
// device (strip_check_kernel: every used bond of a strip's direction equal
// to the strip's value, bit for bit).  When either fails, the fused passes
// run the general kernels of fused.hip, so this kernel changes speed, never
// results.  Reference step: src/evolution.jl:73-78 (U P U' with U from
// myexp!, :93-128); the polynomial is the reference's order-K Taylor
// polynomial evaluated in Horner form.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <set>
#include <type_traits>
#include <vector>

#include "lmhip_internal.h"

namespace lmh {

namespace {

constexpr int kStripThreads = 512;      // 8 waves, 256 VGPRs each: one workgroup per CU
constexpr int kStripLdsBytes = 163840;
constexpr int kStripMaxN = 10240;
constexpr int kDiagAhead = 6;           // diagonal loads in flight ahead of their site
constexpr int kLdsAhead = 2;            // neighbour-strip reads ahead of their site
constexpr int kStripS[] = {4, 6, 8, 10, 12, 16, 18, 20};
// (kStripMaxK: Horner coefficients passed with the launch or read from a
// TaylorOut, lmhip_internal.h)

// one Taylor polynomial p(-i dts H)^reps of a pass
struct StripPoly {
    const double2 *stval;   // [4][nt] bond value per direction: y+, y-, x+, x-
    const double *sdiag;    // [s][nt] diagonal per strip position
    int K;
    int reps;
    double dts;             // dt / s
    double ck[kStripMaxK + 1];  // ck[k] = dts / k for k <= min(K, kStripMaxK)
    const TaylorOut *coef;  // when set: K, reps and ck read from here instead
};

// a device TaylorOut read through the constant address space: wave-uniform
// scalar loads (nothing writes it while a strip launch runs)
typedef __attribute__((address_space(4))) const TaylorOut ConstTaylorOut;

struct StripArgs {
    StripGeom g;
    StripPoly poly[2];      // applied to each column in this order
    int npoly;              // 1, or 2: the second step's first side fused onto the first's second
    const double2 *X;
    double2 *Y;
    int64_t ld;
    int64_t ncols;
    int n;
    int64_t nprog;          // workgroups (a multiple of 8)
    int64_t stagger;        // start delay (shader cycles) of every other group of 8 workgroups
    int xreal;              // batched kernel: the member's x bonds are real
    const TaylorOut *guard; // when set: the launch runs only if strip_coef_ok(*guard, guard_xreal)
    int guard_xreal;
};

__device__ __forceinline__ void cmac(double2 &acc, double2 h, double2 v) {
    acc.x = fma(h.x, v.x, acc.x);
    acc.x = fma(-h.y, v.y, acc.x);
    acc.y = fma(h.x, v.y, acc.y);
    acc.y = fma(h.y, v.x, acc.y);
}

// MI355X deals workgroups round-robin to its 8 XCDs: logical workgroup
// numbers run consecutively on one XCD, so the conjugate-transposed stores of
// neighbouring columns (one 128-byte line holds 8 of them) meet in one L2.
__device__ __forceinline__ int64_t strip_xcd_order(int64_t b, int64_t nprog) {
    const int64_t per = nprog / kNumXCD;
    return (b % kNumXCD) * per + b / kNumXCD;
}

// direct-to-LDS loads of the column group starting at jb: one wave
// instruction per 64-site chunk of an x column
__device__ __forceinline__ void strip_dma(const StripArgs &a, double2 *lds, int64_t jb, int wave, int lane) {
    const StripGeom &g = a.g;
    int64_t ldg = a.ld;
    asm volatile("" : "+s"(ldg));
    const int nch = (g.ly + 63) >> 6;
    const int per_col = g.lx * nch;
    const int npair = g.cols * per_col;
    for (int p = wave; p < npair; p += kStripThreads / 64) {
        const int gg = p / per_col;
        const int rem = p - gg * per_col;
        const int xc = rem / nch, ch = rem - xc * nch;
        const int64_t jj = jb + gg;
        const int y = ch * 64 + lane;
        if (jj < a.ncols && y < g.ly) {
            const double2 *src = a.X + jj * ldg + (int64_t)xc * g.ly + y;
            __builtin_amdgcn_global_load_lds(
                (const void *)src,
                (__attribute__((address_space(3))) void *)(lds + gg * g.region + xc * g.lyp + ch * 64), 16, 0, 0);
        }
    }
}

// plain result of the column group starting at jb from its LDS copy: the
// inverse of strip_dma, one 1 KB contiguous store per wave instruction
__device__ __forceinline__ void strip_store(const StripArgs &a, const double2 *lds, int64_t jb, int wave, int lane) {
    const StripGeom &g = a.g;
    int64_t ldg = a.ld;
    asm volatile("" : "+s"(ldg));
    const int nch = (g.ly + 63) >> 6;
    const int per_col = g.lx * nch;
    const int npair = g.cols * per_col;
    for (int p = wave; p < npair; p += kStripThreads / 64) {
        const int gg = p / per_col;
        const int rem = p - gg * per_col;
        const int xc = rem / nch, ch = rem - xc * nch;
        const int64_t jj = jb + gg;
        const int y = ch * 64 + lane;
        if (jj < a.ncols && y < g.ly)
            a.Y[jj * ldg + (int64_t)xc * g.ly + y] = lds[gg * g.region + xc * g.lyp + y];
    }
}

// acc += h v for a complex h, or for a real h (XR: h.y == 0, two FMAs)
template <bool XR>
__device__ __forceinline__ void hmac(double2 &acc, double2 h, double2 v) {
    if constexpr (XR) {
        acc.x = fma(h.x, v.x, acc.x);
        acc.y = fma(h.x, v.y, acc.y);
    } else {
        cmac(acc, h, v);
    }
}

// One column group of a strip pass, its columns already in LDS (with the
// zero slot set).  TOUT: the result is written conjugate-transposed
// (Y[s, j] = conj(y_s)), else plainly; XR: the x bonds are real (Landau
// gauge, no field).  At the end the next group (na, njb, when next) is
// loaded into LDS with its zero slot, ready for the next call.
// a thread's place in the column group: its strip, its neighbour strips'
// LDS positions and its bond values
struct StripThread {
    int slot, tl, x, y0, skip, base, xp, xm, lo, hi;
    bool act;
    double2 hyp, hym, hxp, hxm;                  // BONDS: the first polynomial's bond values
};

template <int S, bool BONDS = true>
__device__ __forceinline__ StripThread strip_thread(const StripArgs &a) {
    const StripGeom &g = a.g;
    StripThread p;
    const int t = threadIdx.x;
    p.slot = t / g.nt;                           // column of the workgroup this thread works on
    p.act = p.slot < g.cols;
    p.tl = p.act ? t - p.slot * g.nt : 0;        // idle threads mirror thread 0 (no writes)
    const int sl = p.act ? p.slot : 0;
    p.x = p.tl % g.lx;
    const int k = p.tl / g.lx;
    p.y0 = min(k * S, g.ly - S);
    p.skip = k * S - p.y0;                       // sites shared with the previous strip
    const int rb = sl * g.region;
    const int zero = rb + g.lx * g.lyp;          // holds 0: the absent neighbour of an open edge
    p.base = rb + p.x * g.lyp + p.y0;
    // neighbour strips (x +- 1) and the strip ends' outside neighbours
    // (y0 - 1, y0 + S); an absent x neighbour reads the own strip (its bond
    // value is 0), an absent y neighbour the zero entry
    p.xp = p.x + 1 < g.lx ? p.base + g.lyp : (g.px ? rb + p.y0 : p.base);
    p.xm = p.x > 0 ? p.base - g.lyp : (g.px ? rb + (g.lx - 1) * g.lyp + p.y0 : p.base);
    p.lo = p.y0 > 0 ? p.base - 1 : (g.py ? rb + p.x * g.lyp + g.ly - 1 : zero);
    p.hi = p.y0 + S < g.ly ? p.base + S : (g.py ? rb + p.x * g.lyp : zero);
    if constexpr (BONDS) {
        const double2 *sv = a.poly[0].stval;
        p.hyp = sv[p.tl];
        p.hym = sv[g.nt + p.tl];
        p.hxp = sv[2 * g.nt + p.tl];
        p.hxm = sv[3 * g.nt + p.tl];
    }
    return p;
}

// One column group of a strip pass, its columns already in LDS (with the
// zero slot set).  TOUT: the result is written conjugate-transposed
// (Y[s, j] = conj(y_s)), else plainly; XR: the x bonds are real (Landau
// gauge, no field).  TWO: a.npoly polynomials in turn, each with its bond
// values loaded when it starts (else one, its bond values in p).  At the end
// the next group (na, njb, when next) is loaded into LDS with its zero slot,
// ready for the next call.
template <int S, bool TOUT, bool XR, bool TWO>
__device__ __forceinline__ void strip_group(const StripArgs &a, const StripThread &p, double2 *lds, int64_t jb,
                                            const StripArgs &na, int64_t njb, bool next) {
    const StripGeom &g = a.g;
    const int t = threadIdx.x;
    const int slot = p.slot, tl = p.tl, x = p.x, y0 = p.y0, skip = p.skip, base = p.base;
    const int xp = p.xp, xm = p.xm, lo = p.lo, hi = p.hi;
    const bool act = p.act;
    const int wave = t >> 6, lane = t & 63;
    const int64_t j = jb + slot;
    double2 xr[S], yv[S];
#pragma unroll
    for (int m = 0; m < S; ++m) {
        xr[m] = lds[base + m];
        yv[m] = xr[m];
    }
    for (int ph = 0; ph < (TWO ? a.npoly : 1); ++ph) {
    // this polynomial's parameters (selected, not indexed: a dynamic index
    // into the argument block would copy it to scratch memory)
    const double2 *stv = ph ? a.poly[1].stval : a.poly[0].stval;
    const double *sdg = ph ? a.poly[1].sdiag : a.poly[0].sdiag;
    const TaylorOut *pcf = ph ? a.poly[1].coef : a.poly[0].coef;
    ConstTaylorOut *cf = (ConstTaylorOut *)pcf;
    const int PK = pcf ? cf->K : (ph ? a.poly[1].K : a.poly[0].K);
    const int preps = pcf ? cf->reps : (ph ? a.poly[1].reps : a.poly[0].reps);
    const double pdts = ph ? a.poly[1].dts : a.poly[0].dts;
    double2 hyp = p.hyp, hym = p.hym, hxp = p.hxp, hxm = p.hxm;
    if constexpr (TWO) {
        // bond values (x bonds: real parts only when XR); the opaque copy
        // keeps their four addresses from being hoisted out of the loops
        int tq = tl;
        asm volatile("" : "+v"(tq));
        hyp = stv[tq];
        hym = stv[g.nt + tq];
        hxp = XR ? make_double2(stv[2 * g.nt + tq].x, 0.0) : stv[2 * g.nt + tq];
        hxm = XR ? make_double2(stv[3 * g.nt + tq].x, 0.0) : stv[3 * g.nt + tq];
    }
    // One Horner term: publish the input (unless the column just loaded is
    // it), then every site of the strip in turn, upwards (UP) or downwards.
    // A site's new value takes the register its neighbour's old value leaves
    // free, one place over in the sweep's direction; alternating the
    // direction puts the values back in place after two terms instead of
    // moving them all at the loop's end.  Each site's arithmetic is the same
    // either way.
    auto term = [&](int kk, bool publish, auto up) {
        constexpr bool UP = decltype(up)::value;
        if (publish) {
            // the previous term's reads are done; publish this term's input
            __syncthreads();
            if (act) {
#pragma unroll
                for (int m = 0; m < S; ++m) lds[base + m] = yv[m];
            }
            __syncthreads();
        }
        const double c = pcf ? cf->ck[kk]
                             : (kk <= kStripMaxK ? (ph ? a.poly[1].ck[kk] : a.poly[0].ck[kk]) : pdts / (double)kk);
        // opaque copy: keeps the S diagonal addresses from being hoisted
        // out of the term loop into registers
        int ntq = g.nt;
        asm volatile("" : "+s"(ntq));
        const double *sd = sdg + tl;
        // the old value behind the sweep (first: the strip's outside
        // neighbour where the sweep starts) and the one past its end
        double2 back = lds[UP ? lo : hi];
        const double2 endv = lds[UP ? hi : lo];
        double2 av[S], bv[S];
        double dv[S];
#pragma unroll
        for (int i = 0; i < kLdsAhead && i < S; ++i) {
            const int m = UP ? i : S - 1 - i;
            av[m] = lds[xp + m];
            bv[m] = lds[xm + m];
        }
#pragma unroll
        for (int i = 0; i < kDiagAhead && i < S; ++i) {
            const int m = UP ? i : S - 1 - i;
            dv[m] = sd[m * ntq];
        }
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int m = UP ? i : S - 1 - i;
            if (i + kLdsAhead < S) {
                const int mf = UP ? m + kLdsAhead : m - kLdsAhead;
                av[mf] = lds[xp + mf];
                bv[mf] = lds[xm + mf];
            }
            if (i + kDiagAhead < S) {
                const int mf = UP ? m + kDiagAhead : m - kDiagAhead;
                dv[mf] = sd[mf * ntq];
            }
            const double2 ahead = i + 1 < S ? yv[UP ? m + 1 : m - 1] : endv;
            const double2 cur = yv[m];
            const double2 nyp = UP ? ahead : back;     // site y + 1 (old value)
            const double2 nym = UP ? back : ahead;     // site y - 1 (old value)
            double2 acc = make_double2(dv[m] * cur.x, dv[m] * cur.y);
            cmac(acc, hyp, nyp);
            cmac(acc, hym, nym);
            hmac<XR>(acc, hxp, av[m]);
            hmac<XR>(acc, hxm, bv[m]);
            back = cur;
            // y_new = x + (-i c) acc
            yv[m] = make_double2(fma(c, acc.y, xr[m].x), fma(-c, acc.x, xr[m].y));
            // one site at a time: bounds the live neighbour values
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    for (int rr = 0; rr < preps; ++rr) {
        // (the first term reads the loaded column from LDS)
        int kk = PK;
        for (; kk >= 2; kk -= 2) {
            term(kk, ph > 0 || rr > 0 || kk != PK, std::true_type{});
            term(kk - 1, true, std::false_type{});
        }
        // an odd term count ends on a downward sweep (an upward one here
        // needs more registers than the kernel has)
        if (kk == 1) term(1, ph > 0 || rr > 0 || PK != 1, std::false_type{});
        // the next repetition (or polynomial) acts on this one's result
        if (rr + 1 < preps || (TWO && ph + 1 < a.npoly)) {
#pragma unroll
            for (int m = 0; m < S; ++m) xr[m] = yv[m];
        }
    }
    }
    // the last term's LDS reads are done
    __syncthreads();
    if constexpr (TOUT) {
        // the next column group loads while this one's results leave from
        // registers: Y[s, j] = conj(y_s), a column of the result is a row of Y
        if (next) strip_dma(na, lds, njb, wave, lane);
        if (act && j < a.ncols) {
            int64_t ldq = a.ld;
            asm volatile("" : "+s"(ldq));
            double2 *yo = a.Y + ((int64_t)x * g.ly + y0) * ldq + j;
#pragma unroll
            for (int m = 0; m < S; ++m) {
                if (m >= skip) *yo = make_double2(yv[m].x, -yv[m].y);
                yo += ldq;
            }
        }
    } else {
        // plain: the results go through the LDS copy so that each wave store
        // covers 64 consecutive sites (a thread's own 16-byte stores would be
        // one request each); then the next group loads
        if (act) {
#pragma unroll
            for (int m = 0; m < S; ++m) lds[base + m] = yv[m];
        }
        __syncthreads();
        strip_store(a, lds, jb, wave, lane);
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        __syncthreads();
        if (next) strip_dma(na, lds, njb, wave, lane);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    // the next group's zero slots (past its sites: the loads leave them alone)
    if (next && t < na.g.cols) lds[t * na.g.region + na.g.lx * na.g.lyp] = make_double2(0.0, 0.0);
    __syncthreads();
}

template <int S, bool TOUT, bool XR, bool TWO>
__global__ void __launch_bounds__(kStripThreads) strip_kernel(StripArgs a) {
    extern __shared__ double2 lds[];
    // a polynomial the host did not see: nothing to do unless it is the one
    // the launch was chosen for (the host tests the same and redoes the step)
    if (a.guard && !strip_coef_ok(*a.guard, a.guard_xreal != 0)) return;
    const StripGeom &g = a.g;
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    int64_t r = strip_xcd_order(blockIdx.x, a.nprog);
    int64_t jb = r * g.cols;
    // every other group of 8 workgroups (neighbouring columns on one XCD)
    // starts about half a column period later: the workgroups then do not
    // all load and store at the same time nor all compute at the same time
    if (a.stagger > 0 && ((r >> 3) & 1)) {
        const int64_t t0 = (int64_t)clock64();
        while ((int64_t)clock64() - t0 < a.stagger) __builtin_amdgcn_s_sleep(16);
    }
    if (t < g.cols) lds[t * g.region + g.lx * g.lyp] = make_double2(0.0, 0.0);
    if (jb < a.ncols) strip_dma(a, lds, jb, wave, lane);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const StripThread p = strip_thread<S, !TWO>(a);
    while (jb < a.ncols) {
        const int64_t rn = r + a.nprog, jbn = rn * g.cols;
        strip_group<S, TOUT, XR, TWO>(a, p, lds, jb, a, jbn, jbn < a.ncols);
        r = rn;
        jb = jbn;
    }
}

// the batch member that owns work unit u (ustart: prefix sums of the
// members' column-group counts)
__device__ __forceinline__ int strip_member_of(const int64_t *__restrict__ ustart, int nm, int64_t u) {
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ustart[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <bool TOUT, bool XR>
__device__ __forceinline__ void strip_group_s(const StripArgs &a, double2 *lds, int64_t jb, const StripArgs &na,
                                              int64_t njb, bool next) {
    switch (a.g.s) {
#define LMH_STRIP_GROUP(SS) \
    case SS: strip_group<SS, TOUT, XR, false>(a, strip_thread<SS>(a), lds, jb, na, njb, next); break;
        LMH_STRIP_GROUP(4) LMH_STRIP_GROUP(6) LMH_STRIP_GROUP(8) LMH_STRIP_GROUP(10) LMH_STRIP_GROUP(12)
        LMH_STRIP_GROUP(16) LMH_STRIP_GROUP(18) LMH_STRIP_GROUP(20)
#undef LMH_STRIP_GROUP
        default: break;
    }
}

// Many independent density matrices (batch members of any grid size and
// strip length) in one persistent launch: the work units are the members'
// column groups, concatenated; workgroup r takes units r, r + nprog, ...
// (consecutive units, i.e. neighbouring columns of one member, on one XCD).
template <bool TOUT>
__global__ void __launch_bounds__(kStripThreads) batched_strip_kernel(const StripArgs *__restrict__ members,
                                                                      const int64_t *__restrict__ ustart, int nm,
                                                                      int64_t nunits, int64_t nprog) {
    extern __shared__ double2 lds[];
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    int64_t u = strip_xcd_order(blockIdx.x, nprog);
    if (u >= nunits) return;
    int m = strip_member_of(ustart, nm, u);
    int64_t jb = (u - ustart[m]) * members[m].g.cols;
    if (t < members[m].g.cols) lds[t * members[m].g.region + members[m].g.lx * members[m].g.lyp] = make_double2(0.0, 0.0);
    strip_dma(members[m], lds, jb, wave, lane);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    while (u < nunits) {
        const int64_t un = u + nprog;
        const int mn = un < nunits ? strip_member_of(ustart, nm, un) : m;
        const int64_t jbn = un < nunits ? (un - ustart[mn]) * members[mn].g.cols : 0;
        if (members[m].xreal)
            strip_group_s<TOUT, true>(members[m], lds, jb, members[mn], jbn, un < nunits);
        else
            strip_group_s<TOUT, false>(members[m], lds, jb, members[mn], jbn, un < nunits);
        u = un;
        m = mn;
        jb = jbn;
    }
}

// Copies for a deferred pass whose polynomial the host has not seen: the
// strip values and the order record, only when the launch it follows ran
// (the same test as its guard), so that a launch that did nothing leaves the
// earlier pass's copies in place.
__global__ void __launch_bounds__(256) strip_snapshot_kernel(const double2 *__restrict__ sv, const double *__restrict__ sd,
                                                             int64_t nsv, int64_t nsd, const TaylorOut *__restrict__ coef,
                                                             int xreal, double2 *__restrict__ sv_out,
                                                             double *__restrict__ sd_out,
                                                             TaylorOut *__restrict__ coef_out) {
    if (!strip_coef_ok(*coef, xreal != 0)) return;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nsv; e += stride) sv_out[e] = sv[e];
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nsd; e += stride) sd_out[e] = sd[e];
    if (blockIdx.x == 0 && threadIdx.x == 0) *coef_out = *coef;
}

// After build_h: each strip thread's bond value per direction (all used
// bonds of the direction must agree bit for bit, else bit 0 of red[3]; bit 1
// when an x bond is complex) and its diagonal in strip order.  One thread per
// (direction, strip thread); the diagonal one per (position, strip thread).
__global__ void __launch_bounds__(256) strip_check_kernel(DevHam h) {
    const StripGeom &g = h.sg;
    const int64_t idx = blockIdx.x * 256 + threadIdx.x;
    const int S = g.s;
    const int64_t n = h.n;
    if (idx < 4 * g.nt) {
        const int d = (int)(idx / g.nt), tl = (int)(idx - (int64_t)d * g.nt);
        const int x = tl % g.lx, k = tl / g.lx;
        const int y0 = min(k * S, g.ly - S);
        bool bad = false, have = false;
        double2 u = make_double2(0.0, 0.0);
        for (int m = 0; m < S; ++m) {
            bool used;
            if (d == 0) used = m + 1 < S || y0 + S < g.ly || g.py;
            else if (d == 1) used = m > 0 || y0 > 0 || g.py;
            else if (d == 2) used = x + 1 < g.lx || g.px;
            else used = x > 0 || g.px;
            if (!used) continue;
            const int64_t i = (int64_t)x * g.ly + y0 + m;
            const int32_t q = h.sslot[d * n + i];
            if (q < 0) {
                bad = true;
                continue;
            }
            const double2 v = h.val[q];
            if (!have) {
                u = v;
                have = true;
            } else if (v.x != u.x || v.y != u.y) {
                bad = true;
            }
        }
        h.stval[d * g.nt + tl] = u;
        const bool xcomplex = d >= 2 && u.y != 0.0;
        if (bad || xcomplex) atomicOr(&h.red[3], (bad ? 1ull : 0ull) | (xcomplex ? 2ull : 0ull));
    }
    for (int64_t e = idx; e < (int64_t)S * g.nt; e += (int64_t)gridDim.x * 256) {
        const int m = (int)(e / g.nt), tl = (int)(e - (int64_t)m * g.nt);
        const int x = tl % g.lx, k = tl / g.lx;
        const int y0 = min(k * S, g.ly - S);
        h.sdiag[e] = h.diag[(int64_t)x * g.ly + y0 + m];
    }
}

template <int S, bool T, bool XR, bool TWO>
const void *strip_fn() {
    return (const void *)strip_kernel<S, T, XR, TWO>;
}

// the dynamic-LDS attribute, once per device and instantiation
template <int S, bool T, bool XR, bool TWO = false>
hipError_t strip_attr() {
    static std::atomic<uint64_t> done{0};
    return once_per_device(done, [] {
        return hipFuncSetAttribute(strip_fn<S, T, XR, TWO>(), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   kStripLdsBytes);
    });
}

template <int S, bool T, bool XR, bool TWO>
hipError_t launch_s(const StripArgs &a, size_t lds, hipStream_t st) {
    hipError_t e = strip_attr<S, T, XR, TWO>();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((strip_kernel<S, T, XR, TWO>), dim3((unsigned)a.nprog), dim3(kStripThreads), lds, st, a);
    return hipGetLastError();
}

// two polynomials in one launch only with conjugate-transposed stores
template <int S>
hipError_t launch_sx(bool tout, bool xr, const StripArgs &a, size_t lds, hipStream_t st) {
    if (a.npoly > 1)
        return xr ? launch_s<S, true, true, true>(a, lds, st) : launch_s<S, true, false, true>(a, lds, st);
    if (tout) return xr ? launch_s<S, true, true, false>(a, lds, st) : launch_s<S, true, false, false>(a, lds, st);
    return xr ? launch_s<S, false, true, false>(a, lds, st) : launch_s<S, false, false, false>(a, lds, st);
}

template <int S>
hipError_t prepare_sx() {
    hipError_t e = strip_attr<S, true, true>();
    if (e == hipSuccess) e = strip_attr<S, true, false>();
    if (e == hipSuccess) e = strip_attr<S, false, true>();
    if (e == hipSuccess) e = strip_attr<S, false, false>();
    if (e == hipSuccess) e = strip_attr<S, true, true, true>();
    if (e == hipSuccess) e = strip_attr<S, true, false, true>();
    return e;
}

#define LMH_STRIP_LENGTHS(X) X(4) X(6) X(8) X(10) X(12) X(16) X(18) X(20)

hipError_t dispatch_s(int s, bool tout, bool xr, const StripArgs &a, size_t lds, hipStream_t st) {
    switch (s) {
#define LMH_STRIP_CASE(S) \
    case S: return launch_sx<S>(tout, xr, a, lds, st);
        LMH_STRIP_LENGTHS(LMH_STRIP_CASE)
#undef LMH_STRIP_CASE
        default: return hipErrorInvalidValue;
    }
}

// H's row pattern as an lx x ly grid (site i = x ly + y)?  Fills the slot of
// each direction of each site.
bool try_grid(const int32_t *col, int64_t n, int zmax, int64_t lx, int64_t ly, int px, int py, int32_t *sslot) {
    std::fill(sslot, sslot + 4 * n, -1);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t x = i / ly, y = i % ly;
        int64_t nb[4];
        nb[0] = y + 1 < ly ? i + 1 : (py ? i - (ly - 1) : -1);
        nb[1] = y > 0 ? i - 1 : (py ? i + ly - 1 : -1);
        nb[2] = x + 1 < lx ? i + ly : (px ? i - (lx - 1) * ly : -1);
        nb[3] = x > 0 ? i - ly : (px ? i + (lx - 1) * ly : -1);
        for (int u = 0; u < zmax; ++u) {
            const int64_t q = (int64_t)u * n + i;
            const int32_t c = col[q];
            if (c < 0) continue;
            int d = 0;
            while (d < 4 && nb[d] != c) ++d;
            if (d == 4 || sslot[d * n + i] >= 0) return false;
            sslot[d * n + i] = (int32_t)q;
        }
        for (int d = 0; d < 4; ++d)
            if (nb[d] >= 0 && sslot[d * n + i] < 0) return false;
    }
    return true;
}

bool has_entry(const int32_t *col, int64_t n, int zmax, int64_t i, int64_t j) {
    for (int u = 0; u < zmax; ++u)
        if (col[(int64_t)u * n + i] == j) return true;
    return false;
}

}  // namespace

bool strip_detect(const int32_t *col, int64_t n, int zmax, StripGeom *g, int32_t *sslot) {
    *g = StripGeom{};
    if (n < 9 || n > kStripMaxN || zmax < 1 || zmax > 4) return false;
    std::set<int64_t> cands;
    for (int64_t q = 0; q < (int64_t)zmax * n; ++q) {
        if (col[q] < 0) continue;
        const int64_t d = std::llabs((int64_t)col[q] - q % n);
        for (int64_t c : {d, d + 1, n - d})
            if (c >= 3 && c < n && n % c == 0) cands.insert(c);
    }
    for (int64_t ly : cands) {
        const int64_t lx = n / ly;
        if (lx < 2) continue;
        // the edges are periodic when their wrap-around bond is there
        const int py = ly >= 3 && has_entry(col, n, zmax, ly - 1, 0);
        const int px = lx >= 3 && has_entry(col, n, zmax, (lx - 1) * ly, 0);
        if (!try_grid(col, n, zmax, lx, ly, px, py, sslot)) continue;
        // strip length: the least overlap (then the longest strip) that fits
        // 512 threads and one column's LDS copy
        const int lyp = (int)(ly | 1);
        const int64_t region = lx * lyp + 1;
        const int64_t cols_lds = kStripLdsBytes / (16 * region);
        if (cols_lds < 1) continue;
        int best = 0;
        double best_ratio = 1e300;
        for (int s : kStripS) {
            if (s > ly) continue;
            const int64_t qn = (ly + s - 1) / s, nt = lx * qn;
            if (nt > kStripThreads) continue;
            const double ratio = (double)(qn * s) / (double)ly;
            if (ratio < best_ratio - 1e-12 || (ratio <= best_ratio + 1e-12 && s > best)) {
                best = s;
                best_ratio = ratio;
            }
        }
        if (!best) continue;
        g->lx = (int)lx;
        g->ly = (int)ly;
        g->lyp = lyp;
        g->px = px;
        g->py = py;
        g->s = best;
        g->q = (int)((ly + best - 1) / best);
        g->nt = (int)lx * g->q;
        g->cols = (int)std::min<int64_t>(kStripThreads / g->nt, cols_lds);
        g->region = (int)region;
        return true;
    }
    std::fill(sslot, sslot + 4 * n, -1);
    return false;
}

int64_t strip_thread_count(const StripGeom &g) { return g.nt; }

hipError_t launch_strip_check(const DevHam &h, hipStream_t st) {
    if (h.sg.s == 0) return hipSuccess;
    const int64_t work = (int64_t)std::max(4, h.sg.s) * h.sg.nt;
    hipLaunchKernelGGL(strip_check_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, st, h);
    return hipGetLastError();
}

namespace {

StripPoly strip_poly(const double2 *stval, const double *sdiag, double dt_over_s, int K, int reps) {
    StripPoly q{};
    q.stval = stval;
    q.sdiag = sdiag;
    q.K = K;
    q.reps = reps;
    q.dts = dt_over_s;
    for (int k = 0; k <= kStripMaxK; ++k) q.ck[k] = k ? dt_over_s / (double)k : 0.0;
    return q;
}

StripArgs strip_args(const DevHam &h, const double2 *X, double2 *Y, int64_t ld, int64_t ncols, double dt_over_s,
                     int K, int reps) {
    StripArgs a{};
    a.g = h.sg;
    a.poly[0] = strip_poly(h.stval, h.sdiag, dt_over_s, K, reps);
    a.npoly = 1;
    a.X = X;
    a.Y = Y;
    a.ld = ld;
    a.ncols = ncols;
    a.n = (int)h.n;
    a.xreal = h.strip_xreal != 0;
    return a;
}

hipError_t device_cus(int *cus) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev);
    return e;
}

}  // namespace

namespace {

// persistent grid, start stagger and LDS size of a strip pass, then the launch
hipError_t launch_args(const DevHam &h, StripArgs &a, int64_t ncols, bool tout, hipStream_t st) {
    const StripGeom &g = h.sg;
    int cus = 0;
    hipError_t e = device_cus(&cus);
    if (e != hipSuccess) return e;
    const int64_t groups = (ncols + g.cols - 1) / g.cols;
    const int64_t np = std::min<int64_t>(groups, std::max(cus, 1));
    a.nprog = ((np + kNumXCD - 1) / kNumXCD) * kNumXCD;
    // half a column period in shader cycles (measured at N = 10^4, S = 20:
    // ~310 cycles per site and term, ~3.3 cycles per site for the load and
    // store), when each workgroup has several column groups to do
    int64_t terms = 0;
    for (int q = 0; q < a.npoly; ++q) terms += (int64_t)a.poly[q].K * a.poly[q].reps;
    a.stagger = groups >= 4 * a.nprog ? (int64_t)((310.0 * terms * g.s + 3.3 * (double)h.n * g.cols) / 2.0) : 0;
    const size_t lds = (size_t)g.cols * g.region * sizeof(double2);
    return dispatch_s(g.s, tout, a.xreal != 0, a, lds, st);
}

bool strip_shape_ok(const DevHam &h, int64_t ncols, bool tout) {
    const StripGeom &g = h.sg;
    return g.s != 0 && h.n == (int64_t)g.lx * g.ly && (!tout || ncols == h.n);
}

}  // namespace

hipError_t launch_strip(const DevHam &h, const double2 *X, double2 *Y, int64_t ld, int64_t ncols, double dt_over_s,
                        int K, int reps, bool tout, hipStream_t st, const StripPending *second,
                        const TaylorOut *coef) {
    if (!strip_shape_ok(h, ncols, tout) || (second && !tout)) return hipErrorInvalidValue;
    if (ncols == 0) return hipSuccess;
    StripArgs a = strip_args(h, X, Y, ld, ncols, dt_over_s, K, reps);
    a.poly[0].coef = coef;
    a.guard = coef;
    a.guard_xreal = h.strip_xreal != 0;
    if (second) {
        // the previous step's second side first, then this step's first side
        a.poly[1] = a.poly[0];
        a.poly[0] = strip_poly(second->stval, second->sdiag, second->dts, second->K, second->reps);
        a.poly[0].coef = second->coef;
        a.npoly = 2;
        a.xreal = a.xreal && second->xreal;
    }
    return launch_args(h, a, ncols, tout, st);
}

hipError_t launch_strip_pending(const DevHam &h, const StripPending &q, const double2 *X, double2 *Y, int64_t ld,
                                int64_t ncols, bool tout, hipStream_t st) {
    if (!strip_shape_ok(h, ncols, tout)) return hipErrorInvalidValue;
    if (ncols == 0) return hipSuccess;
    StripArgs a = strip_args(h, X, Y, ld, ncols, q.dts, q.K, q.reps);
    a.poly[0].stval = q.stval;
    a.poly[0].sdiag = q.sdiag;
    a.poly[0].coef = q.coef;
    a.xreal = q.xreal;
    return launch_args(h, a, ncols, tout, st);
}

hipError_t launch_strip_snapshot(const DevHam &h, const TaylorOut *coef, bool xreal, double2 *stval, double *sdiag,
                                 TaylorOut *coef_out, hipStream_t st) {
    const int64_t nsv = strip_stval_count(h), nsd = strip_sdiag_count(h);
    const int64_t work = std::max(nsv, nsd);
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, 1024));
    hipLaunchKernelGGL(strip_snapshot_kernel, dim3((unsigned)blocks), dim3(256), 0, st, h.stval, h.sdiag, nsv, nsd,
                       coef, (int)xreal, stval, sdiag, coef_out);
    return hipGetLastError();
}

hipError_t strip_batch_table(const StripMemberDesc *m, int nm, void **d_args, int64_t **d_ustart, int64_t *nunits,
                             size_t *lds) {
    *d_args = nullptr;
    *d_ustart = nullptr;
    std::vector<StripArgs> args(nm);
    std::vector<int64_t> ustart(nm + 1, 0);
    size_t need = 0;
    for (int i = 0; i < nm; ++i) {
        const DevHam &h = *m[i].h;
        if (h.sg.s == 0 || h.n != (int64_t)h.sg.lx * h.sg.ly) return hipErrorInvalidValue;
        args[i] = strip_args(h, m[i].X, m[i].Y, m[i].ld, h.n, m[i].dts, m[i].K, m[i].reps);
        ustart[i + 1] = ustart[i] + (h.n + h.sg.cols - 1) / h.sg.cols;
        need = std::max(need, (size_t)h.sg.cols * h.sg.region * sizeof(double2));
    }
    StripArgs *da = nullptr;
    int64_t *du = nullptr;
    hipError_t e = hipMalloc((void **)&da, sizeof(StripArgs) * std::max(nm, 1));
    if (e == hipSuccess) e = hipMalloc((void **)&du, sizeof(int64_t) * (nm + 1));
    if (e == hipSuccess && nm) e = hipMemcpy(da, args.data(), sizeof(StripArgs) * nm, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(du, ustart.data(), sizeof(int64_t) * (nm + 1), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (da) (void)hipFree(da);
        if (du) (void)hipFree(du);
        return e;
    }
    *d_args = da;
    *d_ustart = du;
    *nunits = ustart[nm];
    *lds = need;
    return hipSuccess;
}

hipError_t launch_batched_strip(const void *d_args, const int64_t *d_ustart, int nm, int64_t nunits, size_t lds,
                                bool tout, hipStream_t st) {
    if (nunits == 0) return hipSuccess;
    static std::atomic<uint64_t> done{0};
    int cus = 0;
    hipError_t e = device_cus(&cus);
    if (e != hipSuccess) return e;
    e = once_per_device(done, [] {
        hipError_t r = hipFuncSetAttribute((const void *)batched_strip_kernel<true>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kStripLdsBytes);
        if (r == hipSuccess)
            r = hipFuncSetAttribute((const void *)batched_strip_kernel<false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kStripLdsBytes);
        return r;
    });
    if (e != hipSuccess) return e;
    const int64_t np = std::min<int64_t>(nunits, std::max(cus, 1));
    const int64_t nprog = ((np + kNumXCD - 1) / kNumXCD) * kNumXCD;
    const auto *args = static_cast<const StripArgs *>(d_args);
    if (tout)
        hipLaunchKernelGGL(batched_strip_kernel<true>, dim3((unsigned)nprog), dim3(kStripThreads), lds, st, args,
                           d_ustart, nm, nunits, nprog);
    else
        hipLaunchKernelGGL(batched_strip_kernel<false>, dim3((unsigned)nprog), dim3(kStripThreads), lds, st, args,
                           d_ustart, nm, nunits, nprog);
    return hipGetLastError();
}

hipError_t prepare_strip(const DevHam &h) {
    switch (h.sg.s) {
        case 0: return hipSuccess;
#define LMH_STRIP_PREP(S) \
    case S: return prepare_sx<S>();
        LMH_STRIP_LENGTHS(LMH_STRIP_PREP)
#undef LMH_STRIP_PREP
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lmh
