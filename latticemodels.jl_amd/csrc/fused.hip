// Fused Taylor propagator: all K Horner terms of y = p(-i dt H / s)^s x for
// whole columns of P held on chip, for N <= 10240 sites (one column of
// complex FP64 fills at most the 160 KB LDS of a CDNA4 CU).
//
// A workgroup of NT threads owns C columns (C * N * 16 B <= 160 KB of LDS)
// and each thread SPT sites of all of them.  Per Horner term
//     y_new(i) = x(i) + (-i dt/(s k)) (H y)(i)
// a thread applies H's row i to its C columns, reading neighbours from LDS;
// the new values wait in registers for the barrier, so one LDS buffer per
// column suffices (x, added back by every term, is re-read from the input
// columns).  HBM traffic per column: read x once, write y once -- instead of
// 3 streams per term for the per-term kernels.  H's row comes from L2 every
// term (its traffic divided by C), or, for one or two sites per thread, is
// held in registers for the whole launch (HREG variants).
//
// Variants (threads, sites/thread, columns, H in registers) trade registers
// for occupancy: with NT threads a thread may hold 128 (1024), 256 (512) or
// 512 (256) VGPRs, and the SPT x C new values must fit without spilling.
//
// The two sides of P <- U P U' use (U P) U' = (U (U P)')': each side is this
// column pass with its result written conjugate-transposed (for a site i
// the C columns of a workgroup are contiguous in the transposed result), or
// written plainly and transposed by conj_transpose_kernel.
//
// ZQ == 2 (value classes): a site's record carries, per ELL slot, a column
// offset id and a class id instead of a column, and H's values come from the
// class table copied into LDS beside the columns: one 16-byte L2 load per
// site and term instead of a record and four values (80 B).
#include "lmhip_internal.h"

namespace lmh {

namespace {

constexpr int kFusedMaxN = 10240;
constexpr int kLdsBytes = 163840;
constexpr int kSiteGroup = 2;        // sites per scheduling group in the term loop

__device__ __forceinline__ int64_t xcd_swizzle(int64_t pid, int64_t nprog) {
    const int64_t per = nprog / kNumXCD;
    return (pid % kNumXCD) * per + pid / kNumXCD;
}

// Body shared by the single-matrix and the batched kernels: columns
// jb .. jb+C-1 of X (n sites) through all K Horner terms, result to Y.
template <int NT, int SPT, int C, int HREG, bool XREG, int ZQ, bool TRANSPOSED_OUT>
__device__ __forceinline__ void fused_columns(
    double2 *__restrict__ ycol, const int32_t *__restrict__ col, const int32_t *__restrict__ col4,
    const double2 *__restrict__ rec, const double2 *__restrict__ dval, const int32_t *__restrict__ doff, int ndict,
    const double2 *__restrict__ val, const double *__restrict__ diag, int n, int zmax, const double2 *__restrict__ X, double2 *__restrict__ Y,
    double2 *__restrict__ scratch, int64_t ld, int64_t ncols, int64_t jb, double dt_over_s, int K, int reps) {
    const int t = threadIdx.x;
    // ZQ == 2: class table and column offsets after the C columns
    double2 *tab = ycol + (size_t)C * n;
    int32_t *offs = reinterpret_cast<int32_t *>(tab + ndict);
    if constexpr (ZQ == 2) {
        for (int e = t; e < ndict; e += NT) tab[e] = dval[e];
        if (t < kDictMaxOffsets) offs[t] = doff[t];
    }
    const double2 *xin = X;
    constexpr int HS = HREG ? SPT : 1;
    constexpr int HZ = HREG ? HREG : 1;
    constexpr int XS = XREG ? SPT : 1;
    constexpr int XC = XREG ? C : 1;
    double2 xr[XS][XC];        // XREG: x held in registers, not re-read every term
    int32_t hc[HS][HZ];
    double2 hv[HS][HZ];
    double hd[HS];
    if constexpr (HREG) {
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int s = t + m * NT;
            hd[m] = s < n ? diag[s] : 0.0;
#pragma unroll
            for (int q = 0; q < HZ; ++q) {
                const int64_t e = (int64_t)q * n + s;
                const bool ok = s < n && q < zmax;
                const int32_t cq = ok ? col[e] : -1;
                hc[m][q] = cq < 0 ? 0 : cq;
                hv[m][q] = cq < 0 ? make_double2(0.0, 0.0) : val[e];
            }
        }
    }
#pragma unroll
    for (int cc = 0; cc < C; ++cc) {
        const int64_t j = jb + cc;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int s = t + m * NT;
            const double2 v = (s < n && j < ncols) ? xin[j * ld + s] : make_double2(0.0, 0.0);
            if (s < n) ycol[cc * n + s] = v;
            if constexpr (XREG) xr[m][cc] = v;
        }
    }
    __syncthreads();
    for (int r = 0; r < reps; ++r) {
        for (int k = K; k >= 1; --k) {
            const double c = dt_over_s / (double)k;
            // opaque copies of n and ld: keeps the compiler from hoisting the
            // SPT x zmax per-site addresses out of the term loop into registers
            int nn = n;
            int64_t ldd = ld;
            asm volatile("" : "+s"(nn), "+s"(ldd));
            double2 ynew[SPT][C];
            // ZQ == 2: the next site's record is loaded while this one is used
            double2 rnext = make_double2(0.0, 0.0);
            if constexpr (ZQ == 2) rnext = rec[t < nn ? t : 0];
            // branch-free: sites beyond n compute on site 0 (discarded), empty
            // ELL slots (column -1) hold the value 0 and read column 0
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int s0 = t + m * NT;
                const int s = s0 < nn ? s0 : 0;
                // ZQ == 1: one 16-byte record holds the diagonal and the four
                // columns (16 bit) of site s
                double2 r1 = make_double2(0.0, 0.0);
                if constexpr (ZQ == 1 && !HREG) r1 = rec[s];
                if constexpr (ZQ == 2) {
                    r1 = rnext;
                    if (m + 1 < SPT) {
                        const int sn = t + (m + 1) * NT;
                        rnext = rec[sn < nn ? sn : 0];
                    }
                }
                const double d = HREG ? hd[m < HS ? m : 0] : (ZQ >= 1 ? r1.x : diag[s]);
                double2 acc[C];
#pragma unroll
                for (int cc = 0; cc < C; ++cc) {
                    const double2 y0 = ycol[cc * nn + s];
                    acc[cc] = make_double2(d * y0.x, d * y0.y);
                }
                if constexpr (HREG) {
#pragma unroll
                    for (int q = 0; q < HZ; ++q) {
                        const int32_t nb = hc[m][q];
                        const double2 hvq = hv[m][q];
#pragma unroll
                        for (int cc = 0; cc < C; ++cc) {
                            const double2 yv = ycol[cc * nn + nb];
                            acc[cc].x += hvq.x * yv.x - hvq.y * yv.y;
                            acc[cc].y += hvq.x * yv.y + hvq.y * yv.x;
                        }
                        // one slot's C neighbour reads in flight at a time
                        if constexpr (C >= 4) __builtin_amdgcn_sched_barrier(0);
                    }
                } else if constexpr (ZQ == 2) {
                    // four slots: offset id | class | conjugate, 16 bits each
                    const unsigned long long pk = (unsigned long long)__double_as_longlong(r1.y);
                    double2 hq[4];
                    int32_t nbs[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const uint32_t sl = (uint32_t)(pk >> (16 * u)) & 0xffffu;
                        nbs[u] = s + offs[sl & 15u];
                        const double2 v = tab[(sl >> 4) & 0x7ffu];
                        hq[u] = make_double2(v.x, (sl >> 15) ? -v.y : v.y);
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
#pragma unroll
                        for (int cc = 0; cc < C; ++cc) {
                            const double2 yv = ycol[cc * nn + nbs[u]];
                            acc[cc].x += hq[u].x * yv.x - hq[u].y * yv.y;
                            acc[cc].y += hq[u].x * yv.y + hq[u].y * yv.x;
                        }
                        if constexpr (C >= 3) __builtin_amdgcn_sched_barrier(0);
                    }
                } else {
                    // one chunk: four ELL slots of site s, their column record
                    // and values loaded together before the LDS gathers (a
                    // slot at a time would serialise the L2 round trips)
                    auto chunk = [&](int qc) {
                        const char *vb = reinterpret_cast<const char *>(val);
                        double2 hq[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            hq[u] = *reinterpret_cast<const double2 *>(vb + (uint32_t)((qc * 4 + u) * nn + s) * 16u);
                        int32_t nbs[4];
                        if constexpr (ZQ == 1) {
                            const unsigned long long cb = (unsigned long long)__double_as_longlong(r1.y);
#pragma unroll
                            for (int u = 0; u < 4; ++u) nbs[u] = (int32_t)((cb >> (16 * u)) & 0xffffu);
                        } else {
                            const char *c4 = reinterpret_cast<const char *>(col4);
                            const int4 nb4 = *reinterpret_cast<const int4 *>(c4 + (uint32_t)(qc * nn + s) * 16u);
                            nbs[0] = nb4.x < 0 ? 0 : nb4.x;
                            nbs[1] = nb4.y < 0 ? 0 : nb4.y;
                            nbs[2] = nb4.z < 0 ? 0 : nb4.z;
                            nbs[3] = nb4.w < 0 ? 0 : nb4.w;
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
#pragma unroll
                            for (int cc = 0; cc < C; ++cc) {
                                const double2 yv = ycol[cc * nn + nbs[u]];
                                acc[cc].x += hq[u].x * yv.x - hq[u].y * yv.y;
                                acc[cc].y += hq[u].x * yv.y + hq[u].y * yv.x;
                            }
                            // several columns: one slot's reads in flight at a time
                            if constexpr (C >= 3) __builtin_amdgcn_sched_barrier(0);
                        }
                    };
                    if constexpr (ZQ == 1) {
                        chunk(0);
                    } else {
                        const int nq = (zmax + 3) >> 2;
                        for (int qc = 0; qc < nq; ++qc) chunk(qc);
                    }
                }
#pragma unroll
                for (int cc = 0; cc < C; ++cc) {
                    double2 xv;
                    if constexpr (XREG) {
                        xv = xr[m][cc];
                    } else {
                        const int64_t j = jb + cc < ncols ? jb + cc : jb;
                        xv = xin[j * ldd + s];
                    }
                    ynew[m][cc] = make_double2(xv.x + c * acc[cc].y, xv.y - c * acc[cc].x);
                    // computed here, not sunk past the barrier into the store
                    // (which would keep every gathered value live across it)
                    if constexpr (ZQ > 0) asm volatile("" : "+v"(ynew[m][cc].x), "+v"(ynew[m][cc].y));
                }
                if constexpr (ZQ > 0) {
                    // a whole chunk per site is in flight: one site at a time
                    __builtin_amdgcn_sched_barrier(0);
                } else if constexpr (SPT >= 4) {
                    if (m % kSiteGroup == kSiteGroup - 1) __builtin_amdgcn_sched_barrier(0);
                }
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int s = t + m * NT;
                if (s < n) {
#pragma unroll
                    for (int cc = 0; cc < C; ++cc) ycol[cc * n + s] = ynew[m][cc];
                }
            }
            __syncthreads();
        }
        if (r + 1 < reps) {
            // x <- this repetition's result, staged in the workgroup's own
            // columns of a free buffer; each thread re-reads only what it wrote
#pragma unroll
            for (int cc = 0; cc < C; ++cc) {
                const int64_t j = jb + cc;
                if (j >= ncols) continue;
#pragma unroll
                for (int m = 0; m < SPT; ++m) {
                    const int s = t + m * NT;
                    if (s < n) scratch[j * ld + s] = ycol[cc * n + s];
                }
            }
            xin = scratch;
            if constexpr (XREG) {
#pragma unroll
                for (int cc = 0; cc < C; ++cc)
#pragma unroll
                    for (int m = 0; m < SPT; ++m) {
                        const int s = t + m * NT;
                        xr[m][cc] = s < n ? ycol[cc * n + s] : make_double2(0.0, 0.0);
                    }
            }
        }
    }
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int s = t + m * NT;
        if (s >= n) continue;
#pragma unroll
        for (int cc = 0; cc < C; ++cc) {
            const int64_t j = jb + cc;
            if (j >= ncols) continue;
            const double2 v = ycol[cc * n + s];
            if (TRANSPOSED_OUT)
                Y[(int64_t)s * ld + j] = make_double2(v.x, -v.y);
            else
                Y[j * ld + s] = v;
        }
    }
}

template <int NT, int SPT, int C, int HREG, bool XREG, int ZQ, bool TRANSPOSED_OUT>
__global__ void __launch_bounds__(NT) fused_left_kernel(
    const int32_t *__restrict__ col, const int32_t *__restrict__ col4, const double2 *__restrict__ rec,
    const double2 *__restrict__ dval, const int32_t *__restrict__ doff, int ndict, const double2 *__restrict__ val, const double *__restrict__ diag, int n, int zmax, const double2 *__restrict__ X, double2 *__restrict__ Y,
    double2 *__restrict__ scratch, int64_t ld, int64_t ncols, double dt_over_s, int K, int reps, int64_t nprog) {
    extern __shared__ double2 ycol[];            // [C][n]
    const int64_t jb = xcd_swizzle(blockIdx.x, nprog) * C;
    if (jb >= ncols) return;
    fused_columns<NT, SPT, C, HREG, XREG, ZQ, TRANSPOSED_OUT>(ycol, col, col4, rec, dval, doff, ndict, val, diag, n, zmax,
                                                         X, Y, scratch, ld, ncols, jb, dt_over_s, K, reps);
}

// Many independent density matrices (realisations of different sizes) in
// one launch: workgroup b belongs to the item whose block range holds b.
template <int NT, int SPT, int C, int HREG, bool XREG, int ZQ>
__global__ void __launch_bounds__(NT) batched_fused_kernel(const BatchItem *__restrict__ items,
                                                           const int64_t *__restrict__ bstart, int nitems,
                                                           int64_t total, int64_t nprog) {
    extern __shared__ double2 ycol[];
    const int64_t b = xcd_swizzle(blockIdx.x, nprog);
    if (b >= total) return;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const BatchItem it = items[lo];
    const int64_t jb = (b - bstart[lo]) * C;
    fused_columns<NT, SPT, C, HREG, XREG, ZQ, true>(ycol, it.col, it.col4, ZQ == 2 ? it.drec : it.rec, it.dval, it.doff,
                                                it.ndict, it.val, it.diag, it.n, it.zmax, it.src, it.dst,
                                                it.scratch, it.ld, it.n, jb, it.dts, it.K, it.reps);
}

// mean(real(diag P)) per item: one workgroup per item.
__global__ void __launch_bounds__(256) batched_mean_density_kernel(const BatchTrace *__restrict__ items,
                                                                   double *__restrict__ out) {
    __shared__ double part[256];
    const BatchTrace it = items[blockIdx.x];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < it.n; i += 256) acc += it.P[i * it.ld + i].x;
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = part[0] / (double)it.n;
}

// B = A^H (conjugate transpose), n x n, 64x64 tiles through LDS (64 KB): a
// wave reads 64 consecutive rows of one column of A and writes 64 consecutive
// rows of one column of B (1 KB each); element (r, c) of the tile sits at
// column c ^ r of its LDS row, which spreads both access directions over the
// banks.  512 threads, 8 elements each: at N = 28800 5.0 ms (5.3 TB/s of
// read + write) against 6.0 ms with 256 threads and 16 elements each.
constexpr int kTransposeThreads = 512;
__global__ void __launch_bounds__(kTransposeThreads) conj_transpose_kernel(const double2 *__restrict__ A,
                                                                           double2 *__restrict__ B, int64_t n,
                                                                           int64_t ld, int64_t tiles_per_dim) {
    constexpr int R = kTransposeThreads / 64, Q = 64 / R;
    __shared__ double2 tile[64][64];
    const int64_t tb = blockIdx.x;
    const int64_t ti = tb % tiles_per_dim, tj = tb / tiles_per_dim;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;     // 64 x R
    const int64_t r0 = ti * 64, c0 = tj * 64;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int64_t r = r0 + tx, c = c0 + ty + R * q;
        if (r < n && c < n) tile[ty + R * q][tx ^ (ty + R * q)] = A[c * ld + r];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        // B[c0 + tx, r0 + ty'] = conj(A[r0 + ty', c0 + tx])
        const int64_t br = c0 + tx, bc = r0 + ty + R * q;
        if (br < n && bc < n) {
            const double2 v = tile[tx][(ty + R * q) ^ tx];
            B[bc * ld + br] = make_double2(v.x, -v.y);
        }
    }
}

template <int NT, int SPT, int C, int HREG, bool XREG, int ZQ, bool T>
hipError_t launch_fused_t(const DevHam &h, const double2 *X, double2 *Y, double2 *scratch, int64_t ld, int64_t ncols,
                          double dt_over_s, int K, int reps, hipStream_t st) {
    const size_t lds = (size_t)C * h.n * sizeof(double2) +
                       (ZQ == 2 ? (size_t)h.ndict * sizeof(double2) + kDictMaxOffsets * sizeof(int32_t) : 0);
    static std::atomic<uint64_t> attr_done{0};
    hipError_t e = once_per_device(attr_done, [] {
        return hipFuncSetAttribute((const void *)fused_left_kernel<NT, SPT, C, HREG, XREG, ZQ, T>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
    });
    if (e != hipSuccess) return e;
    if (ncols == 0) return hipSuccess;
    const int64_t groups = (ncols + C - 1) / C;
    const int64_t nprog = ((groups + kNumXCD - 1) / kNumXCD) * kNumXCD;
    hipLaunchKernelGGL((fused_left_kernel<NT, SPT, C, HREG, XREG, ZQ, T>), dim3((unsigned)nprog), dim3(NT), lds, st, h.col,
                       h.col4, ZQ == 2 ? h.drec : h.rec, h.dval, h.doff, h.ndict, h.val, h.diag, (int)h.n, h.zmax, X, Y,
                       scratch, ld, ncols, dt_over_s, K, reps, nprog);
    return hipGetLastError();
}

template <int NT, int SPT, int C, int HREG, bool XREG, int ZQ>
hipError_t launch_batched_t(const BatchItem *items, const int64_t *bstart, int nitems, int64_t total, int64_t maxn,
                            int maxdict, hipStream_t st) {
    static std::atomic<uint64_t> attr_done{0};
    hipError_t e = once_per_device(attr_done, [] {
        return hipFuncSetAttribute((const void *)batched_fused_kernel<NT, SPT, C, HREG, XREG, ZQ>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
    });
    if (e != hipSuccess) return e;
    if (total == 0) return hipSuccess;
    const int64_t nprog = ((total + kNumXCD - 1) / kNumXCD) * kNumXCD;
    const size_t lds = (size_t)C * maxn * sizeof(double2) +
                       (ZQ == 2 ? (size_t)maxdict * sizeof(double2) + kDictMaxOffsets * sizeof(int32_t) : 0);
    hipLaunchKernelGGL((batched_fused_kernel<NT, SPT, C, HREG, XREG, ZQ>), dim3((unsigned)nprog), dim3(NT), lds, st,
                       items, bstart, nitems, total, nprog);
    return hipGetLastError();
}

// The variant table: {threads, sites per thread, columns, H slots held in
// registers (0: read from L2 every term), x held in registers, 4-slot chunks
// of H per site fixed at compile time (0: ceil(zmax/4) at run time; 2: value
// classes, see above)}.  The batched pass uses the first group only.
struct Variant {
    int nt, spt, c, hreg, xreg, zq;
};
#define LMH_FUSED_CASES(X)                                                                                         \
    X(0, 1024, 1, 8, 0, false, 0) X(1, 1024, 2, 4, 0, false, 0) X(2, 1024, 3, 3, 0, false, 0)                       \
    X(3, 1024, 5, 2, 0, false, 0) X(4, 1024, 10, 1, 0, false, 0) X(5, 1024, 1, 4, 4, false, 0)                      \
    X(6, 1024, 1, 2, 4, false, 0) X(7, 512, 2, 4, 4, false, 0) X(8, 512, 4, 4, 0, false, 0)                         \
    X(9, 512, 6, 3, 0, false, 0) X(10, 512, 10, 2, 0, false, 0) X(11, 512, 20, 1, 0, false, 0)                      \
    X(12, 1024, 10, 1, 0, true, 0) X(13, 512, 20, 1, 0, true, 0) X(14, 1024, 5, 2, 0, true, 0)                      \
    X(15, 512, 10, 2, 0, true, 0) X(16, 1024, 1, 2, 4, true, 0) X(17, 1024, 2, 4, 0, true, 0)                       \
    X(18, 1024, 3, 3, 0, true, 0) X(19, 1024, 1, 8, 0, true, 0) X(20, 1024, 10, 1, 0, true, 1)                      \
    X(21, 512, 20, 1, 0, true, 1) X(22, 1024, 5, 2, 0, true, 1) X(23, 512, 10, 2, 0, true, 1)                       \
    X(24, 1024, 2, 4, 0, true, 1) X(25, 1024, 3, 3, 0, true, 1) X(26, 1024, 1, 8, 0, true, 1)                       \
    X(27, 1024, 10, 1, 0, false, 1) X(28, 512, 20, 1, 0, false, 1)
#define LMH_FUSED_DICT_CASES(X) \
    X(29, 1024, 10, 1, 0, true, 2) X(30, 512, 20, 1, 0, true, 2) X(31, 1024, 5, 2, 0, true, 2)
#define LMH_VARIANT_ROW(idx, nt, spt, c, hreg, xreg, zq) {nt, spt, c, hreg, xreg, zq},
constexpr Variant kVariants[] = {LMH_FUSED_CASES(LMH_VARIANT_ROW) LMH_FUSED_DICT_CASES(LMH_VARIANT_ROW)};
#undef LMH_VARIANT_ROW
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
// the strip kernel (strip.hip) for grid-stencil H with uniform strip bonds
constexpr int kStripVariant = kNumVariants;

template <bool T>
hipError_t dispatch(int v, const DevHam &h, const double2 *X, double2 *Y, double2 *scratch, int64_t ld,
                    int64_t ncols, double dt_over_s, int K, int reps, hipStream_t st) {
    switch (v) {
#define LMH_CASE(idx, nt, spt, c, hreg, xreg, zq) \
    case idx: return launch_fused_t<nt, spt, c, hreg, xreg, zq, T>(h, X, Y, scratch, ld, ncols, dt_over_s, K, reps, st);
        LMH_FUSED_CASES(LMH_CASE)
        LMH_FUSED_DICT_CASES(LMH_CASE)
#undef LMH_CASE
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool fused_supported(const DevHam &h) { return h.n > 0 && h.n <= kFusedMaxN; }

int fused_variant_count() { return kNumVariants + 1; }

int fused_strip_variant() { return kStripVariant; }

bool fused_variant_valid(int v, int64_t n, int zmax) {
    // the strip variant runs where the last H build allows it, else the
    // by-size default: valid wherever a fused pass is
    if (v == kStripVariant) return n > 0 && n <= kFusedMaxN;
    if (v < 0 || v >= kNumVariants) return false;
    const Variant &w = kVariants[v];
    const int64_t extra = w.zq == 2 ? 4 * kDictMaxOffsets : 0;
    return n > 0 && n <= (int64_t)w.nt * w.spt && (int64_t)w.c * n * 16 + extra <= kLdsBytes &&
           (!w.hreg || zmax <= w.hreg) && (w.zq == 2 ? zmax <= 4 : (!w.zq || zmax <= 4 * w.zq));
}

bool fused_variant_is_dict(int v) { return v >= 0 && v < kNumVariants && kVariants[v].zq == 2; }

int fused_variant_fallback(int v) {
    if (!fused_variant_is_dict(v)) return v;
    const Variant &w = kVariants[v];
    for (int u = 0; u < kNumVariants; ++u) {
        const Variant &x = kVariants[u];
        if (x.zq == 1 && x.nt == w.nt && x.spt == w.spt && x.c == w.c && x.hreg == w.hreg && x.xreg == w.xreg) return u;
    }
    return -1;
}

// Defaults that read the class table where it holds: only the one-column
// shape, where the class lookups replace H's L2 stream (variant 29: 3.07
// against 3.41 ms per launch at N = 10^4, Landau gauge).  With two columns per
// workgroup the stream is already shared and the lookups cost more than they
// save (31 against 22: 1.00 against 0.93 ms per step at N = 4096).
int fused_default_dict(int v) { return v == 20 ? 29 : -1; }

bool fused_dict_fits(int v, int64_t n, int ndict) {
    return fused_variant_is_dict(v) && (int64_t)kVariants[v].c * n * 16 + 16 * (int64_t)ndict + 4 * kDictMaxOffsets <= kLdsBytes;
}

int fused_variant_columns(int v) { return v == kStripVariant ? 1 : kVariants[v].c; }

// Default variant by size, from timing every variant on square lattices
// (N = 1024 .. 10^4, K = 8): x held in registers wins at every size; H rows in
// registers for N <= 1024; above that, for zmax <= 4, the packed site record
// (diagonal + 16-bit columns, one 16-byte load) beats separate loads by 5-10%.
int fused_default_variant(int64_t n, int zmax) {
    const bool packed = zmax <= 4;
    if (n <= 1024) return packed ? 16 : 19;
    if (n <= 2048) return packed ? 24 : 17;
    if (n <= 3072) return packed ? 25 : 18;
    if (n <= 5120) return packed ? 22 : 14;
    return packed ? 20 : 12;
}

// Default variant of a batch size class.
int fused_batch_variant(int64_t n, int zmax) { return fused_default_variant(n, zmax); }

int fused_resolved_variant(const DevHam &h, int variant) {
    if ((variant < 0 || variant == kStripVariant) && h.strip_ok && h.sg.s > 0) return kStripVariant;
    int v = variant < 0 || variant == kStripVariant ? fused_default_variant(h.n, h.zmax) : variant;
    if (variant < 0 && fused_default_dict(v) >= 0) v = fused_default_dict(v);
    if (fused_variant_is_dict(v) && !(h.ndict > 0 && h.dict_ok && fused_dict_fits(v, h.n, h.ndict)))
        v = fused_variant_fallback(v);
    return v;
}

hipError_t launch_fused_left(const DevHam &h, const double2 *X, double2 *Y, double2 *scratch, int64_t ld,
                             int64_t ncols, double dt_over_s, int K, int reps, bool transposed_out, int variant,
                             hipStream_t st) {
    if (!fused_supported(h)) return hipErrorInvalidValue;
    // the strip kernel by default wherever the last H build allows it
    if ((variant < 0 || variant == kStripVariant) && h.strip_ok && h.sg.s > 0)
        return launch_strip(h, X, Y, ld, ncols, dt_over_s, K, reps, transposed_out, st);
    if (variant == kStripVariant) variant = -1;
    if (reps > 1 && !scratch) return hipErrorInvalidValue;
    int v = variant < 0 ? fused_default_variant(h.n, h.zmax) : variant;
    if (variant < 0 && fused_default_dict(v) >= 0) v = fused_default_dict(v);
    // the class table only when the last H build matched it and it fits
    if (fused_variant_is_dict(v) && !(h.ndict > 0 && h.dict_ok && fused_dict_fits(v, h.n, h.ndict)))
        v = fused_variant_fallback(v);
    if (!fused_variant_valid(v, h.n, h.zmax)) return hipErrorInvalidValue;
    return transposed_out ? dispatch<true>(v, h, X, Y, scratch, ld, ncols, dt_over_s, K, reps, st)
                          : dispatch<false>(v, h, X, Y, scratch, ld, ncols, dt_over_s, K, reps, st);
}

hipError_t prepare_fused(int variant) {
    if (variant == kStripVariant) return hipSuccess;
    // sets the LDS attribute of every instantiation of this variant before
    // any stream capture (no attribute call happens while capturing)
    DevHam empty;
    empty.n = 1;
    hipError_t e = dispatch<true>(variant, empty, nullptr, nullptr, nullptr, 1, 0, 0.0, 0, 1, nullptr);
    if (e == hipSuccess) e = dispatch<false>(variant, empty, nullptr, nullptr, nullptr, 1, 0, 0.0, 0, 1, nullptr);
    return e;
}

hipError_t launch_batched_fused(int variant, const BatchItem *items, const int64_t *bstart, int nitems, int64_t total,
                                int64_t maxn, int maxdict, hipStream_t st) {
    switch (variant) {
#define LMH_CASE(idx, nt, spt, c, hreg, xreg, zq) \
    case idx: return launch_batched_t<nt, spt, c, hreg, xreg, zq>(items, bstart, nitems, total, maxn, maxdict, st);
        LMH_FUSED_CASES(LMH_CASE)
        LMH_FUSED_DICT_CASES(LMH_CASE)
#undef LMH_CASE
        default: return hipErrorInvalidValue;
    }
}

hipError_t prepare_batched(int variant) {
    return launch_batched_fused(variant, nullptr, nullptr, 0, 0, 0, 0, nullptr);
}

hipError_t launch_batched_mean_density(const BatchTrace *items, int nitems, double *out, hipStream_t st) {
    if (nitems == 0) return hipSuccess;
    hipLaunchKernelGGL(batched_mean_density_kernel, dim3((unsigned)nitems), dim3(256), 0, st, items, out);
    return hipGetLastError();
}

hipError_t launch_conj_transpose(const double2 *A, double2 *B, int64_t n, int64_t ld, hipStream_t st) {
    const int64_t tiles = (n + 63) / 64;
    hipLaunchKernelGGL(conj_transpose_kernel, dim3((unsigned)(tiles * tiles)), dim3(kTransposeThreads), 0, st, A, B, n,
                       ld, tiles);
    return hipGetLastError();
}

}  // namespace lmh
