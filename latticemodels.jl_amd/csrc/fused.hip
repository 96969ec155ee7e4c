// Fused Taylor propagator: all K Horner terms of y = p(-i dt H / s)^s x for
// whole columns of P held on chip, for N <= 10240 sites (one column of
// complex FP64 fills at most the 160 KB LDS of a CDNA4 CU).
//
// A workgroup (1024 threads) owns C columns (C * N * 16 B <= 160 KB of LDS)
// and each thread SPT sites of all of them.  Per Horner term
//     y_new(i) = x(i) + (-i dt/(s k)) (H y)(i)
// a thread reads H's row i once (ELL, from L2) and applies it to its C
// columns, reading neighbours from LDS; the new values wait in registers for
// the barrier, so one LDS buffer per column suffices (x, added back by every
// term, is re-read from the input columns).  HBM traffic per column: read x
// once, write y once -- instead of 3 streams per term for the per-term
// kernels -- and H's L2 traffic is divided by C.
//
// The two sides of P <- U P U' use (U P) U' = (U (U P)')': each side is this
// column pass with its result written conjugate-transposed (for a site i
// the C columns of a workgroup are contiguous in the transposed result), or
// written plainly and transposed by conj_transpose_kernel.
#include "lmhip_internal.h"

namespace lmh {

namespace {

constexpr int kFusedThreads = 1024;
constexpr int kFusedMaxN = 10240;
constexpr int kLdsBytes = 163840;

__device__ __forceinline__ int64_t xcd_swizzle(int64_t pid, int64_t nprog) {
    const int64_t per = nprog / kNumXCD;
    return (pid % kNumXCD) * per + pid / kNumXCD;
}

// Body shared by the single-matrix and the batched kernels: columns
// jb .. jb+C-1 of X (n sites) through all K Horner terms, result to Y.
template <int SPT, int C, bool TRANSPOSED_OUT>
__device__ __forceinline__ void fused_columns(
    double2 *__restrict__ ycol, const int32_t *__restrict__ col, const double2 *__restrict__ val,
    const double *__restrict__ diag, int n, int zmax, const double2 *__restrict__ X, double2 *__restrict__ Y,
    double2 *__restrict__ scratch, int64_t ld, int64_t ncols, int64_t jb, double dt_over_s, int K, int reps) {
    const int t = threadIdx.x;
    const double2 *xin = X;
#pragma unroll
    for (int cc = 0; cc < C; ++cc) {
        const int64_t j = jb + cc;
#pragma unroll
        for (int m = 0; m < SPT; ++m) {
            const int s = t + m * kFusedThreads;
            if (s < n) ycol[cc * n + s] = j < ncols ? xin[j * ld + s] : make_double2(0.0, 0.0);
        }
    }
    __syncthreads();
    for (int r = 0; r < reps; ++r) {
        for (int k = K; k >= 1; --k) {
            const double c = dt_over_s / (double)k;
            double2 ynew[SPT][C];
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int s = t + m * kFusedThreads;
                if (s < n) {
                    const double d = diag[s];
                    double2 acc[C];
#pragma unroll
                    for (int cc = 0; cc < C; ++cc) {
                        const double2 y0 = ycol[cc * n + s];
                        acc[cc] = make_double2(d * y0.x, d * y0.y);
                    }
                    for (int q = 0; q < zmax; ++q) {
                        const int64_t e = (int64_t)q * n + s;
                        const int32_t nb = col[e];
                        if (nb < 0) continue;
                        const double2 hv = val[e];
#pragma unroll
                        for (int cc = 0; cc < C; ++cc) {
                            const double2 yv = ycol[cc * n + nb];
                            acc[cc].x += hv.x * yv.x - hv.y * yv.y;
                            acc[cc].y += hv.x * yv.y + hv.y * yv.x;
                        }
                    }
#pragma unroll
                    for (int cc = 0; cc < C; ++cc) {
                        const int64_t j = jb + cc;
                        const double2 xv = j < ncols ? xin[j * ld + s] : make_double2(0.0, 0.0);
                        ynew[m][cc] = make_double2(xv.x + c * acc[cc].y, xv.y - c * acc[cc].x);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int s = t + m * kFusedThreads;
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
                    const int s = t + m * kFusedThreads;
                    if (s < n) scratch[j * ld + s] = ycol[cc * n + s];
                }
            }
            xin = scratch;
        }
    }
#pragma unroll
    for (int m = 0; m < SPT; ++m) {
        const int s = t + m * kFusedThreads;
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

template <int SPT, int C, bool TRANSPOSED_OUT>
__global__ void __launch_bounds__(kFusedThreads) fused_left_kernel(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag, int n,
    int zmax, const double2 *__restrict__ X, double2 *__restrict__ Y, double2 *__restrict__ scratch, int64_t ld,
    int64_t ncols, double dt_over_s, int K, int reps, int64_t nprog) {
    extern __shared__ double2 ycol[];            // [C][n]
    const int64_t jb = xcd_swizzle(blockIdx.x, nprog) * C;
    if (jb >= ncols) return;
    fused_columns<SPT, C, TRANSPOSED_OUT>(ycol, col, val, diag, n, zmax, X, Y, scratch, ld, ncols, jb, dt_over_s, K,
                                          reps);
}

// Many independent density matrices (realisations of different sizes) in
// one launch: workgroup b belongs to the item whose block range holds b.
template <int SPT, int C>
__global__ void __launch_bounds__(kFusedThreads) batched_fused_kernel(const BatchItem *__restrict__ items,
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
    fused_columns<SPT, C, true>(ycol, it.col, it.val, it.diag, it.n, it.zmax, it.src, it.dst, it.scratch, it.ld,
                                it.n, jb, it.dts, it.K, it.reps);
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

// B = A^H (conjugate transpose), n x n, 32x32 tiles through LDS.
__global__ void __launch_bounds__(256) conj_transpose_kernel(const double2 *__restrict__ A, double2 *__restrict__ B,
                                                             int64_t n, int64_t ld, int64_t tiles_per_dim) {
    __shared__ double2 tile[32][33];
    const int64_t tb = blockIdx.x;
    const int64_t ti = tb % tiles_per_dim, tj = tb / tiles_per_dim;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;     // 32 x 8
    const int64_t r0 = ti * 32, c0 = tj * 32;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t r = r0 + tx, c = c0 + ty + 8 * q;
        if (r < n && c < n) tile[ty + 8 * q][tx] = A[c * ld + r];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        // B[c0 + tx, r0 + ty'] = conj(A[r0 + ty', c0 + tx])
        const int64_t br = c0 + tx, bc = r0 + ty + 8 * q;
        if (br < n && bc < n) {
            const double2 v = tile[tx][ty + 8 * q];
            B[bc * ld + br] = make_double2(v.x, -v.y);
        }
    }
}

template <int SPT, int C, bool T>
hipError_t launch_fused_t(const DevHam &h, const double2 *X, double2 *Y, double2 *scratch, int64_t ld, int64_t ncols,
                          double dt_over_s, int K, int reps, hipStream_t st) {
    const size_t lds = (size_t)C * h.n * sizeof(double2);
    static bool attr_done = false;
    if (!attr_done) {
        hipError_t e = hipFuncSetAttribute((const void *)fused_left_kernel<SPT, C, T>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
        if (e != hipSuccess) return e;
        attr_done = true;
    }
    const int64_t groups = (ncols + C - 1) / C;
    const int64_t nprog = ((groups + kNumXCD - 1) / kNumXCD) * kNumXCD;
    hipLaunchKernelGGL((fused_left_kernel<SPT, C, T>), dim3((unsigned)nprog), dim3(kFusedThreads), lds, st, h.col,
                       h.val, h.diag, (int)h.n, h.zmax, X, Y, scratch, ld, ncols, dt_over_s, K, reps, nprog);
    return hipGetLastError();
}

// (sites per thread, columns per workgroup) by N: C * N * 16 B <= 160 KB,
// capped so the held new values stay in registers.
template <bool T>
hipError_t dispatch(const DevHam &h, const double2 *X, double2 *Y, double2 *scratch, int64_t ld, int64_t ncols,
                    double dt_over_s, int K, int reps, hipStream_t st) {
    const int64_t n = h.n;
    if (n <= 1024) return launch_fused_t<1, 8, T>(h, X, Y, scratch, ld, ncols, dt_over_s, K, reps, st);
    if (n <= 2048) return launch_fused_t<2, 4, T>(h, X, Y, scratch, ld, ncols, dt_over_s, K, reps, st);
    if (n <= 3072) return launch_fused_t<3, 3, T>(h, X, Y, scratch, ld, ncols, dt_over_s, K, reps, st);
    if (n <= 5120) return launch_fused_t<5, 2, T>(h, X, Y, scratch, ld, ncols, dt_over_s, K, reps, st);
    return launch_fused_t<10, 1, T>(h, X, Y, scratch, ld, ncols, dt_over_s, K, reps, st);
}

}  // namespace

bool fused_supported(const DevHam &h) { return h.n > 0 && h.n <= kFusedMaxN; }

hipError_t launch_fused_left(const DevHam &h, const double2 *X, double2 *Y, double2 *scratch, int64_t ld,
                             int64_t ncols, double dt_over_s, int K, int reps, bool transposed_out, hipStream_t st) {
    if (!fused_supported(h)) return hipErrorInvalidValue;
    if (reps > 1 && !scratch) return hipErrorInvalidValue;
    return transposed_out ? dispatch<true>(h, X, Y, scratch, ld, ncols, dt_over_s, K, reps, st)
                          : dispatch<false>(h, X, Y, scratch, ld, ncols, dt_over_s, K, reps, st);
}

int fused_class(int64_t n) {
    if (n <= 1024) return 0;
    if (n <= 2048) return 1;
    if (n <= 3072) return 2;
    if (n <= 5120) return 3;
    return 4;
}

int fused_class_columns(int cls) {
    static const int c[5] = {8, 4, 3, 2, 1};
    return c[cls];
}

template <int SPT, int C>
hipError_t launch_batched_t(const BatchItem *items, const int64_t *bstart, int nitems, int64_t total, int64_t maxn,
                            hipStream_t st) {
    static bool attr_done = false;
    if (!attr_done) {
        hipError_t e = hipFuncSetAttribute((const void *)batched_fused_kernel<SPT, C>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
        if (e != hipSuccess) return e;
        attr_done = true;
    }
    if (total == 0) return hipSuccess;
    const int64_t nprog = ((total + kNumXCD - 1) / kNumXCD) * kNumXCD;
    hipLaunchKernelGGL((batched_fused_kernel<SPT, C>), dim3((unsigned)nprog), dim3(kFusedThreads),
                       (size_t)C * maxn * sizeof(double2), st, items, bstart, nitems, total, nprog);
    return hipGetLastError();
}

hipError_t prepare_batched(int cls) {
    // the launchers set the LDS attribute on first use; do it before any
    // stream capture so no attribute call happens while capturing
    return launch_batched_fused(cls, nullptr, nullptr, 0, 0, 0, nullptr);
}

hipError_t launch_batched_fused(int cls, const BatchItem *items, const int64_t *bstart, int nitems, int64_t total,
                                int64_t maxn, hipStream_t st) {
    switch (cls) {
        case 0: return launch_batched_t<1, 8>(items, bstart, nitems, total, maxn, st);
        case 1: return launch_batched_t<2, 4>(items, bstart, nitems, total, maxn, st);
        case 2: return launch_batched_t<3, 3>(items, bstart, nitems, total, maxn, st);
        case 3: return launch_batched_t<5, 2>(items, bstart, nitems, total, maxn, st);
        default: return launch_batched_t<10, 1>(items, bstart, nitems, total, maxn, st);
    }
}

hipError_t launch_batched_mean_density(const BatchTrace *items, int nitems, double *out, hipStream_t st) {
    if (nitems == 0) return hipSuccess;
    hipLaunchKernelGGL(batched_mean_density_kernel, dim3((unsigned)nitems), dim3(256), 0, st, items, out);
    return hipGetLastError();
}

hipError_t launch_conj_transpose(const double2 *A, double2 *B, int64_t n, int64_t ld, hipStream_t st) {
    const int64_t tiles = (n + 31) / 32;
    hipLaunchKernelGGL(conj_transpose_kernel, dim3((unsigned)(tiles * tiles)), dim3(256), 0, st, A, B, n, ld, tiles);
    return hipGetLastError();
}

}  // namespace lmh
