// Taylor propagator kernels for P <- U P U', U = sum_{n<=K} (-i dt H/s)^n / n!
// (applied s times per side).
//
//  * taylor_order_kernel: the reference's order rule (src/evolution.jl:93-118).
//    Column j of T_n = droptol((H/s) T_{n-1}[:, j] * (-i dt / n), nztol) only
//    involves column j, so one wavefront walks the K-hop neighbourhood of site
//    j with an LDS hash map (global site -> local slot), computes the column
//    exactly as the sparse product would, drops |x| <= nztol, and folds
//    max|T_n| of the column into a global per-n maximum.  K is the first n with
//    max|T_n| <= threshold/s.
//  * horner_left_kernel / horner_right_kernel: one Horner pass each,
//      left : Yout[:, j] = X[:, j] + c_k * (H Yin[:, j]),      c_k = -i dt/(s k)
//      right: Wout[i, :] = Y[i, :] + conj(c_k) * (Win[i, :] H)
//    (W H)[i, j] = diag_j W[i,j] + sum_l W[i, l] conj(H[j, l]) since H is
//    Hermitian.  Complex FP64, HBM-bound streaming passes, no MFMA (the
//    operator is a stencil, not a dense contraction).
#include "lmhip_internal.h"

namespace lmh {

namespace {

// ---------------------------------------------------------------- K selection

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t hash_site(int32_t k, uint32_t mask) {
    return ((uint32_t)k * 2654435761u) & mask;
}

template <int MAXB>
__global__ void __launch_bounds__(kWave) taylor_order_kernel(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag,
    int64_t n, int zmax, double dt, double inv_s, int nmax, double nztol,
    unsigned long long *__restrict__ maxbits, int *__restrict__ overflow) {
    constexpr int HSZ = 2 * MAXB;
    constexpr uint32_t HMASK = HSZ - 1;
    __shared__ int32_t gidx[MAXB];
    __shared__ double2 va[MAXB];
    __shared__ double2 vb[MAXB];
    __shared__ int32_t hkey[HSZ];
    __shared__ int32_t hslot[HSZ];
    __shared__ int32_t s_count;
    __shared__ int32_t s_over;

    const int lane = threadIdx.x;
    for (int64_t j = blockIdx.x; j < n; j += gridDim.x) {
        for (int t = lane; t < HSZ; t += kWave) hkey[t] = -1;
        __syncthreads();
        if (lane == 0) {
            const uint32_t h0 = hash_site((int32_t)j, HMASK);
            hkey[h0] = (int32_t)j;
            hslot[h0] = 0;
            gidx[0] = (int32_t)j;
            va[0] = make_double2(1.0, 0.0);
            s_count = 1;
            s_over = 0;
        }
        __syncthreads();
        int f0 = 0, f1 = 1;      // frontier: nodes added last (distance nt-1)
        double2 *cur = va, *nxt = vb;
        for (int nt = 1; nt <= nmax; ++nt) {
            // 1. add the neighbours of the frontier (distance nt)
            const int work = (f1 - f0) * zmax;
            for (int t = lane; t < work; t += kWave) {
                const int a = f0 + t / zmax;
                const int s = t - (t / zmax) * zmax;
                const int32_t k = col[(int64_t)s * n + gidx[a]];
                if (k < 0) continue;
                uint32_t hh = hash_site(k, HMASK);
                bool placed = false;
                for (int probe = 0; probe < HSZ; ++probe) {      // bounded: table may fill on overflow
                    const int32_t old = atomicCAS(&hkey[hh], -1, k);
                    if (old == -1) {
                        const int32_t slot = atomicAdd(&s_count, 1);
                        if (slot < MAXB) {
                            gidx[slot] = k;
                            cur[slot] = make_double2(0.0, 0.0);
                            hslot[hh] = slot;
                        } else {
                            s_over = 1;
                            hslot[hh] = -1;
                        }
                        placed = true;
                        break;
                    }
                    if (old == k) { placed = true; break; }
                    hh = (hh + 1) & HMASK;
                }
                if (!placed) s_over = 1;
            }
            __syncthreads();
            if (s_over) break;
            const int cnt = s_count;
            f0 = f1;
            f1 = cnt;
            // 2. T_nt[:, j] = droptol((H T_{nt-1}) / s * (-i dt / nt))
            const double cn = dt / (double)nt;
            double lmax = 0.0;
            for (int a = lane; a < cnt; a += kWave) {
                const int32_t g = gidx[a];
                const double d = diag[g];
                double2 acc = make_double2(d * cur[a].x, d * cur[a].y);
                for (int s = 0; s < zmax; ++s) {
                    const int64_t q = (int64_t)s * n + g;
                    const int32_t k = col[q];
                    if (k < 0) continue;
                    uint32_t hh = hash_site(k, HMASK);
                    int32_t sl = -1;
                    for (int probe = 0; probe < HSZ; ++probe) {
                        const int32_t kk = hkey[hh];
                        if (kk == k) { sl = hslot[hh]; break; }
                        if (kk == -1) break;
                        hh = (hh + 1) & HMASK;
                    }
                    if (sl < 0) continue;
                    const double2 hv = val[q];
                    const double2 tv = cur[sl];
                    acc.x += hv.x * tv.x - hv.y * tv.y;
                    acc.y += hv.x * tv.y + hv.y * tv.x;
                }
                // (acc / s) * (0 - i cn) = (im*cn, -re*cn)
                double2 v = make_double2((acc.y * inv_s) * cn, -((acc.x * inv_s) * cn));
                const double m = hypot(v.x, v.y);
                if (!(m > nztol)) {
                    v = make_double2(0.0, 0.0);
                } else {
                    lmax = fmax(lmax, m);
                }
                nxt[a] = v;
            }
            // wave max
            for (int off = kWave / 2; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off));
            if (lane == 0 && lmax > 0.0)
                atomicMax(&maxbits[nt - 1], (unsigned long long)__double_as_longlong(lmax));
            __syncthreads();
            double2 *tmp = cur;
            cur = nxt;
            nxt = tmp;
        }
        if (s_over && lane == 0) atomicExch(overflow, 1);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- Horner passes

constexpr int kRowsPerBlock = 256;
constexpr int kColsPerThread = 4;

// blockIdx -> logical tile so that consecutive logical tiles share an XCD
// (blocks are dealt round-robin to the 8 XCDs, each with its own L2).
__device__ __forceinline__ int64_t xcd_logical_block(int64_t nblocks_real) {
    const int64_t pid = blockIdx.x;
    const int64_t per = (int64_t)gridDim.x / kNumXCD;
    return (pid % kNumXCD) * per + pid / kNumXCD;
}

__global__ void __launch_bounds__(kRowsPerBlock) horner_left_kernel(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag,
    int64_t n, int zmax, const double2 *__restrict__ X, const double2 *__restrict__ Yin,
    double2 *__restrict__ Yout, int64_t ld, int64_t ncols, double c, int first,
    int64_t rblocks, int64_t nreal) {
    const int64_t lb = xcd_logical_block(nreal);
    if (lb >= nreal) return;
    const int64_t rb = lb % rblocks;
    const int64_t cb = lb / rblocks;
    const int64_t i = rb * kRowsPerBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t j0 = cb * kColsPerThread;
    const double d = diag[i];
    double2 acc[kColsPerThread];
#pragma unroll
    for (int t = 0; t < kColsPerThread; ++t) {
        const int64_t j = j0 + t;
        if (j < ncols) {
            const double2 y = Yin[j * ld + i];
            acc[t] = make_double2(d * y.x, d * y.y);
        } else {
            acc[t] = make_double2(0.0, 0.0);
        }
    }
    for (int s = 0; s < zmax; ++s) {
        const int64_t q = (int64_t)s * n + i;
        const int32_t k = col[q];
        if (k < 0) continue;
        const double2 hv = val[q];
#pragma unroll
        for (int t = 0; t < kColsPerThread; ++t) {
            const int64_t j = j0 + t;
            if (j < ncols) {
                const double2 y = Yin[j * ld + k];
                acc[t].x += hv.x * y.x - hv.y * y.y;
                acc[t].y += hv.x * y.y + hv.y * y.x;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < kColsPerThread; ++t) {
        const int64_t j = j0 + t;
        if (j < ncols) {
            const double2 x = first ? Yin[j * ld + i] : X[j * ld + i];
            // x + (-i c) * acc
            Yout[j * ld + i] = make_double2(x.x + c * acc[t].y, x.y - c * acc[t].x);
        }
    }
}

__global__ void __launch_bounds__(kRowsPerBlock) horner_right_kernel(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag,
    int64_t n, int zmax, const double2 *__restrict__ Y, const double2 *__restrict__ Win,
    double2 *__restrict__ Wout, int64_t ld, int64_t nrows, double c, int first,
    int64_t rblocks, int64_t nreal) {
    const int64_t lb = xcd_logical_block(nreal);
    if (lb >= nreal) return;
    const int64_t rb = lb % rblocks;
    const int64_t cb = lb / rblocks;
    const int64_t i = rb * kRowsPerBlock + threadIdx.x;
    if (i >= nrows) return;
    const int64_t j0 = cb * kColsPerThread;
#pragma unroll
    for (int t = 0; t < kColsPerThread; ++t) {
        const int64_t j = j0 + t;
        if (j >= n) break;
        const double d = diag[j];
        const double2 w = Win[j * ld + i];
        double2 acc = make_double2(d * w.x, d * w.y);
        for (int s = 0; s < zmax; ++s) {
            const int64_t q = (int64_t)s * n + j;
            const int32_t l = col[q];
            if (l < 0) continue;
            const double2 hv = val[q];                 // H[j,l]; H[l,j] = conj(hv)
            const double2 wl = Win[(int64_t)l * ld + i];
            acc.x += hv.x * wl.x + hv.y * wl.y;
            acc.y += hv.x * wl.y - hv.y * wl.x;
        }
        const double2 y = first ? w : Y[j * ld + i];
        // y + (+i c) * acc
        Wout[j * ld + i] = make_double2(y.x - c * acc.y, y.y + c * acc.x);
    }
}

int64_t xcd_padded(int64_t nreal) { return ((nreal + kNumXCD - 1) / kNumXCD) * kNumXCD; }

}  // namespace

hipError_t launch_taylor_order(const DevHam &h, double dt, double inv_s, int nmax, double nztol,
                               unsigned long long *maxbits, int *overflow, hipStream_t st) {
    hipError_t e = hipMemsetAsync(maxbits, 0, sizeof(unsigned long long) * kMaxTaylor, st);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(overflow, 0, sizeof(int), st);
    if (e != hipSuccess) return e;
    const int64_t grid = h.n < 8192 ? h.n : 8192;
    hipLaunchKernelGGL(taylor_order_kernel<1024>, dim3((unsigned)grid), dim3(kWave), 0, st, h.col, h.val,
                       h.diag, h.n, h.zmax, dt, inv_s, nmax, nztol, maxbits, overflow);
    return hipGetLastError();
}

hipError_t launch_horner_left(const DevHam &h, const double2 *X, const double2 *Yin, double2 *Yout,
                              int64_t ld, int64_t ncols, double c, int first, hipStream_t st) {
    const int64_t rblocks = (h.n + kRowsPerBlock - 1) / kRowsPerBlock;
    const int64_t cblocks = (ncols + kColsPerThread - 1) / kColsPerThread;
    const int64_t nreal = rblocks * cblocks;
    const int64_t grid = xcd_padded(nreal);
    hipLaunchKernelGGL(horner_left_kernel, dim3((unsigned)grid), dim3(kRowsPerBlock), 0, st, h.col, h.val,
                       h.diag, h.n, h.zmax, X, Yin, Yout, ld, ncols, c, first, rblocks, nreal);
    return hipGetLastError();
}

hipError_t launch_horner_right(const DevHam &h, const double2 *Y, const double2 *Win, double2 *Wout,
                               int64_t ld, int64_t nrows, double c, int first, hipStream_t st) {
    const int64_t rblocks = (nrows + kRowsPerBlock - 1) / kRowsPerBlock;
    const int64_t cblocks = (h.n + kColsPerThread - 1) / kColsPerThread;
    const int64_t nreal = rblocks * cblocks;
    const int64_t grid = xcd_padded(nreal);
    hipLaunchKernelGGL(horner_right_kernel, dim3((unsigned)grid), dim3(kRowsPerBlock), 0, st, h.col, h.val,
                       h.diag, h.n, h.zmax, Y, Win, Wout, ld, nrows, c, first, rblocks, nreal);
    return hipGetLastError();
}

}  // namespace lmh
