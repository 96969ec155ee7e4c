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
// The Horner passes that apply the polynomial are in passes.hip (one term per
// launch) and fused.hip (all terms per launch).
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

}  // namespace lmh
