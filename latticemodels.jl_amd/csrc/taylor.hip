// Taylor propagator kernels for P <- U P U', U = sum_{n<=K} (-i dt H/s)^n / n!
// (applied s times per side).
//
//  * taylor_order_kernel: the reference's order rule (src/evolution.jl:93-118).
//    Column j of T_n = droptol((H/s) T_{n-1}[:, j] * (-i dt / n), nztol) only
//    involves column j, so one wavefront walks the K-hop neighbourhood of site
//    j with an LDS hash map (global site -> local slot), computes the column
//    exactly as the sparse product would, drops |x| <= nztol, and folds
//    max|T_n| of the column into a global per-n maximum.  K is the first n with
//    max|T_n| <= threshold/s.  Neighbourhood tables of 256, 1024 and 2048
//    sites are tried in turn (a table that overflows sets a flag).
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

// LDS ordering among the lanes of one wavefront (no block barrier: the
// waves of a workgroup walk different columns for different numbers of terms)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wavefront per column, WAVES columns per workgroup, each with a
// MAXB-site neighbourhood.  A column stops after term nt once its terms can
// only shrink (growth bound |dt| rowsum / (s (nt+1)) <= 1: every entry of
// T_{nt+1} is at most that factor times max|T_nt|, and droptol only removes
// entries) and max|T_nt| <= thr: it cannot lift the maximum of any later
// term above thr, so the first n with max|T_n| <= thr is unchanged.
template <int MAXB, int WAVES>
__global__ void __launch_bounds__(kWave * WAVES) taylor_order_kernel(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag,
    int64_t n, int zmax, double dt, double inv_s, int nmax, double nztol, double thr, double growth,
    unsigned long long *__restrict__ maxbits, int *__restrict__ overflow) {
    constexpr int HSZ = 2 * MAXB;
    constexpr uint32_t HMASK = HSZ - 1;
    __shared__ int32_t gidx_s[WAVES][MAXB];
    __shared__ double2 va_s[WAVES][MAXB];
    __shared__ double2 vb_s[WAVES][MAXB];
    __shared__ int32_t hkey_s[WAVES][HSZ];
    __shared__ int32_t hslot_s[WAVES][HSZ];
    __shared__ int32_t count_s[WAVES];
    __shared__ int32_t over_s[WAVES];

    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    int32_t *gidx = gidx_s[w];
    int32_t *hkey = hkey_s[w];
    int32_t *hslot = hslot_s[w];
    for (int64_t j = (int64_t)blockIdx.x * WAVES + w; j < n; j += (int64_t)gridDim.x * WAVES) {
        for (int t = lane; t < HSZ; t += kWave) hkey[t] = -1;
        wave_sync();
        if (lane == 0) {
            const uint32_t h0 = hash_site((int32_t)j, HMASK);
            hkey[h0] = (int32_t)j;
            hslot[h0] = 0;
            gidx[0] = (int32_t)j;
            va_s[w][0] = make_double2(1.0, 0.0);
            count_s[w] = 1;
            over_s[w] = 0;
        }
        wave_sync();
        int f0 = 0, f1 = 1;      // frontier: nodes added last (distance nt-1)
        double2 *cur = va_s[w], *nxt = vb_s[w];
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
                for (int probe = 0; probe < HSZ; ++probe) {      // bounded: the table may fill
                    const int32_t old = atomicCAS(&hkey[hh], -1, k);
                    if (old == -1) {
                        const int32_t slot = atomicAdd(&count_s[w], 1);
                        if (slot < MAXB) {
                            gidx[slot] = k;
                            cur[slot] = make_double2(0.0, 0.0);
                            hslot[hh] = slot;
                        } else {
                            over_s[w] = 1;
                            hslot[hh] = -1;
                        }
                        placed = true;
                        break;
                    }
                    if (old == k) { placed = true; break; }
                    hh = (hh + 1) & HMASK;
                }
                if (!placed) over_s[w] = 1;
            }
            wave_sync();
            if (over_s[w]) break;
            const int cnt = count_s[w];
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
            wave_sync();
            double2 *tmp = cur;
            cur = nxt;
            nxt = tmp;
            if (lmax <= thr && growth <= (double)(nt + 1)) break;     // uniform across the wave
        }
        if (over_s[w] && lane == 0) atomicExch(overflow, 1);
        wave_sync();
    }
}

template <int MAXB, int WAVES>
hipError_t launch_order_t(const DevHam &h, double dt, double inv_s, int nmax, double nztol, double thr,
                          double growth, unsigned long long *maxbits, int *overflow, hipStream_t st) {
    const int64_t cols_per_block = WAVES;
    int64_t grid = (h.n + cols_per_block - 1) / cols_per_block;
    if (grid > 16384) grid = 16384;
    hipLaunchKernelGGL((taylor_order_kernel<MAXB, WAVES>), dim3((unsigned)grid), dim3(kWave * WAVES), 0, st, h.col,
                       h.val, h.diag, h.n, h.zmax, dt, inv_s, nmax, nztol, thr, growth, maxbits, overflow);
    return hipGetLastError();
}

}  // namespace

int taylor_order_levels() { return 3; }

hipError_t launch_taylor_order(const DevHam &h, double dt, double inv_s, int nmax, double nztol, double thr,
                               double growth, int level, unsigned long long *maxbits, int *overflow,
                               hipStream_t st) {
    hipError_t e = hipMemsetAsync(maxbits, 0, sizeof(unsigned long long) * kMaxTaylor, st);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(overflow, 0, sizeof(int), st);
    if (e != hipSuccess) return e;
    switch (level) {
        case 0: return launch_order_t<256, 4>(h, dt, inv_s, nmax, nztol, thr, growth, maxbits, overflow, st);
        case 1: return launch_order_t<1024, 1>(h, dt, inv_s, nmax, nztol, thr, growth, maxbits, overflow, st);
        default: return launch_order_t<2048, 1>(h, dt, inv_s, nmax, nztol, thr, growth, maxbits, overflow, st);
    }
}

}  // namespace lmh
