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

// Per-term maxima of the block's waves to partmax[block][0 .. kMaxTaylor):
// plain stores (one row per workgroup), reduced by taylor_max_kernel.  (A
// device-scope atomic on one address per term from every column serialises
// across the chip: ~10^4 of them took ~1 ms.)
template <int WAVES>
__device__ __forceinline__ void block_maxima(double (*tmax)[kMaxTaylor], double *__restrict__ partmax) {
    __syncthreads();
    for (int t = threadIdx.x; t < kMaxTaylor; t += blockDim.x) {
        double m = 0.0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) m = fmax(m, tmax[w][t]);
        partmax[(int64_t)blockIdx.x * kMaxTaylor + t] = m;
    }
}

// maxbits[t] = max over blocks of partmax[b][t] (as the bits of a nonnegative
// double); one workgroup per term
__global__ void __launch_bounds__(256) taylor_max_kernel(const double *__restrict__ partmax, int64_t nblocks,
                                                         unsigned long long *__restrict__ maxbits) {
    __shared__ double red[256];
    const int t = blockIdx.x;
    double m = 0.0;
    for (int64_t b = threadIdx.x; b < nblocks; b += 256) m = fmax(m, partmax[b * kMaxTaylor + t]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) maxbits[t] = (unsigned long long)__double_as_longlong(red[0]);
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
    double *__restrict__ partmax, int *__restrict__ overflow) {
    constexpr int HSZ = 2 * MAXB;
    constexpr uint32_t HMASK = HSZ - 1;
    __shared__ int32_t gidx_s[WAVES][MAXB];
    __shared__ double2 va_s[WAVES][MAXB];
    __shared__ double2 vb_s[WAVES][MAXB];
    __shared__ int32_t hkey_s[WAVES][HSZ];
    __shared__ int32_t hslot_s[WAVES][HSZ];
    __shared__ int32_t count_s[WAVES];
    __shared__ int32_t over_s[WAVES];
    __shared__ double tmax_s[WAVES][kMaxTaylor];      // per-term maxima over this wave's columns

    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    for (int t = lane; t < kMaxTaylor; t += kWave) tmax_s[w][t] = 0.0;
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
            if (lane == 0) tmax_s[w][nt - 1] = fmax(tmax_s[w][nt - 1], lmax);
            wave_sync();
            double2 *tmp = cur;
            cur = nxt;
            nxt = tmp;
            if (lmax <= thr && growth <= (double)(nt + 1)) break;     // uniform across the wave
        }
        if (over_s[w] && lane == 0) atomicExch(overflow, 1);
        wave_sync();
    }
    block_maxima<WAVES>(tmax_s, partmax);
}

// The same rule for rows of at most 4 ELL slots (square lattices), with each
// neighbourhood node's H row (diagonal, columns, values) loaded into LDS once
// per column when the node joins, and its neighbours' local slots resolved
// once when the node leaves the frontier; nodes that joined this term look
// their neighbours up in the hash table.  The sums run over the same slots in
// the same order as taylor_order_kernel.
template <int MAXB, int WAVES>
__global__ void __launch_bounds__(kWave * WAVES) taylor_order_rows_kernel(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag,
    int64_t n, int zmax, double dt, double inv_s, int nmax, double nztol, double thr, double growth,
    double *__restrict__ partmax, int *__restrict__ overflow) {
    constexpr int ZM = 4;
    constexpr int HSZ = 2 * MAXB;
    constexpr uint32_t HMASK = HSZ - 1;
    __shared__ int32_t gidx_s[WAVES][MAXB];
    __shared__ double2 va_s[WAVES][MAXB];
    __shared__ double2 vb_s[WAVES][MAXB];
    __shared__ int32_t hkey_s[WAVES][HSZ];
    __shared__ int32_t hslot_s[WAVES][HSZ];
    __shared__ double rd_s[WAVES][MAXB];              // H[a, a]
    __shared__ double2 rv_s[WAVES][ZM][MAXB];         // H row values per slot
    __shared__ int32_t rc_s[WAVES][ZM][MAXB];         // global columns per slot (-1 empty)
    __shared__ int16_t nb_s[WAVES][ZM][MAXB];         // local slot of the neighbour (-1: none)
    __shared__ int32_t count_s[WAVES];
    __shared__ int32_t over_s[WAVES];
    __shared__ double tmax_s[WAVES][kMaxTaylor];      // per-term maxima over this wave's columns

    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    for (int t = lane; t < kMaxTaylor; t += kWave) tmax_s[w][t] = 0.0;
    int32_t *gidx = gidx_s[w];
    int32_t *hkey = hkey_s[w];
    int32_t *hslot = hslot_s[w];
    auto lookup = [&](int32_t k) -> int32_t {
        uint32_t hh = hash_site(k, HMASK);
        for (int probe = 0; probe < HSZ; ++probe) {
            const int32_t kk = hkey[hh];
            if (kk == k) return hslot[hh];
            if (kk == -1) return -1;
            hh = (hh + 1) & HMASK;
        }
        return -1;
    };
    auto load_row = [&](int slot, int32_t g) {
        rd_s[w][slot] = diag[g];
#pragma unroll
        for (int s = 0; s < ZM; ++s) {
            const int32_t k = s < zmax ? col[(int64_t)s * n + g] : -1;
            rc_s[w][s][slot] = k;
            rv_s[w][s][slot] = k >= 0 ? val[(int64_t)s * n + g] : make_double2(0.0, 0.0);
        }
    };
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
            load_row(0, (int32_t)j);
        }
        wave_sync();
        int f0 = 0, f1 = 1;      // frontier: nodes added last (distance nt-1)
        double2 *cur = va_s[w], *nxt = vb_s[w];
        for (int nt = 1; nt <= nmax; ++nt) {
            // 1. add the neighbours of the frontier (distance nt), loading each
            //    new node's row
            const int work = (f1 - f0) * zmax;
            for (int t = lane; t < work; t += kWave) {
                const int a = f0 + t / zmax;
                const int s = t - (t / zmax) * zmax;
                const int32_t k = rc_s[w][s][a];
                if (k < 0) continue;
                uint32_t hh = hash_site(k, HMASK);
                bool placed = false;
                for (int probe = 0; probe < HSZ; ++probe) {
                    const int32_t old = atomicCAS(&hkey[hh], -1, k);
                    if (old == -1) {
                        const int32_t slot = atomicAdd(&count_s[w], 1);
                        if (slot < MAXB) {
                            gidx[slot] = k;
                            cur[slot] = make_double2(0.0, 0.0);
                            hslot[hh] = slot;
                            load_row(slot, k);
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
            // 1b. the frontier's neighbours are all present now: resolve them once
            for (int t = lane; t < (f1 - f0) * ZM; t += kWave) {
                const int a = f0 + t / ZM;
                const int s = t - (t / ZM) * ZM;
                const int32_t k = rc_s[w][s][a];
                nb_s[w][s][a] = (int16_t)(k < 0 ? -1 : lookup(k));
            }
            wave_sync();
            const int fr = f1;       // nodes below fr have resolved neighbours
            f0 = f1;
            f1 = cnt;
            // 2. T_nt[:, j] = droptol((H T_{nt-1}) / s * (-i dt / nt))
            const double cn = dt / (double)nt;
            double lmax = 0.0;
            for (int a = lane; a < cnt; a += kWave) {
                const double d = rd_s[w][a];
                double2 acc = make_double2(d * cur[a].x, d * cur[a].y);
#pragma unroll
                for (int s = 0; s < ZM; ++s) {
                    const int32_t k = rc_s[w][s][a];
                    if (k < 0) continue;
                    const int32_t sl = a < fr ? (int32_t)nb_s[w][s][a] : lookup(k);
                    if (sl < 0) continue;
                    const double2 hv = rv_s[w][s][a];
                    const double2 tv = cur[sl];
                    acc.x += hv.x * tv.x - hv.y * tv.y;
                    acc.y += hv.x * tv.y + hv.y * tv.x;
                }
                double2 v = make_double2((acc.y * inv_s) * cn, -((acc.x * inv_s) * cn));
                const double m = hypot(v.x, v.y);
                if (!(m > nztol)) {
                    v = make_double2(0.0, 0.0);
                } else {
                    lmax = fmax(lmax, m);
                }
                nxt[a] = v;
            }
            for (int off = kWave / 2; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off));
            if (lane == 0) tmax_s[w][nt - 1] = fmax(tmax_s[w][nt - 1], lmax);
            wave_sync();
            double2 *tmp = cur;
            cur = nxt;
            nxt = tmp;
            if (lmax <= thr && growth <= (double)(nt + 1)) break;
        }
        if (over_s[w] && lane == 0) atomicExch(overflow, 1);
        wave_sync();
    }
    block_maxima<WAVES>(tmax_s, partmax);
}

constexpr int64_t kOrderBlocks = 2048;    // persistent grid: each wave walks several columns

template <int MAXB, int WAVES>
hipError_t launch_rows_t(const DevHam &h, double dt, double inv_s, int nmax, double nztol, double thr,
                         double growth, double *partmax, int *overflow, int64_t *grid_out, hipStream_t st) {
    int64_t grid = (h.n + WAVES - 1) / WAVES;
    if (grid > kOrderBlocks) grid = kOrderBlocks;
    *grid_out = grid;
    hipLaunchKernelGGL((taylor_order_rows_kernel<MAXB, WAVES>), dim3((unsigned)grid), dim3(kWave * WAVES), 0, st,
                       h.col, h.val, h.diag, h.n, h.zmax, dt, inv_s, nmax, nztol, thr, growth, partmax, overflow);
    return hipGetLastError();
}

template <int MAXB, int WAVES>
hipError_t launch_order_t(const DevHam &h, double dt, double inv_s, int nmax, double nztol, double thr,
                          double growth, double *partmax, int *overflow, int64_t *grid_out, hipStream_t st) {
    int64_t grid = (h.n + WAVES - 1) / WAVES;
    if (grid > kOrderBlocks) grid = kOrderBlocks;
    *grid_out = grid;
    hipLaunchKernelGGL((taylor_order_kernel<MAXB, WAVES>), dim3((unsigned)grid), dim3(kWave * WAVES), 0, st, h.col,
                       h.val, h.diag, h.n, h.zmax, dt, inv_s, nmax, nztol, thr, growth, partmax, overflow);
    return hipGetLastError();
}

}  // namespace

int taylor_order_levels() { return 3; }

int64_t taylor_order_scratch() { return kOrderBlocks * kMaxTaylor; }

hipError_t launch_taylor_order(const DevHam &h, double dt, double inv_s, int nmax, double nztol, double thr,
                               double growth, int level, double *partmax, unsigned long long *maxbits, int *overflow,
                               hipStream_t st) {
    hipError_t e = hipMemsetAsync(overflow, 0, sizeof(int), st);
    if (e != hipSuccess) return e;
    int64_t grid = 0;
    if (h.zmax <= 4 && level == 0)
        e = launch_rows_t<256, 2>(h, dt, inv_s, nmax, nztol, thr, growth, partmax, overflow, &grid, st);
    else if (level == 0)
        e = launch_order_t<256, 4>(h, dt, inv_s, nmax, nztol, thr, growth, partmax, overflow, &grid, st);
    else if (level == 1)
        e = launch_order_t<1024, 1>(h, dt, inv_s, nmax, nztol, thr, growth, partmax, overflow, &grid, st);
    else
        e = launch_order_t<2048, 1>(h, dt, inv_s, nmax, nztol, thr, growth, partmax, overflow, &grid, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(taylor_max_kernel, dim3((unsigned)kMaxTaylor), dim3(256), 0, st, partmax, grid, maxbits);
    return hipGetLastError();
}

}  // namespace lmh
