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
#include <atomic>

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

__global__ void clear_flag_kernel(int *flag) {
    if (threadIdx.x == 0) *flag = 0;
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
    int64_t n, int zmax, double dt, double nztol, const TaylorRun *__restrict__ run,
    double *__restrict__ partmax, int *__restrict__ overflow) {
    const double inv_s = run->inv_s, thr = run->thr, growth = run->growth;
    const int nmax = run->nmax;
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
    int64_t n, int zmax, double dt, double nztol, const TaylorRun *__restrict__ run,
    double *__restrict__ partmax, int *__restrict__ overflow) {
    const double inv_s = run->inv_s, thr = run->thr, growth = run->growth;
    const int nmax = run->nmax;
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
hipError_t launch_rows_t(const DevHam &h, double dt, double nztol, const TaylorRun *run, double *partmax,
                         int *overflow, int64_t *grid_out, hipStream_t st) {
    int64_t grid = (h.n + WAVES - 1) / WAVES;
    if (grid > kOrderBlocks) grid = kOrderBlocks;
    *grid_out = grid;
    hipLaunchKernelGGL((taylor_order_rows_kernel<MAXB, WAVES>), dim3((unsigned)grid), dim3(kWave * WAVES), 0, st,
                       h.col, h.val, h.diag, h.n, h.zmax, dt, nztol, run, partmax, overflow);
    return hipGetLastError();
}

template <int MAXB, int WAVES>
hipError_t launch_order_t(const DevHam &h, double dt, double nztol, const TaylorRun *run, double *partmax,
                          int *overflow, int64_t *grid_out, hipStream_t st) {
    int64_t grid = (h.n + WAVES - 1) / WAVES;
    if (grid > kOrderBlocks) grid = kOrderBlocks;
    *grid_out = grid;
    hipLaunchKernelGGL((taylor_order_kernel<MAXB, WAVES>), dim3((unsigned)grid), dim3(kWave * WAVES), 0, st, h.col,
                       h.val, h.diag, h.n, h.zmax, dt, nztol, run, partmax, overflow);
    return hipGetLastError();
}

// s, 1/s, threshold/s, the growth bound and the order bound from H's
// reductions, as the host rule did (src/evolution.jl:93-104: s =
// nextpow(2, max|H_ij| |dt|), Julia's nextpow being 1 for arguments <= 1;
// the order bound is the first n with (|dt| rowsum / s)^n / n! <= thr).
__global__ void taylor_setup_kernel(const unsigned long long *__restrict__ red, double dt, double threshold,
                                    TaylorRun *__restrict__ run, int *__restrict__ overflow) {
    if (threadIdx.x != 0) return;
    const double maxabs = __longlong_as_double((long long)red[0]);
    const double rowsum = __longlong_as_double((long long)red[1]);
    TaylorRun r;
    r.s = 1.0;
    r.nmax = 0;
    r.pad = 0;
    r.thr = threshold;
    r.growth = 0.0;
    r.dt = dt;
    if (maxabs != 0.0 && dt != 0.0) {
        const double adt = fabs(dt);
        const double x = maxabs * adt;
        double p = 1.0;
        if (x > 1.0)
            while (p < x) p *= 2.0;
        r.s = p;
        r.growth = adt * rowsum / p;
        r.thr = threshold / p;
        double bound = 1.0;
        for (int nn = 1; nn <= kMaxTaylor; ++nn) {
            bound = __dmul_rn(bound, r.growth) / (double)nn;
            r.nmax = nn;
            if (bound <= r.thr) break;
        }
    }
    r.inv_s = 1.0 / r.s;
    *run = r;
    *overflow = 0;
}

// K = the first n <= nmax whose max|T_n| is at most threshold/s; the
// reductions travel back with it, and the step's Horner coefficients
// (dt / s, then (dt / s) / k: the divisions the host would do)
__global__ void taylor_finish_kernel(const unsigned long long *__restrict__ red, const TaylorRun *__restrict__ run,
                                     const unsigned long long *__restrict__ maxbits, const int *__restrict__ overflow,
                                     TaylorOut *__restrict__ out) {
    if (threadIdx.x != 0) return;
    TaylorOut o;
    for (int q = 0; q < 4; ++q) o.red[q] = red[q];
    o.s = run->s;
    o.nmax = run->nmax;
    o.overflow = *overflow;
    o.reps = o.s <= 1073741824.0 ? (int)o.s : 0;
    o.dts = run->dt / o.s;
    o.ck[0] = 0.0;
    for (int k = 1; k <= kStripMaxK; ++k) o.ck[k] = o.dts / (double)k;
    int K = 0;
    if (run->nmax > 0) {
        K = -1;
        for (int nn = 1; nn <= run->nmax; ++nn)
            if (__longlong_as_double((long long)maxbits[nn - 1]) <= run->thr) {
                K = nn;
                break;
            }
    }
    o.K = K;
    *out = o;
}

// The order rule for grid-stencil H (StripGeom): T_n of column j is
// supported within n lattice hops of site j (H couples a site to itself and
// its four grid neighbours), so column j's terms live in the diamond of
// sites within Manhattan distance R of j, indexed directly instead of
// through a hash table.  The diamond is stored ring by ring (distance 0, 1,
// ..., R): the cells of term nt are the first 2 nt^2 + 2 nt + 1, and a
// cell's neighbours are found from their offsets (diamond_index).  A cell's
// sum keeps the ELL slot order; a neighbour beyond distance R is not stored
// (it holds no term: the terms computed reach distance nmax <= R only).  One
// wavefront per column.  Needs nmax <= R and, along a periodic axis,
// 2 nmax + 1 <= L (the diamond must not wrap onto itself); otherwise it flags
// overflow and the hash-table kernels run.
__device__ __forceinline__ int diamond_index(int du, int dv) {
    const int r = abs(du) + abs(dv);
    if (r == 0) return 0;
    // ring r holds 4r cells after the 2r^2 - 2r + 1 closer ones, in four
    // runs of r: (r - t, t), (-t, r - t), (t - r, -t), (t, t - r), t < r
    int seg, t;
    if (du > 0 && dv >= 0) seg = 0, t = dv;
    else if (du <= 0 && dv > 0) seg = 1, t = -du;
    else if (du < 0 && dv <= 0) seg = 2, t = -dv;
    else seg = 3, t = du;
    return 2 * r * r - 2 * r + 1 + seg * r + t;
}

__device__ __forceinline__ void diamond_cell(int a, int &du, int &dv) {
    if (a == 0) {
        du = dv = 0;
        return;
    }
    // the ring: largest r with 2r^2 - 2r + 1 <= a
    int r = (int)((1.0f + sqrtf((float)(2 * a - 1))) * 0.5f);
    while (2 * r * r - 2 * r + 1 > a) --r;
    while (2 * (r + 1) * (r + 1) - 2 * (r + 1) + 1 <= a) ++r;
    const int q = a - (2 * r * r - 2 * r + 1), seg = q / r, t = q - seg * r;
    if (seg == 0) du = r - t, dv = t;
    else if (seg == 1) du = -t, dv = r - t;
    else if (seg == 2) du = t - r, dv = -t;
    else du = t, dv = t - r;
}

template <int R, int WAVES>
__global__ void __launch_bounds__(kWave * WAVES, R <= 9 ? 4 : 2) taylor_grid_kernel(
    const uint32_t *__restrict__ sdir, const double2 *__restrict__ val, const double *__restrict__ diag, int64_t n,
    int zmax, StripGeom g, double dt, double nztol, const TaylorRun *__restrict__ run, double *__restrict__ partmax,
    int *__restrict__ overflow) {
    constexpr int D = 2 * R * R + 2 * R + 1;          // cells within distance R
    constexpr int SL = (D + kWave - 1) / kWave;       // cells per lane
    extern __shared__ double2 win[];                  // [WAVES][2][D]
    __shared__ double tmax_s[WAVES][kMaxTaylor];
    const double inv_s = run->inv_s, thr = run->thr, growth = run->growth;
    const int nmax = run->nmax;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    for (int t = lane; t < kMaxTaylor; t += kWave) tmax_s[w][t] = 0.0;
    const bool fits = nmax <= R && zmax <= 4 && (!g.px || 2 * nmax + 1 <= g.lx) && (!g.py || 2 * nmax + 1 <= g.ly);
    if (!fits) {
        if (threadIdx.x == 0) atomicExch(overflow, 1);
    } else {
        double2 *A = win + (size_t)w * 2 * D, *B = A + D;
        const int reach = 2 * nmax * nmax + 2 * nmax + 1;   // cells any term of this search touches
        for (int64_t j = (int64_t)blockIdx.x * WAVES + w; j < n; j += (int64_t)gridDim.x * WAVES) {
            const int x0 = (int)(j / g.ly), y0 = (int)(j - (int64_t)x0 * g.ly);
            // this lane's cells a = lane + 64 k: a term's cells are a prefix
            // of the next term's, so each cell's H (diagonal, bond values,
            // the cells its bonds reach, -1 for none; a cell that is not a
            // site is marked by site[k] = false and stays 0) is loaded once
            // per column, before the terms
            bool site[SL];
            double dg[SL];
            double2 hv[SL][4];
            int nb[SL][4];
#pragma unroll
            for (int k = 0; k < SL; ++k) {
                site[k] = false;
                dg[k] = 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    hv[k][q] = make_double2(0.0, 0.0);
                    nb[k][q] = -1;
                }
                const int a = lane + k * kWave;
                if (a >= reach) continue;
                int du, dv;
                diamond_cell(a, du, dv);
                int xs = x0 + du, ys = y0 + dv;
                if (g.px) xs = xs < 0 ? xs + g.lx : (xs >= g.lx ? xs - g.lx : xs);
                else if (xs < 0 || xs >= g.lx) continue;
                if (g.py) ys = ys < 0 ? ys + g.ly : (ys >= g.ly ? ys - g.ly : ys);
                else if (ys < 0 || ys >= g.ly) continue;
                const int64_t i = (int64_t)xs * g.ly + ys;
                site[k] = true;
                dg[k] = diag[i];
                const uint32_t dirs = sdir[i];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (q >= zmax) continue;
                    hv[k][q] = val[(int64_t)q * n + i];
                    if (!((dirs >> (8 + q)) & 1u)) continue;                // empty slot
                    const uint32_t dr = (dirs >> (2 * q)) & 3u;             // 0 y+, 1 y-, 2 x+, 3 x-
                    const int ndu = du + (dr == 2 ? 1 : (dr == 3 ? -1 : 0));
                    const int ndv = dv + (dr == 0 ? 1 : (dr == 1 ? -1 : 0));
                    if (abs(ndu) + abs(ndv) > R) continue;                // beyond the terms: 0
                    nb[k][q] = diamond_index(ndu, ndv);
                }
            }
            for (int c = lane; c < D; c += kWave) {
                A[c] = make_double2(0.0, 0.0);
                B[c] = make_double2(0.0, 0.0);
            }
            wave_sync();
            if (lane == 0) A[0] = make_double2(1.0, 0.0);
            wave_sync();
            double2 *cur = A, *nxt = B;
            for (int nt = 1; nt <= nmax; ++nt) {
                const double cn = dt / (double)nt;
                const int cells = 2 * nt * nt + 2 * nt + 1;
                double lmax = 0.0;
#pragma unroll
                for (int k = 0; k < SL; ++k) {
                    const int a = lane + k * kWave;
                    if (a >= cells || !site[k]) continue;         // past this term, or not a site
                    const double2 tc = cur[a];
                    double2 acc = make_double2(dg[k] * tc.x, dg[k] * tc.y);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (nb[k][q] < 0) continue;
                        const double2 tv = cur[nb[k][q]];
                        acc.x += hv[k][q].x * tv.x - hv[k][q].y * tv.y;
                        acc.y += hv[k][q].x * tv.y + hv[k][q].y * tv.x;
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
            wave_sync();
        }
    }
    block_maxima<WAVES>(tmax_s, partmax);
}

// the two window sizes: R = 9 (2 x 181 x 16 B = 5.8 KB of LDS per wave) and
// R = 11 (8.5 KB per wave), 8 waves per workgroup
constexpr int kGridWaves = 8;

template <int R>
hipError_t launch_grid_t(const DevHam &h, double dt, double nztol, const TaylorRun *run, double *partmax,
                         int *overflow, int64_t *grid_out, hipStream_t st) {
    constexpr int D = 2 * R * R + 2 * R + 1;
    const size_t lds = (size_t)kGridWaves * 2 * D * sizeof(double2);
    static std::atomic<uint64_t> done{0};
    hipError_t e = once_per_device(done, [lds] {
        return hipFuncSetAttribute((const void *)taylor_grid_kernel<R, kGridWaves>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    });
    if (e != hipSuccess) return e;
    int64_t grid = (h.n + kGridWaves - 1) / kGridWaves;
    if (grid > kOrderBlocks) grid = kOrderBlocks;
    *grid_out = grid;
    hipLaunchKernelGGL((taylor_grid_kernel<R, kGridWaves>), dim3((unsigned)grid), dim3(kWave * kGridWaves), lds, st,
                       h.sdir, h.val, h.diag, h.n, h.zmax, h.sg, dt, nztol, run, partmax, overflow);
    return hipGetLastError();
}

}  // namespace

int taylor_order_levels() { return 4; }

int64_t taylor_order_scratch() { return kOrderBlocks * kMaxTaylor; }

hipError_t launch_taylor_setup(const unsigned long long *red, double dt, double threshold, TaylorRun *run,
                               int *overflow, hipStream_t st) {
    hipLaunchKernelGGL(taylor_setup_kernel, dim3(1), dim3(64), 0, st, red, dt, threshold, run, overflow);
    return hipGetLastError();
}

// Grid-stencil H: level 0 the R = 9 diamond, 1 the R = 11 diamond, 2 and 3
// the hash kernel with 1024 and 2048 entries.  Other H: level 0 the
// row-cached hash kernel (zmax <= 4) or the general one, then 1024 and 2048
// entries (level 3 as level 2).  A level is tried after the one below it
// overflowed.  The overflow flag is cleared by the setup kernel; a level >= 1
// run clears it itself.
hipError_t launch_taylor_order(const DevHam &h, double dt, double nztol, const TaylorRun *run, int level,
                               double *partmax, unsigned long long *maxbits, int *overflow, hipStream_t st) {
    hipError_t e = hipSuccess;
    if (level > 0) {
        hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(64), 0, st, overflow);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    int64_t grid = 0;
    const bool gridh = h.sg.s > 0 && h.sdir;
    if (gridh && level == 0)
        e = launch_grid_t<9>(h, dt, nztol, run, partmax, overflow, &grid, st);
    else if (gridh && level == 1)
        e = launch_grid_t<11>(h, dt, nztol, run, partmax, overflow, &grid, st);
    else if (!gridh && level == 0 && h.zmax <= 4)
        e = launch_rows_t<256, 2>(h, dt, nztol, run, partmax, overflow, &grid, st);
    else if (!gridh && level == 0)
        e = launch_order_t<256, 4>(h, dt, nztol, run, partmax, overflow, &grid, st);
    else if (level == (gridh ? 2 : 1))
        e = launch_order_t<1024, 1>(h, dt, nztol, run, partmax, overflow, &grid, st);
    else
        e = launch_order_t<2048, 1>(h, dt, nztol, run, partmax, overflow, &grid, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(taylor_max_kernel, dim3((unsigned)kMaxTaylor), dim3(256), 0, st, partmax, grid, maxbits);
    return hipGetLastError();
}

hipError_t launch_taylor_finish(const unsigned long long *red, const TaylorRun *run,
                                const unsigned long long *maxbits, const int *overflow, TaylorOut *out,
                                hipStream_t st) {
    hipLaunchKernelGGL(taylor_finish_kernel, dim3(1), dim3(64), 0, st, red, run, maxbits, overflow, out);
    return hipGetLastError();
}

}  // namespace lmh
