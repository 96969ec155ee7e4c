// One Horner term per launch (any N): the streaming form of
//   left : Yout[:, j] = X[:, j] + c_k (H Yin[:, j])            c_k = -i dt/(s k)
//   right: Wout[i, :] = Y[i, :] + conj(c_k) (Win[i, :] H)
// with (W H)[i, j] = diag_j W[i, j] + sum_l W[i, l] conj(H[j, l]) (H Hermitian).
// Complex FP64, HBM-bound; no MFMA (the operator is a stencil).
//
// Left pass: a thread owns one row i and keeps H's row i (ELL) in registers
// while it sweeps a block of columns, so H is read once per block instead of
// once per element; neighbouring rows come through L1/L2.
// Right pass: a block owns 64 rows of a group of 32 columns whose sites are
// close on the lattice (Z-order of the site coordinates).  It stages the
// group's union of columns and their neighbours (64 rows each, 1 KB
// contiguous) in LDS once, then every output reads its neighbours from LDS;
// H's row of each column is wave-uniform and comes through scalar loads.
#include "lmhip_internal.h"

namespace lmh {

namespace {

constexpr int kLeftThreads = 256;
constexpr int kLeftCols = 32;        // columns swept per thread
constexpr int kRightRows = 64;       // rows per block (one wavefront wide)
constexpr int kRightThreads = 256;   // 4 waves, each handles every 4th column of the group
constexpr int kRightCols3 = 16;      // columns swept per block by the unstaged right pass


// Results of a pass are read back only by the next launch: write them past
// the caches' normal allocation so they do not evict the gathered inputs.
typedef double v2d __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_stream(double2 *p, double2 v) {
    v2d w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<v2d *>(p));
}

__device__ __forceinline__ int64_t xcd_logical(int64_t pid, int64_t nprog) {
    const int64_t per = nprog / kNumXCD;
    return (pid % kNumXCD) * per + pid / kNumXCD;
}

template <int ZR, bool TOUT>
__global__ void __launch_bounds__(kLeftThreads) horner_left_v2(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag, int64_t n,
    int zmax, const double2 *__restrict__ X, const double2 *__restrict__ Yin, double2 *__restrict__ Yout, int64_t ld,
    int64_t ncols, double c, int first, int lcols, int64_t rblocks, int64_t nreal, int64_t nprog) {
    const int64_t lb = xcd_logical(blockIdx.x, nprog);
    if (lb >= nreal) return;
    const int64_t rb = lb % rblocks, cb = lb / rblocks;
    const int64_t i = rb * kLeftThreads + threadIdx.x;
    if (i >= n) return;
    int32_t nb[ZR];
    double2 hv[ZR];
#pragma unroll
    for (int q = 0; q < ZR; ++q) {
        nb[q] = -1;
        hv[q] = make_double2(0.0, 0.0);
        if (q < zmax) {
            nb[q] = col[(int64_t)q * n + i];
            hv[q] = val[(int64_t)q * n + i];
        }
    }
    const double d = diag[i];
    const int64_t j0 = cb * lcols;
    const int64_t j1 = j0 + lcols < ncols ? j0 + lcols : ncols;
    for (int64_t j = j0; j < j1; ++j) {
        const double2 *yc = Yin + j * ld;
        const double2 y = yc[i];
        double2 acc = make_double2(d * y.x, d * y.y);
#pragma unroll
        for (int q = 0; q < ZR; ++q) {
            if (nb[q] >= 0) {
                const double2 v = yc[nb[q]];
                acc.x += hv[q].x * v.x - hv[q].y * v.y;
                acc.y += hv[q].x * v.y + hv[q].y * v.x;
            }
        }
        const double2 x = first ? y : X[j * ld + i];
        const double2 r = make_double2(x.x + c * acc.y, x.y - c * acc.x);
        if (TOUT)
            Yout[i * ld + j] = make_double2(r.x, -r.y);
        else
            store_stream(Yout + j * ld + i, r);
    }
}

// Left pass for a banded H (every neighbour within `band` rows): a block owns
// kBandRows consecutive rows; per column it stages rows [r0 - band, r0 +
// kBandRows + band) of Yin in LDS with one coalesced read (double-buffered,
// the next column's segment loaded into registers while the current one is
// used) and gathers every neighbour from LDS, so each element of Yin comes
// from L2 (1 + 2 band / kBandRows) times instead of once per neighbour.
// TOUT: the result is written conjugate-transposed (Yout[j, i] = conj(y_ij)),
// for the last term of a side computed as left passes on (U P)'.
constexpr int kBandRows = 1024;
constexpr int kBandStage = 3;        // staged elements per thread: band <= kBandRows

template <int ZR, bool TOUT>
__global__ void __launch_bounds__(kBandRows) horner_left_band(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag, int64_t n,
    int zmax, int band, const double2 *__restrict__ X, const double2 *__restrict__ Yin, double2 *__restrict__ Yout,
    int64_t ld, int64_t ncols, double c, int first, int lcols, int64_t rblocks, int64_t nreal, int64_t nprog) {
    extern __shared__ double2 tile[];          // [2][span]
    const int64_t lb = xcd_logical(blockIdx.x, nprog);
    if (lb >= nreal) return;
    const int64_t rb = lb % rblocks, cb = lb / rblocks;
    const int t = threadIdx.x;
    const int64_t r0 = rb * kBandRows;
    const int64_t i = r0 + t;
    const bool live = i < n;
    const int span = kBandRows + 2 * band;
    const int64_t g0 = r0 - band;              // row of tile slot 0
    const int self = band + t;
    // H row i in registers, neighbours as tile slots (empty slots: own slot, value 0)
    int nbs[ZR];
    double2 hv[ZR];
#pragma unroll
    for (int q = 0; q < ZR; ++q) {
        const int32_t cq = (live && q < zmax) ? col[(int64_t)q * n + i] : -1;
        nbs[q] = cq >= 0 ? (int)(cq - g0) : self;
        hv[q] = cq >= 0 ? val[(int64_t)q * n + i] : make_double2(0.0, 0.0);
    }
    const double d = live ? diag[i] : 0.0;
    const int64_t j0 = cb * lcols;
    const int64_t j1 = j0 + lcols < ncols ? j0 + lcols : ncols;
    double2 pre[kBandStage];
    auto stage_load = [&](int64_t j) {
#pragma unroll
        for (int p = 0; p < kBandStage; ++p) {
            const int e = t + p * kBandRows;
            const int64_t g = g0 + e;
            pre[p] = (e < span && g >= 0 && g < n) ? Yin[j * ld + g] : make_double2(0.0, 0.0);
        }
    };
    if (j0 < j1) stage_load(j0);
    int buf = 0;
    for (int64_t j = j0; j < j1; ++j) {
        double2 *tb = tile + buf * span;
#pragma unroll
        for (int p = 0; p < kBandStage; ++p) {
            const int e = t + p * kBandRows;
            if (e < span) tb[e] = pre[p];
        }
        // one barrier per column: with two buffers, a thread writing buffer b
        // again has passed the barrier that every reader of b reached after
        // its last read
        __syncthreads();
        if (j + 1 < j1) stage_load(j + 1);
        const double2 xv = (!first && live) ? X[j * ld + i] : make_double2(0.0, 0.0);
        const double2 y = tb[self];
        double2 acc = make_double2(d * y.x, d * y.y);
#pragma unroll
        for (int q = 0; q < ZR; ++q) {
            const double2 v = tb[nbs[q]];
            acc.x += hv[q].x * v.x - hv[q].y * v.y;
            acc.y += hv[q].x * v.y + hv[q].y * v.x;
        }
        const double2 x = first ? y : xv;
        const double2 r = make_double2(x.x + c * acc.y, x.y - c * acc.x);
        if (live) {
            if (TOUT)
                Yout[i * ld + j] = make_double2(r.x, -r.y);
            else
                store_stream(Yout + j * ld + i, r);
        }
        buf ^= 1;
    }
}

// Two Horner terms per launch for a banded H:
//   y_k = X + c1 (-i) H y_{k+1},   y_{k-1} = X + c2 (-i) H y_k
// A block owns kBandRows rows; per column it stages y_{k+1} on its rows +- 2
// band, computes y_k on its rows +- band into LDS (the halo rows' H rows are
// held in LDS for the whole column sweep, the block's own rows' H in
// registers), then y_{k-1} on its rows.  HBM per element and pair of terms:
// read y_{k+1} and X once, write y_{k-1} once (half the per-term bytes).
// LDS: stage [span2] | mid [span1] | halo values [ZR][2 band] | halo diagonal
// [2 band] | halo stage slots (16 bit) [ZR][2 band].
size_t band2_lds_bytes(int band, int zr) {
    const size_t span2 = kBandRows + 4 * (size_t)band, span1 = kBandRows + 2 * (size_t)band, nh = 2 * (size_t)band;
    return 16 * (span2 + span1 + zr * nh) + 8 * nh + 2 * zr * nh;
}

template <int ZR>
__global__ void __launch_bounds__(kBandRows) horner_left_band2(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag, int64_t n,
    int zmax, int band, const double2 *__restrict__ X, const double2 *__restrict__ Yin, double2 *__restrict__ Yout,
    int64_t ld, int64_t ncols, double c1, double c2, int first, int lcols, int64_t rblocks, int64_t nreal,
    int64_t nprog) {
    extern __shared__ double2 lds[];
    const int64_t lb = xcd_logical(blockIdx.x, nprog);
    if (lb >= nreal) return;
    const int64_t rb = lb % rblocks, cb = lb / rblocks;
    const int t = threadIdx.x;
    const int b = band, nh = 2 * band;
    const int span2 = kBandRows + 4 * b, span1 = kBandRows + 2 * b;
    double2 *stg = lds;                                                   // row g2 + e
    double2 *mid = stg + span2;                                           // row g1 + e
    double2 *hval = mid + span1;
    double *hdiag = reinterpret_cast<double *>(hval + (size_t)ZR * nh);
    int16_t *hnb = reinterpret_cast<int16_t *>(hdiag + nh);
    const int64_t r0 = rb * kBandRows;
    const int64_t g2 = r0 - 2 * b, g1 = r0 - b;
    const int64_t i = r0 + t;
    const bool live = i < n;
    // own row: neighbours as stage slots (mid slot = stage slot - band); empty
    // slots point at the row itself with value 0
    const int own2 = 2 * b + t;
    int nbs[ZR];
    double2 hv[ZR];
#pragma unroll
    for (int q = 0; q < ZR; ++q) {
        const int32_t cq = (live && q < zmax) ? col[(int64_t)q * n + i] : -1;
        nbs[q] = cq >= 0 ? (int)(cq - g2) : own2;
        hv[q] = cq >= 0 ? val[(int64_t)q * n + i] : make_double2(0.0, 0.0);
    }
    const double d = live ? diag[i] : 0.0;
    // halo row of thread t < 2 band: mid slot eh (before or after the own rows)
    const bool has_halo = t < nh;
    const int eh = t < b ? t : kBandRows + t;
    const int64_t gh = g1 + eh;
    const bool hlive = has_halo && gh >= 0 && gh < n;
    if (has_halo) {
#pragma unroll
        for (int q = 0; q < ZR; ++q) {
            const int32_t cq = (hlive && q < zmax) ? col[(int64_t)q * n + gh] : -1;
            hnb[q * nh + t] = (int16_t)(cq >= 0 ? (int)(cq - g2) : eh + b);
            hval[q * nh + t] = cq >= 0 ? val[(int64_t)q * n + gh] : make_double2(0.0, 0.0);
        }
        hdiag[t] = hlive ? diag[gh] : 0.0;
    }
    const int64_t j0 = cb * lcols;
    const int64_t j1 = j0 + lcols < ncols ? j0 + lcols : ncols;
    double2 pre[kBandStage];
    auto stage_load = [&](int64_t j) {
#pragma unroll
        for (int p = 0; p < kBandStage; ++p) {
            const int e = t + p * kBandRows;
            const int64_t g = g2 + e;
            pre[p] = (e < span2 && g >= 0 && g < n) ? Yin[j * ld + g] : make_double2(0.0, 0.0);
        }
    };
    if (j0 < j1) stage_load(j0);
    for (int64_t j = j0; j < j1; ++j) {
        // the stage is free: every thread passed the previous column's second
        // barrier after its last stage read (and the halo setup before the first)
#pragma unroll
        for (int p = 0; p < kBandStage; ++p) {
            const int e = t + p * kBandRows;
            if (e < span2) stg[e] = pre[p];
        }
        __syncthreads();
        if (j + 1 < j1) stage_load(j + 1);
        const double2 xo = (!first && live) ? X[j * ld + i] : make_double2(0.0, 0.0);
        const double2 xh = (!first && hlive) ? X[j * ld + gh] : make_double2(0.0, 0.0);
        // y_k on the own row
        const double2 yo = stg[own2];
        double2 acc = make_double2(d * yo.x, d * yo.y);
#pragma unroll
        for (int q = 0; q < ZR; ++q) {
            const double2 v = stg[nbs[q]];
            acc.x += hv[q].x * v.x - hv[q].y * v.y;
            acc.y += hv[q].x * v.y + hv[q].y * v.x;
        }
        const double2 x_own = first ? yo : xo;
        mid[b + t] = make_double2(x_own.x + c1 * acc.y, x_own.y - c1 * acc.x);
        // y_k on the halo row (zero outside the lattice)
        if (has_halo) {
            const double2 yh = stg[eh + b];
            const double dh = hdiag[t];
            double2 ah = make_double2(dh * yh.x, dh * yh.y);
            for (int q = 0; q < zmax; ++q) {
                const double2 hq = hval[q * nh + t];
                const double2 v = stg[hnb[q * nh + t]];
                ah.x += hq.x * v.x - hq.y * v.y;
                ah.y += hq.x * v.y + hq.y * v.x;
            }
            const double2 x_h = first ? yh : xh;
            mid[eh] = hlive ? make_double2(x_h.x + c1 * ah.y, x_h.y - c1 * ah.x) : make_double2(0.0, 0.0);
        }
        __syncthreads();
        // y_{k-1} on the own row
        const double2 ym = mid[b + t];
        double2 acc2 = make_double2(d * ym.x, d * ym.y);
#pragma unroll
        for (int q = 0; q < ZR; ++q) {
            const double2 v = mid[nbs[q] - b];
            acc2.x += hv[q].x * v.x - hv[q].y * v.y;
            acc2.y += hv[q].x * v.y + hv[q].y * v.x;
        }
        if (live) store_stream(Yout + j * ld + i, make_double2(x_own.x + c2 * acc2.y, x_own.y - c2 * acc2.x));
    }
}

// Right pass with the group's columns staged in LDS.
//   gcols : columns of the group (Z-order), group g = [g*G, min(g*G+G, n))
//   ucols : union of columns + neighbours, group g = [uptr[g], uptr[g+1])
//   umap  : per group member m, slot 0 = LDS slot of its own column, slots
//           1..zmax = LDS slots of its ELL neighbours (-1 if none)
__global__ void __launch_bounds__(kRightThreads) horner_right_v2(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag, int64_t n,
    int zmax, const int32_t *__restrict__ gcols, const int32_t *__restrict__ ucols, const int64_t *__restrict__ uptr,
    const int16_t *__restrict__ umap, int G, const double2 *__restrict__ Y, const double2 *__restrict__ Win,
    double2 *__restrict__ Wout, int64_t ld, int64_t nrows, double c, int first, int64_t rblocks, int64_t nreal,
    int64_t nprog) {
    extern __shared__ double2 tile[];          // [U][kRightRows]
    const int64_t lb = xcd_logical(blockIdx.x, nprog);
    if (lb >= nreal) return;
    const int64_t rb = lb % rblocks, g = lb / rblocks;
    const int64_t r0 = rb * kRightRows;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t row = r0 + lane;
    const bool rok = row < nrows;
    const int64_t u0 = uptr[g], U = uptr[g + 1] - u0;
    for (int64_t u = wave; u < U; u += kRightThreads / 64) {
        const int64_t cu = ucols[u0 + u];
        tile[u * kRightRows + lane] = rok ? Win[cu * ld + row] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int64_t m0 = g * G;
    const int64_t m1 = m0 + G < n ? m0 + G : n;
    const int zs = zmax + 1;
    for (int64_t m = m0 + wave; m < m1; m += kRightThreads / 64) {
        const int64_t j = gcols[m];
        const int16_t *mp = umap + m * zs;
        const double d = diag[j];
        const double2 w = tile[(int64_t)mp[0] * kRightRows + lane];
        double2 acc = make_double2(d * w.x, d * w.y);
        for (int q = 0; q < zmax; ++q) {
            const int slot = mp[1 + q];
            if (slot < 0) continue;
            const double2 hv = val[(int64_t)q * n + j];     // H[j, l]; H[l, j] = conj(hv)
            const double2 wl = tile[(int64_t)slot * kRightRows + lane];
            acc.x += hv.x * wl.x + hv.y * wl.y;
            acc.y += hv.x * wl.y - hv.y * wl.x;
        }
        if (rok) {
            const double2 y = first ? w : Y[j * ld + row];
            Wout[j * ld + row] = make_double2(y.x - c * acc.y, y.y + c * acc.x);
        }
    }
}

// Right pass without LDS staging: a thread owns one row, the block sweeps the
// columns of one group (Z-order, so consecutive columns share neighbours and
// their rows stay in L1/L2); H's row j is uniform across the block.
__global__ void __launch_bounds__(256) horner_right_v3(
    const int32_t *__restrict__ col, const double2 *__restrict__ val, const double *__restrict__ diag, int64_t n,
    int zmax, const int32_t *__restrict__ gcols, int G, const double2 *__restrict__ Y,
    const double2 *__restrict__ Win, double2 *__restrict__ Wout, int64_t ld, int64_t nrows, double c, int first,
    int64_t rblocks, int64_t nreal, int64_t nprog, int group_fastest) {
    const int64_t lb = xcd_logical(blockIdx.x, nprog);
    if (lb >= nreal) return;
    // group_fastest: an XCD walks the column groups of one row block in
    // Z-order, so the neighbour columns shared by nearby groups are re-read
    // from its L2; otherwise it walks the row blocks of one group
    const int64_t ngroups = nreal / rblocks;
    const int64_t rb = group_fastest ? lb / ngroups : lb % rblocks;
    const int64_t g = group_fastest ? lb % ngroups : lb / rblocks;
    const int64_t row = rb * 256 + threadIdx.x;
    if (row >= nrows) return;
    const int64_t m0 = g * G;
    const int64_t m1 = m0 + G < n ? m0 + G : n;
    for (int64_t m = m0; m < m1; ++m) {
        const int64_t j = gcols[m];
        const double d = diag[j];
        const double2 w = Win[j * ld + row];
        double2 acc = make_double2(d * w.x, d * w.y);
        for (int q = 0; q < zmax; ++q) {
            const int32_t l = col[(int64_t)q * n + j];
            if (l < 0) continue;
            const double2 hv = val[(int64_t)q * n + j];     // H[j, l]; H[l, j] = conj(hv)
            const double2 wl = Win[(int64_t)l * ld + row];
            acc.x += hv.x * wl.x + hv.y * wl.y;
            acc.y += hv.x * wl.y - hv.y * wl.x;
        }
        const double2 y = first ? w : Y[j * ld + row];
        store_stream(Wout + j * ld + row, make_double2(y.x - c * acc.y, y.y + c * acc.x));
    }
}

int64_t xcd_pad(int64_t nreal) { return ((nreal + kNumXCD - 1) / kNumXCD) * kNumXCD; }

}  // namespace

bool left_band_fits(const DevHam &h) { return h.band >= 0 && h.band <= kBandRows; }

hipError_t launch_horner_left(const DevHam &h, const double2 *X, const double2 *Yin, double2 *Yout, int64_t ld,
                              int64_t ncols, double c, int first, int lcols, hipStream_t st, int variant,
                              bool transposed_out) {
    if (lcols <= 0) lcols = kLeftCols;
    const int z = h.zmax;
    if (transposed_out && ncols != h.n) return hipErrorInvalidValue;
    // variant 3 (two terms per launch) is a step schedule; single terms of it are banded
    const bool band = variant == 2 || variant == 3 || (variant == 0 && left_band_fits(h));
    if (band && !left_band_fits(h)) return hipErrorInvalidValue;
    if (band) {
        const int64_t rblocks = (h.n + kBandRows - 1) / kBandRows;
        const int64_t cblocks = (ncols + lcols - 1) / lcols;
        const int64_t nreal = rblocks * cblocks, nprog = xcd_pad(nreal);
        const size_t lds = 2 * (size_t)(kBandRows + 2 * h.band) * sizeof(double2);
#define LMH_BAND(ZR, T)                                                                                          \
    hipLaunchKernelGGL((horner_left_band<ZR, T>), dim3((unsigned)nprog), dim3(kBandRows), lds, st, h.col, h.val,   \
                       h.diag, h.n, h.zmax, (int)h.band, X, Yin, Yout, ld, ncols, c, first, lcols, rblocks, nreal, \
                       nprog)
#define LMH_BAND_T(ZR) \
    if (transposed_out) LMH_BAND(ZR, true); else LMH_BAND(ZR, false);
        if (z <= 4) {
            LMH_BAND_T(4)
        } else if (z <= 8) {
            LMH_BAND_T(8)
        } else if (z <= 9) {          // honeycomb with next-nearest neighbours (Haldane)
            LMH_BAND_T(9)
        } else if (z <= 12) {
            LMH_BAND_T(12)
        } else {
            return hipErrorInvalidValue;
        }
#undef LMH_BAND_T
#undef LMH_BAND
        return hipGetLastError();
    }
    const int64_t rblocks = (h.n + kLeftThreads - 1) / kLeftThreads;
    const int64_t cblocks = (ncols + lcols - 1) / lcols;
    const int64_t nreal = rblocks * cblocks, nprog = xcd_pad(nreal);
#define LMH_LEFT(ZR, T)                                                                                            \
    hipLaunchKernelGGL((horner_left_v2<ZR, T>), dim3((unsigned)nprog), dim3(kLeftThreads), 0, st, h.col, h.val,     \
                       h.diag, h.n, h.zmax, X, Yin, Yout, ld, ncols, c, first, lcols, rblocks, nreal, nprog)
#define LMH_LEFT_T(ZR) \
    if (transposed_out) LMH_LEFT(ZR, true); else LMH_LEFT(ZR, false);
    if (z <= 4) {
        LMH_LEFT_T(4)
    } else if (z <= 8) {
        LMH_LEFT_T(8)
    } else if (z <= 12) {
        LMH_LEFT_T(12)
    } else if (z <= 24) {
        LMH_LEFT_T(24)
    } else {
        return hipErrorInvalidValue;
    }
#undef LMH_LEFT_T
#undef LMH_LEFT
    return hipGetLastError();
}

bool left_band2_fits(const DevHam &h) {
    const int zr = h.zmax <= 4 ? 4 : h.zmax <= 8 ? 8 : h.zmax <= 9 ? 9 : h.zmax <= 12 ? 12 : 0;
    return zr > 0 && h.band >= 0 && 2 * h.band <= kBandRows && band2_lds_bytes((int)h.band, zr) <= 163840;
}

hipError_t launch_horner_left2(const DevHam &h, const double2 *X, const double2 *Yin, double2 *Yout, int64_t ld,
                               int64_t ncols, double c1, double c2, int first, int lcols, hipStream_t st) {
    if (!left_band2_fits(h)) return hipErrorInvalidValue;
    if (lcols <= 0) lcols = kLeftCols;
    const int64_t rblocks = (h.n + kBandRows - 1) / kBandRows;
    const int64_t cblocks = (ncols + lcols - 1) / lcols;
    const int64_t nreal = rblocks * cblocks, nprog = xcd_pad(nreal);
    const int z = h.zmax;
#define LMH_BAND2(ZR)                                                                                            \
    hipLaunchKernelGGL(horner_left_band2<ZR>, dim3((unsigned)nprog), dim3(kBandRows), band2_lds_bytes((int)h.band, ZR), \
                       st, h.col, h.val, h.diag, h.n, h.zmax, (int)h.band, X, Yin, Yout, ld, ncols, c1, c2, first,  \
                       lcols, rblocks, nreal, nprog)
    if (z <= 4)
        LMH_BAND2(4);
    else if (z <= 8)
        LMH_BAND2(8);
    else if (z <= 9)
        LMH_BAND2(9);
    else
        LMH_BAND2(12);
#undef LMH_BAND2
    return hipGetLastError();
}

hipError_t launch_horner_right(const DevHam &h, const double2 *Y, const double2 *Win, double2 *Wout, int64_t ld,
                               int64_t nrows, double c, int first, int variant, int rcols, hipStream_t st) {
    if (rcols <= 0) rcols = kRightCols3;
    if (!h.gcols) return hipErrorInvalidValue;
    if (variant == 1 || variant == 5) variant = 3;   // 1, 5 (left passes on the transpose) are step schedules
    if (variant == 2) {
        const int64_t rblocks = (nrows + kRightRows - 1) / kRightRows;
        const int64_t nreal = rblocks * h.ngroups, nprog = xcd_pad(nreal);
        const size_t lds = (size_t)h.umax * kRightRows * sizeof(double2);
        hipLaunchKernelGGL(horner_right_v2, dim3((unsigned)nprog), dim3(kRightThreads), lds, st, h.col, h.val, h.diag,
                           h.n, h.zmax, h.gcols, h.ucols, h.uptr, h.umap, h.group, Y, Win, Wout, ld, nrows, c, first,
                           rblocks, nreal, nprog);
    } else {
        const int64_t rblocks = (nrows + 255) / 256;
        const int64_t ngroups = (h.n + rcols - 1) / rcols;
        const int64_t nreal = rblocks * ngroups, nprog = xcd_pad(nreal);
        hipLaunchKernelGGL(horner_right_v3, dim3((unsigned)nprog), dim3(256), 0, st, h.col, h.val, h.diag, h.n, h.zmax,
                           h.gcols, rcols, Y, Win, Wout, ld, nrows, c, first, rblocks, nreal, nprog,
                           variant == 4 ? 1 : 0);
    }
    return hipGetLastError();
}

hipError_t prepare_passes(hipStream_t) {
    static std::atomic<uint64_t> done{0};
    return once_per_device(done, [] {
        hipError_t e =
            hipFuncSetAttribute((const void *)horner_right_v2, hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
        const void *band[] = {(const void *)horner_left_band<4, false>, (const void *)horner_left_band<4, true>,
                              (const void *)horner_left_band<8, false>, (const void *)horner_left_band<8, true>,
                              (const void *)horner_left_band<9, false>, (const void *)horner_left_band<9, true>,
                              (const void *)horner_left_band2<4>, (const void *)horner_left_band2<8>,
                              (const void *)horner_left_band2<9>, (const void *)horner_left_band2<12>,
                              (const void *)horner_left_band<12, false>, (const void *)horner_left_band<12, true>};
        for (const void *f : band)
            if (e == hipSuccess) e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
        return e;
    });
}

int right_pass_group_size() { return 32; }
int64_t right_pass_max_union() { return 163840 / (kRightRows * (int64_t)sizeof(double2)); }

}  // namespace lmh
