// Internal declarations shared by the liblmhip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>

namespace lmh {

// Runs set() once per device (the current one) for the flag word `done`:
// one-time launch attributes such as the dynamic-LDS limit are per device,
// and several threads may race to set them.
template <class F>
hipError_t once_per_device(std::atomic<uint64_t> &done, F set) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const uint64_t bit = dev < 64 ? (1ull << dev) : 0;
    if (bit && (done.load() & bit)) return hipSuccess;
    e = set();
    if (e == hipSuccess && bit) done.fetch_or(bit);
    return e;
}

constexpr int kMaxTaylor = 64;      // upper bound on Taylor terms per application
constexpr int kNumXCD = 8;          // MI355X: 8 accelerator complex dies

// One gauge-field component (LMH_FIELD_* kind) with 3 parameters.
struct FieldComp {
    int32_t kind;
    int32_t pad;
    double a;     // B (Landau / symmetric) or flux (point fluxes)
    double px;
    double py;
};

// Grid-stencil geometry of H for the strip kernel (strip.hip): site
// i = x * ly + y on an lx x ly grid, every stored entry of a row being the
// diagonal or one of the four grid neighbours (y +- 1, x +- 1; across the
// edges when px / py).  A thread owns a strip of s consecutive sites of one
// x column (y0 .. y0 + s - 1; the last strip of a column overlaps its
// neighbour when s does not divide ly), a workgroup `cols` columns of P.
struct StripGeom {
    int lx = 0, ly = 0;
    int lyp = 0;                // LDS length of an x column (ly rounded up to odd)
    int px = 0, py = 0;         // periodic along x / y
    int s = 0;                  // sites per strip (0: H is not a grid stencil)
    int q = 0;                  // strips per x column
    int nt = 0;                 // threads per column of P = lx * q
    int cols = 0;               // columns of P per workgroup
    int region = 0;             // LDS entries per column: lx * lyp + 1 (the last holds 0)
};

// Geometry + Hamiltonian in ELL layout on the device.
//   slot q = s * n + row  (s < zmax): column index col[q] (-1 empty) and value val[q]
//   contributions of bonds to slot q: cbond[cptr[q] .. cptr[q+1]) with the
//   sign bit of the entry marking the Hermitian-partner (conjugated) direction.
struct DevHam {
    int64_t n = 0;
    int64_t nbonds = 0;
    int zmax = 0;
    int64_t band = 0;           // max |column - row| over the stored entries of H
    // bond data (device)
    double2 *ri = nullptr, *rj = nullptr, *amp = nullptr, *twist = nullptr;
    int64_t *bslot = nullptr;   // ELL slot of H[i_b, j_b] per bond
    int32_t *bi = nullptr, *bj = nullptr;
    int64_t *cptr = nullptr;    // [zmax*n + 1]
    int64_t *cbond = nullptr;   // bond index, ~b for the conjugated partner
    // Hamiltonian (device)
    int32_t *col = nullptr;     // [zmax*n]
    int32_t *col4 = nullptr;    // [ceil(zmax/4)][n][4]: a site's columns four to a 16-byte record
    double2 *val = nullptr;     // [zmax*n] (allocated to whole chunks of col4, padding zero)
    double *diag = nullptr;     // [n]
    double2 *rec = nullptr;     // [n] {diag, four 16-bit columns (empty: 0)}; zmax <= 4, n < 65536 only
    double *onsite = nullptr;   // [n]
    // field (device)
    FieldComp *field = nullptr;
    int64_t nfield = 0;
    int64_t field_cap = 0;
    // reductions: [0] max |H_ij| bits, [1] max row abs-sum bits, [2] nonzero
    // when a slot's value differs from its value class (see below), [3]
    // bit 0 when a strip's bonds of one direction differ, bit 1 when an x
    // bond is complex (strip kernel)
    unsigned long long *red = nullptr;
    // Value classes (zmax <= 4, n <= 10240): slots whose values are equal, or
    // conjugate, for every field amplitude share one class, so a fused pass
    // can read H's values from a table in LDS instead of from L2.  Per site a
    // 16-byte record {diag, four 16-bit slots: offset id (4 bit) | class
    // (11 bit) | conjugate (1 bit)}; column = site + doff[offset id].
    int ndict = 0;              // classes (0: no table for this geometry and field)
    int noff = 0;               // column offsets (id 0: the site itself, for empty slots)
    int dict_ok = 0;            // the last H build matched every slot to its class
    uint64_t *dpack = nullptr;  // [n] packed slots of a site
    double2 *drec = nullptr;    // [n] {diag, dpack}, written with H
    int32_t *dcls = nullptr;    // [zmax*n] class | conjugate << 15 (empty slot: 0)
    int32_t *drep = nullptr;    // [ndict] slot whose value defines the class (-1: class 0, value 0)
    double2 *dval = nullptr;    // [ndict] class values, written with H
    int32_t *doff = nullptr;    // [16] column offsets
    // strip kernel: grid geometry, the ELL slot of each direction of each
    // site ([4][n], -1: no bond), and per strip thread the value of each
    // direction ([4][nt]) and the diagonal ([s][nt]), rewritten after each H
    // build; red[3] != 0 when a thread's bonds of one direction differ
    StripGeom sg;
    int32_t *sslot = nullptr;
    uint32_t *sdir = nullptr;   // [n] direction (2 bits) of each ELL slot of a site (grid-window Taylor search)
    double2 *stval = nullptr;
    double *sdiag = nullptr;
    int strip_ok = 0;           // the last H build gave every strip uniform bonds
    int strip_xreal = 0;        // ... and real x bonds
    // right-pass column groups (sites close on the lattice), see passes.hip
    int32_t *gcols = nullptr;   // [n] columns in group order
    int32_t *ucols = nullptr;   // unions of each group's columns and neighbours
    int64_t *uptr = nullptr;    // [ngroups + 1]
    int16_t *umap = nullptr;    // [n * (zmax + 1)] LDS slots per group member
    int64_t ngroups = 0;
    int64_t umax = 0;
    int group = 0;
};

// kernels/launchers (hamiltonian.hip)
// builds val/diag/rec (and drec/dval with the class check when h.ndict > 0)
hipError_t launch_build_h(const DevHam &h, hipStream_t st);
// CSC values of a fixed pattern -> bond amplitudes and on-site energies;
// flag: bit 0 not Hermitian, bit 1 complex diagonal, bit 2 not finite
hipError_t launch_csc_gather(int64_t nbonds, int64_t n, const int64_t *kup, const int64_t *klo, const int64_t *kd,
                             const double2 *nz, double2 *amp, double *onsite, int *flag, hipStream_t st);
// out[p * zmax * n + q] = value of slot q for field amplitudes probe[p * nfield ..]
// (same kinds and points as h.field), p < nprobe
hipError_t launch_slot_probe(const DevHam &h, const FieldComp *probe, int nprobe, double2 *out, hipStream_t st);
constexpr int kDictMaxClasses = 2048;
constexpr int kDictMaxOffsets = 16;

// The order rule's device records (taylor.hip).  TaylorRun is derived on the
// device from H's reductions (red[0] max|H_ij|, red[1] max row abs-sum) by
// launch_taylor_setup; the order kernels read it, and launch_taylor_finish
// writes TaylorOut: the result with a copy of red[], and the Horner
// coefficients of the step, so that a strip launch can take its term count
// and coefficients from device memory instead of from the host.
constexpr int kStripMaxK = 31;          // Horner coefficients carried per polynomial
struct TaylorRun {
    double s, inv_s;            // s = nextpow(2, max|H| |dt|)
    double thr;                 // threshold / s
    double growth;              // |dt| rowsum / s: terms can only shrink from n + 1 > growth on
    double dt;
    int nmax;                   // first n with growth^n / n! <= thr (0: U = I)
    int pad;
};
struct TaylorOut {
    unsigned long long red[4];
    double s;
    int K;                      // the order (0: U = I; -1: threshold not reached within nmax terms)
    int overflow;               // a column's neighbourhood did not fit the kernel's table
    int nmax;
    int reps;                   // s as an integer (0 when s > 2^30)
    double dts;                 // dt / s
    double ck[kStripMaxK + 1];  // ck[k] = dts / k, k >= 1
};
// A strip launch whose polynomial comes from a TaylorOut it did not see on
// the host runs only when that result is what the host assumed when it chose
// the launch: a strip-form H (bit 0 of red[3] clear) whose x bonds are real
// exactly when xreal, and an order 1 <= K <= kStripMaxK that was found.  The
// same test on the host tells whether the launch ran.
__host__ __device__ inline bool strip_coef_ok(const TaylorOut &o, bool xreal) {
    return o.overflow == 0 && o.K >= 1 && o.K <= kStripMaxK && o.reps >= 1 && (o.red[3] & 1ull) == 0 &&
           ((o.red[3] & 2ull) == 0) == xreal;
}

// kernels/launchers (strip.hip): the fused Taylor pass for grid-stencil H.
// strip_detect: from H's ELL columns (host), the grid and the slot of each
// direction ([4][n]); false when H's pattern is not a grid stencil the
// strip kernel can run.  The per-strip check runs inside launch_build_h.
bool strip_detect(const int32_t *col, int64_t n, int zmax, StripGeom *g, int32_t *sslot);
int64_t strip_thread_count(const StripGeom &g);
hipError_t launch_strip_check(const DevHam &h, hipStream_t st);
// a strip polynomial waiting to be applied: the second side of a step whose
// launch was deferred, with its H's strip values kept apart (the context's
// are rebuilt by the next H)
// (coef: its K, reps and coefficients are read from this device record when
// the launch runs; else dts, K and reps are the ones given)
struct StripPending {
    const double2 *stval;
    const double *sdiag;
    double dts;
    int K;
    int reps;
    int xreal;
    const TaylorOut *coef;
};
// one side of a step: Y = p(-i dt/s H)^reps X column by column; tout: Y
// holds the conjugate transpose of the result (square, ncols == n).  With
// second: Y = p(H) q X, q the pending polynomial (the previous step's
// second side), in the same launch.  With coef (device): this side's K, reps
// and coefficients come from coef when the launch runs (dt_over_s, K and
// reps are then estimates for the launch shape only), and the launch does
// nothing unless strip_coef_ok(*coef, h.strip_xreal) holds on the device.
hipError_t launch_strip(const DevHam &h, const double2 *X, double2 *Y, int64_t ld, int64_t ncols, double dt_over_s,
                        int K, int reps, bool tout, hipStream_t st, const StripPending *second = nullptr,
                        const TaylorOut *coef = nullptr);
// the strip values of h's current build and the record coef copied for a
// deferred pass, when strip_coef_ok(*coef, xreal) holds on the device (else
// the destinations keep what they hold)
hipError_t launch_strip_snapshot(const DevHam &h, const TaylorOut *coef, bool xreal, double2 *stval, double *sdiag,
                                 TaylorOut *coef_out, hipStream_t st);
// the pending polynomial alone (with the strip geometry of h)
hipError_t launch_strip_pending(const DevHam &h, const StripPending &q, const double2 *X, double2 *Y, int64_t ld,
                                int64_t ncols, bool tout, hipStream_t st);
// strip values of h's current build: [4][nt] bond values, [s][nt] diagonal
inline int64_t strip_stval_count(const DevHam &h) { return 4 * (int64_t)h.sg.nt; }
inline int64_t strip_sdiag_count(const DevHam &h) { return (int64_t)h.sg.s * h.sg.nt; }
hipError_t prepare_strip(const DevHam &h);
// batched strip passes: one persistent launch over the column groups of
// many members (each a grid-stencil H with its strip geometry checked)
struct StripMemberDesc {
    const DevHam *h;
    const double2 *X;
    double2 *Y;
    int64_t ld;
    double dts;
    int K;
    int reps;
};
// device table for one pass over the members (hipFree both pointers)
hipError_t strip_batch_table(const StripMemberDesc *m, int nm, void **d_args, int64_t **d_ustart, int64_t *nunits,
                             size_t *lds);
hipError_t launch_batched_strip(const void *d_args, const int64_t *d_ustart, int nm, int64_t nunits, size_t lds,
                                bool tout, hipStream_t st);

// kernels/launchers (taylor.hip): the reference's order rule on the device
// (TaylorRun and TaylorOut are declared before the strip kernel's launchers).
int taylor_order_levels();
int64_t taylor_order_scratch();
hipError_t launch_taylor_setup(const unsigned long long *red, double dt, double threshold, TaylorRun *run,
                               int *overflow, hipStream_t st);
// level 0..taylor_order_levels()-1 picks the neighbourhood table size (the
// grid-window kernel at level 0 when H is a grid stencil)
hipError_t launch_taylor_order(const DevHam &h, double dt, double nztol, const TaylorRun *run, int level,
                               double *partmax, unsigned long long *maxbits, int *overflow, hipStream_t st);
hipError_t launch_taylor_finish(const unsigned long long *red, const TaylorRun *run,
                                const unsigned long long *maxbits, const int *overflow, TaylorOut *out,
                                hipStream_t st);

// kernels/launchers (passes.hip)
// lcols: columns swept per left-pass thread or block (0 = default); variant:
// 0 = banded when H's band fits the LDS tile, else gather; 1 = gather through
// L1/L2; 2 = banded.  transposed_out: write the conjugate transpose (square).
bool left_band_fits(const DevHam &h);
hipError_t launch_horner_left(const DevHam &h, const double2 *X, const double2 *Yin, double2 *Yout,
                              int64_t ld, int64_t ncols, double c, int first, int lcols, hipStream_t st,
                              int variant = 0, bool transposed_out = false);
// two Horner terms per launch (banded H; see left_band2_fits): y_{k-1} from
// y_{k+1} with coefficients c1 = dt/(s k), c2 = dt/(s (k-1))
bool left_band2_fits(const DevHam &h);
hipError_t launch_horner_left2(const DevHam &h, const double2 *X, const double2 *Yin, double2 *Yout, int64_t ld,
                               int64_t ncols, double c1, double c2, int first, int lcols, hipStream_t st);
// variant 2: group columns staged in LDS; 3: unstaged sweep of rcols Z-ordered
// columns per block (0 and 1 = 3)
hipError_t launch_horner_right(const DevHam &h, const double2 *Y, const double2 *Win, double2 *Wout,
                               int64_t ld, int64_t nrows, double c, int first, int variant, int rcols,
                               hipStream_t st);
hipError_t prepare_passes(hipStream_t st);
int right_pass_group_size();
int64_t right_pass_max_union();

// kernels/launchers (observables.hip)
// currents from region src into region dst (dst null: the complement of src):
// per site out[j] (may be null), per-block partial sums part[region_currents_parts],
// their total into *total (device, may be null)
// conj (also launch_observables): the buffer holds P^H, not P
hipError_t launch_region_currents(const DevHam &h, const double2 *P, int64_t ld, const uint8_t *src,
                                  const uint8_t *dst, double *out, double *part, double *total, hipStream_t st,
                                  int conj = 0);
int64_t region_currents_parts(const DevHam &h);
// the same with the partial sums in the library's static device scratch (no
// allocation; up to region_currents_static_max_sites() sites); prepare looks
// the scratch up on the current device
int64_t region_currents_static_max_sites();
hipError_t prepare_region_currents_static();
hipError_t launch_region_currents_static(const DevHam &h, const double2 *P, int64_t ld, const uint8_t *src,
                                         const uint8_t *dst, double *out, double *total, hipStream_t st, int conj);
hipError_t launch_series_differentiate(double *v, int64_t width, int64_t len, const double *a, const double *b,
                                       hipStream_t st);
hipError_t launch_series_integrate(double *v, int64_t width, int64_t len, const double *hh, hipStream_t st);
hipError_t launch_observables(const DevHam &h, const double2 *P, int64_t ld, double *rho, double *J,
                              hipStream_t st, int conj = 0);
hipError_t launch_copy(const double2 *src, double2 *dst, int64_t count, hipStream_t st);
hipError_t launch_scatter(double2 *P, int64_t ld, int64_t count, const int64_t *rows, const int64_t *cols,
                          const double2 *vals, hipStream_t st);
hipError_t launch_gather(const double2 *P, int64_t ld, int64_t count, const int64_t *rows, const int64_t *cols,
                         double2 *out, hipStream_t st);
hipError_t launch_gather_h(const DevHam &h, double *diag, double2 *hij, hipStream_t st);
hipError_t launch_ket_observables(const DevHam &h, const double2 *psi, double *rho, double *J, hipStream_t st);

// kernels/launchers (krylov.hip): Lanczos over the columns of a ket block.
// Basis vector k of column c at V + k * ld * m + c * ld; per-column arrays
// (alpha, beta) hold m doubles per Lanczos index; part_* nb * m partial sums.
// Iteration j: v_j = W / |W| (|W|^2 from part_w), Wn = H v_j - beta_j v_{j-1}
// - alpha_j v_j, |Wn|^2 partials into part_wn; beta_j and alpha_j recorded.
int lanczos_blocks(int64_t n);
hipError_t launch_lanczos_start(const double2 *X, int64_t n, int64_t ld, int64_t m, double *part, hipStream_t st);
hipError_t launch_lanczos_iteration(const DevHam &h, const double2 *W, const double *part_w, double2 *Vj,
                                    const double2 *Vprev, double2 *Wn, int64_t ld, int64_t m, double *part_re,
                                    double *part_im, double *part_wn, double *alpha_out, double *beta_out,
                                    hipStream_t st);
// beta_out[c] = sqrt(sum of column c's partials)
hipError_t launch_lanczos_beta(const double *part, int64_t n, int64_t m, double *beta_out, hipStream_t st);
// Y[:, c] = sum_k coef[c * kmax + k] V_k[:, c] for k < len[c]; len[c] < 0: Y[:, c] = X[:, c]
hipError_t launch_krylov_combine(const double2 *V, int64_t vstride, const double2 *coef, int kmax, const int *len,
                                 const double2 *X, double2 *Y, int64_t n, int64_t ld, int64_t m, hipStream_t st);

// one density matrix of a batch, for one side of a step (batched fused pass)
struct BatchItem {
    const int32_t *col;
    const int32_t *col4;
    const double2 *rec;
    const double2 *val;
    const double *diag;
    const double2 *src;
    double2 *dst;
    double2 *scratch;
    int64_t ld;
    int n;
    int zmax;
    int K;
    int reps;
    double dts;
    // value-class variants: the member's records and class table
    const double2 *drec;
    const double2 *dval;
    const int32_t *doff;
    int ndict;
};
struct BatchTrace {
    const double2 *P;
    int64_t ld;
    int64_t n;
};

// kernels/launchers (fused.hip): variants = (threads, sites/thread, columns,
// H row in registers), see fused.hip
int fused_variant_count();
bool fused_variant_valid(int v, int64_t n, int zmax);
// variants that read H's values from the class table (h.ndict > 0 and
// h.dict_ok), and the variant of the same shape they fall back to otherwise
bool fused_variant_is_dict(int v);
int fused_variant_fallback(int v);
// the class table of ndict entries fits beside the variant's columns in LDS
bool fused_dict_fits(int v, int64_t n, int ndict);
int fused_variant_columns(int v);
int fused_default_variant(int64_t n, int zmax);
int fused_batch_variant(int64_t n, int zmax);
hipError_t prepare_fused(int variant);
// maxdict: the largest class table among the items (value-class variants)
hipError_t launch_batched_fused(int variant, const BatchItem *items, const int64_t *bstart, int nitems, int64_t total,
                                int64_t maxn, int maxdict, hipStream_t st);
// the value-class variant a by-size default v switches to where the
// classes hold (-1: keep v)
int fused_default_dict(int v);
hipError_t launch_batched_mean_density(const BatchTrace *items, int nitems, double *out, hipStream_t st);
hipError_t prepare_batched(int variant);
bool fused_supported(const DevHam &h);
// variant < 0: fused_default_variant
hipError_t launch_fused_left(const DevHam &h, const double2 *X, double2 *Y, double2 *scratch, int64_t ld,
                             int64_t ncols, double dt_over_s, int K, int reps, bool transposed_out, int variant,
                             hipStream_t st);
hipError_t launch_conj_transpose(const double2 *A, double2 *B, int64_t n, int64_t ld, hipStream_t st);
// the variant number of the strip kernel, and the variant a fused launch with
// `variant` (-1: default) runs for the current H
int fused_strip_variant();
int fused_resolved_variant(const DevHam &h, int variant);

// kernels/launchers (adjoint.hip): one side of a reverse step, see adjoint.hip
struct AdjointSide {
    const double2 *X;       // x source columns
    int x_given;            // 1: x = X[:, c] (checkpoint); 0: x = U^H X[:, c]
    double2 *Xout;          // x written here (nullptr: not written)
    int xout_t;             // ... conjugate-transposed
    const double2 *G;       // cotangent columns of U x
    double2 *Gout;          // U^H G written here
    int gout_t;             // ... conjugate-transposed
    int64_t ld;
    double dt;
    int K;
    double2 *partial;       // [nparts][(zmax + 1) * n] F partial sums
    int nparts;
    int accumulate = 0;     // add to the partial sums already there
};
bool adjoint_supported(const DevHam &h, int K);
int adjoint_parts(const DevHam &h);
int64_t adjoint_reduce_blocks(const DevHam &h);
hipError_t launch_adjoint_side(const DevHam &h, const AdjointSide &a, hipStream_t st);
// F = scale * (sum of the partials): 2 when only one side's sweep was summed
hipError_t launch_adjoint_reduce(const DevHam &h, const double2 *partial, int nparts, double scale,
                                 double *gonsite, double *gfield, hipStream_t st);
hipError_t launch_bond_derivative(const DevHam &h, int64_t comp, double2 *out, hipStream_t st);
// per-term reverse Horner over all columns (adjoint for any N); partial holds
// reverse_term_parts(h) blocks of (zmax + 1) * n sums, accumulated
int64_t reverse_term_parts(const DevHam &h);
bool reverse_term_supported(const DevHam &h);
// Term m of K: xbar = g_1 + ... + g_{K+1} is built in XB two terms at a time
// (odd m adds g_m + g_{m+1}, the last term g_{K+1}; term 1 writes XB without
// reading it, so XB needs no clearing).  variant 1: neighbours gathered
// through L1/L2; otherwise the banded kernel (rows + band staged in LDS per
// column) when H's band is <= 512 rows
hipError_t launch_reverse_term(const DevHam &h, const double2 *G, const double2 *Y, double2 *XB, double2 *Gout,
                               int m, int K, int64_t ld, double cr, double2 *partial, hipStream_t st,
                               int variant = 0);
hipError_t launch_hermitian_part(const double2 *src, double2 *dst, int64_t n, int64_t ld, hipStream_t st);

// kernels/launchers (spectrum.hip)
hipError_t launch_zero(double2 *A, int64_t count, hipStream_t st);
hipError_t launch_dense_from_ell(const DevHam &h, double2 *A, int64_t lda, hipStream_t st);
hipError_t launch_scale_columns(double2 *V, int64_t lda, int64_t n, const double *w, int64_t m, hipStream_t st);
hipError_t launch_herk(const double2 *W, int64_t ldw, int64_t n, int64_t m, double2 *P, int64_t ldp, hipStream_t st,
                       const double *sgn = nullptr);
hipError_t launch_nonfinite(const double2 *A, int64_t n, int64_t ld, int *flag, hipStream_t st);
// max |P P - P| with W an n x ld scratch matrix and part[nblocks] (0 on success)
int idempotency(hipStream_t st, const double2 *P, int64_t n, int64_t ld, double2 *W, double *part, int nblocks,
                double *out);
// Hermitian eigensolver (rocSOLVER): method 0 zheevd, 1 zheev (QR), 2 zheevdj (Jacobi)
int eigh(hipStream_t st, int method, double2 *A, int64_t n, int64_t lda, double *E, double *work, int *info);
// real symmetric H: flag[0] = 1 if any stored hopping has a nonzero imaginary
// part; the real dense matrix; its eigenvectors widened to complex; the real
// eigensolver (method 0 dsyevd, 1 dsyev (QR), 2 dsyevdj (Jacobi))
hipError_t launch_imag_check(const DevHam &h, int *flag, hipStream_t st);
hipError_t launch_dense_real_from_ell(const DevHam &h, double *A, int64_t lda, hipStream_t st);
hipError_t launch_widen(const double *R, int64_t ldr, double2 *A, int64_t lda, int64_t n, hipStream_t st);
int eigh_real(hipStream_t st, int method, double *A, int64_t n, int64_t lda, double *E, double *work, int *info);

}  // namespace lmh
