// Internal declarations shared by the C-ABI translation units (api_*.hip):
// the definitions of the library's opaque handles (contexts, states, kets,
// series, batches) and the helpers the entry points share.  The public
// declarations and the reference interfaces they replace are in
// include/lmhip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/lmhip.h"
#include "lmhip_internal.h"
#include "tridiag.h"

namespace lmhapi {

extern thread_local std::string g_err;   // lmh_last_error
int fail(int code, const char *fmt, ...);
int hipfail(hipError_t e, const char *what);

}  // namespace lmhapi

#define HIP_TRY(expr, what)                    \
    do {                                       \
        hipError_t e_ = (expr);                \
        if (e_ != hipSuccess) return hipfail(e_, what); \
    } while (0)

namespace lmhapi {

template <class T>
hipError_t dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void **)p, count * sizeof(T));
}

template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

inline bool isapprox(double a, double b) {
    // Julia's default `≈` (rtol = sqrt(eps), atol = 0)
    return a == b || std::fabs(a - b) <= std::sqrt(2.220446049250313e-16) * std::max(std::fabs(a), std::fabs(b));
}

template <class T>
hipError_t grow(T *&p, size_t &cap, size_t count) {
    if (count <= cap) return hipSuccess;
    dfree(p);
    cap = 0;
    hipError_t e = dalloc(&p, count);
    if (e == hipSuccess) cap = count;
    return e;
}

}  // namespace lmhapi

struct lmh_state;

struct lmh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    lmh_progress_fn progress = nullptr;   // lmh_set_progress
    void *progress_user = nullptr;
    lmh::DevHam h;
    bool geom_set = false;
    // Hamiltonian given as the values of a fixed CSC pattern (lmh_set_csc_pattern /
    // lmh_set_csc_values): the pattern's entries above the diagonal are the
    // bonds, its values their amplitudes, no field
    bool csc_mode = false, csc_values = false;
    int64_t csc_nnz = 0;
    int64_t *d_kup = nullptr, *d_klo = nullptr, *d_kd = nullptr;
    double2 *d_nz = nullptr;
    int *d_cscflag = nullptr;
    std::vector<double> onsite_host;
    bool onsite_set = false;
    std::vector<lmh::FieldComp> field_host;
    uint64_t param_version = 1;
    uint64_t built_version = 0;
    // solver
    bool solver_ready = false;
    double dt = 0.0, s = 1.0, threshold = 1e-10, nztol = 1e-14;
    int K = 0;
    uint64_t solver_version = 0;
    uint64_t solver_serial = 0;       // bumped by every recompute of K, s and dt (lmh_update_solver)
    // scratch
    unsigned long long *d_maxbits = nullptr;
    double *d_partmax = nullptr;      // Taylor-order partial maxima per workgroup
    int *d_overflow = nullptr;
    lmh::TaylorRun *d_run = nullptr;  // order-rule parameters derived on the device
    lmh::TaylorOut *d_tout = nullptr; // its result, read back once per update
    double *d_rho = nullptr, *d_J = nullptr;
    double *d_hdiag = nullptr;
    double2 *d_hij = nullptr;
    std::set<lmh_state *> states;
    std::set<lmh_kets *> kets;
    std::set<lmh_series *> series;
    // profiling: event pairs per kernel kind, resolved on query
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_pending[LMH_NUM_KERNEL_KINDS];
    int64_t prof_count[LMH_NUM_KERNEL_KINDS] = {0};
    double prof_ms[LMH_NUM_KERNEL_KINDS] = {0};
    std::vector<hipEvent_t> event_pool;
    int step_mode = LMH_STEP_AUTO;
    int right_variant = 3;
    int left_variant = 0;             // LMH_OPT_LEFT_PASS
    // eigensystem of lmh_diagonalize (eigenvectors N x ld on the device)
    double2 *eigvec = nullptr;
    int64_t eig_n = 0, eig_ld = 0;
    std::vector<double> eig_E;
    int left_cols = 32;
    int right_cols = 16;
    int64_t fused_max_n = 10240;      // AUTO: fused column passes up to this N
    std::vector<int32_t> col_host;   // ELL columns (host copy, for column grouping)
    int eigensolver = 0;              // LMH_OPT_EIGENSOLVER
    int eig_real = -1, eig_method = -1;  // solver of the last eigendecomposition (LMH_OPT_EIGEN_USED)
    int batch_strip = 1;              // LMH_OPT_BATCH_STRIP
    int defer = 1;                    // LMH_OPT_DEFER
    // K on the device (LMH_OPT_ASYNC_ORDER): an update whose caller does not
    // ask for K leaves the order search's result on the device, and strip
    // steps read their polynomial from it (ctx->K and ctx->s are then the
    // last known ones, estimates for launch shapes).  The result comes back
    // into pinned memory behind ev_order and is taken up (ctx_resolve) by
    // the next call that needs it; the steps launched meanwhile are kept in
    // the journal, so that those whose launches found a result other than
    // the one assumed (and so did nothing) can be run again.
    int async_order = 1;
    bool order_async = false;         // an update's result not yet taken up
    int async_xreal = 0;              // the x-bond form the strip launches assumed
    int async_level = 0;              // the table level its order search started at
    // the order search starts at the table level the last one for this
    // geometry needed (0: the grid window; see launch_taylor_order)
    int order_level = 0;
    uint64_t order_level_geom = 0;    // geometry version + 1 it belongs to
    uint64_t flags_geom = 0;          // geometry version of the last result taken up (+1; 0: none)
    lmh::TaylorOut *h_tout = nullptr; // pinned
    hipEvent_t ev_order = nullptr;
    lmh::FieldComp *field_stage = nullptr;    // pinned staging of field uploads (h.field_cap entries)
    char *rb_stage = nullptr;                 // pinned staging of read-backs up to 64 MB (readback_sync)
    size_t rb_cap = 0;
    hipEvent_t ev_field = nullptr;            // after the last field upload
    struct JournalEntry {
        lmh_state *st;
        int cur, conj;
        bool pending;
        lmh::StripPending pend;
    };
    std::vector<JournalEntry> journal;
    int64_t stat_steps = 0, stat_terms = 0;   // LMH_OPT_STEP_COUNT, LMH_OPT_STEP_TERMS
    int64_t stat_async = 0, stat_redone = 0;  // LMH_OPT_ASYNC_STEPS, LMH_OPT_ASYNC_REDONE
    // reusable device memory: N x ld buffers (checkpoints, adjoint work),
    // adjoint partial sums, recorded field parameters, small transfers
    std::vector<double2 *> pool;
    int64_t pool_elems = 0;
    double2 *adj_partial = nullptr;
    size_t adj_partial_cap = 0;
    lmh::FieldComp *d_tape_fields = nullptr;
    size_t tape_fields_cap = 0;
    char *xfer = nullptr;
    size_t xfer_cap = 0;
    int fused_variant = -1;           // LMH_OPT_FUSED_VARIANT (-1: by size)
    int fused_launched = -1;          // LMH_OPT_FUSED_LAUNCHED
    int adjoint_mode = 0;             // LMH_OPT_ADJOINT_MODE
    uint64_t geom_version = 0;        // bumped by lmh_set_geometry
    uint64_t onsite_version = 0;      // bumped when the on-site energies change
    // recording of steps for lmh_tape_backward
    struct TapeRec {
        std::vector<lmh::FieldComp> field;
        double dt;
        int K;
        double s;
    };
    struct Tape {
        bool active = false;
        lmh_state *st = nullptr;
        int64_t every = 1;
        uint64_t geom_version = 0, onsite_version = 0;
        std::vector<TapeRec> recs;
        std::vector<double2 *> checkpoints;
        std::vector<char> cp_host;    // checkpoint in pinned host memory
    } tape;
    int checkpoint_place = 0;         // LMH_OPT_CHECKPOINTS
    int krylov_graphs = 1;            // LMH_OPT_KRYLOV_GRAPHS
    int dict_probes = 3;              // LMH_OPT_VALUE_CLASSES: field probes that define H's value classes
    // value classes were found for this geometry, field structure and probe count
    uint64_t dict_geom = 0;
    int dict_probes_built = -1;
    std::vector<double> dict_sig;
};

// a time series of per-step observables on the device (entry k at buf[k * width])
struct lmh_series {
    lmh_context *ctx = nullptr;
    int kind = 0;
    int64_t width = 0, cap = 0, len = 0;
    uint64_t geom_version = 0;
    double *buf = nullptr;
    std::vector<double> times;
    uint8_t *src = nullptr, *dst = nullptr;   // region masks (LMH_SERIES_REGION)
    double *part = nullptr;                   // region partial sums
    double *coef = nullptr;                   // 2 (cap - 1) coefficients of differentiate / integrate
};

struct lmh_kets {
    lmh_context *ctx = nullptr;
    double2 *buf[3] = {nullptr, nullptr, nullptr};
    int cur = 0;
    int64_t n = 0, m = 0, ld = 0;
    // lmh_krylov_step: Lanczos basis, partial sums, recurrence coefficients
    double2 *basis = nullptr;
    size_t basis_cap = 0;
    double *lz = nullptr;
    size_t lz_cap = 0;
    double2 *coef = nullptr;
    size_t coef_cap = 0;
    int *len = nullptr;
    size_t len_cap = 0;
    int krylov_len = 0;               // basis length the last step converged at
    // captured first chunks of a step (start, `target` iterations, the next
    // beta) keyed by source buffer and length; valid for the pointers, sizes
    // and geometry they were captured with
    struct ChunkGraph {
        const double2 *src;
        int target;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<ChunkGraph> graphs;
    const void *graph_basis = nullptr, *graph_lz = nullptr;
    int graph_kd = 0;
    uint64_t graph_geom = 0;
    // host copies of the coefficients and basis lengths
    std::vector<std::complex<double>> h_coef;
    std::vector<int> h_len;
    // pinned staging: the recurrence coefficients read back, then the
    // coefficients and lengths uploaded (written only after a wait on the
    // stream, so an upload in flight is never overwritten)
    char *h_pin = nullptr;
    size_t h_pin_cap = 0;
};

struct lmh_state {
    lmh_context *ctx = nullptr;
    double2 *buf[3] = {nullptr, nullptr, nullptr};
    int cur = 0;
    // buf[cur] holds P^H instead of P: a fused step computes U (U S)^H =
    // (U S U^H)^H from the stored S, so consecutive steps alternate between
    // the two forms with one transposition per step; everything that hands
    // the buffer out or reads it as P either accounts for the flag or brings
    // the state back to P first (state_normalize)
    int conj = 0;
    // a deferred pass (LMH_OPT_DEFER): the second side of the last fused
    // strip step, U applied to buf[cur] from the left, not yet launched; the
    // next fused step runs it in the same launch as its own first side, and
    // everything that reads or hands out the state runs it first
    // (state_settle).  Its H's strip values are kept in pend_stval/pend_sdiag.
    bool pending = false;
    lmh::StripPending pend{};
    double2 *pend_stval = nullptr;
    double *pend_sdiag = nullptr;
    lmh::TaylorOut *pend_coef = nullptr;   // its polynomial, when the host has not seen it
    int64_t pend_cap_stval = 0, pend_cap_sdiag = 0;
    int64_t n = 0;
    int64_t ld = 0;
};

struct lmh_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<lmh_context *> ctxs;
    std::vector<lmh_state *> states;
    std::vector<uint64_t> versions;      // member solver versions captured at creation
    std::vector<uint64_t> serials;       // member solver serials (K, s, dt) the tables hold
    std::vector<uint64_t> geoms;         // member geometry versions at creation
    std::vector<int> cur0;               // member buffer roles at creation
    struct Class {
        int cls = 0;
        std::vector<int> members;
        int64_t total = 0, maxn = 0;
        int maxdict = 0;                 // largest class table (value-class variants)
        lmh::BatchItem *items[6] = {};   // (phase 0..2) x (side 0, 1)
        int64_t *bstart = nullptr;
    };
    std::vector<Class> classes;
    // members with a grid-stencil H whose strips hold (strip_ok): one
    // persistent strip launch over all of them per step, plus one per call.
    // A call of n steps on Z (= P) launches (U Z)^H, then n - 1 times
    // (U U Y)^H -- the previous step's second side and the next step's
    // first side in one launch (the polynomial applied twice) -- then U Y,
    // conjugate-transposed when n is odd, so that every call ends with P
    // stored.  Each launch moves the state one buffer on; tables per
    // (rotation phase, single or doubled polynomial)
    std::vector<int> strip_members;
    int64_t strip_rot = 0;               // strip launches since creation
    void *strip_args[6] = {};
    int64_t *strip_ustart[6] = {};
    int64_t strip_units = 0;
    size_t strip_lds = 0;
    std::vector<int> member_kernel;      // per member: kStripVariant or its fused class
    int64_t steps_done = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    lmh::BatchTrace *d_trace = nullptr;
    double *d_out = nullptr;
};

namespace lmhapi {

void pool_free(lmh_context *ctx);
hipError_t pool_take(lmh_context *ctx, int64_t elems, double2 **out);
void pool_give(lmh_context *ctx, double2 *p);
void tape_clear(lmh_context *ctx);
hipError_t checkpoint_alloc(lmh_context *ctx, int64_t elems, double2 **out, bool *host);
int tape_record(lmh_context *ctx, lmh_state *st);
void free_dict(lmh::DevHam &h);
std::vector<double> dict_signature(const lmh_context *ctx);
hipError_t build_dict(lmh_context *ctx);
void free_geometry(lmh::DevHam &h);
int build_right_groups(lmh::DevHam &h, const std::vector<int32_t> &col, const std::vector<int32_t> &order);
std::vector<int32_t> zorder(const double *xy, int64_t n);
int ctx_resolve(lmh_context *ctx);
// the context's device current; without_resolve: an order result still on
// the device stays there (lmh_step), else it is taken up first
int check_ctx(lmh_context *ctx, bool without_resolve = false);
hipEvent_t take_event(lmh_context *ctx);
void resolve_profile(lmh_context *ctx);
int state_settle(lmh_state *st);
int ctx_settle_all(lmh_context *ctx);
int state_normalize(lmh_state *st);
void csc_release(lmh_context *ctx);
int order_finish(lmh_context *ctx, lmh::TaylorOut &out, int level0);
bool strip_step_path(const lmh_context *ctx);
bool order_may_stay(const lmh_context *ctx);
int check_entries(const lmh_state *st, int64_t count, const int64_t *rows, const int64_t *cols);
int step_impl(lmh_context *ctx, lmh_state *st);
bool step_may_stay(const lmh_context *ctx, const lmh_state *st);
int step_async(lmh_context *ctx, lmh_state *st);
int ensure_built(lmh_context *ctx);
int upload_mask(const uint8_t *mask, int64_t n, uint8_t **dev);
int eigensolve(lmh_context *ctx, double2 *A, int64_t lda, std::vector<double> &E);
int project_weights(lmh_context *ctx, double2 *W, int64_t ldw, const double *f, double2 *P, int64_t ldp,
                    int64_t *ncols);
void batch_free(lmh_batch *b);
hipError_t batch_launch_step(lmh_batch *b, int64_t p);
int batch_expected_cur(const lmh_batch *b, size_t r);
std::vector<lmh::BatchItem> batch_items(const lmh_batch *b, const lmh_batch::Class &c, int p, int side);
hipError_t batch_strip_tables(lmh_batch *b);
int batch_refresh(lmh_batch *b);
void krylov_graphs_free(lmh_kets *k);
// Device-to-host copies on the context's stream, then one wait for the
// stream: arrays of up to 64 MB go through a pinned staging buffer and a host
// copy (a copy into pageable memory is slower and, when small, waits on its
// own), larger ones straight to their destination.  Null entries are skipped.
struct Readback {
    void *dst;
    const void *src;
    size_t bytes;
};
hipError_t readback_sync(lmh_context *ctx, const Readback *r, int count);

// Records an event pair around one launch when profiling is on.
struct ProfScope {
    lmh_context *ctx;
    int kind;
    hipEvent_t a = nullptr;
    ProfScope(lmh_context *c, int k) : ctx(c), kind(k) {
        if (ctx->profiling) {
            a = take_event(ctx);
            if (a) (void)hipEventRecord(a, ctx->stream);
        }
    }
    ~ProfScope() {
        if (a) {
            hipEvent_t b = take_event(ctx);
            if (b) {
                (void)hipEventRecord(b, ctx->stream);
                ctx->prof_pending[kind].emplace_back(a, b);
            } else {
                ctx->event_pool.push_back(a);
            }
        }
    }
};

}  // namespace lmhapi
