// C ABI of liblmhip.so: contexts (the EvolutionSolver), geometry/field/on-site
// parameters, states (density matrices with their caches) and the step.
// The declarations and the reference interfaces they replace are in
// include/lmhip.h.
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

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hipfail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? LMH_ENOMEM : LMH_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                    \
    do {                                       \
        hipError_t e_ = (expr);                \
        if (e_ != hipSuccess) return hipfail(e_, what); \
    } while (0)

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

bool isapprox(double a, double b) {
    // Julia's default `≈` (rtol = sqrt(eps), atol = 0)
    return a == b || std::fabs(a - b) <= std::sqrt(2.220446049250313e-16) * std::max(std::fabs(a), std::fabs(b));
}

}  // namespace

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
    // host sources of the coefficient uploads (lmh_krylov_step synchronises
    // the stream after each upload)
    std::vector<std::complex<double>> h_coef;
    std::vector<int> h_len;
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

namespace {

void pool_free(lmh_context *ctx) {
    for (double2 *&p : ctx->pool) dfree(p);
    ctx->pool.clear();
}

// an N x ld buffer from the context's pool (allocated when the pool is empty)
hipError_t pool_take(lmh_context *ctx, int64_t elems, double2 **out) {
    if (elems != ctx->pool_elems) {
        pool_free(ctx);
        ctx->pool_elems = elems;
    }
    if (!ctx->pool.empty()) {
        *out = ctx->pool.back();
        ctx->pool.pop_back();
        return hipSuccess;
    }
    return dalloc(out, (size_t)elems);
}

void pool_give(lmh_context *ctx, double2 *p) {
    if (p) ctx->pool.push_back(p);
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

void tape_clear(lmh_context *ctx) {
    // device checkpoints go back to the pool (same size as the states' buffers)
    auto &tp = ctx->tape;
    for (size_t c = 0; c < tp.checkpoints.size(); ++c) {
        if (tp.cp_host[c])
            (void)hipHostFree(tp.checkpoints[c]);
        else
            pool_give(ctx, tp.checkpoints[c]);
    }
    ctx->tape = lmh_context::Tape();
}

// A new checkpoint of `elems` complex entries: in device memory while what
// the backward sweep will need beside the checkpoints still fits (the four
// pooled work buffers, the costate's three buffers and K - 1 Horner
// iterates, with 8 GiB to spare; pooled buffers count as free), else in
// pinned host memory -- or as LMH_OPT_CHECKPOINTS forces.
hipError_t checkpoint_alloc(lmh_context *ctx, int64_t elems, double2 **out, bool *host) {
    const size_t bytes = sizeof(double2) * (size_t)elems;
    bool want_host = ctx->checkpoint_place == 2;
    if (ctx->checkpoint_place == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
            const size_t pooled = elems == ctx->pool_elems ? ctx->pool.size() * bytes : 0;
            const size_t reserve = bytes * (size_t)(7 + std::max(ctx->K - 1, 0)) + ((size_t)8 << 30);
            want_host = free_b + pooled < bytes + reserve;
        }
    }
    if (want_host) {
        void *p = nullptr;
        hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e == hipSuccess) {
            *out = static_cast<double2 *>(p);
            *host = true;
            return hipSuccess;
        }
        (void)hipGetLastError();
        if (ctx->checkpoint_place == 2) return e;
    }
    *host = false;
    return pool_take(ctx, elems, out);
}

// Called by lmh_step before it changes the recorded state: keeps P_k every
// `every` steps and the parameters of step k.
int tape_record(lmh_context *ctx, lmh_state *st) {
    auto &tp = ctx->tape;
    if (tp.geom_version != ctx->geom_version || tp.onsite_version != ctx->onsite_version)
        return fail(LMH_ESTATE, "geometry or on-site energies changed while a tape is recording");
    if (!tp.recs.empty()) {
        const auto &f0 = tp.recs.front().field;
        bool same = f0.size() == ctx->field_host.size();
        for (size_t c = 0; same && c < f0.size(); ++c)
            same = f0[c].kind == ctx->field_host[c].kind;
        if (!same) return fail(LMH_EINVAL, "field components changed kind or number while a tape is recording");
    }
    if ((int64_t)tp.recs.size() % tp.every == 0) {
        double2 *cp = nullptr;
        bool host = false;
        HIP_TRY(checkpoint_alloc(ctx, st->ld * st->n, &cp, &host), "checkpoint allocation");
        tp.checkpoints.push_back(cp);
        tp.cp_host.push_back(host);
        // stream-ordered: the copy ends before the next step overwrites the buffer
        if (host)
            HIP_TRY(hipMemcpyAsync(cp, st->buf[st->cur], sizeof(double2) * st->ld * st->n, hipMemcpyDeviceToHost,
                                   ctx->stream), "checkpoint copy to host");
        else
            HIP_TRY(lmh::launch_copy(st->buf[st->cur], cp, st->ld * st->n, ctx->stream), "checkpoint copy");
    }
    tp.recs.push_back({ctx->field_host, ctx->dt, ctx->K, ctx->s});
    return LMH_OK;
}

void free_dict(lmh::DevHam &h) {
    dfree(h.dpack);
    dfree(h.drec);
    dfree(h.dcls);
    dfree(h.drep);
    dfree(h.dval);
    dfree(h.doff);
    h.ndict = h.noff = h.dict_ok = 0;
}

// H's value classes for the fused passes (zmax <= 4, n <= 10240): every slot
// is evaluated for `dict_probes` sets of field amplitudes (the field's kinds
// and points kept; probe 0 has all amplitudes zero), and slots whose values
// agree bit for bit at every probe -- directly or conjugated -- share a class.
// Each H build re-checks the classes against the actual values (red[2]); a
// mismatch makes the fused passes read val instead, so the classes change
// speed, never results.  Column offsets (column - site) must take at most 16
// values and the classes at most 2048.
std::vector<double> dict_signature(const lmh_context *ctx) {
    std::vector<double> sig;
    for (const auto &f : ctx->field_host) {
        sig.push_back((double)f.kind);
        sig.push_back(f.px);
        sig.push_back(f.py);
    }
    return sig;
}

hipError_t build_dict(lmh_context *ctx) {
    lmh::DevHam &h = ctx->h;
    free_dict(h);
    ctx->dict_geom = ctx->geom_version;
    ctx->dict_probes_built = ctx->dict_probes;
    ctx->dict_sig = dict_signature(ctx);
    ctx->built_version = 0;                  // the next update rebuilds H with the classes
    const int64_t n = h.n, z = h.zmax, nslots = z * n;
    if (ctx->dict_probes <= 0 || !lmh::fused_supported(h) || z == 0 || z > 4) return hipSuccess;
    const int np = ctx->dict_probes;
    const int64_t nf = h.nfield;
    std::vector<lmh::FieldComp> pf((size_t)np * nf);
    for (int p = 0; p < np; ++p)
        for (int64_t c = 0; c < nf; ++c) {
            lmh::FieldComp f = ctx->field_host[c];
            f.a = p == 0 ? 0.0 : (p == 1 ? 0.3183098861837907 * (1.0 + 0.37 * c) : 0.1414213562373095 * (1.0 + 0.61 * c));
            pf[(size_t)p * nf + c] = f;
        }
    lmh::FieldComp *dpf = nullptr;
    double2 *dprobe = nullptr;
    std::vector<double2> probe((size_t)np * nslots);
    std::vector<int32_t> col(nslots);
    hipError_t e = hipSuccess;
    if (nf) e = dalloc(&dpf, pf.size());
    if (e == hipSuccess) e = dalloc(&dprobe, probe.size());
    if (e == hipSuccess && nf) e = hipMemcpyAsync(dpf, pf.data(), pf.size() * sizeof(lmh::FieldComp), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = lmh::launch_slot_probe(h, dpf, np, dprobe, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(probe.data(), dprobe, probe.size() * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(col.data(), h.col, 4 * nslots, hipMemcpyDeviceToHost, ctx->stream);
    const hipError_t es = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = es;
    dfree(dpf);
    dfree(dprobe);
    if (e != hipSuccess) return e;
    struct Key {
        uint64_t w[6];
        bool operator==(const Key &o) const { return std::memcmp(w, o.w, sizeof w) == 0; }
    };
    struct KeyHash {
        size_t operator()(const Key &k) const {
            uint64_t x = 0x9e3779b97f4a7c15ull;
            for (uint64_t v : k.w) x = (x ^ v) * 0x100000001b3ull;
            return (size_t)(x ^ (x >> 29));
        }
    };
    auto bits = [](double x) {
        x += 0.0;                              // -0.0 and 0.0 are one value
        uint64_t u;
        std::memcpy(&u, &x, 8);
        return u;
    };
    std::unordered_map<Key, int32_t, KeyHash> classes;
    std::vector<int32_t> rep{-1};            // class 0: the value 0 of empty slots
    std::vector<int32_t> cls(nslots, 0);
    for (int64_t q = 0; q < nslots; ++q) {
        if (col[q] < 0) continue;
        Key k{}, kc{};
        for (int p = 0; p < np && p < 3; ++p) {
            const double2 v = probe[(size_t)p * nslots + q];
            k.w[2 * p] = kc.w[2 * p] = bits(v.x);
            k.w[2 * p + 1] = bits(v.y);
            kc.w[2 * p + 1] = bits(-v.y);
        }
        auto it = classes.find(k);
        if (it != classes.end()) {
            cls[q] = it->second;
        } else if ((it = classes.find(kc)) != classes.end()) {
            cls[q] = it->second | (1 << 15);
        } else {
            if ((int)rep.size() >= lmh::kDictMaxClasses) return hipSuccess;   // too many: no classes
            classes.emplace(k, (int32_t)rep.size());
            cls[q] = (int32_t)rep.size();
            rep.push_back((int32_t)q);
        }
    }
    std::vector<int32_t> offs{0};
    std::vector<uint64_t> pack(n, 0);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t u = 0; u < z; ++u) {
            const int64_t q = u * n + i;
            if (col[q] < 0) continue;
            const int32_t delta = col[q] - (int32_t)i;
            int oid = (int)(std::find(offs.begin(), offs.end(), delta) - offs.begin());
            if (oid == (int)offs.size()) {
                if ((int)offs.size() >= lmh::kDictMaxOffsets) return hipSuccess;   // no classes
                offs.push_back(delta);
            }
            const uint64_t sl = (uint64_t)oid | (uint64_t)(cls[q] & 0x7ff) << 4 | (uint64_t)((cls[q] >> 15) & 1) << 15;
            pack[i] |= sl << (16 * u);
        }
    const int noff = (int)offs.size();
    offs.resize(lmh::kDictMaxOffsets, 0);
    if (e == hipSuccess) e = dalloc(&h.dpack, n);
    if (e == hipSuccess) e = dalloc(&h.drec, n);
    if (e == hipSuccess) e = dalloc(&h.dcls, nslots);
    if (e == hipSuccess) e = dalloc(&h.drep, rep.size());
    if (e == hipSuccess) e = dalloc(&h.dval, rep.size());
    if (e == hipSuccess) e = dalloc(&h.doff, offs.size());
    if (e == hipSuccess) e = hipMemcpy(h.dpack, pack.data(), 8 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h.dcls, cls.data(), 4 * nslots, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h.drep, rep.data(), 4 * rep.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h.doff, offs.data(), 4 * offs.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_dict(h);
        return e;
    }
    h.ndict = (int)rep.size();
    h.noff = noff;
    return hipSuccess;
}

void free_geometry(lmh::DevHam &h) {
    dfree(h.ri);
    dfree(h.rj);
    dfree(h.amp);
    dfree(h.twist);
    dfree(h.bslot);
    dfree(h.bi);
    dfree(h.bj);
    dfree(h.cptr);
    dfree(h.cbond);
    dfree(h.col);
    dfree(h.col4);
    dfree(h.val);
    dfree(h.diag);
    dfree(h.rec);
    dfree(h.onsite);
    dfree(h.gcols);
    dfree(h.ucols);
    dfree(h.uptr);
    dfree(h.umap);
    h.ngroups = h.umax = 0;
    h.group = 0;
    dfree(h.sslot);
    dfree(h.sdir);
    dfree(h.stval);
    dfree(h.sdiag);
    h.sg = lmh::StripGeom{};
    h.strip_ok = h.strip_xreal = 0;
    free_dict(h);
}

// Column groups of the right pass: consecutive entries of `order` in groups of
// G, the union of each group's columns with their ELL neighbours, and per
// member the LDS slots of itself and its neighbours.  G shrinks until every
// union fits the pass's LDS tile.
int build_right_groups(lmh::DevHam &h, const std::vector<int32_t> &col, const std::vector<int32_t> &order) {
    const int64_t n = h.n;
    const int zs = h.zmax + 1;
    for (int G = lmh::right_pass_group_size(); G >= 1; G /= 2) {
        const int64_t ng = (n + G - 1) / G;
        std::vector<int32_t> ucols;
        std::vector<int64_t> uptr(ng + 1, 0);
        std::vector<int16_t> umap((size_t)n * zs, -1);
        int64_t umax = 0;
        bool fits = true;
        std::vector<int32_t> u;
        for (int64_t g = 0; g < ng && fits; ++g) {
            const int64_t m0 = g * G, m1 = std::min<int64_t>(m0 + G, n);
            u.clear();
            for (int64_t m = m0; m < m1; ++m) {
                const int32_t j = order[m];
                u.push_back(j);
                for (int q = 0; q < h.zmax; ++q) {
                    const int32_t l = col[(int64_t)q * n + j];
                    if (l >= 0) u.push_back(l);
                }
            }
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
            if ((int64_t)u.size() > lmh::right_pass_max_union()) {
                fits = false;
                break;
            }
            umax = std::max<int64_t>(umax, (int64_t)u.size());
            auto slot = [&](int32_t c) { return (int16_t)(std::lower_bound(u.begin(), u.end(), c) - u.begin()); };
            for (int64_t m = m0; m < m1; ++m) {
                const int32_t j = order[m];
                umap[(size_t)m * zs] = slot(j);
                for (int q = 0; q < h.zmax; ++q) {
                    const int32_t l = col[(int64_t)q * n + j];
                    umap[(size_t)m * zs + 1 + q] = l >= 0 ? slot(l) : (int16_t)-1;
                }
            }
            uptr[g] = (int64_t)ucols.size();
            ucols.insert(ucols.end(), u.begin(), u.end());
        }
        if (!fits) continue;
        uptr[ng] = (int64_t)ucols.size();
        dfree(h.gcols);
        dfree(h.ucols);
        dfree(h.uptr);
        dfree(h.umap);
        hipError_t e = dalloc(&h.gcols, n);
        if (e == hipSuccess) e = dalloc(&h.ucols, ucols.size());
        if (e == hipSuccess) e = dalloc(&h.uptr, ng + 1);
        if (e == hipSuccess) e = dalloc(&h.umap, umap.size());
        if (e == hipSuccess) e = hipMemcpy(h.gcols, order.data(), 4 * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h.ucols, ucols.data(), 4 * ucols.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h.uptr, uptr.data(), 8 * (ng + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h.umap, umap.data(), 2 * umap.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = lmh::prepare_passes(nullptr);
        if (e != hipSuccess) return hipfail(e, "right-pass groups");
        h.ngroups = ng;
        h.umax = umax;
        h.group = G;
        return LMH_OK;
    }
    return fail(LMH_ERANGE, "right-pass column groups do not fit the LDS tile");
}

// Z-order (Morton) permutation of sites by their coordinates rounded to
// lattice units; ties keep index order.
std::vector<int32_t> zorder(const double *xy, int64_t n) {
    double xmin = xy[0], ymin = xy[1];
    for (int64_t i = 0; i < n; ++i) {
        xmin = std::min(xmin, xy[2 * i]);
        ymin = std::min(ymin, xy[2 * i + 1]);
    }
    auto spread = [](uint64_t v) {
        v &= 0xffffffffull;
        v = (v | (v << 16)) & 0x0000ffff0000ffffull;
        v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
        v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
        v = (v | (v << 2)) & 0x3333333333333333ull;
        v = (v | (v << 1)) & 0x5555555555555555ull;
        return v;
    };
    std::vector<std::pair<uint64_t, int32_t>> keys(n);
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t qx = (uint64_t)std::llround(xy[2 * i] - xmin);
        const uint64_t qy = (uint64_t)std::llround(xy[2 * i + 1] - ymin);
        keys[i] = {spread(qx) | (spread(qy) << 1), (int32_t)i};
    }
    std::sort(keys.begin(), keys.end());
    std::vector<int32_t> order(n);
    for (int64_t i = 0; i < n; ++i) order[i] = keys[i].second;
    return order;
}

int ctx_resolve(lmh_context *ctx);

// the context's device current; without_resolve: an order result still on
// the device stays there (lmh_step), else it is taken up first
int check_ctx(lmh_context *ctx, bool without_resolve = false) {
    if (!ctx) return fail(LMH_EINVAL, "null context");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hipfail(e, "hipSetDevice");
    return without_resolve ? LMH_OK : ctx_resolve(ctx);
}

hipEvent_t take_event(lmh_context *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

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

void resolve_profile(lmh_context *ctx) {
    for (int k = 0; k < LMH_NUM_KERNEL_KINDS; ++k) {
        for (auto &pr : ctx->prof_pending[k]) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
                ctx->prof_ms[k] += ms;
                ctx->prof_count[k] += 1;
            }
            ctx->event_pool.push_back(pr.first);
            ctx->event_pool.push_back(pr.second);
        }
        ctx->prof_pending[k].clear();
    }
}

}  // namespace

// the state's buffer back to P (one conjugate transpose into a free buffer)
// run a deferred pass: buf[cur] <- U buf[cur] (plain stores), cur advances;
// the stored form (conj) already counts the step
static int state_settle(lmh_state *st) {
    if (!st->pending) return LMH_OK;
    lmh_context *ctx = st->ctx;
    const int dst = (st->cur + 1) % 3;
    {
        ProfScope ps(ctx, LMH_KERNEL_FUSED);
        HIP_TRY(lmh::launch_strip_pending(ctx->h, st->pend, st->buf[st->cur], st->buf[dst], st->ld, st->n, false,
                                          ctx->stream),
                "deferred pass launch");
    }
    st->cur = dst;
    st->pending = false;
    return LMH_OK;
}

static int ctx_settle_all(lmh_context *ctx) {
    for (lmh_state *st : ctx->states) {
        const int rc = state_settle(st);
        if (rc) return rc;
    }
    return LMH_OK;
}

static int state_normalize(lmh_state *st) {
    const int rc = state_settle(st);
    if (rc) return rc;
    if (!st->conj) return LMH_OK;
    lmh_context *ctx = st->ctx;
    const int dst = (st->cur + 1) % 3;
    {
        ProfScope ps(ctx, LMH_KERNEL_TRANSPOSE);
        HIP_TRY(lmh::launch_conj_transpose(st->buf[st->cur], st->buf[dst], st->n, st->ld, ctx->stream),
                "transpose launch");
    }
    st->cur = dst;
    st->conj = 0;
    return LMH_OK;
}

extern "C" {

int lmh_version(void) { return 100; }

const char *lmh_last_error(void) { return g_err.c_str(); }

int lmh_device_count(int *count) {
    if (!count) return fail(LMH_EINVAL, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *count = 0;
        return hipfail(e, "hipGetDeviceCount");
    }
    *count = c;
    return LMH_OK;
}

int lmh_create(lmh_context **out, int device) {
    if (!out) return fail(LMH_EINVAL, "null output pointer");
    *out = nullptr;
    int nd = 0;
    HIP_TRY(hipGetDeviceCount(&nd), "hipGetDeviceCount");
    if (device < 0 || device >= nd) return fail(LMH_EINVAL, "device %d out of range (%d devices)", device, nd);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    lmh_context *ctx = new (std::nothrow) lmh_context();
    if (!ctx) return fail(LMH_ENOMEM, "host allocation failed");
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = dalloc(&ctx->d_maxbits, lmh::kMaxTaylor);
    if (e == hipSuccess) e = dalloc(&ctx->d_partmax, (size_t)lmh::taylor_order_scratch());
    if (e == hipSuccess) e = dalloc(&ctx->d_overflow, 1);
    if (e == hipSuccess) e = dalloc(&ctx->d_run, 1);
    if (e == hipSuccess) e = dalloc(&ctx->d_tout, 1);
    if (e == hipSuccess) e = dalloc(&ctx->h.red, 4);
    if (e == hipSuccess) e = lmh::prepare_region_currents_static();
    if (e != hipSuccess) {
        int rc = hipfail(e, "context allocation");
        lmh_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return LMH_OK;
}

namespace {

void csc_release(lmh_context *ctx) {
    dfree(ctx->d_kup);
    dfree(ctx->d_klo);
    dfree(ctx->d_kd);
    dfree(ctx->d_nz);
    dfree(ctx->d_cscflag);
    ctx->csc_mode = ctx->csc_values = false;
    ctx->csc_nnz = 0;
}

}  // namespace

int lmh_destroy(lmh_context *ctx) {
    if (!ctx) return LMH_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->order_async = false;                // the states go with the context
    ctx->journal.clear();
    for (lmh_state *st : std::vector<lmh_state *>(ctx->states.begin(), ctx->states.end())) lmh_state_destroy(st);
    for (lmh_kets *k : std::vector<lmh_kets *>(ctx->kets.begin(), ctx->kets.end())) lmh_kets_destroy(k);
    for (lmh_series *q : std::vector<lmh_series *>(ctx->series.begin(), ctx->series.end())) lmh_series_destroy(q);
    tape_clear(ctx);
    pool_free(ctx);
    csc_release(ctx);
    dfree(ctx->adj_partial);
    dfree(ctx->d_tape_fields);
    dfree(ctx->xfer);
    free_geometry(ctx->h);
    dfree(ctx->h.field);
    dfree(ctx->h.red);
    dfree(ctx->d_maxbits);
    dfree(ctx->d_partmax);
    dfree(ctx->d_overflow);
    dfree(ctx->d_run);
    dfree(ctx->d_tout);
    dfree(ctx->d_rho);
    dfree(ctx->d_J);
    dfree(ctx->d_hdiag);
    dfree(ctx->d_hij);
    dfree(ctx->eigvec);
    for (int k = 0; k < LMH_NUM_KERNEL_KINDS; ++k)
        for (auto &pr : ctx->prof_pending[k]) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->ev_order) (void)hipEventDestroy(ctx->ev_order);
    if (ctx->h_tout) (void)hipHostFree(ctx->h_tout);
    if (ctx->ev_field) (void)hipEventDestroy(ctx->ev_field);
    if (ctx->field_stage) (void)hipHostFree(ctx->field_stage);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LMH_OK;
}

int lmh_set_geometry(lmh_context *ctx, int64_t nsites, int64_t nbonds, const int64_t *bond_i,
                     const int64_t *bond_j, const double *r_i, const double *r_j, const double *amp,
                     const double *twist) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (nsites <= 0 || nsites >= (int64_t(1) << 31)) return fail(LMH_EINVAL, "invalid number of sites %lld", (long long)nsites);
    if (nbonds < 0) return fail(LMH_EINVAL, "negative number of bonds");
    if (nbonds > 0 && (!bond_i || !bond_j || !r_i || !r_j || !amp))
        return fail(LMH_EINVAL, "null bond array");
    for (lmh_state *st : ctx->states)
        if (st->n != nsites) return fail(LMH_ESTATE, "existing states have %lld sites", (long long)st->n);
    for (int64_t b = 0; b < nbonds; ++b) {
        if (bond_i[b] < 0 || bond_i[b] >= nsites || bond_j[b] < 0 || bond_j[b] >= nsites)
            return fail(LMH_EINVAL, "bond %lld has a site index out of range", (long long)b);
        if (bond_i[b] == bond_j[b])
            return fail(LMH_EINVAL, "bond %lld connects site %lld to itself", (long long)b, (long long)bond_i[b]);
    }
    // deferred passes run with the geometry they were made for
    rc = ctx_settle_all(ctx);
    if (rc) return rc;
    // ELL slots: distinct columns of each row, sorted; contributions in bond order
    const int64_t n = nsites;
    std::vector<std::vector<int64_t>> rowcols(n);
    for (int64_t b = 0; b < nbonds; ++b) {
        rowcols[bond_i[b]].push_back(bond_j[b]);
        rowcols[bond_j[b]].push_back(bond_i[b]);
    }
    int zmax = 0;
    for (auto &rc_ : rowcols) {
        std::sort(rc_.begin(), rc_.end());
        rc_.erase(std::unique(rc_.begin(), rc_.end()), rc_.end());
        zmax = std::max<int>(zmax, (int)rc_.size());
    }
    auto slot_of = [&](int64_t r, int64_t c) -> int64_t {
        const auto &v = rowcols[r];
        int64_t s = std::lower_bound(v.begin(), v.end(), c) - v.begin();
        return s * n + r;
    };
    const int64_t nslots = (int64_t)zmax * n;
    std::vector<int32_t> col(nslots, -1);
    for (int64_t r = 0; r < n; ++r)
        for (size_t s = 0; s < rowcols[r].size(); ++s) col[(int64_t)s * n + r] = (int32_t)rowcols[r][s];
    int64_t band = 0;
    for (int64_t q = 0; q < nslots; ++q)
        if (col[q] >= 0) band = std::max<int64_t>(band, std::llabs((int64_t)col[q] - q % n));
    std::vector<std::vector<int64_t>> contrib(nslots);
    std::vector<int64_t> bslot(nbonds);
    for (int64_t b = 0; b < nbonds; ++b) {
        const int64_t q = slot_of(bond_i[b], bond_j[b]);
        contrib[q].push_back(b);
        bslot[b] = q;
        contrib[slot_of(bond_j[b], bond_i[b])].push_back(~b);
    }
    std::vector<int64_t> cptr(nslots + 1, 0), cbond;
    cbond.reserve(2 * nbonds);
    for (int64_t q = 0; q < nslots; ++q) {
        cptr[q] = (int64_t)cbond.size();
        cbond.insert(cbond.end(), contrib[q].begin(), contrib[q].end());
    }
    cptr[nslots] = (int64_t)cbond.size();
    std::vector<int32_t> bi(nbonds), bj(nbonds);
    std::vector<double> tw(2 * nbonds);
    for (int64_t b = 0; b < nbonds; ++b) {
        bi[b] = (int32_t)bond_i[b];
        bj[b] = (int32_t)bond_j[b];
        tw[2 * b] = twist ? twist[2 * b] : 1.0;
        tw[2 * b + 1] = twist ? twist[2 * b + 1] : 0.0;
    }
    // (re)allocate
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    csc_release(ctx);
    std::vector<double> onsite_keep = ctx->onsite_host;
    tape_clear(ctx);
    pool_free(ctx);
    free_geometry(ctx->h);
    lmh::DevHam &h = ctx->h;
    h.n = n;
    h.nbonds = nbonds;
    h.zmax = zmax;
    h.band = band;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = dalloc(&h.ri, nbonds);
    if (e == hipSuccess) e = dalloc(&h.rj, nbonds);
    if (e == hipSuccess) e = dalloc(&h.amp, nbonds);
    if (e == hipSuccess) e = dalloc(&h.twist, nbonds);
    if (e == hipSuccess) e = dalloc(&h.bslot, nbonds);
    if (e == hipSuccess) e = dalloc(&h.bi, nbonds);
    if (e == hipSuccess) e = dalloc(&h.bj, nbonds);
    if (e == hipSuccess) e = dalloc(&h.cptr, nslots + 1);
    if (e == hipSuccess) e = dalloc(&h.cbond, cbond.size());
    const int64_t zq = (zmax + 3) / 4;
    std::vector<int32_t> col4(zq * n * 4, -1);
    for (int q = 0; q < zmax; ++q)
        for (int64_t i = 0; i < n; ++i) col4[((q / 4) * n + i) * 4 + q % 4] = col[q * n + i];
    if (e == hipSuccess) e = dalloc(&h.col, nslots);
    if (e == hipSuccess) e = dalloc(&h.col4, col4.size());
    // val padded to whole 4-slot chunks of col4; the padding stays zero
    if (e == hipSuccess) e = dalloc(&h.val, zq * 4 * n);
    if (e == hipSuccess) e = hipMemset(h.val, 0, sizeof(double2) * zq * 4 * n);
    if (e == hipSuccess) e = dalloc(&h.diag, n);
    if (e == hipSuccess && zmax <= 4 && n < 65536) e = dalloc(&h.rec, n);
    if (e != hipSuccess) {
        free_geometry(ctx->h);
        ctx->geom_set = false;
        return hipfail(e, "geometry allocation");
    }
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (bytes && e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    };
    up(h.ri, r_i, 16 * nbonds);
    up(h.rj, r_j, 16 * nbonds);
    up(h.amp, amp, 16 * nbonds);
    up(h.twist, tw.data(), 16 * nbonds);
    up(h.bslot, bslot.data(), 8 * nbonds);
    up(h.bi, bi.data(), 4 * nbonds);
    up(h.bj, bj.data(), 4 * nbonds);
    up(h.cptr, cptr.data(), 8 * (nslots + 1));
    up(h.cbond, cbond.data(), 8 * cbond.size());
    up(h.col, col.data(), 4 * nslots);
    up(h.col4, col4.data(), 4 * col4.size());
    if (e == hipSuccess) e = lmh::prepare_passes(nullptr);
    // grid-stencil pattern: the strip kernel's geometry and direction slots
    if (e == hipSuccess && n <= 10240) {
        std::vector<int32_t> sslot(4 * n);
        lmh::StripGeom sg;
        if (lmh::strip_detect(col.data(), n, zmax, &sg, sslot.data())) {
            // per site: the direction of each ELL slot (2 bits) and which
            // slots hold a bond (bits 8 + slot), for the grid-window order search
            std::vector<uint32_t> sdir(n, 0);
            for (int d = 0; d < 4; ++d)
                for (int64_t i = 0; i < n; ++i) {
                    const int32_t q = sslot[d * n + i];
                    if (q < 0) continue;
                    const int u = (int)(q / n);
                    sdir[i] |= (uint32_t)d << (2 * u) | 1u << (8 + u);
                }
            e = dalloc(&h.sdir, n);
            if (e == hipSuccess) e = hipMemcpy(h.sdir, sdir.data(), 4 * n, hipMemcpyHostToDevice);
            if (e == hipSuccess) e = dalloc(&h.sslot, sslot.size());
            if (e == hipSuccess) e = dalloc(&h.stval, (size_t)4 * sg.nt);
            if (e == hipSuccess) e = dalloc(&h.sdiag, (size_t)sg.s * sg.nt);
            if (e == hipSuccess) e = hipMemcpy(h.sslot, sslot.data(), 4 * sslot.size(), hipMemcpyHostToDevice);
            if (e == hipSuccess) {
                h.sg = sg;
                e = lmh::prepare_strip(h);
            }
        }
    }
    if (e != hipSuccess) {
        free_geometry(ctx->h);
        ctx->geom_set = false;
        return hipfail(e, "geometry upload");
    }
    ctx->geom_set = true;
    ctx->param_version++;
    ctx->geom_version++;
    // schedule options pinned (or autotuned) for the previous geometry that do
    // not fit this one go back to their by-size defaults: they only choose
    // kernels, never results
    if (ctx->fused_variant >= 0 && !lmh::fused_variant_valid(ctx->fused_variant, h.n, h.zmax)) ctx->fused_variant = -1;
    if ((ctx->left_variant == 2 && !lmh::left_band_fits(h)) || (ctx->left_variant == 3 && !lmh::left_band2_fits(h)))
        ctx->left_variant = 0;
    ctx->solver_ready = false;
    dfree(ctx->d_rho);
    dfree(ctx->d_J);
    dfree(ctx->d_hdiag);
    dfree(ctx->d_hij);
    if (ctx->onsite_set && (int64_t)onsite_keep.size() == n) {
        HIP_TRY(dalloc(&h.onsite, n), "on-site allocation");
        HIP_TRY(hipMemcpy(h.onsite, onsite_keep.data(), 8 * n, hipMemcpyHostToDevice), "on-site upload");
    } else {
        ctx->onsite_set = false;
        ctx->onsite_host.clear();
    }
    ctx->col_host = col;
    std::vector<int32_t> order(n);
    for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
    return build_right_groups(h, ctx->col_host, order);
}

int lmh_set_site_coords(lmh_context *ctx, const double *xy) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "lmh_set_site_coords before lmh_set_geometry");
    if (!xy) return fail(LMH_EINVAL, "null coordinates");
    for (int64_t i = 0; i < 2 * ctx->h.n; ++i)
        if (!std::isfinite(xy[i])) return fail(LMH_EINVAL, "coordinate %lld is not finite", (long long)i);
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return build_right_groups(ctx->h, ctx->col_host, zorder(xy, ctx->h.n));
}

int lmh_set_onsite(lmh_context *ctx, const double *diag) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "lmh_set_onsite before lmh_set_geometry");
    if (ctx->csc_mode) return fail(LMH_ESTATE, "H is given as CSC values: its diagonal comes with lmh_set_csc_values");
    const int64_t n = ctx->h.n;
    if (!diag) {
        if (ctx->onsite_set) {
            HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
            dfree(ctx->h.onsite);
            ctx->onsite_set = false;
            ctx->onsite_host.clear();
            ctx->param_version++;
            ctx->onsite_version++;
        }
        return LMH_OK;
    }
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(diag[i])) return fail(LMH_EINVAL, "on-site energy %lld is not finite", (long long)i);
    if (ctx->onsite_set && std::memcmp(ctx->onsite_host.data(), diag, 8 * n) == 0) return LMH_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (!ctx->h.onsite) HIP_TRY(dalloc(&ctx->h.onsite, n), "on-site allocation");
    ctx->onsite_host.assign(diag, diag + n);
    HIP_TRY(hipMemcpy(ctx->h.onsite, diag, 8 * n, hipMemcpyHostToDevice), "on-site upload");
    ctx->onsite_set = true;
    ctx->param_version++;
    ctx->onsite_version++;
    return LMH_OK;
}

int lmh_set_field(lmh_context *ctx, int64_t ncomp, const int32_t *kind, const double *params) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ncomp < 0) return fail(LMH_EINVAL, "negative number of field components");
    if (ctx->csc_mode && ncomp > 0)
        return fail(LMH_ESTATE, "H is given as CSC values: the field's phases are part of them");
    if (ncomp > 0 && (!kind || !params)) return fail(LMH_EINVAL, "null field arrays");
    std::vector<lmh::FieldComp> f(ncomp);
    for (int64_t c = 0; c < ncomp; ++c) {
        if (kind[c] < LMH_FIELD_LANDAU || kind[c] > LMH_FIELD_POINT_SINGULAR)
            return fail(LMH_EINVAL, "invalid field kind %d", kind[c]);
        for (int q = 0; q < 3; ++q)
            if (!std::isfinite(params[3 * c + q])) return fail(LMH_EINVAL, "field parameter not finite");
        f[c].kind = kind[c];
        f[c].pad = 0;
        f[c].a = params[3 * c];
        f[c].px = params[3 * c + 1];
        f[c].py = params[3 * c + 2];
    }
    if (f.size() == ctx->field_host.size() &&
        (f.empty() || std::memcmp(f.data(), ctx->field_host.data(), f.size() * sizeof(lmh::FieldComp)) == 0))
        return LMH_OK;
    if (ncomp > ctx->h.field_cap) {
        HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        dfree(ctx->h.field);
        HIP_TRY(dalloc(&ctx->h.field, ncomp), "field allocation");
        if (ctx->field_stage) (void)hipHostFree(ctx->field_stage);
        ctx->field_stage = nullptr;
        HIP_TRY(hipHostMalloc((void **)&ctx->field_stage, ncomp * sizeof(lmh::FieldComp), 0), "field staging");
        ctx->h.field_cap = ncomp;
    }
    if (ncomp) {
        // in stream order (the launches already queued read the earlier
        // values), through pinned memory: no wait for the GPU here, beyond
        // the previous upload having left the staging buffer
        if (!ctx->ev_field) HIP_TRY(hipEventCreateWithFlags(&ctx->ev_field, hipEventDisableTiming), "hipEventCreate");
        else HIP_TRY(hipEventSynchronize(ctx->ev_field), "field staging");
        std::memcpy(ctx->field_stage, f.data(), ncomp * sizeof(lmh::FieldComp));
        HIP_TRY(hipMemcpyAsync(ctx->h.field, ctx->field_stage, ncomp * sizeof(lmh::FieldComp), hipMemcpyHostToDevice,
                               ctx->stream),
                "field upload");
        HIP_TRY(hipEventRecord(ctx->ev_field, ctx->stream), "hipEventRecord");
    }
    ctx->h.nfield = ncomp;
    ctx->field_host = f;
    ctx->param_version++;
    return LMH_OK;
}

int lmh_set_csc_pattern(lmh_context *ctx, int64_t n, const int64_t *colptr, const int64_t *rowval, int32_t base) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (base != 0 && base != 1) return fail(LMH_EINVAL, "index base must be 0 or 1");
    if (n <= 0 || n >= (int64_t(1) << 31)) return fail(LMH_EINVAL, "invalid matrix size %lld", (long long)n);
    if (!colptr) return fail(LMH_EINVAL, "null colptr");
    if (colptr[0] != base) return fail(LMH_EINVAL, "colptr[0] = %lld, expected the index base %d", (long long)colptr[0], base);
    for (int64_t c = 0; c < n; ++c)
        if (colptr[c + 1] < colptr[c]) return fail(LMH_EINVAL, "colptr decreases at column %lld", (long long)c);
    const int64_t nnz = colptr[n] - base;
    if (nnz > 0 && !rowval) return fail(LMH_EINVAL, "null rowval");
    for (int64_t c = 0; c < n; ++c)
        for (int64_t k = colptr[c] - base; k < colptr[c + 1] - base; ++k) {
            const int64_t r = rowval[k] - base;
            if (r < 0 || r >= n) return fail(LMH_EINVAL, "row index %lld out of range in column %lld", (long long)rowval[k], (long long)c);
            if (k > colptr[c] - base && r <= rowval[k - 1] - base)
                return fail(LMH_EINVAL, "row indices of column %lld are not strictly increasing", (long long)c);
        }
    // position of entry (r, c) in the value array, -1 when not stored
    auto find = [&](int64_t r, int64_t c) -> int64_t {
        const int64_t *lo = rowval + (colptr[c] - base), *hi = rowval + (colptr[c + 1] - base);
        const int64_t *p = std::lower_bound(lo, hi, r + base);
        return (p != hi && *p == r + base) ? (int64_t)(p - rowval) : -1;
    };
    std::vector<int64_t> bi, bj, kup, klo, kd(n, -1);
    for (int64_t c = 0; c < n; ++c)
        for (int64_t k = colptr[c] - base; k < colptr[c + 1] - base; ++k) {
            const int64_t r = rowval[k] - base;
            if (r == c) {
                kd[c] = k;
                continue;
            }
            const int64_t kt = find(c, r);
            if (kt < 0)
                return fail(LMH_EINVAL, "entry (%lld, %lld) is stored but (%lld, %lld) is not: a Hermitian H has a "
                                        "symmetric pattern", (long long)r, (long long)c, (long long)c, (long long)r);
            if (r < c) {
                bi.push_back(r);
                bj.push_back(c);
                kup.push_back(k);
                klo.push_back(kt);
            }
        }
    const int64_t nb = (int64_t)bi.size();
    std::vector<double> zero(2 * std::max<int64_t>(nb, 1), 0.0);
    rc = lmh_set_geometry(ctx, n, nb, bi.data(), bj.data(), zero.data(), zero.data(), zero.data(), nullptr);
    if (rc) return rc;
    // no field and no parametric on-site terms: the values carry both
    ctx->field_host.clear();
    ctx->h.nfield = 0;
    dfree(ctx->h.onsite);
    ctx->onsite_host.clear();
    hipError_t e = dalloc(&ctx->h.onsite, n);
    if (e == hipSuccess) e = dalloc(&ctx->d_kup, std::max<int64_t>(nb, 1));
    if (e == hipSuccess) e = dalloc(&ctx->d_klo, std::max<int64_t>(nb, 1));
    if (e == hipSuccess) e = dalloc(&ctx->d_kd, n);
    if (e == hipSuccess) e = dalloc(&ctx->d_nz, std::max<int64_t>(nnz, 1));
    if (e == hipSuccess) e = dalloc(&ctx->d_cscflag, 1);
    if (e == hipSuccess && nb) e = hipMemcpy(ctx->d_kup, kup.data(), 8 * nb, hipMemcpyHostToDevice);
    if (e == hipSuccess && nb) e = hipMemcpy(ctx->d_klo, klo.data(), 8 * nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ctx->d_kd, kd.data(), 8 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(ctx->h.onsite, 0, 8 * n);
    if (e != hipSuccess) {
        csc_release(ctx);
        ctx->onsite_set = false;
        return hipfail(e, "CSC pattern upload");
    }
    ctx->onsite_set = true;
    ctx->csc_mode = true;
    ctx->csc_values = false;
    ctx->csc_nnz = nnz;
    ctx->param_version++;
    ctx->onsite_version++;
    return LMH_OK;
}

int lmh_set_csc_values(lmh_context *ctx, int64_t nnz, const double *nzval) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->csc_mode) return fail(LMH_ESTATE, "lmh_set_csc_values before lmh_set_csc_pattern");
    if (nnz != ctx->csc_nnz)
        return fail(LMH_EINVAL, "%lld values for a pattern of %lld entries", (long long)nnz, (long long)ctx->csc_nnz);
    if (nnz > 0 && !nzval) return fail(LMH_EINVAL, "null values");
    lmh::DevHam &h = ctx->h;
    int flag = 0;
    hipError_t e = nnz ? hipMemcpyAsync(ctx->d_nz, nzval, 16 * nnz, hipMemcpyHostToDevice, ctx->stream) : hipSuccess;
    if (e == hipSuccess)
        e = lmh::launch_csc_gather(h.nbonds, h.n, ctx->d_kup, ctx->d_klo, ctx->d_kd, ctx->d_nz, h.amp, h.onsite,
                                   ctx->d_cscflag, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, ctx->d_cscflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    ctx->param_version++;
    ctx->onsite_version++;
    ctx->csc_values = false;
    if (e != hipSuccess) return hipfail(e, "CSC value upload");
    if (flag & 4) return fail(LMH_EINVAL, "H has a value that is not finite");
    if (flag & 1) return fail(LMH_EINVAL, "H is not Hermitian: a stored H[j,i] is not the conjugate of H[i,j]");
    if (flag & 2) return fail(LMH_EINVAL, "H is not Hermitian: a diagonal entry has a nonzero imaginary part");
    ctx->csc_values = true;
    return LMH_OK;
}

namespace {

// Takes up an order result: the larger neighbourhood tables when a column's
// did not fit (one level at a time, each read back), then the value-class
// and strip checks that travel with it, K and s.  An error leaves the solver
// not ready (the next lmh_update_solver recomputes).
int order_finish(lmh_context *ctx, lmh::TaylorOut &out, int level0) {
    lmh::DevHam &h = ctx->h;
    int level = level0;
    while (out.overflow && level + 1 < lmh::taylor_order_levels()) {
        ++level;
        {
            ProfScope ps(ctx, LMH_KERNEL_TAYLOR_ORDER);
            HIP_TRY(lmh::launch_taylor_order(h, ctx->dt, ctx->nztol, ctx->d_run, level, ctx->d_partmax,
                                             ctx->d_maxbits, ctx->d_overflow, ctx->stream),
                    "taylor_order launch");
            HIP_TRY(lmh::launch_taylor_finish(h.red, ctx->d_run, ctx->d_maxbits, ctx->d_overflow, ctx->d_tout,
                                              ctx->stream),
                    "taylor finish launch");
        }
        HIP_TRY(hipMemcpyAsync(&out, ctx->d_tout, sizeof out, hipMemcpyDeviceToHost, ctx->stream), "order readback");
        HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    }
    if (!out.overflow) {
        ctx->order_level = level;
        ctx->order_level_geom = ctx->geom_version + 1;
    }
    const unsigned long long *red = out.red;
    h.dict_ok = h.ndict > 0 && red[2] == 0;
    h.strip_ok = h.sg.s > 0 && (red[3] & 1) == 0;
    h.strip_xreal = h.strip_ok && (red[3] & 2) == 0;
    ctx->flags_geom = ctx->geom_version + 1;
    if (out.overflow) {
        ctx->solver_ready = false;
        return fail(LMH_ERANGE, "Taylor-order search: neighbourhood of %d terms exceeds the kernel capacity",
                    out.nmax);
    }
    if (out.K < 0) {
        ctx->solver_ready = false;
        return fail(LMH_ERANGE, "Taylor series did not reach the threshold within %d terms", out.nmax);
    }
    ctx->K = out.K;
    ctx->s = out.s;
    return LMH_OK;
}

// the step that follows runs on the strip kernel with a deferred second pass
// (the path that can take its polynomial from the device), by the flags of
// the last result taken up for this geometry
bool strip_step_path(const lmh_context *ctx) {
    const lmh::DevHam &h = ctx->h;
    int mode = ctx->step_mode;
    if (mode == LMH_STEP_AUTO)
        mode = (lmh::fused_supported(h) && h.n <= ctx->fused_max_n) ? LMH_STEP_FUSED_T : LMH_STEP_PER_PASS;
    return mode == LMH_STEP_FUSED_T && lmh::fused_supported(h) && ctx->defer && h.strip_ok &&
           lmh::fused_resolved_variant(h, ctx->fused_variant) == lmh::fused_strip_variant();
}

bool order_may_stay(const lmh_context *ctx) {
    return ctx->async_order && ctx->flags_geom == ctx->geom_version + 1 && !ctx->tape.active && strip_step_path(ctx);
}

}  // namespace

int lmh_update_solver(lmh_context *ctx, double dt, int force, double threshold, double nztol, int32_t *K_out,
                      double *s_out) {
    // (an earlier update's result still on the device is taken up here: by
    // now the launches in front of it have normally finished)
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "lmh_update_solver before lmh_set_geometry");
    if (ctx->csc_mode && !ctx->csc_values)
        return fail(LMH_ESTATE, "lmh_update_solver without valid lmh_set_csc_values for the CSC pattern");
    if (!std::isfinite(dt) || !(threshold > 0) || !(nztol >= 0))
        return fail(LMH_EINVAL, "invalid dt / threshold / nztol");
    const bool same = ctx->solver_ready && !force && ctx->solver_version == ctx->param_version &&
                      isapprox(-dt, -ctx->dt) && threshold == ctx->threshold && nztol == ctx->nztol;
    if (!same) {
        lmh::DevHam &h = ctx->h;
        if (ctx->dict_geom != ctx->geom_version || ctx->dict_probes_built != ctx->dict_probes ||
            ctx->dict_sig != dict_signature(ctx))
            HIP_TRY(build_dict(ctx), "value classes");
        if (ctx->built_version != ctx->param_version) {
            ProfScope ps(ctx, LMH_KERNEL_BUILD_H);
            HIP_TRY(lmh::launch_build_h(h, ctx->stream), "build_h launch");
            ctx->built_version = ctx->param_version;
        }
        // the order rule on the device: s and the order bound from H's
        // reductions, the per-column term maxima, K and the Horner
        // coefficients
        const int level0 = ctx->order_level_geom == ctx->geom_version + 1 ? ctx->order_level : 0;
        {
            ProfScope ps(ctx, LMH_KERNEL_TAYLOR_ORDER);
            HIP_TRY(lmh::launch_taylor_setup(h.red, dt, threshold, ctx->d_run, ctx->d_overflow, ctx->stream),
                    "taylor setup launch");
            HIP_TRY(lmh::launch_taylor_order(h, dt, nztol, ctx->d_run, level0, ctx->d_partmax, ctx->d_maxbits,
                                             ctx->d_overflow, ctx->stream),
                    "taylor_order launch");
            HIP_TRY(lmh::launch_taylor_finish(h.red, ctx->d_run, ctx->d_maxbits, ctx->d_overflow, ctx->d_tout,
                                              ctx->stream),
                    "taylor finish launch");
        }
        ctx->dt = dt;
        ctx->threshold = threshold;
        ctx->nztol = nztol;
        if (!K_out && !s_out && order_may_stay(ctx)) {
            // K stays on the device: the strip steps read it from there; the
            // result comes back behind an event and is taken up later
            if (!ctx->h_tout) HIP_TRY(hipHostMalloc((void **)&ctx->h_tout, sizeof(lmh::TaylorOut), 0), "pinned result");
            if (!ctx->ev_order)
                HIP_TRY(hipEventCreateWithFlags(&ctx->ev_order, hipEventDisableTiming), "hipEventCreate");
            HIP_TRY(hipMemcpyAsync(ctx->h_tout, ctx->d_tout, sizeof(lmh::TaylorOut), hipMemcpyDeviceToHost,
                                   ctx->stream),
                    "order readback");
            HIP_TRY(hipEventRecord(ctx->ev_order, ctx->stream), "hipEventRecord");
            ctx->order_async = true;
            ctx->async_xreal = h.strip_xreal ? 1 : 0;
            ctx->async_level = level0;
            ctx->solver_version = ctx->param_version;
            ctx->solver_serial++;
            ctx->solver_ready = true;
            return LMH_OK;
        }
        // one readback of the result (with the reductions and the value-class
        // / strip checks)
        lmh::TaylorOut out;
        HIP_TRY(hipMemcpyAsync(&out, ctx->d_tout, sizeof out, hipMemcpyDeviceToHost, ctx->stream), "order readback");
        HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        rc = order_finish(ctx, out, level0);
        if (rc) return rc;
        ctx->solver_version = ctx->param_version;
        ctx->solver_serial++;
        ctx->solver_ready = true;
    }
    if (K_out) *K_out = ctx->K;
    if (s_out) *s_out = ctx->s;
    return LMH_OK;
}

int lmh_solver_order(lmh_context *ctx, int32_t *K, double *s) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->solver_ready) return fail(LMH_ESTATE, "lmh_solver_order before lmh_update_solver");
    if (K) *K = ctx->K;
    if (s) *s = ctx->s;
    return LMH_OK;
}

int lmh_state_create(lmh_context *ctx, lmh_state **out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out) return fail(LMH_EINVAL, "null output pointer");
    *out = nullptr;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "lmh_state_create before lmh_set_geometry");
    lmh_state *st = new (std::nothrow) lmh_state();
    if (!st) return fail(LMH_ENOMEM, "host allocation failed");
    st->ctx = ctx;
    st->n = ctx->h.n;
    st->ld = ((ctx->h.n + 7) / 8) * 8;
    for (int b = 0; b < 3; ++b) {
        hipError_t e = dalloc(&st->buf[b], (size_t)st->ld * st->n);
        if (e != hipSuccess) {
            for (int q = 0; q < 3; ++q) dfree(st->buf[q]);
            delete st;
            return hipfail(e, "state allocation");
        }
    }
    ctx->states.insert(st);
    *out = st;
    return LMH_OK;
}

int lmh_state_destroy(lmh_state *st) {
    if (!st) return LMH_OK;
    if (st->ctx) {
        (void)hipSetDevice(st->ctx->device);
        (void)ctx_resolve(st->ctx);          // the journal may name this state
        (void)hipStreamSynchronize(st->ctx->stream);
        st->ctx->states.erase(st);
        if (st->ctx->tape.st == st) tape_clear(st->ctx);
    }
    for (int q = 0; q < 3; ++q) dfree(st->buf[q]);
    dfree(st->pend_stval);
    dfree(st->pend_sdiag);
    dfree(st->pend_coef);
    delete st;
    return LMH_OK;
}

int lmh_state_upload(lmh_state *st, const double *P, int64_t ld) {
    if (!st) return fail(LMH_EINVAL, "null state");
    int rc = check_ctx(st->ctx);
    if (rc) return rc;
    if (!P || ld < st->n) return fail(LMH_EINVAL, "invalid matrix or leading dimension");
    HIP_TRY(hipMemcpy2DAsync(st->buf[st->cur], st->ld * 16, P, ld * 16, st->n * 16, st->n, hipMemcpyHostToDevice,
                             st->ctx->stream),
            "state upload");
    st->conj = 0;
    st->pending = false;
    HIP_TRY(hipStreamSynchronize(st->ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

int lmh_state_download(lmh_state *st, double *P, int64_t ld) {
    if (!st) return fail(LMH_EINVAL, "null state");
    int rc = check_ctx(st->ctx);
    if (rc) return rc;
    if (!P || ld < st->n) return fail(LMH_EINVAL, "invalid matrix or leading dimension");
    rc = state_normalize(st);
    if (rc) return rc;
    HIP_TRY(hipMemcpy2DAsync(P, ld * 16, st->buf[st->cur], st->ld * 16, st->n * 16, st->n, hipMemcpyDeviceToHost,
                             st->ctx->stream),
            "state download");
    HIP_TRY(hipStreamSynchronize(st->ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

int lmh_state_copy(lmh_state *dst, const lmh_state *src) {
    if (!dst || !src) return fail(LMH_EINVAL, "null state");
    if (dst->ctx != src->ctx || dst->n != src->n) return fail(LMH_EINVAL, "states of different contexts or sizes");
    int rc = check_ctx(dst->ctx);
    if (rc) return rc;
    // the source's deferred pass runs first (its value is P, not the pending form)
    rc = state_settle(const_cast<lmh_state *>(src));
    if (rc) return rc;
    HIP_TRY(lmh::launch_copy(src->buf[src->cur], dst->buf[dst->cur], src->ld * src->n, dst->ctx->stream), "state copy");
    dst->conj = src->conj;
    dst->pending = false;
    return LMH_OK;
}

namespace {

int check_entries(const lmh_state *st, int64_t count, const int64_t *rows, const int64_t *cols) {
    if (!st) return fail(LMH_EINVAL, "null state");
    int rc = check_ctx(st->ctx);
    if (rc) return rc;
    if (count < 0 || (count > 0 && (!rows || !cols))) return fail(LMH_EINVAL, "invalid entry list");
    for (int64_t k = 0; k < count; ++k)
        if (rows[k] < 0 || rows[k] >= st->n || cols[k] < 0 || cols[k] >= st->n)
            return fail(LMH_EINVAL, "entry %lld (%lld, %lld) out of range", (long long)k, (long long)rows[k],
                        (long long)cols[k]);
    return LMH_OK;
}

}  // namespace

int lmh_state_set_entries(lmh_state *st, int64_t count, const int64_t *rows, const int64_t *cols, const double *vals) {
    int rc = check_entries(st, count, rows, cols);
    if (rc) return rc;
    if (count > 0 && !vals) return fail(LMH_EINVAL, "null values");
    lmh_context *ctx = st->ctx;
    // sum repeated (row, col) pairs so the scatter writes each entry once
    std::vector<int64_t> order(count);
    for (int64_t k = 0; k < count; ++k) order[k] = k;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return cols[a] != cols[b] ? cols[a] < cols[b] : (rows[a] != rows[b] ? rows[a] < rows[b] : a < b);
    });
    std::vector<int64_t> r, c;
    std::vector<double> v;
    for (int64_t t = 0; t < count; ++t) {
        const int64_t k = order[t];
        if (!r.empty() && r.back() == rows[k] && c.back() == cols[k]) {
            v[v.size() - 2] += vals[2 * k];
            v[v.size() - 1] += vals[2 * k + 1];
        } else {
            r.push_back(rows[k]);
            c.push_back(cols[k]);
            v.push_back(vals[2 * k]);
            v.push_back(vals[2 * k + 1]);
        }
    }
    const int64_t m = (int64_t)r.size();
    const size_t bytes = (size_t)m * 32;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    HIP_TRY(grow(ctx->xfer, ctx->xfer_cap, std::max<size_t>(bytes, 1)), "transfer buffer");
    int64_t *dr = reinterpret_cast<int64_t *>(ctx->xfer);
    int64_t *dc = dr + m;
    double2 *dv = reinterpret_cast<double2 *>(dc + m);
    double2 *P = st->buf[st->cur];
    st->conj = 0;
    st->pending = false;
    HIP_TRY(lmh::launch_zero(P, st->ld * st->n, ctx->stream), "zero state");
    if (m) {
        HIP_TRY(hipMemcpyAsync(dr, r.data(), 8 * m, hipMemcpyHostToDevice, ctx->stream), "entry upload");
        HIP_TRY(hipMemcpyAsync(dc, c.data(), 8 * m, hipMemcpyHostToDevice, ctx->stream), "entry upload");
        HIP_TRY(hipMemcpyAsync(dv, v.data(), 16 * m, hipMemcpyHostToDevice, ctx->stream), "entry upload");
        HIP_TRY(lmh::launch_scatter(P, st->ld, m, dr, dc, dv, ctx->stream), "scatter");
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

int lmh_state_get_entries(lmh_state *st, int64_t count, const int64_t *rows, const int64_t *cols, double *out) {
    int rc = check_entries(st, count, rows, cols);
    if (rc) return rc;
    if (count > 0 && !out) return fail(LMH_EINVAL, "null output");
    if (count == 0) return LMH_OK;
    lmh_context *ctx = st->ctx;
    rc = state_normalize(st);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    HIP_TRY(grow(ctx->xfer, ctx->xfer_cap, (size_t)count * 32), "transfer buffer");
    int64_t *dr = reinterpret_cast<int64_t *>(ctx->xfer);
    int64_t *dc = dr + count;
    double2 *dv = reinterpret_cast<double2 *>(dc + count);
    HIP_TRY(hipMemcpyAsync(dr, rows, 8 * count, hipMemcpyHostToDevice, ctx->stream), "entry upload");
    HIP_TRY(hipMemcpyAsync(dc, cols, 8 * count, hipMemcpyHostToDevice, ctx->stream), "entry upload");
    HIP_TRY(lmh::launch_gather(st->buf[st->cur], st->ld, count, dr, dc, dv, ctx->stream), "gather");
    HIP_TRY(hipMemcpyAsync(out, dv, 16 * count, hipMemcpyDeviceToHost, ctx->stream), "entry readback");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

int lmh_state_idempotency(lmh_state *st, double *max_err) {
    if (!st || !max_err) return fail(LMH_EINVAL, "null argument");
    lmh_context *ctx = st->ctx;
    int rc = check_ctx(ctx);
    if (rc) return rc;
    rc = state_settle(st);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    double2 *W = nullptr;
    HIP_TRY(pool_take(ctx, st->ld * st->n, &W), "idempotency scratch");
    constexpr int kBlocks = 1024;
    hipError_t e = grow(ctx->xfer, ctx->xfer_cap, (size_t)8 * kBlocks);
    int s = e == hipSuccess ? lmh::idempotency(ctx->stream, st->buf[st->cur], st->n, st->ld, W,
                                                 reinterpret_cast<double *>(ctx->xfer), kBlocks, max_err)
                            : -3;
    pool_give(ctx, W);
    if (s != 0) return fail(LMH_EDEVICE, "idempotency check failed (%d)", s);
    return LMH_OK;
}

int lmh_state_device_ptr(lmh_state *st, void **dptr, int64_t *ld) {
    if (!st || !dptr || !ld) return fail(LMH_EINVAL, "null argument");
    int rc = check_ctx(st->ctx);
    if (rc) return rc;
    rc = state_normalize(st);
    if (rc) return rc;
    *dptr = st->buf[st->cur];
    *ld = st->ld;
    return LMH_OK;
}

namespace {

int step_impl(lmh_context *ctx, lmh_state *st);

// a strip step whose polynomial is still on the device (ctx->order_async):
// the state's own deferred-pass buffers are in place (allocated by an
// earlier step) and the step runs on the strip path
bool step_may_stay(const lmh_context *ctx, const lmh_state *st) {
    const lmh::DevHam &h = ctx->h;
    return strip_step_path(ctx) && !(ctx->tape.active && ctx->tape.st == st) && st->pend_coef &&
           st->pend_cap_stval >= lmh::strip_stval_count(h) && st->pend_cap_sdiag >= lmh::strip_sdiag_count(h);
}

// The step as lmh_step runs it when K is on the device: one strip launch
// (the pending pass, then this step's first side) whose K, s and
// coefficients come from the order result, and the copies for its deferred
// second pass; both do nothing unless the result is the one assumed.  The
// state's bookkeeping before the step goes into the journal.
int step_async(lmh_context *ctx, lmh_state *st) {
    const lmh::DevHam &h = ctx->h;
    ctx->journal.push_back({st, st->cur, st->conj, st->pending, st->pend});
    ctx->stat_async++;
    const int a = st->cur, b = (a + 1) % 3;
    // the last known K and s shape the launch (its start stagger) only
    const int K = std::max(ctx->K, 1);
    const int reps = std::max<int>(1, (int)std::lround(ctx->s));
    {
        ProfScope ps(ctx, LMH_KERNEL_FUSED);
        HIP_TRY(lmh::launch_strip(h, st->buf[a], st->buf[b], st->ld, st->n, ctx->dt / ctx->s, K, reps, true,
                                  ctx->stream, st->pending ? &st->pend : nullptr, ctx->d_tout),
                "fused launch");
    }
    HIP_TRY(lmh::launch_strip_snapshot(h, ctx->d_tout, ctx->async_xreal != 0, st->pend_stval, st->pend_sdiag,
                                       st->pend_coef, ctx->stream),
            "deferred pass values");
    st->pend = {st->pend_stval, st->pend_sdiag, ctx->dt / ctx->s, K, reps, ctx->async_xreal, st->pend_coef};
    st->pending = true;
    st->cur = b;
    st->conj ^= 1;
    ctx->fused_launched = lmh::fused_strip_variant();
    return LMH_OK;
}

}  // namespace

int lmh_step(lmh_context *ctx, lmh_state *st) {
    int rc = check_ctx(ctx, true);
    if (rc) return rc;
    if (!st || st->ctx != ctx) return fail(LMH_EINVAL, "state does not belong to this context");
    if (!ctx->solver_ready || ctx->solver_version != ctx->param_version)
        return fail(LMH_ESTATE, "lmh_step without an up-to-date lmh_update_solver");
    if (ctx->order_async) {
        if (step_may_stay(ctx, st)) return step_async(ctx, st);
        rc = ctx_resolve(ctx);
        if (rc) return rc;
    }
    if (ctx->tape.active && ctx->tape.st == st) {
        rc = tape_record(ctx, st);
        if (rc) return rc;
    }
    return step_impl(ctx, st);
}

namespace {

// Takes up an order result left on the device by lmh_update_solver (waits
// for it; by the next update the launches in front of it have normally
// finished).  When the journal's launches assumed it rightly they have done
// the steps; else they did nothing, and the states go back to their
// bookkeeping before those steps, which run again on the path the result
// calls for (H's values still those of the update: nothing rebuilds H
// before this runs).
int ctx_resolve(lmh_context *ctx) {
    if (!ctx->order_async) return LMH_OK;
    ctx->order_async = false;
    std::vector<lmh_context::JournalEntry> jr;
    jr.swap(ctx->journal);
    HIP_TRY(hipEventSynchronize(ctx->ev_order), "order result");
    lmh::TaylorOut out = *ctx->h_tout;
    const bool ran = lmh::strip_coef_ok(out, ctx->async_xreal != 0);
    if (!ran) {
        std::set<lmh_state *> seen;
        for (const auto &e : jr)
            if (seen.insert(e.st).second) {
                e.st->cur = e.cur;
                e.st->conj = e.conj;
                e.st->pending = e.pending;
                e.st->pend = e.pend;
            }
    }
    int rc = order_finish(ctx, out, ctx->async_level);
    if (rc) return rc;
    if (ran) {
        ctx->stat_steps += (int64_t)jr.size();
        ctx->stat_terms += (int64_t)jr.size() * ctx->K;
        return LMH_OK;
    }
    ctx->stat_redone += (int64_t)jr.size();
    for (const auto &e : jr) {
        rc = step_impl(ctx, e.st);
        if (rc) return rc;
    }
    return LMH_OK;
}

int step_impl(lmh_context *ctx, lmh_state *st) {
    const int K = ctx->K;
    ctx->stat_steps++;
    ctx->stat_terms += K;
    if (K == 0) return LMH_OK;    // U = I (a deferred pass stays deferred)
    int rc = LMH_OK;
    const int reps = (int)std::lround(ctx->s);
    const lmh::DevHam &h = ctx->h;
    int mode = ctx->step_mode;
    if (mode == LMH_STEP_AUTO)
        mode = (lmh::fused_supported(h) && h.n <= ctx->fused_max_n) ? LMH_STEP_FUSED_T : LMH_STEP_PER_PASS;
    if ((mode == LMH_STEP_FUSED_T || mode == LMH_STEP_FUSED_N) && !lmh::fused_supported(h))
        return fail(LMH_ERANGE, "fused step needs N <= 10240 sites (N = %lld)", (long long)h.n);
    if (mode == LMH_STEP_FUSED_T || mode == LMH_STEP_FUSED_N) {
        // Two column passes from the stored matrix S (P, or P^H when
        // st->conj): Y = (U S)^H, then U Y = U S^H U^H = (U S U^H)^H.  The
        // second pass stores plainly and the state changes form, so a step
        // costs one transposition; applied to S = P^H the same two passes
        // give U P U^H.  A state recorded by a tape keeps the P form (its
        // checkpoints are copies of the buffer): its second pass is
        // transposed as well.  FUSED_N does the transposition with the tiled
        // transpose kernel between two plainly stored passes.
        const bool keep_form = ctx->tape.active && ctx->tape.st == st;
        const double dts = ctx->dt / ctx->s;
        const int v = ctx->fused_variant;
        ctx->fused_launched = lmh::fused_resolved_variant(h, v);
        const bool defer = mode == LMH_STEP_FUSED_T && ctx->defer && !keep_form &&
                           ctx->fused_launched == lmh::fused_strip_variant();
        if (!defer) {
            // a deferred pass of an earlier step runs first
            rc = state_settle(st);
            if (rc) return rc;
        }
        const int a = st->cur, b = (a + 1) % 3, c = (a + 2) % 3;
        // Strip kernel with LMH_OPT_DEFER: only the first pass runs now; the
        // second, U Y, is deferred and runs in the same launch as the next
        // step's first pass, (U' U Y)^H, one memory round trip per step --
        // or alone when the state is read (state_settle).
        if (defer) {
            const int64_t nsv = lmh::strip_stval_count(h), nsd = lmh::strip_sdiag_count(h);
            if (st->pend_cap_stval < nsv) {
                HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
                dfree(st->pend_stval);
                st->pend_cap_stval = 0;
                HIP_TRY(dalloc(&st->pend_stval, (size_t)nsv), "deferred pass allocation");
                st->pend_cap_stval = nsv;
            }
            if (!st->pend_coef) HIP_TRY(dalloc(&st->pend_coef, 1), "deferred pass allocation");
            if (st->pend_cap_sdiag < nsd) {
                HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
                dfree(st->pend_sdiag);
                st->pend_cap_sdiag = 0;
                HIP_TRY(dalloc(&st->pend_sdiag, (size_t)nsd), "deferred pass allocation");
                st->pend_cap_sdiag = nsd;
            }
            {
                ProfScope ps(ctx, LMH_KERNEL_FUSED);
                HIP_TRY(lmh::launch_strip(h, st->buf[a], st->buf[b], st->ld, st->n, dts, K, reps, true, ctx->stream,
                                          st->pending ? &st->pend : nullptr),
                        "fused launch");
            }
            // this step's strip values, for its deferred second pass (after
            // the launch, which may still read the previous ones)
            HIP_TRY(hipMemcpyAsync(st->pend_stval, h.stval, sizeof(double2) * nsv, hipMemcpyDeviceToDevice,
                                   ctx->stream),
                    "deferred pass values");
            HIP_TRY(hipMemcpyAsync(st->pend_sdiag, h.sdiag, sizeof(double) * nsd, hipMemcpyDeviceToDevice,
                                   ctx->stream),
                    "deferred pass values");
            st->pend = {st->pend_stval, st->pend_sdiag, dts, K, reps, h.strip_xreal};
            st->pending = true;
            st->cur = b;
            st->conj ^= 1;
            return LMH_OK;
        }

        auto pass = [&](int from, int to, int scratch, bool tout) -> hipError_t {
            ProfScope ps(ctx, LMH_KERNEL_FUSED);
            return lmh::launch_fused_left(h, st->buf[from], st->buf[to], st->buf[scratch], st->ld, st->n, dts, K, reps,
                                          tout, v, ctx->stream);
        };
        auto transpose = [&](int from, int to) -> hipError_t {
            ProfScope ps(ctx, LMH_KERNEL_TRANSPOSE);
            return lmh::launch_conj_transpose(st->buf[from], st->buf[to], st->n, st->ld, ctx->stream);
        };
        if (mode == LMH_STEP_FUSED_T) {
            HIP_TRY(pass(a, b, c, true), "fused launch");
            HIP_TRY(pass(b, c, a, keep_form), "fused launch");
            st->cur = c;
        } else {
            HIP_TRY(pass(a, b, c, false), "fused launch");     // W = U S
            HIP_TRY(transpose(b, c), "transpose launch");       // W^H
            HIP_TRY(pass(c, b, a, false), "fused launch");     // U W^H
            st->cur = b;
            if (keep_form) {
                HIP_TRY(transpose(b, a), "transpose launch");
                st->cur = a;
            }
        }
        if (!keep_form) st->conj ^= 1;
        return LMH_OK;
    }
    // the per-term passes work on P itself
    rc = state_normalize(st);
    if (rc) return rc;
    // right_variant 1: both sides as left passes, each side's last term
    // written conjugate-transposed: (U (U P)')' = U P U' for any P;
    // right_variant 5: the same with each side's result transposed by the
    // tiled transpose kernel instead (coalesced writes)
    const bool tr = ctx->right_variant == 1 || ctx->right_variant == 5;
    const bool tr_kernel = ctx->right_variant == 5;
    int x = st->cur;
    for (int side = 0; side < 2; ++side) {
        for (int r = 0; r < reps; ++r) {
            int others[2], m = 0;
            for (int q = 0; q < 3; ++q)
                if (q != x) others[m++] = q;
            int in = x, turn = 0;
            for (int k = K; k >= 1;) {
                const int out = others[turn];
                turn ^= 1;
                const double c = ctx->dt / (ctx->s * k);
                const bool left = side == 0 || tr;
                const bool tout = tr && !tr_kernel && r == reps - 1 && k == 1;
                // left variant 3: terms k and k-1 in one launch (not the
                // conjugate-transposed last term)
                const bool pair = left && ctx->left_variant == 3 && k >= 2 && !(tr && !tr_kernel && r == reps - 1 && k == 2);
                ProfScope ps(ctx, left ? LMH_KERNEL_HORNER_LEFT : LMH_KERNEL_HORNER_RIGHT);
                hipError_t e;
                if (pair)
                    e = lmh::launch_horner_left2(h, st->buf[x], st->buf[in], st->buf[out], st->ld, st->n, c,
                                                 ctx->dt / (ctx->s * (k - 1)), k == K, ctx->left_cols, ctx->stream);
                else if (left)
                    e = lmh::launch_horner_left(h, st->buf[x], st->buf[in], st->buf[out], st->ld, st->n, c, k == K,
                                                ctx->left_cols, ctx->stream, ctx->left_variant == 3 ? 2 : ctx->left_variant,
                                                tout);
                else
                    e = lmh::launch_horner_right(h, st->buf[x], st->buf[in], st->buf[out], st->ld, st->n, c, k == K,
                                                 ctx->right_variant, ctx->right_cols, ctx->stream);
                if (e != hipSuccess) return hipfail(e, "horner launch");
                in = out;
                k -= pair ? 2 : 1;
            }
            x = in;
        }
        if (tr_kernel) {
            const int dst = (x + 1) % 3;          // both other buffers are free after the side
            ProfScope ps(ctx, LMH_KERNEL_TRANSPOSE);
            HIP_TRY(lmh::launch_conj_transpose(st->buf[x], st->buf[dst], st->n, st->ld, ctx->stream),
                    "transpose launch");
            x = dst;
        }
    }
    st->cur = x;
    return LMH_OK;
}

}  // namespace

int lmh_observables(lmh_context *ctx, lmh_state *st, double *rho, double *J) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!st || st->ctx != ctx) return fail(LMH_EINVAL, "state does not belong to this context");
    if (!ctx->geom_set) return fail(LMH_ESTATE, "no geometry");
    rc = state_settle(st);
    if (rc) return rc;
    if (ctx->built_version != ctx->param_version) {
        // the Hamiltonian of the current parameters: after update + step that
        // is the H of the step; before the first step (the iterator's first
        // timestamp yields H(t0) with the unevolved P) it is built here
        ProfScope ps(ctx, LMH_KERNEL_BUILD_H);
        HIP_TRY(lmh::launch_build_h(ctx->h, ctx->stream), "build_h launch");
        ctx->built_version = ctx->param_version;
    }
    const lmh::DevHam &h = ctx->h;
    if (rho && !ctx->d_rho) HIP_TRY(dalloc(&ctx->d_rho, h.n), "observable allocation");
    if (J && h.nbonds && !ctx->d_J) HIP_TRY(dalloc(&ctx->d_J, h.nbonds), "observable allocation");
    double *dr = rho ? ctx->d_rho : nullptr;
    double *dj = (J && h.nbonds) ? ctx->d_J : nullptr;
    {
        ProfScope ps(ctx, LMH_KERNEL_OBSERVABLES);
        HIP_TRY(lmh::launch_observables(h, st->buf[st->cur], st->ld, dr, dj, ctx->stream, st->conj),
                "observables launch");
    }
    if (dr) HIP_TRY(hipMemcpyAsync(rho, dr, 8 * h.n, hipMemcpyDeviceToHost, ctx->stream), "observables readback");
    if (dj) HIP_TRY(hipMemcpyAsync(J, dj, 8 * h.nbonds, hipMemcpyDeviceToHost, ctx->stream), "observables readback");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

namespace {

// the H of the current parameters built (as lmh_observables does)
int ensure_built(lmh_context *ctx) {
    if (ctx->built_version != ctx->param_version) {
        ProfScope ps(ctx, LMH_KERNEL_BUILD_H);
        HIP_TRY(lmh::launch_build_h(ctx->h, ctx->stream), "build_h launch");
        ctx->built_version = ctx->param_version;
    }
    return LMH_OK;
}

int upload_mask(const uint8_t *mask, int64_t n, uint8_t **dev) {
    HIP_TRY(dalloc(dev, (size_t)n), "mask allocation");
    std::vector<uint8_t> m(n);
    for (int64_t k = 0; k < n; ++k) m[k] = mask[k] ? 1 : 0;
    HIP_TRY(hipMemcpy(*dev, m.data(), n, hipMemcpyHostToDevice), "mask upload");
    return LMH_OK;
}

}  // namespace

int lmh_currents_region(lmh_context *ctx, lmh_state *st, const uint8_t *src, const uint8_t *dst, double *per_site,
                        double *total) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!st || st->ctx != ctx) return fail(LMH_EINVAL, "state does not belong to this context");
    if (!ctx->geom_set) return fail(LMH_ESTATE, "no geometry");
    rc = state_settle(st);
    if (rc) return rc;
    if (!src) return fail(LMH_EINVAL, "null source region");
    rc = ensure_built(ctx);
    if (rc) return rc;
    const lmh::DevHam &h = ctx->h;
    const int64_t n = h.n, nb = lmh::region_currents_parts(h);
    uint8_t *dsrc = nullptr, *ddst = nullptr;
    double *dout = nullptr, *dpart = nullptr;
    rc = upload_mask(src, n, &dsrc);
    if (!rc && dst) rc = upload_mask(dst, n, &ddst);
    hipError_t e = hipSuccess;
    if (!rc) e = dalloc(&dout, (size_t)n + 1);
    if (!rc && e == hipSuccess) e = dalloc(&dpart, (size_t)std::max<int64_t>(nb, 1));
    if (!rc && e == hipSuccess) {
        ProfScope ps(ctx, LMH_KERNEL_OBSERVABLES);
        e = lmh::launch_region_currents(h, st->buf[st->cur], st->ld, dsrc, ddst, dout, dpart, dout + n, ctx->stream,
                                        st->conj);
    }
    if (!rc && e == hipSuccess && per_site)
        e = hipMemcpyAsync(per_site, dout, 8 * n, hipMemcpyDeviceToHost, ctx->stream);
    if (!rc && e == hipSuccess && total) e = hipMemcpyAsync(total, dout + n, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(dsrc);
    dfree(ddst);
    dfree(dout);
    dfree(dpart);
    if (rc) return rc;
    if (e != hipSuccess) return hipfail(e, "region currents");
    return LMH_OK;
}

int lmh_currents_region_noalloc(lmh_context *ctx, lmh_state *st, const uint8_t *src, const uint8_t *dst,
                                double *per_site, double *total) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "no geometry");
    const lmh::DevHam &h = ctx->h;
    if (h.n > lmh::region_currents_static_max_sites())
        return fail(LMH_ERANGE, "allocation-free region currents take at most %lld sites (N = %lld)",
                    (long long)lmh::region_currents_static_max_sites(), (long long)h.n);
    if (!st || st->ctx != ctx) return fail(LMH_EINVAL, "state does not belong to this context");
    if (!src) return fail(LMH_EINVAL, "null source region");
    if (!per_site && !total) return LMH_OK;
    rc = ensure_built(ctx);
    if (rc) return rc;
    // a deferred pass runs first (a launch, no allocation)
    rc = state_settle(st);
    if (rc) return rc;
    ProfScope ps(ctx, LMH_KERNEL_OBSERVABLES);
    HIP_TRY(lmh::launch_region_currents_static(h, st->buf[st->cur], st->ld, src, dst, per_site, total, ctx->stream,
                                               st->conj),
            "region currents");
    return LMH_OK;
}

int lmh_series_create(lmh_context *ctx, int kind, int64_t capacity, const uint8_t *src, const uint8_t *dst,
                      lmh_series **out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out) return fail(LMH_EINVAL, "null output");
    if (!ctx->geom_set) return fail(LMH_ESTATE, "no geometry");
    if (kind < LMH_SERIES_DENSITY || kind > LMH_SERIES_REGION) return fail(LMH_EINVAL, "unknown series kind %d", kind);
    if (capacity < 1) return fail(LMH_EINVAL, "capacity must be positive");
    if (kind == LMH_SERIES_REGION && !src) return fail(LMH_EINVAL, "region series needs a source region");
    const lmh::DevHam &h = ctx->h;
    auto *q = new lmh_series;
    q->ctx = ctx;
    q->kind = kind;
    q->cap = capacity;
    q->width = kind == LMH_SERIES_DENSITY ? h.n : kind == LMH_SERIES_CURRENTS ? h.nbonds : 1;
    q->geom_version = ctx->geom_version;
    hipError_t e = dalloc(&q->buf, (size_t)std::max<int64_t>(q->width * capacity, 1));
    if (e == hipSuccess) e = dalloc(&q->coef, (size_t)std::max<int64_t>(2 * capacity, 2));
    if (e == hipSuccess && kind == LMH_SERIES_REGION)
        e = dalloc(&q->part, (size_t)std::max<int64_t>(lmh::region_currents_parts(h), 1));
    rc = e == hipSuccess ? LMH_OK : hipfail(e, "series allocation");
    if (!rc && kind == LMH_SERIES_REGION) rc = upload_mask(src, h.n, &q->src);
    if (!rc && kind == LMH_SERIES_REGION && dst) rc = upload_mask(dst, h.n, &q->dst);
    if (rc) {
        dfree(q->buf);
        dfree(q->coef);
        dfree(q->part);
        dfree(q->src);
        dfree(q->dst);
        delete q;
        return rc;
    }
    ctx->series.insert(q);
    *out = q;
    return LMH_OK;
}

int lmh_series_destroy(lmh_series *q) {
    if (!q) return LMH_OK;
    if (q->ctx) {
        (void)hipStreamSynchronize(q->ctx->stream);
        q->ctx->series.erase(q);
    }
    dfree(q->buf);
    dfree(q->coef);
    dfree(q->part);
    dfree(q->src);
    dfree(q->dst);
    delete q;
    return LMH_OK;
}

int lmh_series_record(lmh_series *q, lmh_state *st, double t) {
    if (!q) return fail(LMH_EINVAL, "null series");
    lmh_context *ctx = q->ctx;
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!st || st->ctx != ctx) return fail(LMH_EINVAL, "state does not belong to the series' context");
    if (ctx->geom_version != q->geom_version) return fail(LMH_ESTATE, "geometry changed since lmh_series_create");
    if (!std::isfinite(t) || (q->len > 0 && !(t > q->times.back())))
        return fail(LMH_EINVAL, "series times must increase (got %g after %g)", t, q->len ? q->times.back() : 0.0);
    if (q->len >= q->cap) return fail(LMH_ERANGE, "series full (capacity %lld)", (long long)q->cap);
    rc = ensure_built(ctx);
    if (rc) return rc;
    rc = state_settle(st);
    if (rc) return rc;
    const lmh::DevHam &h = ctx->h;
    double *slot = q->buf + q->len * q->width;
    {
        ProfScope ps(ctx, LMH_KERNEL_OBSERVABLES);
        hipError_t e;
        if (q->kind == LMH_SERIES_DENSITY)
            e = lmh::launch_observables(h, st->buf[st->cur], st->ld, slot, nullptr, ctx->stream, st->conj);
        else if (q->kind == LMH_SERIES_CURRENTS)
            e = lmh::launch_observables(h, st->buf[st->cur], st->ld, nullptr, h.nbonds ? slot : nullptr, ctx->stream,
                                        st->conj);
        else
            e = lmh::launch_region_currents(h, st->buf[st->cur], st->ld, q->src, q->dst, nullptr, q->part, slot,
                                            ctx->stream, st->conj);
        if (e != hipSuccess) return hipfail(e, "series record");
    }
    q->times.push_back(t);
    q->len++;
    return LMH_OK;
}

int lmh_series_length(lmh_series *q, int64_t *len, int64_t *width) {
    if (!q) return fail(LMH_EINVAL, "null series");
    if (len) *len = q->len;
    if (width) *width = q->width;
    return LMH_OK;
}

int lmh_series_differentiate(lmh_series *q) {
    if (!q) return fail(LMH_EINVAL, "null series");
    int rc = check_ctx(q->ctx);
    if (rc) return rc;
    if (q->len < 2) return fail(LMH_ESTATE, "Cannot differentiate a series of length %lld", (long long)q->len);
    const int64_t m = q->len - 1;
    std::vector<double> ab(2 * m);
    std::vector<double> &td = q->times;
    for (int64_t i = 1; i < q->len; ++i) {
        const double dt = td[i] - td[i - 1];
        ab[i - 1] = 1.0 / dt;
        ab[m + i - 1] = -1.0 / dt;
        td[i - 1] += dt / 2;
    }
    td.pop_back();
    HIP_TRY(hipMemcpyAsync(q->coef, ab.data(), 8 * ab.size(), hipMemcpyHostToDevice, q->ctx->stream), "coefficients");
    HIP_TRY(lmh::launch_series_differentiate(q->buf, q->width, q->len, q->coef, q->coef + m, q->ctx->stream),
            "differentiate");
    HIP_TRY(hipStreamSynchronize(q->ctx->stream), "hipStreamSynchronize");
    q->len--;
    return LMH_OK;
}

int lmh_series_integrate(lmh_series *q) {
    if (!q) return fail(LMH_EINVAL, "null series");
    int rc = check_ctx(q->ctx);
    if (rc) return rc;
    if (q->len < 2) return fail(LMH_ESTATE, "Cannot integrate a series of length %lld", (long long)q->len);
    std::vector<double> hh(q->len - 1);
    for (int64_t i = 1; i < q->len; ++i) hh[i - 1] = (q->times[i] - q->times[i - 1]) / 2;
    HIP_TRY(hipMemcpyAsync(q->coef, hh.data(), 8 * hh.size(), hipMemcpyHostToDevice, q->ctx->stream), "coefficients");
    HIP_TRY(lmh::launch_series_integrate(q->buf, q->width, q->len, q->coef, q->ctx->stream), "integrate");
    HIP_TRY(hipStreamSynchronize(q->ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

int lmh_series_download(lmh_series *q, double *times, double *values) {
    if (!q) return fail(LMH_EINVAL, "null series");
    int rc = check_ctx(q->ctx);
    if (rc) return rc;
    if (times && q->len) std::memcpy(times, q->times.data(), 8 * q->len);
    if (values && q->len * q->width)
        HIP_TRY(hipMemcpyAsync(values, q->buf, 8 * q->len * q->width, hipMemcpyDeviceToHost, q->ctx->stream),
                "series readback");
    HIP_TRY(hipStreamSynchronize(q->ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

int lmh_get_hamiltonian(lmh_context *ctx, double *diag, double *hij) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "no geometry");
    lmh::DevHam &h = ctx->h;
    if (ctx->built_version != ctx->param_version) {
        HIP_TRY(lmh::launch_build_h(h, ctx->stream), "build_h launch");
        ctx->built_version = ctx->param_version;
    }
    if (diag && !ctx->d_hdiag) HIP_TRY(dalloc(&ctx->d_hdiag, h.n), "allocation");
    if (hij && h.nbonds && !ctx->d_hij) HIP_TRY(dalloc(&ctx->d_hij, h.nbonds), "allocation");
    double *dd = diag ? ctx->d_hdiag : nullptr;
    double2 *dh = (hij && h.nbonds) ? ctx->d_hij : nullptr;
    HIP_TRY(lmh::launch_gather_h(h, dd, dh, ctx->stream), "gather launch");
    if (dd) HIP_TRY(hipMemcpyAsync(diag, dd, 8 * h.n, hipMemcpyDeviceToHost, ctx->stream), "readback");
    if (dh) HIP_TRY(hipMemcpyAsync(hij, dh, 16 * h.nbonds, hipMemcpyDeviceToHost, ctx->stream), "readback");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

}  // extern "C"

namespace {

// Eigenvectors of the dense H of the current parameters into A (n x lda,
// column-major) and eigenvalues (ascending) into E.  When every stored
// hopping is real (no magnetic field), H is real symmetric and goes to the
// real solvers (dsyevd, then dsyevdj, then dsyev), its eigenvectors widened to
// complex; otherwise, and if none of those converges, to the complex ones.
// rocsolver_zheevd has returned NaN eigenvectors (info = 0) for real
// symmetric input from N ~ 150 up, so every result is checked for finite
// eigenvalues and eigenvectors and the next solver takes over if needed
// (complex order: zheevd, zheevdj, zheev; LMH_OPT_EIGENSOLVER picks the first
// solver of either kind).
int eigensolve(lmh_context *ctx, double2 *A, int64_t lda, std::vector<double> &E) {
    lmh::DevHam &h = ctx->h;
    if (ctx->built_version != ctx->param_version) {
        ProfScope ps(ctx, LMH_KERNEL_BUILD_H);
        HIP_TRY(lmh::launch_build_h(h, ctx->stream), "build_h launch");
        ctx->built_version = ctx->param_version;
    }
    const int64_t n = h.n;
    double *dE = nullptr, *dwork = nullptr, *R = nullptr;
    int *dinfo = nullptr;
    hipError_t e = dalloc(&dE, n);
    if (e == hipSuccess) e = dalloc(&dwork, n);
    if (e == hipSuccess) e = dalloc(&dinfo, 1);
    auto cleanup = [&]() {
        dfree(dE);
        dfree(dwork);
        dfree(dinfo);
        dfree(R);
    };
    if (e != hipSuccess) {
        cleanup();
        return hipfail(e, "eigensolver allocation");
    }
    int cplx = 1;
    e = lmh::launch_imag_check(h, dinfo, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&cplx, dinfo, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        cleanup();
        return hipfail(e, "real-H check");
    }
    const int64_t ldr = std::max<int64_t>(n, 1);
    if (!cplx) {
        e = dalloc(&R, (size_t)(ldr * n));
        if (e != hipSuccess) {
            cleanup();
            return hipfail(e, "real eigensolver allocation");
        }
    }
    const int order[3][3] = {{0, 2, 1}, {1, 2, 0}, {2, 1, 0}};
    E.assign(n, 0.0);
    int info = 0, used = -1;
    for (int attempt = 0; attempt < 6 && used < 0; ++attempt) {
        const bool real = attempt < 3;
        if (real && cplx) continue;
        const int meth = order[ctx->eigensolver][attempt % 3];
        int zs;
        if (real) {
            e = lmh::launch_dense_real_from_ell(h, R, ldr, ctx->stream);
            if (e != hipSuccess) break;
            zs = lmh::eigh_real(ctx->stream, meth, R, n, ldr, dE, dwork, dinfo);
            if (zs == 0) e = lmh::launch_widen(R, ldr, A, lda, n, ctx->stream);
            if (e != hipSuccess) break;
        } else {
            e = lmh::launch_dense_from_ell(h, A, lda, ctx->stream);
            if (e != hipSuccess) break;
            zs = lmh::eigh(ctx->stream, meth, A, n, lda, dE, dwork, dinfo);
        }
        if (zs != 0) {
            cleanup();
            return fail(LMH_EDEVICE, "rocSOLVER %s eigensolver %d failed (status %d)", real ? "real" : "complex",
                        meth, zs);
        }
        int bad = 0;
        e = hipMemcpy(E.data(), dE, 8 * n, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = lmh::launch_nonfinite(A, n, lda, dinfo, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&bad, dinfo, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) break;
        for (int64_t k = 0; k < n && !bad; ++k) bad = !std::isfinite(E[k]);
        if (info == 0 && !bad) {
            used = meth;
            ctx->eig_real = real ? 1 : 0;
            ctx->eig_method = meth;
        }
    }
    cleanup();
    if (e != hipSuccess) return hipfail(e, "eigenvalue readback");
    if (used < 0) return fail(LMH_EDEVICE, "no rocSOLVER eigensolver converged to finite eigenvectors (last info %d)", info);
    return LMH_OK;
}

// P = V diag(f) V^H from the eigenvectors in W (overwritten: columns scaled by
// sqrt|f|), weights of either sign; only the columns between the first and the
// last nonzero weight enter the product.  *ncols: that column count.
int project_weights(lmh_context *ctx, double2 *W, int64_t ldw, const double *f, double2 *P, int64_t ldp,
                    int64_t *ncols) {
    const int64_t n = ctx->h.n;
    int64_t lo = n, hi = 0;
    bool neg = false;
    std::vector<double> w(n), sg(n);
    for (int64_t k = 0; k < n; ++k) {
        if (!std::isfinite(f[k])) return fail(LMH_EINVAL, "weight %lld is not finite", (long long)k);
        w[k] = std::sqrt(std::fabs(f[k]));
        sg[k] = f[k] < 0.0 ? -1.0 : 1.0;
        neg = neg || f[k] < 0.0;
        if (f[k] != 0.0) {
            lo = std::min(lo, k);
            hi = k + 1;
        }
    }
    const int64_t m = hi > lo ? hi - lo : 0;
    double *dw = nullptr, *dsg = nullptr;
    hipError_t e = dalloc(&dw, std::max<int64_t>(n, 1));
    if (e == hipSuccess) e = dalloc(&dsg, std::max<int64_t>(n, 1));
    if (e == hipSuccess) e = hipMemcpy(dw, w.data(), 8 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dsg, sg.data(), 8 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess && m > 0) e = lmh::launch_scale_columns(W + lo * ldw, ldw, n, dw + lo, m, ctx->stream);
    if (e == hipSuccess) {
        ProfScope ps(ctx, LMH_KERNEL_PROJECTOR);
        if (m > 0)
            e = lmh::launch_herk(W + lo * ldw, ldw, n, m, P, ldp, ctx->stream, neg ? dsg + lo : nullptr);
        else
            e = lmh::launch_zero(P, ldp * n, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(dw);
    dfree(dsg);
    if (e != hipSuccess) return hipfail(e, "projector");
    if (ncols) *ncols = m;
    return LMH_OK;
}

}  // namespace

extern "C" {

int lmh_densitymatrix(lmh_context *ctx, lmh_state *st, double mu, double T, int64_t *nocc, double *E_out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!st || st->ctx != ctx) return fail(LMH_EINVAL, "state does not belong to this context");
    if (!ctx->geom_set) return fail(LMH_ESTATE, "no geometry");
    if (!std::isfinite(mu) || !std::isfinite(T) || T < 0) return fail(LMH_EINVAL, "invalid mu / T");
    const int64_t n = ctx->h.n;
    double2 *A = st->buf[(st->cur + 1) % 3];
    std::vector<double> E;
    rc = eigensolve(ctx, A, st->ld, E);
    if (rc) return rc;
    // occupations f(E) (densfun with Fermi-Dirac statistics, src/spectrum.jl:230-232)
    std::vector<double> f(n);
    for (int64_t k = 0; k < n; ++k) {
        const double x = E[k] - mu;
        if (T == 0.0)
            f[k] = (x == 0.0 || x < 0.0) ? 1.0 : 0.0;
        else
            f[k] = 1.0 / (std::exp(x / T) + 1.0);
    }
    int64_t m = 0;
    rc = project_weights(ctx, A, st->ld, f.data(), st->buf[st->cur], st->ld, &m);
    st->conj = 0;
    st->pending = false;
    if (rc) return rc;
    if (nocc) *nocc = m;
    if (E_out) std::memcpy(E_out, E.data(), 8 * n);
    return LMH_OK;
}

int lmh_diagonalize(lmh_context *ctx, double *E_out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "no geometry");
    const int64_t n = ctx->h.n, ld = ((n + 7) / 8) * 8;
    if (ctx->eig_n != n || !ctx->eigvec) {
        dfree(ctx->eigvec);
        ctx->eig_n = 0;
        HIP_TRY(dalloc(&ctx->eigvec, (size_t)(ld * n)), "eigenvector allocation");
        ctx->eig_ld = ld;
    }
    ctx->eig_n = 0;
    rc = eigensolve(ctx, ctx->eigvec, ctx->eig_ld, ctx->eig_E);
    if (rc) return rc;
    ctx->eig_n = n;
    if (E_out) std::memcpy(E_out, ctx->eig_E.data(), 8 * n);
    return LMH_OK;
}

int lmh_projector(lmh_context *ctx, lmh_state *st, const double *f, int64_t *ncols) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!st || st->ctx != ctx) return fail(LMH_EINVAL, "state does not belong to this context");
    if (!f) return fail(LMH_EINVAL, "null weights");
    if (!ctx->eigvec || ctx->eig_n == 0) return fail(LMH_ESTATE, "lmh_projector without an lmh_diagonalize");
    if (ctx->eig_n != st->n) return fail(LMH_EINVAL, "state size %lld differs from the eigensystem's %lld",
                                         (long long)st->n, (long long)ctx->eig_n);
    // the state's free cache holds the scaled copy of V
    double2 *W = st->buf[(st->cur + 1) % 3];
    HIP_TRY(hipMemcpy2DAsync(W, sizeof(double2) * st->ld, ctx->eigvec, sizeof(double2) * ctx->eig_ld,
                             sizeof(double2) * st->n, st->n, hipMemcpyDeviceToDevice, ctx->stream),
            "eigenvector copy");
    st->conj = 0;
    st->pending = false;
    return project_weights(ctx, W, st->ld, f, st->buf[st->cur], st->ld, ncols);
}

int lmh_eigensystem_free(lmh_context *ctx) {
    if (!ctx) return fail(LMH_EINVAL, "null context");
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    dfree(ctx->eigvec);
    ctx->eig_n = 0;
    ctx->eig_E.clear();
    return LMH_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batches

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

namespace {

void batch_free(lmh_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto &c : b->classes) {
        for (auto &p : c.items) dfree(p);
        dfree(c.bstart);
    }
    for (int q = 0; q < 6; ++q) {
        if (b->strip_args[q]) (void)hipFree(b->strip_args[q]);
        dfree(b->strip_ustart[q]);
    }
    dfree(b->d_trace);
    dfree(b->d_out);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

// one step of every size-class member at rotation phase p = steps_done % 3
// (two conjugate-transposed passes: U (U P)^H)^H = U P U^H)
hipError_t batch_launch_step(lmh_batch *b, int64_t p) {
    for (int side = 0; side < 2; ++side) {
        for (auto &c : b->classes) {
            hipError_t e = lmh::launch_batched_fused(c.cls, c.items[2 * p + side], c.bstart, (int)c.members.size(),
                                                     c.total, c.maxn, c.maxdict, b->stream);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// expected buffer of member r after the steps and launches so far
int batch_expected_cur(const lmh_batch *b, size_t r) {
    const bool strip = b->member_kernel[r] == lmh::fused_strip_variant();
    return (int)((b->cur0[r] + (strip ? b->strip_rot : 2 * b->steps_done)) % 3);
}

}  // namespace

extern "C" {

int lmh_tape_begin(lmh_context *ctx, lmh_state *st, int64_t checkpoint_every) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!st || st->ctx != ctx) return fail(LMH_EINVAL, "state does not belong to this context");
    if (checkpoint_every < 1) return fail(LMH_EINVAL, "checkpoint interval must be >= 1");
    if (ctx->csc_mode)
        return fail(LMH_ESTATE, "adjoint tapes differentiate the parametric Hamiltonian (lmh_set_geometry / "
                                "lmh_set_onsite / lmh_set_field); this context's H is given as CSC values");
    rc = state_normalize(st);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    tape_clear(ctx);
    auto &tp = ctx->tape;
    tp.active = true;
    tp.st = st;
    tp.every = checkpoint_every;
    tp.geom_version = ctx->geom_version;
    tp.onsite_version = ctx->onsite_version;
    return LMH_OK;
}

int lmh_tape_end(lmh_context *ctx) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    tape_clear(ctx);
    return LMH_OK;
}

int lmh_tape_length(lmh_context *ctx, int64_t *nsteps, int64_t *ncomp) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    const auto &tp = ctx->tape;
    if (nsteps) *nsteps = tp.active ? (int64_t)tp.recs.size() : 0;
    if (ncomp) *ncomp = tp.recs.empty() ? (int64_t)ctx->field_host.size() : (int64_t)tp.recs[0].field.size();
    return LMH_OK;
}

int lmh_value_classes(lmh_context *ctx, int64_t *nclasses, int64_t *matched) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (nclasses) *nclasses = ctx->h.ndict;
    if (matched) *matched = ctx->h.dict_ok;
    return LMH_OK;
}

int lmh_tape_checkpoints(lmh_context *ctx, int64_t *on_device, int64_t *on_host) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    int64_t nh = 0;
    for (char h : ctx->tape.cp_host) nh += h ? 1 : 0;
    if (on_device) *on_device = (int64_t)ctx->tape.cp_host.size() - nh;
    if (on_host) *on_host = nh;
    return LMH_OK;
}

int lmh_tape_backward(lmh_context *ctx, lmh_state *lam, double *grad_field, double *grad_onsite) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    auto &tp = ctx->tape;
    if (!tp.active || tp.recs.empty()) return fail(LMH_ESTATE, "no recorded steps (lmh_tape_begin, then lmh_step)");
    if (!lam || lam->ctx != ctx || lam == tp.st) return fail(LMH_EINVAL, "costate must be another state of this context");
    rc = state_normalize(lam);
    if (rc) return rc;
    if (tp.geom_version != ctx->geom_version || tp.onsite_version != ctx->onsite_version)
        return fail(LMH_ESTATE, "geometry or on-site energies changed since lmh_tape_begin");
    lmh::DevHam &h = ctx->h;
    const int64_t M = (int64_t)tp.recs.size();
    const int64_t ncomp = (int64_t)tp.recs[0].field.size();
    bool fused = ctx->adjoint_mode != 2 && ctx->adjoint_mode != 3;
    for (int64_t k = 0; k < M; ++k) {
        const auto &r = tp.recs[k];
        if (r.K > 0 && std::lround(r.s) != 1)
            return fail(LMH_ERANGE, "step %lld: the adjoint needs s = 1 (max|H_ij| |dt| <= 1), got s = %g", (long long)k, r.s);
        if (r.K > 0 && !lmh::adjoint_supported(h, r.K)) fused = false;
    }
    if (ctx->adjoint_mode == 1 && !fused)
        return fail(LMH_ERANGE, "fused adjoint keeps K columns of N sites in LDS (needs K N <= 10240, N <= 2048)");
    if (!fused && !lmh::reverse_term_supported(h)) return fail(LMH_ERANGE, "adjoint needs at most 16 ELL slots per row");
    const int64_t n = h.n, ld = tp.st->ld;
    const int64_t nslots = (int64_t)(h.zmax + 1) * n;
    const int64_t nblk = lmh::adjoint_reduce_blocks(h);
    const int nparts = fused ? lmh::adjoint_parts(h) : (int)lmh::reverse_term_parts(h);
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    // work buffers: pooled N x ld buffers, plus the free caches of the recorded
    // state and the costate's input buffer (after its Hermitian part is taken)
    const int lm = (lam->cur + 1) % 3, yb = (lam->cur + 2) % 3;
    double2 *Lm = lam->buf[lm], *Ybt = lam->buf[yb], *Lin = lam->buf[lam->cur];
    double2 *spare1 = tp.st->buf[(tp.st->cur + 1) % 3], *spare2 = tp.st->buf[(tp.st->cur + 2) % 3];
    std::vector<double2 *> taken;
    double *gV = nullptr, *gF = nullptr;
    hipError_t e = hipSuccess;
    const int npool = fused ? 2 : 4;
    for (int q = 0; q < npool && e == hipSuccess; ++q) {
        double2 *p = nullptr;
        e = pool_take(ctx, ld * n, &p);
        if (e == hipSuccess) taken.push_back(p);
    }
    if (e == hipSuccess) e = grow(ctx->adj_partial, ctx->adj_partial_cap, (size_t)nparts * nslots);
    if (e == hipSuccess && ncomp) e = grow(ctx->d_tape_fields, ctx->tape_fields_cap, (size_t)M * ncomp);
    if (e == hipSuccess) e = dalloc(&gV, (size_t)n);
    if (e == hipSuccess) e = dalloc(&gF, (size_t)std::max<int64_t>(1, M * nblk * ncomp));
    auto cleanup = [&]() {
        for (double2 *p : taken) pool_give(ctx, p);
        dfree(gV);
        dfree(gF);
    };
    if (e != hipSuccess) {
        cleanup();
        return hipfail(e, "adjoint work buffers");
    }
    double2 *partial = ctx->adj_partial;
    hipStream_t st = ctx->stream;
    // every recorded step's field parameters in one upload
    std::vector<lmh::FieldComp> allf((size_t)M * ncomp);
    for (int64_t k = 0; k < M; ++k)
        for (int64_t c = 0; c < ncomp; ++c) allf[(size_t)k * ncomp + c] = tp.recs[k].field[c];
    if (ncomp)
        e = hipMemcpyAsync(ctx->d_tape_fields, allf.data(), allf.size() * sizeof(lmh::FieldComp), hipMemcpyHostToDevice,
                           st);
    double2 *Pa = taken[0];
    if (e == hipSuccess) e = hipMemsetAsync(gV, 0, 8 * n, st);
    if (e == hipSuccess) e = hipMemsetAsync(gF, 0, 8 * std::max<int64_t>(1, M * nblk * ncomp), st);
    if (e == hipSuccess) e = lmh::launch_copy(tp.st->buf[tp.st->cur], Pa, ld * n, st);
    if (e == hipSuccess) e = lmh::launch_hermitian_part(Lin, Lm, n, ld, st);
    double2 *Lcur = Lm;                      // the current costate Lambda_{k+1}

    // per-term helpers (any N): y = p(-i a H) x from the left, the result in `out`
    // (with left variant 3, terms k and k-1 in one launch)
    auto left_poly = [&](const lmh::DevHam &hh, const double2 *X, double2 *out, double2 *tmp, double a, int K) {
        const bool pairs = ctx->left_variant == 3 && lmh::left_band2_fits(hh);
        const int nl = pairs ? K / 2 + K % 2 : K;           // launches; the last one writes `out`
        double2 *b[2] = {(nl % 2) ? out : tmp, (nl % 2) ? tmp : out};
        const double2 *yin = X;
        hipError_t ee = hipSuccess;
        for (int k = K, j = 0; k >= 1 && ee == hipSuccess; ++j) {
            ProfScope ps(ctx, LMH_KERNEL_ADJOINT);
            if (pairs && k >= 2) {
                ee = lmh::launch_horner_left2(hh, X, yin, b[j % 2], ld, n, a / (double)k, a / (double)(k - 1), k == K,
                                              ctx->left_cols, st);
                k -= 2;
            } else {
                ee = lmh::launch_horner_left(hh, X, yin, b[j % 2], ld, n, a / (double)k, k == K, ctx->left_cols, st,
                                             ctx->left_variant);
                k -= 1;
            }
            yin = b[j % 2];
        }
        return ee;
    };
    // W = Y p(+i a H) from the right (a = -dt: Y U), the result in `out`
    auto right_poly = [&](const lmh::DevHam &hh, const double2 *Y, double2 *out, double2 *tmp, double a, int K) {
        double2 *b[2] = {(K % 2) ? out : tmp, (K % 2) ? tmp : out};
        const double2 *win = Y;
        hipError_t ee = hipSuccess;
        for (int j = 0; j < K && ee == hipSuccess; ++j) {
            ProfScope ps(ctx, LMH_KERNEL_ADJOINT);
            ee = lmh::launch_horner_right(hh, Y, win, b[j % 2], ld, n, a / (double)(K - j), j == 0,
                                          ctx->right_variant, ctx->right_cols, st);
            win = b[j % 2];
        }
        return ee;
    };
    // the K-1 Horner iterates of a side, kept when HBM allows (else each
    // reverse term recomputes its iterate from x)
    int kmax = 0;
    for (const auto &rec : tp.recs) kmax = std::max(kmax, rec.K);
    std::vector<double2 *> iters;
    if (!fused && kmax > 1 && ctx->adjoint_mode != 3) {
        const size_t bytes = sizeof(double2) * (size_t)(ld * n);
        for (int q = 0; q < kmax - 1; ++q) {
            double2 *p = nullptr;
            if (ctx->pool.empty() || ctx->pool_elems != ld * n) {
                size_t free_b = 0, total_b = 0;
                if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b < bytes + ((size_t)4 << 30)) break;
            }
            if (pool_take(ctx, ld * n, &p) != hipSuccess) break;
            iters.push_back(p);
        }
        if ((int)iters.size() < kmax - 1) {
            for (double2 *p : iters) pool_give(ctx, p);
            iters.clear();
        }
    }
    // reverse sweep of one side: x (columns), cotangent G of U x; xbar into XB
    auto side_sweep = [&](const lmh::DevHam &hh, const double2 *x, const double2 *G, double2 *XB, double2 *G1,
                          double2 *G2, double2 *I1, double2 *I2, double dt, int K) {
        hipError_t ee = hipSuccess;          // term 1 writes XB without reading it
        if (!iters.empty()) {
            // iters[j] = y_{K-j}: Horner terms K .. K-j of U x
            const double2 *y = x;
            for (int k = K, j = 0; k >= 2 && ee == hipSuccess; --k, ++j) {
                ProfScope ps(ctx, LMH_KERNEL_ADJOINT);
                ee = lmh::launch_horner_left(hh, x, y, iters[j], ld, n, dt / (double)k, k == K, ctx->left_cols, st,
                                             ctx->left_variant);
                y = iters[j];
            }
        }
        const double2 *gsrc = G;
        for (int m = 1; m <= K && ee == hipSuccess; ++m) {
            // y_{m+1}: Horner terms K .. m+1 of U x, kept or recomputed from x
            const double2 *y = x;
            if (!iters.empty()) {
                if (m < K) y = iters[K - 1 - m];
            } else {
                for (int k = K, j = 0; k >= m + 1 && ee == hipSuccess; --k, ++j) {
                    double2 *out = (j % 2) ? I2 : I1;
                    ProfScope ps(ctx, LMH_KERNEL_ADJOINT);
                    ee = lmh::launch_horner_left(hh, x, y, out, ld, n, dt / (double)k, k == K, ctx->left_cols, st,
                                                 ctx->left_variant);
                    y = out;
                }
            }
            double2 *gdst = m < K ? ((m % 2) ? G1 : G2) : nullptr;
            if (ee == hipSuccess) {
                ProfScope ps(ctx, LMH_KERNEL_ADJOINT);
                ee = lmh::launch_reverse_term(hh, gsrc, y, XB, gdst, m, K, ld, dt / (double)m, partial, st,
                                              ctx->left_variant == 1 ? 1 : 0);
            }
            gsrc = gdst;
        }
        return ee;
    };

    for (int64_t k = M - 1; k >= 0 && e == hipSuccess; --k) {
        const auto &r = tp.recs[k];
        lmh::DevHam hk = h;                  // H of step k: the recorded field parameters
        hk.field = ncomp ? ctx->d_tape_fields + (size_t)k * ncomp : h.field;
        hk.nfield = ncomp;
        if (e == hipSuccess) {
            ProfScope ps(ctx, LMH_KERNEL_BUILD_H);
            e = lmh::launch_build_h(hk, st);
        }
        const bool cp = k % tp.every == 0;
        // a checkpoint in host memory comes back through a device buffer
        const bool cp_host = cp && tp.cp_host[k / tp.every];
        const double2 *ckpt = cp ? tp.checkpoints[k / tp.every] : nullptr;
        if (fused) {
            double2 *V1t = taken[1];
            // B adds its sums onto A's (same workgroups): nparts partial rows
            lmh::AdjointSide A{Pa, 0, V1t, 1, Lcur, Ybt, 1, ld, r.dt, r.K, partial, nparts, 0};
            lmh::AdjointSide B{cp ? (cp_host ? V1t : ckpt) : V1t, cp ? 1 : 0, Pa, 0, Ybt, Lcur, 0, ld, r.dt,
                               r.K, partial, nparts, 1};
            if (e == hipSuccess) {
                ProfScope ps(ctx, LMH_KERNEL_ADJOINT);
                e = lmh::launch_adjoint_side(hk, A, st);
            }
            // at a checkpoint B reads P_k itself, so A's x in V1t is not needed
            if (e == hipSuccess && cp_host)
                e = hipMemcpyAsync(V1t, ckpt, sizeof(double2) * ld * n, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) {
                ProfScope ps(ctx, LMH_KERNEL_ADJOINT);
                e = lmh::launch_adjoint_side(hk, B, st);
            }
        } else if (r.K > 0) {
            // P_{k+1} = U P_k U^H with P_k and Lambda_{k+1} Hermitian: the right
            // factor's term Re tr(Lambda U P dU^H) is the conjugate of the left
            // one's, Re tr(Lambda dU P U^H), for every dH, so one side's sweep
            // with its F doubled gives dL, and Lambda_k = U^H Lambda_{k+1} U
            // needs no second sweep: U^H (U^H Lambda_{k+1})^H.
            double2 *G1 = taken[1], *G2 = taken[2], *XB = taken[3];
            double2 *XA = spare1, *T2 = spare2, *I2 = Lin;
            if (e == hipSuccess) e = lmh::launch_zero(partial, (int64_t)nparts * nslots, st);
            // x = U^H P_{k+1} (= P_k U^H), cotangent Lambda_{k+1} -> xbar = U^H Lambda_{k+1}
            if (e == hipSuccess) e = left_poly(hk, Pa, XA, T2, -r.dt, r.K);
            if (e == hipSuccess) e = side_sweep(hk, XA, Lcur, XB, G1, G2, T2, I2, r.dt, r.K);
            if (e == hipSuccess) e = lmh::launch_conj_transpose(XB, Ybt, n, ld, st);
            // Lambda_k = U^H (Lambda_{k+1} U), into XB's buffer
            if (e == hipSuccess) e = left_poly(hk, Ybt, XB, T2, -r.dt, r.K);
            // P_k = U^H (P_{k+1} U) with P_{k+1} U = (U^H P_{k+1})^H, or the checkpoint
            if (e == hipSuccess && cp)
                e = cp_host ? hipMemcpyAsync(Pa, ckpt, sizeof(double2) * ld * n, hipMemcpyHostToDevice, st)
                            : lmh::launch_copy(ckpt, Pa, ld * n, st);
            if (e == hipSuccess && !cp) {
                if (ctx->right_variant == 1 || ctx->right_variant == 5) {
                    e = lmh::launch_conj_transpose(XA, G1, n, ld, st);
                    if (e == hipSuccess) e = left_poly(hk, G1, Pa, T2, -r.dt, r.K);
                } else {
                    e = right_poly(hk, XA, Pa, T2, -r.dt, r.K);
                }
            }
            std::swap(taken[3], Lcur);       // Lambda_k now in XB's buffer; the old costate buffer is free
        }
        if (e == hipSuccess && r.K > 0) {
            ProfScope ps(ctx, LMH_KERNEL_ADJOINT_REDUCE);
            e = lmh::launch_adjoint_reduce(hk, partial, nparts, fused ? 1.0 : 2.0, gV, gF + (size_t)k * nblk * ncomp,
                                           st);
        }
        if (e == hipSuccess && ctx->progress && cp) {
            // a checkpoint segment is done once the device has run it
            e = hipStreamSynchronize(st);
            if (e == hipSuccess) ctx->progress(ctx->progress_user, M - k, M);
        }
    }
    // Lambda_0 into the costate's own buffer
    int out_idx = -1;
    for (int q = 0; q < 3; ++q)
        if (lam->buf[q] == Lcur) out_idx = q;
    if (out_idx < 0 && e == hipSuccess) {
        // a pooled buffer holds it, and one of the costate's buffers is among `taken`
        for (size_t q = 0; q < taken.size() && out_idx < 0; ++q)
            for (int b = 0; b < 3; ++b)
                if (taken[q] == lam->buf[b]) {
                    e = lmh::launch_copy(Lcur, lam->buf[b], ld * n, st);
                    taken[q] = Lcur;         // the pooled buffer goes back to the pool
                    out_idx = b;
                    break;
                }
    }
    std::vector<double> hF((size_t)M * nblk * ncomp);
    if (e == hipSuccess && grad_onsite) e = hipMemcpyAsync(grad_onsite, gV, 8 * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && !hF.empty()) e = hipMemcpyAsync(hF.data(), gF, 8 * hF.size(), hipMemcpyDeviceToHost, st);
    // the context's own field is untouched; H was rebuilt from the recorded
    // parameters, so it and the solver are rebuilt on the next update
    ctx->param_version++;
    hipError_t es = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es;
    for (double2 *p : iters) pool_give(ctx, p);
    cleanup();
    if (e != hipSuccess) return hipfail(e, "adjoint sweep");
    if (out_idx < 0) return fail(LMH_ESTATE, "adjoint: costate buffer lost");
    lam->cur = out_idx;
    if (grad_field)
        for (int64_t k = 0; k < M; ++k)
            for (int64_t c = 0; c < ncomp; ++c) {
                double acc = 0.0;
                for (int64_t b = 0; b < nblk; ++b) acc += hF[((size_t)k * nblk + b) * ncomp + c];
                grad_field[k * ncomp + c] = acc;
            }
    return LMH_OK;
}

int lmh_hamiltonian_derivative(lmh_context *ctx, int64_t comp, double *dh) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "lmh_hamiltonian_derivative before lmh_set_geometry");
    if (comp < 0 || comp >= ctx->h.nfield) return fail(LMH_EINVAL, "field component %lld out of range", (long long)comp);
    if (!dh) return fail(LMH_EINVAL, "null output");
    double2 *d = nullptr;
    HIP_TRY(dalloc(&d, (size_t)std::max<int64_t>(1, ctx->h.nbonds)), "derivative buffer");
    hipError_t e = lmh::launch_bond_derivative(ctx->h, comp, d, ctx->stream);
    if (e == hipSuccess && ctx->h.nbonds)
        e = hipMemcpyAsync(dh, d, 16 * ctx->h.nbonds, hipMemcpyDeviceToHost, ctx->stream);
    hipError_t es = hipStreamSynchronize(ctx->stream);
    dfree(d);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return hipfail(e, "bond derivative");
    return LMH_OK;
}

namespace {

// one size class's work items for buffer-rotation phase p and side, from
// the members' current H, K, s and dt
std::vector<lmh::BatchItem> batch_items(const lmh_batch *b, const lmh_batch::Class &c, int p, int side) {
    std::vector<lmh::BatchItem> it(c.members.size());
    for (size_t q = 0; q < c.members.size(); ++q) {
        const int r = c.members[q];
        const lmh_context *cx = b->ctxs[r];
        const lmh_state *s = b->states[r];
        const int a = (b->cur0[r] + 2 * p) % 3;
        const int src = side == 0 ? a : (a + 1) % 3;
        const int dst = side == 0 ? (a + 1) % 3 : (a + 2) % 3;
        const int scr = 3 - src - dst;
        it[q] = {cx->h.col, cx->h.col4, cx->h.rec, cx->h.val, cx->h.diag, s->buf[src], s->buf[dst], s->buf[scr], s->ld,
                 (int)cx->h.n, cx->h.zmax, cx->K, (int)std::lround(cx->s), cx->dt / cx->s,
                 cx->h.drec, cx->h.dval, cx->h.doff, cx->h.ndict};
    }
    return it;
}

// the strip members' tables per (rotation phase, single or doubled
// polynomial), from their current H, K, s and dt (earlier tables freed)
hipError_t batch_strip_tables(lmh_batch *b) {
    hipError_t e = hipSuccess;
    for (int t = 0; t < 6; ++t) {
        dfree(b->strip_args[t]);
        dfree(b->strip_ustart[t]);
    }
    for (int q = 0; q < 3 && e == hipSuccess && !b->strip_members.empty(); ++q)
        for (int dbl = 0; dbl < 2 && e == hipSuccess; ++dbl) {
            std::vector<lmh::StripMemberDesc> md;
            for (int r : b->strip_members) {
                lmh_context *cx = b->ctxs[r];
                lmh_state *st = b->states[r];
                const int src = (b->cur0[r] + q) % 3, dst = (src + 1) % 3;
                md.push_back({&cx->h, st->buf[src], st->buf[dst], st->ld, cx->dt / cx->s, cx->K,
                              (int)std::lround(cx->s) * (dbl ? 2 : 1)});
            }
            e = lmh::strip_batch_table(md.data(), (int)md.size(), &b->strip_args[2 * q + dbl],
                                       &b->strip_ustart[2 * q + dbl], &b->strip_units, &b->strip_lds);
        }
    return e;
}

// A member's H, K, s or dt changed since the tables were made (a time-
// dependent H: the reference evaluates h(t) at every step of every
// realisation, src/evolution.jl:238-250).  The kernel chosen for it at
// creation must still suit it (strip form for the strip launch; value
// classes that hold and fit for a value-class variant); then the tables are
// refilled in place, so that a captured graph stays valid.
int batch_refresh(lmh_batch *b) {
    for (size_t r = 0; r < b->ctxs.size(); ++r) {
        const lmh::DevHam &h = b->ctxs[r]->h;
        const int v = b->member_kernel[r];
        bool ok;
        if (v == lmh::fused_strip_variant()) {
            ok = h.strip_ok && h.sg.s > 0;
        } else if (lmh::fused_variant_is_dict(v)) {
            ok = h.ndict > 0 && h.dict_ok;
            for (const auto &c : b->classes)
                if (c.cls == v) ok = ok && h.ndict <= c.maxdict;
        } else {
            ok = true;
        }
        if (!ok)
            return fail(LMH_ESTATE,
                        "member %zu: its new H no longer suits the kernel chosen at lmh_batch_create (%s); create the "
                        "batch again",
                        r, v == lmh::fused_strip_variant() ? "strip form" : "value classes");
    }
    // the launches that read the tables are done before they change
    HIP_TRY(hipStreamSynchronize(b->stream), "batch sync");
    for (auto &c : b->classes)
        for (int p = 0; p < 3; ++p)
            for (int side = 0; side < 2; ++side) {
                const std::vector<lmh::BatchItem> it = batch_items(b, c, p, side);
                HIP_TRY(hipMemcpy(c.items[2 * p + side], it.data(), sizeof(lmh::BatchItem) * it.size(),
                                  hipMemcpyHostToDevice),
                        "batch tables");
            }
    HIP_TRY(batch_strip_tables(b), "batch strip tables");
    for (size_t r = 0; r < b->ctxs.size(); ++r) {
        b->versions[r] = b->ctxs[r]->solver_version;
        b->serials[r] = b->ctxs[r]->solver_serial;
    }
    return LMH_OK;
}

}  // namespace

int lmh_batch_create(lmh_batch **out, int64_t nmembers, lmh_context *const *ctxs, lmh_state *const *states) {
    if (!out || nmembers <= 0 || !ctxs || !states) return fail(LMH_EINVAL, "invalid batch arguments");
    *out = nullptr;
    const int dev = ctxs[0] ? ctxs[0]->device : -1;
    for (int64_t r = 0; r < nmembers; ++r) {
        lmh_context *c = ctxs[r];
        lmh_state *s = states[r];
        if (!c || !s || s->ctx != c) return fail(LMH_EINVAL, "member %lld: state does not belong to its context", (long long)r);
        if (c->device != dev) return fail(LMH_EINVAL, "batch members must share one device");
        const int rc = check_ctx(c);          // the member's K taken up from the device
        if (rc) return rc;
        if (!c->solver_ready || c->solver_version != c->param_version)
            return fail(LMH_ESTATE, "member %lld: call lmh_update_solver before lmh_batch_create", (long long)r);
        if (!lmh::fused_supported(c->h))
            return fail(LMH_ERANGE, "member %lld: batched step needs N <= 10240", (long long)r);
        if (c->tape.active && c->tape.st == s)
            return fail(LMH_ESTATE, "member %lld: its state is being recorded by lmh_tape_begin", (long long)r);
        for (int64_t q = 0; q < r; ++q)
            if (states[q] == s) return fail(LMH_EINVAL, "state listed twice in a batch");
    }
    HIP_TRY(hipSetDevice(dev), "hipSetDevice");
    for (int64_t r = 0; r < nmembers; ++r) {
        int rc = state_normalize(states[r]);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(ctxs[r]->stream), "member sync");
    }
    lmh_batch *b = new (std::nothrow) lmh_batch();
    if (!b) return fail(LMH_ENOMEM, "host allocation failed");
    b->device = dev;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (int64_t r = 0; r < nmembers; ++r) {
        b->ctxs.push_back(ctxs[r]);
        b->states.push_back(states[r]);
        b->versions.push_back(ctxs[r]->solver_version);
        b->serials.push_back(ctxs[r]->solver_serial);
        b->geoms.push_back(ctxs[r]->geom_version);
        b->cur0.push_back(states[r]->cur);
    }
    // size classes, each with item tables for the 3 buffer-rotation phases
    // a member whose value classes hold (checked by its last H build) joins
    // the value-class variant of its size's shape; a class whose largest
    // column and largest table do not fit the LDS together runs the plain one
    std::map<int, std::vector<int>> byclass;
    b->member_kernel.assign(nmembers, -1);
    const bool use_strip = ctxs[0]->batch_strip != 0;
    for (int64_t r = 0; r < nmembers; ++r) {
        const lmh::DevHam &hr = ctxs[r]->h;
        if (use_strip && hr.strip_ok && hr.sg.s > 0) {
            b->strip_members.push_back((int)r);
            b->member_kernel[r] = lmh::fused_strip_variant();
            continue;
        }
        const int v = lmh::fused_batch_variant(hr.n, hr.zmax), dv = lmh::fused_default_dict(v);
        const bool dict = dv >= 0 && hr.ndict > 0 && hr.dict_ok && lmh::fused_dict_fits(dv, hr.n, hr.ndict);
        byclass[dict ? dv : v].push_back((int)r);
    }
    for (auto it = byclass.begin(); it != byclass.end();) {
        if (!lmh::fused_variant_is_dict(it->first)) {
            ++it;
            continue;
        }
        int64_t mn = 0;
        int md = 0;
        for (int r : it->second) {
            mn = std::max(mn, ctxs[r]->h.n);
            md = std::max(md, ctxs[r]->h.ndict);
        }
        if (lmh::fused_dict_fits(it->first, mn, md)) {
            ++it;
            continue;
        }
        auto &plain = byclass[lmh::fused_variant_fallback(it->first)];
        plain.insert(plain.end(), it->second.begin(), it->second.end());
        std::sort(plain.begin(), plain.end());
        it = byclass.erase(it);
    }
    for (auto &kv : byclass) {
        if (e != hipSuccess) break;
        lmh_batch::Class c;
        c.cls = kv.first;
        c.members = kv.second;
        for (int r : c.members) b->member_kernel[r] = c.cls;
        const int C = lmh::fused_variant_columns(c.cls);
        std::vector<int64_t> bstart(c.members.size() + 1, 0);
        for (size_t q = 0; q < c.members.size(); ++q) {
            const int64_t n = ctxs[c.members[q]]->h.n;
            bstart[q + 1] = bstart[q] + (n + C - 1) / C;
            c.maxn = std::max(c.maxn, n);
            if (lmh::fused_variant_is_dict(c.cls)) c.maxdict = std::max(c.maxdict, ctxs[c.members[q]]->h.ndict);
        }
        c.total = bstart.back();
        e = dalloc(&c.bstart, bstart.size());
        if (e == hipSuccess) e = hipMemcpy(c.bstart, bstart.data(), 8 * bstart.size(), hipMemcpyHostToDevice);
        for (int p = 0; p < 3 && e == hipSuccess; ++p)
            for (int side = 0; side < 2 && e == hipSuccess; ++side) {
                const std::vector<lmh::BatchItem> it = batch_items(b, c, p, side);
                e = dalloc(&c.items[2 * p + side], it.size());
                if (e == hipSuccess)
                    e = hipMemcpy(c.items[2 * p + side], it.data(), sizeof(lmh::BatchItem) * it.size(),
                                  hipMemcpyHostToDevice);
            }
        if (e == hipSuccess) e = lmh::prepare_batched(c.cls);
        b->classes.push_back(c);
    }
    if (e == hipSuccess) e = batch_strip_tables(b);
    if (e == hipSuccess) e = dalloc(&b->d_trace, nmembers);
    if (e == hipSuccess) e = dalloc(&b->d_out, nmembers);
    if (e != hipSuccess) {
        batch_free(b);
        delete b;
        return hipfail(e, "batch setup");
    }
    *out = b;
    return LMH_OK;
}

int lmh_batch_destroy(lmh_batch *b) {
    if (!b) return LMH_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    delete b;
    return LMH_OK;
}

int lmh_batch_step(lmh_batch *b, int64_t nsteps, int use_graph) {
    if (!b || nsteps < 0) return fail(LMH_EINVAL, "invalid batch step");
    HIP_TRY(hipSetDevice(b->device), "hipSetDevice");
    bool changed = false;
    for (size_t r = 0; r < b->ctxs.size(); ++r) {
        lmh_context *c = b->ctxs[r];
        const int rc = check_ctx(c);
        if (rc) return rc;
        if (c->geom_version != b->geoms[r])
            return fail(LMH_ESTATE, "member %zu: its geometry changed since lmh_batch_create", r);
        if (!c->solver_ready || c->solver_version != c->param_version)
            return fail(LMH_ESTATE, "member %zu: its parameters changed without lmh_update_solver", r);
        if (c->solver_version != b->versions[r] || c->solver_serial != b->serials[r]) changed = true;
        if (c->tape.active && c->tape.st == b->states[r])
            return fail(LMH_ESTATE, "member %zu: its state is being recorded by lmh_tape_begin", r);
        if (b->states[r]->cur != batch_expected_cur(b, r) || b->states[r]->pending || b->states[r]->conj)
            return fail(LMH_ESTATE, "member %zu was stepped outside the batch", r);
        HIP_TRY(hipStreamSynchronize(c->stream), "member sync");
    }
    if (changed) {
        const int rc = batch_refresh(b);
        if (rc) return rc;
    }
    // strip members: n + 1 launches (see lmh_batch)
    if (!b->strip_members.empty() && nsteps > 0)
        for (int64_t i = 0; i <= nsteps; ++i) {
            const bool dbl = i > 0 && i < nsteps;
            const bool tout = i < nsteps || (nsteps & 1);
            const int q = (int)(b->strip_rot % 3);
            HIP_TRY(lmh::launch_batched_strip(b->strip_args[2 * q + dbl], b->strip_ustart[2 * q + dbl],
                                              (int)b->strip_members.size(), b->strip_units, b->strip_lds, tout,
                                              b->stream),
                    "batched strip launch");
            ++b->strip_rot;
        }
    const int64_t steps0 = b->steps_done;
    int64_t left = b->classes.empty() ? 0 : nsteps;
    // size classes: eager steps up to the next multiple of 3 (the graph
    // covers one rotation period of the three buffers)
    while (left > 0 && (!use_graph || b->steps_done % 3 != 0 || left < 3)) {
        HIP_TRY(batch_launch_step(b, b->steps_done % 3), "batched step");
        ++b->steps_done;
        --left;
    }
    if (left >= 3) {
        if (!b->exec) {
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "graph capture");
            hipError_t e = hipSuccess;
            for (int p = 0; p < 3 && e == hipSuccess; ++p) e = batch_launch_step(b, p);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(b->stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return hipfail(e != hipSuccess ? e : e2, "graph capture");
            }
            b->graph = g;
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
        }
        while (left >= 3) {
            HIP_TRY(hipGraphLaunch(b->exec, b->stream), "graph launch");
            b->steps_done += 3;
            left -= 3;
        }
        while (left > 0) {
            HIP_TRY(batch_launch_step(b, b->steps_done % 3), "batched step");
            ++b->steps_done;
            --left;
        }
    }
    b->steps_done = steps0 + nsteps;
    HIP_TRY(hipStreamSynchronize(b->stream), "batch sync");
    for (size_t r = 0; r < b->states.size(); ++r) b->states[r]->cur = batch_expected_cur(b, r);
    return LMH_OK;
}

int lmh_batch_member_kernel(lmh_batch *b, int64_t member, int32_t *kernel) {
    if (!b || !kernel) return fail(LMH_EINVAL, "null argument");
    if (member < 0 || member >= (int64_t)b->member_kernel.size()) return fail(LMH_EINVAL, "no member %lld", (long long)member);
    *kernel = b->member_kernel[member];
    return LMH_OK;
}

int lmh_batch_mean_density(lmh_batch *b, double *out) {
    if (!b || !out) return fail(LMH_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(b->device), "hipSetDevice");
    const size_t m = b->states.size();
    std::vector<lmh::BatchTrace> tr(m);
    for (size_t r = 0; r < m; ++r) tr[r] = {b->states[r]->buf[b->states[r]->cur], b->states[r]->ld, b->states[r]->n};
    HIP_TRY(hipMemcpyAsync(b->d_trace, tr.data(), sizeof(lmh::BatchTrace) * m, hipMemcpyHostToDevice, b->stream),
            "trace upload");
    HIP_TRY(lmh::launch_batched_mean_density(b->d_trace, (int)m, b->d_out, b->stream), "mean density");
    HIP_TRY(hipMemcpyAsync(out, b->d_out, 8 * m, hipMemcpyDeviceToHost, b->stream), "readback");
    HIP_TRY(hipStreamSynchronize(b->stream), "batch sync");
    return LMH_OK;
}

// ------------------------------------------------------------------ kets

static void krylov_graphs_free(lmh_kets *k);

int lmh_kets_create(lmh_context *ctx, int64_t m, lmh_kets **out) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out || m <= 0) return fail(LMH_EINVAL, "invalid ket block");
    *out = nullptr;
    if (!ctx->geom_set) return fail(LMH_ESTATE, "lmh_kets_create before lmh_set_geometry");
    lmh_kets *k = new (std::nothrow) lmh_kets();
    if (!k) return fail(LMH_ENOMEM, "host allocation failed");
    k->ctx = ctx;
    k->n = ctx->h.n;
    k->m = m;
    k->ld = ((ctx->h.n + 7) / 8) * 8;
    for (int b = 0; b < 3; ++b) {
        hipError_t e = dalloc(&k->buf[b], (size_t)k->ld * m);
        if (e != hipSuccess) {
            for (int q = 0; q < 3; ++q) dfree(k->buf[q]);
            delete k;
            return hipfail(e, "ket allocation");
        }
    }
    ctx->kets.insert(k);
    *out = k;
    return LMH_OK;
}

int lmh_kets_destroy(lmh_kets *k) {
    if (!k) return LMH_OK;
    if (k->ctx) {
        (void)hipSetDevice(k->ctx->device);
        (void)hipStreamSynchronize(k->ctx->stream);
        k->ctx->kets.erase(k);
    }
    for (int q = 0; q < 3; ++q) dfree(k->buf[q]);
    krylov_graphs_free(k);
    dfree(k->basis);
    dfree(k->lz);
    dfree(k->coef);
    dfree(k->len);
    delete k;
    return LMH_OK;
}

int lmh_kets_upload(lmh_kets *k, const double *psi, int64_t ld) {
    if (!k || !psi || ld < k->n) return fail(LMH_EINVAL, "invalid ket upload");
    int rc = check_ctx(k->ctx);
    if (rc) return rc;
    HIP_TRY(hipMemcpy2DAsync(k->buf[k->cur], k->ld * 16, psi, ld * 16, k->n * 16, k->m, hipMemcpyHostToDevice,
                             k->ctx->stream),
            "ket upload");
    HIP_TRY(hipStreamSynchronize(k->ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

int lmh_kets_download(lmh_kets *k, double *psi, int64_t ld) {
    if (!k || !psi || ld < k->n) return fail(LMH_EINVAL, "invalid ket download");
    int rc = check_ctx(k->ctx);
    if (rc) return rc;
    HIP_TRY(hipMemcpy2DAsync(psi, ld * 16, k->buf[k->cur], k->ld * 16, k->n * 16, k->m, hipMemcpyDeviceToHost,
                             k->ctx->stream),
            "ket download");
    HIP_TRY(hipStreamSynchronize(k->ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

int lmh_kets_step(lmh_context *ctx, lmh_kets *k) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!k || k->ctx != ctx) return fail(LMH_EINVAL, "kets do not belong to this context");
    if (!ctx->solver_ready || ctx->solver_version != ctx->param_version)
        return fail(LMH_ESTATE, "lmh_kets_step without an up-to-date lmh_update_solver");
    const int K = ctx->K;
    if (K == 0) return LMH_OK;
    const int reps = (int)std::lround(ctx->s);
    const lmh::DevHam &h = ctx->h;
    const int a = k->cur, b = (a + 1) % 3, c = (a + 2) % 3;
    if (lmh::fused_supported(h) && h.n <= ctx->fused_max_n && ctx->step_mode != LMH_STEP_PER_PASS) {
        ProfScope ps(ctx, LMH_KERNEL_FUSED);
        ctx->fused_launched = lmh::fused_resolved_variant(h, ctx->fused_variant);
        HIP_TRY(lmh::launch_fused_left(h, k->buf[a], k->buf[b], k->buf[c], k->ld, k->m, ctx->dt / ctx->s, K, reps,
                                       false, ctx->fused_variant, ctx->stream),
                "fused ket launch");
        k->cur = b;
        return LMH_OK;
    }
    int x = a;
    for (int r = 0; r < reps; ++r) {
        int others[2], q = 0;
        for (int i = 0; i < 3; ++i)
            if (i != x) others[q++] = i;
        int in = x;
        for (int kk = K; kk >= 1; --kk) {
            const int o = others[(K - kk) & 1];
            ProfScope ps(ctx, LMH_KERNEL_HORNER_LEFT);
            HIP_TRY(lmh::launch_horner_left(h, k->buf[x], k->buf[in], k->buf[o], k->ld, k->m, ctx->dt / (ctx->s * kk),
                                            kk == K, ctx->left_cols, ctx->stream, ctx->left_variant),
                    "ket horner launch");
            in = o;
        }
        x = in;
    }
    k->cur = x;
    return LMH_OK;
}

static void krylov_graphs_free(lmh_kets *k) {
    for (auto &g : k->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    k->graphs.clear();
}

// Krylov propagation of a ket block (KrylovKit.exponentiate semantics, see
// include/lmhip.h).  Every column runs its own Lanczos recurrence; the device
// advances all of them together and the host checks convergence (the
// eigen-decomposition of T_j per column) first at the basis length the
// previous step needed, then every two iterations.
int lmh_krylov_step(lmh_context *ctx, lmh_kets *k, double dt, int krylovdim, double tol, int maxiter,
                    int64_t *matvecs) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!k || k->ctx != ctx) return fail(LMH_EINVAL, "kets do not belong to this context");
    if (!std::isfinite(dt)) return fail(LMH_EINVAL, "dt must be finite");
    if (krylovdim < 1 || krylovdim > 512) return fail(LMH_EINVAL, "krylovdim must be 1 .. 512");
    if (!(tol > 0.0)) return fail(LMH_EINVAL, "tol must be positive");
    if (maxiter < 1) return fail(LMH_EINVAL, "maxiter must be positive");
    if (k->m > 65535) return fail(LMH_EINVAL, "at most 65535 kets per block");
    if (ctx->h.zmax > 16) return fail(LMH_ERANGE, "the Lanczos kernels take at most 16 ELL slots per row");
    if (matvecs) *matvecs = 0;
    rc = ensure_built(ctx);
    if (rc) return rc;
    const lmh::DevHam &h = ctx->h;
    const int64_t n = k->n, ld = k->ld, m = k->m;
    const int kd = krylovdim;
    const int nb = lmh::lanczos_blocks(n);
    const int64_t vstride = ld * m;
    hipStream_t st = ctx->stream;
    // basis v_0 .. v_{kd-1}, then two scratch blocks for the unnormalised w
    HIP_TRY(grow(k->basis, k->basis_cap, (size_t)(kd + 2) * vstride), "Krylov basis allocation");
    // alpha [kd][m] and beta [kd + 1][m] (contiguous: one readback), three partial arrays [m][nb]
    const size_t nalpha = (size_t)kd * m, nbeta = (size_t)(kd + 1) * m, npart = (size_t)m * nb;
    HIP_TRY(grow(k->lz, k->lz_cap, nalpha + nbeta + 3 * npart), "Krylov work allocation");
    HIP_TRY(grow(k->coef, k->coef_cap, (size_t)m * kd), "Krylov work allocation");
    HIP_TRY(grow(k->len, k->len_cap, (size_t)m), "Krylov work allocation");
    double *d_alpha = k->lz, *d_beta = d_alpha + nalpha, *pre = d_beta + nbeta, *pim = pre + npart, *pw = pim + npart;
    auto V = [&](int j) { return k->basis + (size_t)j * vstride; };
    auto S = [&](int j) { return k->basis + (size_t)(kd + (j & 1)) * vstride; };

    const double T = std::fabs(dt), sgn = dt < 0.0 ? -1.0 : 1.0;
    std::vector<double> rem(m, T);             // time left per ket
    std::vector<char> done(m, T == 0.0 ? 1 : 0);
    std::vector<double> hab(nalpha + nbeta);
    const double *ha = hab.data(), *hb = hab.data() + nalpha;
    std::vector<std::complex<double>> &coef = k->h_coef, y(kd);
    std::vector<int> &len = k->h_len;
    coef.assign((size_t)m * kd, 0.0);
    len.assign(m, 0);
    std::vector<double> taken(m);
    std::vector<double> lam, z;
    int64_t mv = 0;
    int restarts = 0;
    int src = k->cur;
    // T_j of column c from the recorded recurrence coefficients
    auto eigen_of = [&](int64_t c, int j) {
        lam.resize(j);
        std::vector<double> off(j > 1 ? j - 1 : 0);
        for (int q = 0; q < j; ++q) lam[q] = ha[(size_t)q * m + c];
        for (int q = 1; q < j; ++q) off[q - 1] = hb[(size_t)q * m + c];
        return lmh::tridiag_eigen(j, lam, off, z);
    };
    auto any_left = [&]() {
        for (int64_t c = 0; c < m; ++c)
            if (!done[c]) return true;
        return false;
    };
    while (any_left()) {
        if (restarts >= maxiter) return fail(LMH_ERANGE, "Krylov propagation did not converge in %d restarts", maxiter);
        ++restarts;
        for (int64_t c = 0; c < m; ++c) len[c] = done[c] ? -1 : 0;
        int j = 0, checked = 0;
        int target = std::max(1, std::min(kd, k->krylov_len > 0 ? k->krylov_len : 8));
        auto iteration = [&](int jj) {
            return lmh::launch_lanczos_iteration(h, jj == 0 ? k->buf[src] : S(jj), pw, V(jj), jj > 0 ? V(jj - 1) : nullptr,
                                                 S(jj + 1), ld, m, pre, pim, pw, d_alpha + (size_t)jj * m,
                                                 d_beta + (size_t)jj * m, st);
        };
        // the first chunk (start, `target` iterations, beta_target): replayed
        // from a graph when one was captured for this source buffer and length
        auto first_chunk = [&]() {
            hipError_t e = lmh::launch_lanczos_start(k->buf[src], n, ld, m, pw, st);
            for (int jj = 0; jj < target && e == hipSuccess; ++jj) e = iteration(jj);
            if (e == hipSuccess) e = lmh::launch_lanczos_beta(pw, n, m, d_beta + (size_t)target * m, st);
            return e;
        };
        if (ctx->krylov_graphs) {
            if (k->graph_basis != k->basis || k->graph_lz != k->lz || k->graph_kd != kd ||
                k->graph_geom != ctx->geom_version) {
                krylov_graphs_free(k);
                k->graph_basis = k->basis;
                k->graph_lz = k->lz;
                k->graph_kd = kd;
                k->graph_geom = ctx->geom_version;
            }
            lmh_kets::ChunkGraph *g = nullptr;
            for (auto &x : k->graphs)
                if (x.src == k->buf[src] && x.target == target) g = &x;
            if (!g) {
                lmh_kets::ChunkGraph x{k->buf[src], target, nullptr, nullptr};
                HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "Krylov graph capture");
                const hipError_t e = first_chunk();
                const hipError_t e2 = hipStreamEndCapture(st, &x.graph);
                if (e != hipSuccess || e2 != hipSuccess) {
                    if (x.graph) (void)hipGraphDestroy(x.graph);
                    return hipfail(e != hipSuccess ? e : e2, "Krylov graph capture");
                }
                const hipError_t e3 = hipGraphInstantiate(&x.exec, x.graph, nullptr, nullptr, 0);
                if (e3 != hipSuccess) {
                    (void)hipGraphDestroy(x.graph);
                    return hipfail(e3, "Krylov graph instantiate");
                }
                k->graphs.push_back(x);
                g = &k->graphs.back();
            }
            ProfScope ps(ctx, LMH_KERNEL_KRYLOV);
            HIP_TRY(hipGraphLaunch(g->exec, st), "Krylov graph launch");
        } else {
            ProfScope ps(ctx, LMH_KERNEL_KRYLOV);
            HIP_TRY(first_chunk(), "Lanczos first chunk");
        }
        j = target;
        mv += target;
        bool beta_ready = true;                 // beta_j of the current j is on the device
        for (;;) {
            for (; j < target; ++j) {
                ProfScope ps(ctx, LMH_KERNEL_KRYLOV);
                HIP_TRY(iteration(j), "Lanczos iteration");
                ++mv;
                beta_ready = false;
            }
            if (!beta_ready) HIP_TRY(lmh::launch_lanczos_beta(pw, n, m, d_beta + (size_t)j * m, st), "Lanczos norm");
            beta_ready = false;
            HIP_TRY(hipMemcpyAsync(hab.data(), d_alpha, 8 * (nalpha + (size_t)(j + 1) * m), hipMemcpyDeviceToHost, st),
                    "Lanczos readback");
            HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
            bool open = false;
            for (int64_t c = 0; c < m; ++c) {
                if (len[c] != 0) continue;
                const double b0 = hb[c];
                if (b0 == 0.0) {                // the zero ket stays zero
                    len[c] = 1;
                    coef[(size_t)c * kd] = 0.0;
                    taken[c] = rem[c];
                    continue;
                }
                const double tau = rem[c];
                for (int jj = checked + 1; jj <= j && len[c] == 0; ++jj) {
                    if (!eigen_of(c, jj)) return fail(LMH_ERANGE, "Lanczos matrix eigen-decomposition failed");
                    lmh::tridiag_expv(jj, lam, z, sgn * tau, b0, y.data());
                    const double err = hb[(size_t)jj * m + c] * std::abs(y[jj - 1]);
                    if (err <= tol * tau * b0) {
                        len[c] = jj;
                        taken[c] = tau;
                        for (int q = 0; q < jj; ++q) coef[(size_t)c * kd + q] = y[q];
                    }
                }
                if (len[c] == 0) open = true;
            }
            checked = j;
            if (!open || j >= kd) break;
            target = std::min(kd, j + 2);
        }
        // kets the whole basis did not converge for: the longest sub-step that
        // meets the tolerance (halving the time left)
        int longest = 0;
        for (int64_t c = 0; c < m; ++c) {
            if (len[c] == 0) {
                const double b0 = hb[c];
                if (!eigen_of(c, kd)) return fail(LMH_ERANGE, "Lanczos matrix eigen-decomposition failed");
                double tau = rem[c];
                bool ok = false;
                for (int halve = 0; halve < 200 && !ok; ++halve) {
                    lmh::tridiag_expv(kd, lam, z, sgn * tau, b0, y.data());
                    ok = hb[(size_t)kd * m + c] * std::abs(y[kd - 1]) <= tol * tau * b0;
                    if (!ok) tau *= 0.5;
                }
                if (!ok) return fail(LMH_ERANGE, "Krylov propagation did not converge (no sub-step meets the tolerance)");
                len[c] = kd;
                taken[c] = tau;
                for (int q = 0; q < kd; ++q) coef[(size_t)c * kd + q] = y[q];
            }
            longest = std::max(longest, len[c]);
        }
        k->krylov_len = longest;
        HIP_TRY(hipMemcpyAsync(k->coef, coef.data(), sizeof(double2) * (size_t)m * kd, hipMemcpyHostToDevice, st),
                "Krylov coefficient upload");
        HIP_TRY(hipMemcpyAsync(k->len, len.data(), sizeof(int) * (size_t)m, hipMemcpyHostToDevice, st),
                "Krylov coefficient upload");
        const int dst = (src + 1) % 3;
        {
            ProfScope ps(ctx, LMH_KERNEL_KRYLOV);
            HIP_TRY(lmh::launch_krylov_combine(k->basis, vstride, k->coef, kd, k->len, k->buf[src], k->buf[dst], n, ld,
                                               m, st),
                    "Krylov combine");
        }
        src = dst;
        for (int64_t c = 0; c < m; ++c) {
            if (done[c]) continue;
            rem[c] -= taken[c];
            if (rem[c] <= 1e-15 * T) done[c] = 1;
        }
        // coef and len are read by the uploads above: nothing rewrites them
        // (a restart, the next call) before they are done
        HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    k->cur = src;
    if (matvecs) *matvecs = mv;
    return LMH_OK;
}

int lmh_kets_observables(lmh_context *ctx, lmh_kets *k, int64_t col, double *rho, double *J) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!k || k->ctx != ctx) return fail(LMH_EINVAL, "kets do not belong to this context");
    if (col < 0 || col >= k->m) return fail(LMH_EINVAL, "ket index %lld out of range", (long long)col);
    if (ctx->built_version != ctx->param_version) {
        ProfScope ps(ctx, LMH_KERNEL_BUILD_H);
        HIP_TRY(lmh::launch_build_h(ctx->h, ctx->stream), "build_h launch");
        ctx->built_version = ctx->param_version;
    }
    const lmh::DevHam &h = ctx->h;
    if (rho && !ctx->d_rho) HIP_TRY(dalloc(&ctx->d_rho, h.n), "observable allocation");
    if (J && h.nbonds && !ctx->d_J) HIP_TRY(dalloc(&ctx->d_J, h.nbonds), "observable allocation");
    double *dr = rho ? ctx->d_rho : nullptr;
    double *dj = (J && h.nbonds) ? ctx->d_J : nullptr;
    {
        ProfScope ps(ctx, LMH_KERNEL_OBSERVABLES);
        HIP_TRY(lmh::launch_ket_observables(h, k->buf[k->cur] + col * k->ld, dr, dj, ctx->stream),
                "ket observables launch");
    }
    if (dr) HIP_TRY(hipMemcpyAsync(rho, dr, 8 * h.n, hipMemcpyDeviceToHost, ctx->stream), "observables readback");
    if (dj) HIP_TRY(hipMemcpyAsync(J, dj, 8 * h.nbonds, hipMemcpyDeviceToHost, ctx->stream), "observables readback");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

int lmh_set_progress(lmh_context *ctx, lmh_progress_fn fn, void *user) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    ctx->progress = fn;
    ctx->progress_user = user;
    return LMH_OK;
}

int lmh_set_profiling(lmh_context *ctx, int enable) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    resolve_profile(ctx);
    for (int k = 0; k < LMH_NUM_KERNEL_KINDS; ++k) {
        ctx->prof_count[k] = 0;
        ctx->prof_ms[k] = 0.0;
    }
    ctx->stat_steps = ctx->stat_terms = ctx->stat_async = ctx->stat_redone = 0;
    ctx->profiling = enable != 0;
    return LMH_OK;
}

int lmh_get_profile(lmh_context *ctx, int kind, int64_t *launches, double *total_ms) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (kind < 0 || kind >= LMH_NUM_KERNEL_KINDS || !launches || !total_ms) return fail(LMH_EINVAL, "invalid profile query");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    resolve_profile(ctx);
    *launches = ctx->prof_count[kind];
    *total_ms = ctx->prof_ms[kind];
    return LMH_OK;
}

int lmh_set_step_mode(lmh_context *ctx, int mode) {
    if (!ctx) return fail(LMH_EINVAL, "null context");
    if (mode < LMH_STEP_AUTO || mode > LMH_STEP_FUSED_N) return fail(LMH_EINVAL, "invalid step mode %d", mode);
    if (mode == ctx->step_mode) return LMH_OK;
    const int rc = check_ctx(ctx);
    if (rc) return rc;
    ctx->step_mode = mode;
    return LMH_OK;
}

int lmh_set_option(lmh_context *ctx, int key, int64_t value) {
    if (!ctx) return fail(LMH_EINVAL, "null context");
    if (key != LMH_OPT_STEP_MODE) {
        // steps launched on an order result still on the device run on the
        // options they were launched with
        const int rc = check_ctx(ctx);
        if (rc) return rc;
    }
    switch (key) {
        case LMH_OPT_STEP_MODE:
            return lmh_set_step_mode(ctx, (int)value);
        case LMH_OPT_RIGHT_PASS:
            if (value < 1 || value > 5) return fail(LMH_EINVAL, "right-pass variant must be 1 .. 5");
            ctx->right_variant = (int)value;
            return LMH_OK;
        case LMH_OPT_LEFT_PASS:
            if (value < 0 || value > 3) return fail(LMH_EINVAL, "left-pass variant must be 0 .. 3");
            if (value == 2 && ctx->geom_set && !lmh::left_band_fits(ctx->h))
                return fail(LMH_ERANGE, "banded left pass needs H's band <= 1024 rows (band = %lld)",
                            (long long)ctx->h.band);
            if (value == 3 && ctx->geom_set && !lmh::left_band2_fits(ctx->h))
                return fail(LMH_ERANGE, "two-term banded left pass needs H's band <= 512 rows and its halo in LDS "
                            "(band = %lld, zmax = %d)", (long long)ctx->h.band, ctx->h.zmax);
            ctx->left_variant = (int)value;
            return LMH_OK;
        case LMH_OPT_LEFT_COLS:
            if (value < 1 || value > 1024) return fail(LMH_EINVAL, "left columns out of range");
            ctx->left_cols = (int)value;
            return LMH_OK;
        case LMH_OPT_RIGHT_COLS:
            if (value < 1 || value > 1024) return fail(LMH_EINVAL, "right columns out of range");
            ctx->right_cols = (int)value;
            return LMH_OK;
        case LMH_OPT_FUSED_MAX_N:
            if (value < 0) return fail(LMH_EINVAL, "negative size");
            ctx->fused_max_n = value;
            return LMH_OK;
        case LMH_OPT_EIGENSOLVER:
            if (value < 0 || value > 2) return fail(LMH_EINVAL, "eigensolver must be 0, 1 or 2");
            ctx->eigensolver = (int)value;
            return LMH_OK;
        case LMH_OPT_ADJOINT_MODE:
            if (value < 0 || value > 3) return fail(LMH_EINVAL, "adjoint mode must be 0 .. 3");
            ctx->adjoint_mode = (int)value;
            return LMH_OK;
        case LMH_OPT_CHECKPOINTS:
            if (value < 0 || value > 2) return fail(LMH_EINVAL, "checkpoint placement must be 0 .. 2");
            ctx->checkpoint_place = (int)value;
            return LMH_OK;
        case LMH_OPT_KRYLOV_GRAPHS:
            if (value < 0 || value > 1) return fail(LMH_EINVAL, "Krylov graphs must be 0 or 1");
            ctx->krylov_graphs = (int)value;
            return LMH_OK;
        case LMH_OPT_BATCH_STRIP:
            if (value != 0 && value != 1) return fail(LMH_EINVAL, "batch strip option must be 0 or 1");
            ctx->batch_strip = (int)value;
            return LMH_OK;
        case LMH_OPT_DEFER:
            if (value != 0 && value != 1) return fail(LMH_EINVAL, "defer option must be 0 or 1");
            if (!value) {
                // nothing stays deferred once the option is off
                const int rs = ctx_settle_all(ctx);
                if (rs) return rs;
            }
            ctx->defer = (int)value;
            return LMH_OK;
        case LMH_OPT_ASYNC_ORDER:
            if (value != 0 && value != 1) return fail(LMH_EINVAL, "async order option must be 0 or 1");
            ctx->async_order = (int)value;
            return LMH_OK;
        case LMH_OPT_VALUE_CLASSES:
            if (value < 0 || value > 3) return fail(LMH_EINVAL, "value-class probes must be 0..3");
            if (ctx->dict_probes != (int)value) ctx->solver_ready = false;
            ctx->dict_probes = (int)value;
            return LMH_OK;
        case LMH_OPT_FUSED_VARIANT:
            if (value != -1 && (value < 0 || value >= lmh::fused_variant_count()))
                return fail(LMH_EINVAL, "fused variant must be -1 or 0..%d", lmh::fused_variant_count() - 1);
            if (value >= 0 && ctx->geom_set && !lmh::fused_variant_valid((int)value, ctx->h.n, ctx->h.zmax))
                return fail(LMH_ERANGE, "fused variant %lld does not fit N = %lld, zmax = %d", (long long)value,
                            (long long)ctx->h.n, ctx->h.zmax);
            ctx->fused_variant = (int)value;
            return LMH_OK;
        default:
            return fail(LMH_EINVAL, "unknown option %d", key);
    }
}

int lmh_get_option(lmh_context *ctx, int key, int64_t *value) {
    if (!ctx || !value) return fail(LMH_EINVAL, "null argument");
    const int rc = check_ctx(ctx);
    if (rc) return rc;
    switch (key) {
        case LMH_OPT_STEP_MODE: *value = ctx->step_mode; return LMH_OK;
        case LMH_OPT_RIGHT_PASS: *value = ctx->right_variant; return LMH_OK;
        case LMH_OPT_LEFT_COLS: *value = ctx->left_cols; return LMH_OK;
        case LMH_OPT_RIGHT_COLS: *value = ctx->right_cols; return LMH_OK;
        case LMH_OPT_FUSED_MAX_N: *value = ctx->fused_max_n; return LMH_OK;
        case LMH_OPT_EIGENSOLVER: *value = ctx->eigensolver; return LMH_OK;
        case LMH_OPT_FUSED_VARIANT: *value = ctx->fused_variant; return LMH_OK;
        case LMH_OPT_ADJOINT_MODE: *value = ctx->adjoint_mode; return LMH_OK;
        case LMH_OPT_CHECKPOINTS: *value = ctx->checkpoint_place; return LMH_OK;
        case LMH_OPT_KRYLOV_GRAPHS: *value = ctx->krylov_graphs; return LMH_OK;
        case LMH_OPT_VALUE_CLASSES: *value = ctx->dict_probes; return LMH_OK;
        case LMH_OPT_LEFT_PASS: *value = ctx->left_variant; return LMH_OK;
        case LMH_OPT_FUSED_LAUNCHED: *value = ctx->fused_launched; return LMH_OK;
        case LMH_OPT_EIGEN_USED: *value = ctx->eig_method < 0 ? -1 : ctx->eig_method + 4 * ctx->eig_real; return LMH_OK;
        case LMH_OPT_BATCH_STRIP: *value = ctx->batch_strip; return LMH_OK;
        case LMH_OPT_DEFER: *value = ctx->defer; return LMH_OK;
        case LMH_OPT_ASYNC_ORDER: *value = ctx->async_order; return LMH_OK;
        case LMH_OPT_STEP_COUNT: *value = ctx->stat_steps; return LMH_OK;
        case LMH_OPT_STEP_TERMS: *value = ctx->stat_terms; return LMH_OK;
        case LMH_OPT_ASYNC_STEPS: *value = ctx->stat_async; return LMH_OK;
        case LMH_OPT_ASYNC_REDONE: *value = ctx->stat_redone; return LMH_OK;
        default: return fail(LMH_EINVAL, "unknown option %d", key);
    }
}

// Autotuner: times lmh_step on a scratch copy of `st` for each candidate
// configuration (step mode x tile shapes) and keeps the fastest.  K and s are
// the solver's (the reference rule's); only the schedule is tuned.
int lmh_autotune(lmh_context *ctx, lmh_state *st, int32_t reps, double *best_ms) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!st || st->ctx != ctx) return fail(LMH_EINVAL, "state does not belong to this context");
    if (!ctx->solver_ready || ctx->solver_version != ctx->param_version)
        return fail(LMH_ESTATE, "lmh_autotune needs an up-to-date lmh_update_solver");
    if (reps < 1) reps = 1;
    lmh_state *tmp = nullptr;
    rc = lmh_state_create(ctx, &tmp);
    if (rc) return rc;
    rc = lmh_state_copy(tmp, st);
    if (rc) {
        lmh_state_destroy(tmp);
        return rc;
    }
    struct Cfg { int mode, rv, lc, rc_, fv, lv; };
    std::vector<Cfg> cands;
    const bool band = lmh::left_band_fits(ctx->h);
    for (int lc : {16, 32, 64}) {
        // both sides as left passes (banded and gather kernels when the band fits)
        for (int rv : {1, 5}) {
            cands.push_back({LMH_STEP_PER_PASS, rv, lc, 16, -1, 0});
            if (band) cands.push_back({LMH_STEP_PER_PASS, rv, lc, 16, -1, 1});
            if (lmh::left_band2_fits(ctx->h)) cands.push_back({LMH_STEP_PER_PASS, rv, lc, 16, -1, 3});
        }
        for (int rv : {3, 4})
            for (int rcol : {4, 8, 16, 32}) cands.push_back({LMH_STEP_PER_PASS, rv, lc, rcol, -1, 0});
    }
    if (lmh::fused_supported(ctx->h)) {
        for (int v = 0; v < lmh::fused_variant_count(); ++v) {
            // value-class variants only where they would run as themselves
            const lmh::DevHam &hh = ctx->h;
            if (lmh::fused_variant_is_dict(v) &&
                !(hh.ndict > 0 && hh.dict_ok && lmh::fused_dict_fits(v, hh.n, hh.ndict)))
                continue;
            if (v == lmh::fused_strip_variant() && !(hh.strip_ok && hh.sg.s > 0)) continue;
            if (lmh::fused_variant_valid(v, ctx->h.n, ctx->h.zmax))
                cands.push_back({LMH_STEP_FUSED_T, 3, 32, 16, v, 0});
        }
        cands.push_back({LMH_STEP_FUSED_N, 3, 32, 16, -1, 0});
    }
    const Cfg saved = {ctx->step_mode, ctx->right_variant, ctx->left_cols, ctx->right_cols, ctx->fused_variant,
                       ctx->left_variant};
    auto apply = [&](const Cfg &cf) {
        ctx->step_mode = cf.mode;
        ctx->right_variant = cf.rv;
        ctx->left_cols = cf.lc;
        ctx->right_cols = cf.rc_;
        ctx->fused_variant = cf.fv;
        ctx->left_variant = cf.lv;
    };
    const bool prof = ctx->profiling;
    ctx->profiling = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    double best = 1e300;
    Cfg bestc = saved;
    for (const Cfg &cf : cands) {
        if (he != hipSuccess) break;
        apply(cf);
        rc = lmh_step(ctx, tmp);            // warm-up (first launch of a variant)
        if (rc) break;
        (void)hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < reps && !rc; ++r) rc = lmh_step(ctx, tmp);
        if (rc) break;
        (void)hipEventRecord(e1, ctx->stream);
        he = hipEventSynchronize(e1);
        float ms = 0.f;
        if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
        const double per = ms / reps;
        if (he == hipSuccess && per < best) {
            best = per;
            bestc = cf;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    lmh_state_destroy(tmp);
    ctx->profiling = prof;
    if (rc || he != hipSuccess) {
        apply(saved);
        return rc ? rc : hipfail(he, "autotune timing");
    }
    apply(bestc);
    if (best_ms) *best_ms = best;
    return LMH_OK;
}

int lmh_get_stream(lmh_context *ctx, void **stream) {
    if (!ctx || !stream) return fail(LMH_EINVAL, "null argument");
    *stream = (void *)ctx->stream;
    return LMH_OK;
}

int lmh_sync(lmh_context *ctx) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    rc = ctx_settle_all(ctx);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return LMH_OK;
}

}  // extern "C"
