// C ABI of liblmhip.so, core: errors, device memory pools, contexts
// (create, destroy, options, profiling, the autotuner, sync) and the
// context checks every entry point runs.

#include "api_internal.h"

using namespace lmhapi;

namespace lmhapi {
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

hipError_t readback_sync(lmh_context *ctx, const Readback *r, int count) {
    constexpr size_t kHi = (size_t)64 << 20;
    size_t need = 0;
    for (int q = 0; q < count; ++q)
        if (r[q].dst && r[q].bytes <= kHi) need += (r[q].bytes + 255) & ~(size_t)255;
    hipError_t e = hipSuccess;
    if (need > ctx->rb_cap) {
        // (every earlier use ended with a wait: nothing reads the old buffer)
        if (ctx->rb_stage) (void)hipHostFree(ctx->rb_stage);
        ctx->rb_stage = nullptr;
        ctx->rb_cap = 0;
        e = hipHostMalloc((void **)&ctx->rb_stage, need, 0);
        if (e != hipSuccess) return e;
        ctx->rb_cap = need;
    }
    size_t off = 0;
    for (int q = 0; q < count && e == hipSuccess; ++q) {
        if (!r[q].dst || r[q].bytes == 0) continue;
        if (r[q].bytes <= kHi) {
            e = hipMemcpyAsync(ctx->rb_stage + off, r[q].src, r[q].bytes, hipMemcpyDeviceToHost, ctx->stream);
            off += (r[q].bytes + 255) & ~(size_t)255;
        } else {
            e = hipMemcpyAsync(r[q].dst, r[q].src, r[q].bytes, hipMemcpyDeviceToHost, ctx->stream);
        }
    }
    // wait also after a failed copy: the ones issued before it still write
    // into the staging buffer, which the next call may free or reuse
    const hipError_t es = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return e;
    off = 0;
    for (int q = 0; q < count; ++q) {
        if (!r[q].dst || r[q].bytes == 0 || r[q].bytes > kHi) continue;
        std::memcpy(r[q].dst, ctx->rb_stage + off, r[q].bytes);
        off += (r[q].bytes + 255) & ~(size_t)255;
    }
    return hipSuccess;
}

void pool_give(lmh_context *ctx, double2 *p) {
    if (p) ctx->pool.push_back(p);
}

int ctx_resolve(lmh_context *ctx);

// the context's device current; without_resolve: an order result still on
// the device stays there (lmh_step), else it is taken up first
int check_ctx(lmh_context *ctx, bool without_resolve) {
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

}  // namespace lmhapi

// the state's buffer back to P (one conjugate transpose into a free buffer)
// run a deferred pass: buf[cur] <- U buf[cur] (plain stores), cur advances;
// the stored form (conj) already counts the step
namespace lmhapi {
int state_settle(lmh_state *st) {
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

int ctx_settle_all(lmh_context *ctx) {
    for (lmh_state *st : ctx->states) {
        const int rc = state_settle(st);
        if (rc) return rc;
    }
    return LMH_OK;
}

int state_normalize(lmh_state *st) {
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
}  // namespace lmhapi

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

}  // extern "C"
namespace lmhapi {

void csc_release(lmh_context *ctx) {
    dfree(ctx->d_kup);
    dfree(ctx->d_klo);
    dfree(ctx->d_kd);
    dfree(ctx->d_nz);
    dfree(ctx->d_cscflag);
    ctx->csc_mode = ctx->csc_values = false;
    ctx->csc_nnz = 0;
}

}  // namespace lmhapi
extern "C" {

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
    if (ctx->rb_stage) (void)hipHostFree(ctx->rb_stage);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
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
