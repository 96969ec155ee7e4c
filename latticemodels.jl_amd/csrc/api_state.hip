// C ABI of liblmhip.so, density matrices: states, the step (with deferred
// passes and device-held orders), observables, region currents and series.

#include "api_internal.h"

using namespace lmhapi;

extern "C" {
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

}  // extern "C"
namespace lmhapi {

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

}  // namespace lmhapi
extern "C" {

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

}  // extern "C"
namespace lmhapi {

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
    const lmh_context::JournalEntry before{st, st->cur, st->conj, st->pending, st->pend};
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
    // (journaled once both launches are queued: a step that failed to launch
    // is not one to run again)
    ctx->journal.push_back(before);
    ctx->stat_async++;
    st->pend = {st->pend_stval, st->pend_sdiag, ctx->dt / ctx->s, K, reps, ctx->async_xreal, st->pend_coef};
    st->pending = true;
    st->cur = b;
    st->conj ^= 1;
    ctx->fused_launched = lmh::fused_strip_variant();
    return LMH_OK;
}

}  // namespace lmhapi
extern "C" {

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

}  // extern "C"
namespace lmhapi {

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

}  // namespace lmhapi
extern "C" {

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
    const Readback rb[2] = {{dr ? rho : nullptr, dr, 8 * (size_t)h.n}, {dj ? J : nullptr, dj, 8 * (size_t)h.nbonds}};
    HIP_TRY(readback_sync(ctx, rb, 2), "observables readback");
    return LMH_OK;
}

}  // extern "C"
namespace lmhapi {

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

}  // namespace lmhapi
extern "C" {

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
