// C ABI of liblmhip.so, batches: many realisations advanced together (the
// persistent strip launch, per-class launches, captured graphs).

#include "api_internal.h"

using namespace lmhapi;

namespace lmhapi {

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

}  // namespace lmhapi
extern "C" {

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

}  // extern "C"
