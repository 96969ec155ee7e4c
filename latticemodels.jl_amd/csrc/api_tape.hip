// C ABI of liblmhip.so, gradients: the tape of recorded steps with its
// checkpoints and the reverse sweep (adjoint) through them.

#include "api_internal.h"

using namespace lmhapi;

namespace lmhapi {
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

}  // namespace lmhapi

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

}  // extern "C"
