// C ABI of liblmhip.so, kets: Taylor and Krylov (Lanczos) propagation of
// blocks of state vectors and their observables.

#include <cstring>

#include "api_internal.h"

using namespace lmhapi;

extern "C" {

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
    if (k->h_pin) (void)hipHostFree(k->h_pin);
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

}  // extern "C"
namespace lmhapi {
void krylov_graphs_free(lmh_kets *k) {
    for (auto &g : k->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    k->graphs.clear();
}
}  // namespace lmhapi
extern "C" {

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
    // alpha [kd][m] and beta [kd + 1][m] (contiguous: one readback), three partial arrays [m][nb]
    const size_t nalpha = (size_t)kd * m, nbeta = (size_t)(kd + 1) * m, npart = (size_t)m * nb;
    // (the previous call's combination may still run: buffers are replaced
    // only after it)
    const auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t pin_ab = al(8 * (nalpha + nbeta)), pin_coef = al(16 * (size_t)m * kd), pin_need = pin_ab + pin_coef + al(4 * (size_t)m);
    if ((size_t)(kd + 2) * vstride > k->basis_cap || nalpha + nbeta + 3 * npart > k->lz_cap ||
        (size_t)m * kd > k->coef_cap || (size_t)m > k->len_cap || pin_need > k->h_pin_cap)
        HIP_TRY(hipStreamSynchronize(st), "hipStreamSynchronize");
    // basis v_0 .. v_{kd-1}, then two scratch blocks for the unnormalised w
    HIP_TRY(grow(k->basis, k->basis_cap, (size_t)(kd + 2) * vstride), "Krylov basis allocation");
    HIP_TRY(grow(k->lz, k->lz_cap, nalpha + nbeta + 3 * npart), "Krylov work allocation");
    HIP_TRY(grow(k->coef, k->coef_cap, (size_t)m * kd), "Krylov work allocation");
    HIP_TRY(grow(k->len, k->len_cap, (size_t)m), "Krylov work allocation");
    if (pin_need > k->h_pin_cap) {
        if (k->h_pin) (void)hipHostFree(k->h_pin);
        k->h_pin = nullptr;
        k->h_pin_cap = 0;
        HIP_TRY(hipHostMalloc((void **)&k->h_pin, pin_need, 0), "Krylov staging allocation");
        k->h_pin_cap = pin_need;
    }
    double *const pin_abv = reinterpret_cast<double *>(k->h_pin);
    std::complex<double> *const pin_coefv = reinterpret_cast<std::complex<double> *>(k->h_pin + pin_ab);
    int *const pin_lenv = reinterpret_cast<int *>(k->h_pin + pin_ab + pin_coef);
    double *d_alpha = k->lz, *d_beta = d_alpha + nalpha, *pre = d_beta + nbeta, *pim = pre + npart, *pw = pim + npart;
    auto V = [&](int j) { return k->basis + (size_t)j * vstride; };
    auto S = [&](int j) { return k->basis + (size_t)(kd + (j & 1)) * vstride; };

    const double T = std::fabs(dt), sgn = dt < 0.0 ? -1.0 : 1.0;
    std::vector<double> rem(m, T);             // time left per ket
    std::vector<char> done(m, T == 0.0 ? 1 : 0);
    const double *ha = pin_abv, *hb = pin_abv + nalpha;
    std::vector<std::complex<double>> &coef = k->h_coef, y(kd);
    std::vector<int> &len = k->h_len;
    coef.assign((size_t)m * kd, 0.0);
    len.assign(m, 0);
    std::vector<double> taken(m);
    std::vector<double> lam, z;
    int64_t mv = 0;
    int restarts = 0;
    int src = k->cur;
    // T_j of column c from the recorded recurrence coefficients: its
    // eigen-decomposition (ENDS: eigenvalues and the eigenvectors' first and
    // last components, all the error estimate needs)
    auto eigen_of = [&](int64_t c, int j, bool ends) {
        lam.resize(j);
        std::vector<double> off(j > 1 ? j - 1 : 0);
        for (int q = 0; q < j; ++q) lam[q] = ha[(size_t)q * m + c];
        for (int q = 1; q < j; ++q) off[q - 1] = hb[(size_t)q * m + c];
        return ends ? lmh::tridiag_eigen_ends(j, lam, off, z) : lmh::tridiag_eigen(j, lam, off, z);
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
            HIP_TRY(hipMemcpyAsync(pin_abv, d_alpha, 8 * (nalpha + (size_t)(j + 1) * m), hipMemcpyDeviceToHost, st),
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
                    // the estimate from T_jj's eigenvalues and end rows (the
                    // same bits as from the whole decomposition)
                    if (!eigen_of(c, jj, true)) return fail(LMH_ERANGE, "Lanczos matrix eigen-decomposition failed");
                    const double err = hb[(size_t)jj * m + c] * std::abs(lmh::tridiag_expv_last(jj, lam, z, sgn * tau, b0));
                    if (err <= tol * tau * b0) {
                        if (!eigen_of(c, jj, false)) return fail(LMH_ERANGE, "Lanczos matrix eigen-decomposition failed");
                        lmh::tridiag_expv(jj, lam, z, sgn * tau, b0, y.data());
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
                if (!eigen_of(c, kd, false)) return fail(LMH_ERANGE, "Lanczos matrix eigen-decomposition failed");
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
        // through the pinned staging (its previous uploads ended before the
        // wait above)
        std::memcpy(pin_coefv, coef.data(), sizeof(double2) * (size_t)m * kd);
        std::memcpy(pin_lenv, len.data(), sizeof(int) * (size_t)m);
        HIP_TRY(hipMemcpyAsync(k->coef, pin_coefv, sizeof(double2) * (size_t)m * kd, hipMemcpyHostToDevice, st),
                "Krylov coefficient upload");
        HIP_TRY(hipMemcpyAsync(k->len, pin_lenv, sizeof(int) * (size_t)m, hipMemcpyHostToDevice, st),
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
        // (no wait here: the combination runs on while the caller queues
        // what follows; a restart waits for its readback first)
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
    const Readback rb[2] = {{dr ? rho : nullptr, dr, 8 * (size_t)h.n}, {dj ? J : nullptr, dj, 8 * (size_t)h.nbonds}};
    HIP_TRY(readback_sync(ctx, rb, 2), "observables readback");
    return LMH_OK;
}

}  // extern "C"
