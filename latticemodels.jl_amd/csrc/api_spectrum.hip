// C ABI of liblmhip.so, spectra: the eigensolver guard and the T = 0 / finite
// T density matrix (Fermi projector) built from the eigensystem.

#include "api_internal.h"

using namespace lmhapi;

namespace lmhapi {

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

}  // namespace lmhapi

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
