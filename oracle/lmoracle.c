/*
 * lmoracle.c -- CPU oracle numerics for density-matrix time evolution.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker / CPU baseline.  The product
 * (liblmhip.so) never links or calls it.
 *
 * Restates LatticeModels.jl 1.0.7 (paths relative to its source tree):
 *   src/evolution.jl:93-128   myexp!: U = exp(factor*H) for sparse H by
 *       (1) s = nextpow(2, max|H_ij| * |factor|), A = H/s;
 *       (2) T_n = droptol!((A*T_{n-1}) .* (factor/n), nztol)   (T_0 = I);
 *       (3) stop after the first n with max|T_n| <= threshold/s;
 *       (4) U = I + sum T_n (in order), then U = U*U, log2(s) times.
 *   src/evolution.jl:73-78    step!(::CachedExp, ::AbstractMatrix): c2 = U*P
 *       (SparseMatrixCSC x Matrix), P = c2*U' (Matrix x Adjoint(Sparse)).
 * The loop orders follow Julia's SparseArrays kernels for those calls, so each
 * output entry is summed in the same order:
 *   sparse*sparse : column j of C from columns k of B in stored order, rows of
 *                   A[:,k] in stored order, dense accumulator per column;
 *   sparse*dense  : C[i,k] += A[i,col]*B[col,k], col ascending;
 *   dense*sparse' : C[r,i] += X[r,c]*conj(A[i,c]), c ascending.
 * norm(A, Inf) of a SparseMatrixCSC is the max |entry| over stored values, and
 * droptol! keeps entries with |x| > tol (|x| = hypot(re, im)).
 *
 * Pins: tests/test_oracle_numerics.py checks U against scipy.linalg.expm
 * (same 1e-10 bar as test/test_timedeps.jl:42-68), K against a brute-force
 * dense count of the stopping rule, unitarity, and the step against dense
 * U P U^H.  The droptol boundary itself is "parity unpinned" (no reference run
 * is possible here).
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b) {
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx cconj(cplx a) { cplx r = {a.re, -a.im}; return r; }
static inline double cabs_(cplx a) { return hypot(a.re, a.im); }

typedef struct {
    int64_t n, nnz, cap;
    int64_t *colptr, *rowval;
    cplx *nzval;
} lmo_csc;

static lmo_csc *csc_alloc(int64_t n, int64_t cap) {
    lmo_csc *m = (lmo_csc *)calloc(1, sizeof(lmo_csc));
    m->n = n;
    m->cap = cap > 0 ? cap : 1;
    m->colptr = (int64_t *)calloc((size_t)(n + 1), sizeof(int64_t));
    m->rowval = (int64_t *)malloc((size_t)m->cap * sizeof(int64_t));
    m->nzval = (cplx *)malloc((size_t)m->cap * sizeof(cplx));
    return m;
}

static void csc_reserve(lmo_csc *m, int64_t need) {
    if (need <= m->cap) return;
    int64_t c = m->cap;
    while (c < need) c *= 2;
    m->rowval = (int64_t *)realloc(m->rowval, (size_t)c * sizeof(int64_t));
    m->nzval = (cplx *)realloc(m->nzval, (size_t)c * sizeof(cplx));
    m->cap = c;
}

void lmo_csc_free(lmo_csc *m) {
    if (!m) return;
    free(m->colptr);
    free(m->rowval);
    free(m->nzval);
    free(m);
}

lmo_csc *lmo_csc_new(int64_t n, const int64_t *colptr, const int64_t *rowval, const double *nzval) {
    int64_t nnz = colptr[n];
    lmo_csc *m = csc_alloc(n, nnz);
    memcpy(m->colptr, colptr, (size_t)(n + 1) * sizeof(int64_t));
    memcpy(m->rowval, rowval, (size_t)nnz * sizeof(int64_t));
    memcpy(m->nzval, nzval, (size_t)nnz * sizeof(cplx));
    m->nnz = nnz;
    return m;
}

int64_t lmo_csc_nnz(const lmo_csc *m) { return m->nnz; }

void lmo_csc_export(const lmo_csc *m, int64_t *colptr, int64_t *rowval, double *nzval) {
    memcpy(colptr, m->colptr, (size_t)(m->n + 1) * sizeof(int64_t));
    memcpy(rowval, m->rowval, (size_t)m->nnz * sizeof(int64_t));
    memcpy(nzval, m->nzval, (size_t)m->nnz * sizeof(cplx));
}

static lmo_csc *csc_identity(int64_t n) {
    lmo_csc *m = csc_alloc(n, n);
    for (int64_t j = 0; j < n; j++) {
        m->colptr[j] = j;
        m->rowval[j] = j;
        m->nzval[j].re = 1.0;
        m->nzval[j].im = 0.0;
    }
    m->colptr[n] = n;
    m->nnz = n;
    return m;
}

static int cmp_i64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* C = A * B (both CSC, n x n), Julia spmatmul summation order. */
static lmo_csc *csc_matmul(const lmo_csc *A, const lmo_csc *B) {
    int64_t n = A->n;
    lmo_csc *C = csc_alloc(n, A->nnz + B->nnz + n);
    cplx *acc = (cplx *)calloc((size_t)n, sizeof(cplx));
    int64_t *mark = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *rows = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) mark[i] = -1;
    int64_t nz = 0;
    for (int64_t j = 0; j < n; j++) {
        int64_t cnt = 0;
        C->colptr[j] = nz;
        for (int64_t kb = B->colptr[j]; kb < B->colptr[j + 1]; kb++) {
            int64_t k = B->rowval[kb];
            cplx b = B->nzval[kb];
            for (int64_t ka = A->colptr[k]; ka < A->colptr[k + 1]; ka++) {
                int64_t i = A->rowval[ka];
                cplx p = cmul(A->nzval[ka], b);
                if (mark[i] != j) {
                    mark[i] = j;
                    rows[cnt++] = i;
                    acc[i] = p;
                } else {
                    acc[i] = cadd(acc[i], p);
                }
            }
        }
        qsort(rows, (size_t)cnt, sizeof(int64_t), cmp_i64);
        csc_reserve(C, nz + cnt);
        for (int64_t t = 0; t < cnt; t++) {
            C->rowval[nz] = rows[t];
            C->nzval[nz] = acc[rows[t]];
            nz++;
        }
    }
    C->colptr[n] = nz;
    C->nnz = nz;
    free(acc);
    free(mark);
    free(rows);
    return C;
}

/* C = A + B elementwise (structural union), values A + B. */
static lmo_csc *csc_add(const lmo_csc *A, const lmo_csc *B) {
    int64_t n = A->n;
    lmo_csc *C = csc_alloc(n, A->nnz + B->nnz);
    int64_t nz = 0;
    for (int64_t j = 0; j < n; j++) {
        C->colptr[j] = nz;
        int64_t a = A->colptr[j], ae = A->colptr[j + 1];
        int64_t b = B->colptr[j], be = B->colptr[j + 1];
        while (a < ae || b < be) {
            if (b >= be || (a < ae && A->rowval[a] < B->rowval[b])) {
                C->rowval[nz] = A->rowval[a];
                C->nzval[nz++] = A->nzval[a++];
            } else if (a >= ae || B->rowval[b] < A->rowval[a]) {
                C->rowval[nz] = B->rowval[b];
                C->nzval[nz++] = B->nzval[b++];
            } else {
                C->rowval[nz] = A->rowval[a];
                C->nzval[nz++] = cadd(A->nzval[a++], B->nzval[b++]);
            }
        }
    }
    C->colptr[n] = nz;
    C->nnz = nz;
    return C;
}

/* in place: values *= f, then droptol (keep |x| > tol); returns max |x| kept */
static double csc_scale_droptol(lmo_csc *m, cplx f, double tol, int do_drop) {
    int64_t nz = 0;
    double mx = 0.0;
    for (int64_t j = 0; j < m->n; j++) {
        int64_t s = m->colptr[j], e = m->colptr[j + 1];
        m->colptr[j] = nz;
        for (int64_t k = s; k < e; k++) {
            cplx v = cmul(m->nzval[k], f);
            double a = cabs_(v);
            if (!do_drop || a > tol) {
                m->rowval[nz] = m->rowval[k];
                m->nzval[nz++] = v;
                if (a > mx) mx = a;
            }
        }
    }
    m->colptr[m->n] = nz;
    m->nnz = nz;
    return mx;
}

static double nextpow2(double x) {
    /* Julia nextpow(2, x) for x > 0: 1 if x <= 1, else smallest 2^n >= x */
    if (x <= 1.0) return 1.0;
    double p = 1.0;
    while (p < x) p *= 2.0;
    return p;
}

/*
 * myexp!(U, H, factor=-i*dt; threshold, nztol).  Writes the number of Taylor
 * terms added (K), the scaling factor s, and max|T_n| for n = 1..K into
 * deltas[0..K-1] (at most max_terms).  Returns NULL for a zero operator or
 * dt == 0 (U = I then).  use_droptol = 0 gives the same series without drops.
 */
lmo_csc *lmo_myexp(const lmo_csc *H, double dt, double threshold, double nztol, int use_droptol,
                   int *K_out, double *s_out, double *deltas, int max_terms) {
    int64_t n = H->n;
    double mat_norm = 0.0;
    for (int64_t k = 0; k < H->nnz; k++) {
        double a = cabs_(H->nzval[k]);
        if (a > mat_norm) mat_norm = a;
    }
    *K_out = 0;
    *s_out = 1.0;
    cplx factor = {0.0, -dt};
    double absf = fabs(dt);
    if (mat_norm == 0.0 || dt == 0.0) return csc_identity(n);
    double s = nextpow2(mat_norm * absf);
    *s_out = s;
    lmo_csc *A = lmo_csc_new(n, H->colptr, H->rowval, (const double *)H->nzval);
    for (int64_t k = 0; k < A->nnz; k++) {
        A->nzval[k].re /= s;
        A->nzval[k].im /= s;
    }
    lmo_csc *U = csc_identity(n);
    lmo_csc *T = csc_identity(n);
    double delta = 1.0;
    int nterm = 1;
    while (delta > threshold / s) {
        lmo_csc *T2 = csc_matmul(A, T);
        lmo_csc_free(T);
        T = T2;
        cplx fn = {factor.re / nterm, factor.im / nterm};
        delta = csc_scale_droptol(T, fn, nztol, use_droptol);
        if (nterm - 1 < max_terms) deltas[nterm - 1] = delta;
        lmo_csc *U2 = csc_add(U, T);
        lmo_csc_free(U);
        U = U2;
        *K_out = nterm;
        nterm++;
        if (nterm > 1000) break;
    }
    lmo_csc_free(T);
    lmo_csc_free(A);
    for (double q = s; q > 1.0; q /= 2.0) {
        lmo_csc *U2 = csc_matmul(U, U);
        lmo_csc_free(U);
        U = U2;
    }
    return U;
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* C[:, cols] = U * P[:, cols] for columns c0..c1-1 (C zeroed there first). */
void lmo_left_cols(const lmo_csc *U, int64_t n, const double *Pd, double *Cd, int64_t c0, int64_t c1, int nthreads) {
    const cplx *P = (const cplx *)Pd;
    cplx *C = (cplx *)Cd;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t k = c0; k < c1; k++) {
        cplx *ck = C + k * n;
        const cplx *pk = P + k * n;
        for (int64_t i = 0; i < n; i++) ck[i].re = ck[i].im = 0.0;
        for (int64_t col = 0; col < n; col++) {
            cplx b = pk[col];
            for (int64_t j = U->colptr[col]; j < U->colptr[col + 1]; j++) {
                int64_t r = U->rowval[j];
                ck[r] = cadd(ck[r], cmul(U->nzval[j], b));
            }
        }
    }
}

/* C[rows r0..r1-1, :] = X[rows, :] * U'  (C zeroed on those rows first). */
void lmo_right_rows(const lmo_csc *U, int64_t n, const double *Xd, double *Cd, int64_t r0, int64_t r1, int nthreads) {
    const cplx *X = (const cplx *)Xd;
    cplx *C = (cplx *)Cd;
    set_threads(nthreads);
    const int64_t RB = 64;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t rb = r0; rb < r1; rb += RB) {
        int64_t re = rb + RB < r1 ? rb + RB : r1;
        for (int64_t c = 0; c < n; c++)
            for (int64_t r = rb; r < re; r++) C[c * n + r].re = C[c * n + r].im = 0.0;
        for (int64_t colB = 0; colB < n; colB++) {
            for (int64_t k = U->colptr[colB]; k < U->colptr[colB + 1]; k++) {
                int64_t rowB = U->rowval[k];
                cplx w = cconj(U->nzval[k]);
                cplx *cdst = C + rowB * n;
                const cplx *xsrc = X + colB * n;
                for (int64_t r = rb; r < re; r++) cdst[r] = cadd(cdst[r], cmul(xsrc[r], w));
            }
        }
    }
}

/* Y = U * X for a thin X (n x m, column-major), same summation order as
 * lmo_left_cols. */
void lmo_spmm(const lmo_csc *U, int64_t n, int64_t m, const double *Xd, double *Yd, int nthreads) {
    const cplx *X = (const cplx *)Xd;
    cplx *Y = (cplx *)Yd;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t k = 0; k < m; k++) {
        cplx *yk = Y + k * n;
        const cplx *xk = X + k * n;
        for (int64_t i = 0; i < n; i++) yk[i].re = yk[i].im = 0.0;
        for (int64_t col = 0; col < n; col++) {
            cplx b = xk[col];
            for (int64_t j = U->colptr[col]; j < U->colptr[col + 1]; j++) {
                int64_t r = U->rowval[j];
                yk[r] = cadd(yk[r], cmul(U->nzval[j], b));
            }
        }
    }
}

/* step!(::CachedExp, P, (c1, c2)): P <- (U*P)*U'.  work holds c2 (n*n complex). */
void lmo_step(const lmo_csc *U, int64_t n, double *P, double *work, int nthreads) {
    lmo_left_cols(U, n, P, work, 0, n, nthreads);
    lmo_right_rows(U, n, work, P, 0, n, nthreads);
}

/*
 * Horner application of the same truncated series without droptol:
 *   left : Y_K = X; Y_{k-1} = X + (f/k) A Y_k;   right: same on the row index
 * with conj(f) and A' = A (Hermitian), repeated s times per side for s > 1.
 * Algorithm-matched CPU variant for timing and for the droptol study.
 */
static void horner_left(const lmo_csc *H, int64_t n, double scale, double dt, int K, int reps,
                        const cplx *X, cplx *Y, cplx *T, int64_t k) {
    cplx *yk = Y + k * n;
    cplx *tk = T + k * n;
    cplx *xk = (cplx *)malloc((size_t)n * sizeof(cplx));
    for (int64_t i = 0; i < n; i++) xk[i] = X[k * n + i];
    for (int r = 0; r < reps; r++) {
        for (int64_t i = 0; i < n; i++) yk[i] = xk[i];
        for (int t = K; t >= 1; t--) {
            for (int64_t i = 0; i < n; i++) tk[i].re = tk[i].im = 0.0;
            for (int64_t col = 0; col < n; col++) {
                cplx b = yk[col];
                for (int64_t j = H->colptr[col]; j < H->colptr[col + 1]; j++) {
                    int64_t rr = H->rowval[j];
                    tk[rr] = cadd(tk[rr], cmul(H->nzval[j], b));
                }
            }
            double c = dt / (scale * t);   /* (f/k) * acc with f = -i dt: (c*acc.im, -c*acc.re) */
            for (int64_t i = 0; i < n; i++) {
                yk[i].re = xk[i].re + c * tk[i].im;
                yk[i].im = xk[i].im - c * tk[i].re;
            }
        }
        for (int64_t i = 0; i < n; i++) xk[i] = yk[i];
    }
    free(xk);
}

/*
 * P <- U P U' with U = p(-i dt H / s)^s applied as Horner passes: the left
 * side is a column-independent pass W = U P; the right side uses
 * (W U') = (U W')' so it is the same column pass on W' followed by a
 * conjugate transpose.
 */
void lmo_horner_step(const lmo_csc *H, int64_t n, double dt, int K, double s, double *Pd,
                     double *w1d, double *w2d, int nthreads) {
    cplx *P = (cplx *)Pd, *W1 = (cplx *)w1d, *W2 = (cplx *)w2d;
    set_threads(nthreads);
    int reps = (int)(s + 0.5);
    for (int side = 0; side < 2; side++) {
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t k = 0; k < n; k++) horner_left(H, n, s, dt, K, reps, P, W1, W2, k);
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < n; j++)
            for (int64_t i = 0; i < n; i++) P[j * n + i] = cconj(W1[i * n + j]);
    }
}
