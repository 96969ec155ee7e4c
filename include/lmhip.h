/*
 * lmhip.h -- C ABI of liblmhip.so, the MI355X (gfx950) density-matrix time
 * evolution library for tight-binding lattice models.
 *
 * It replaces the hot path of LatticeModels.jl 1.0.7's
 * `Evolution(t -> H(t), P0)` with the default CachedExp solver
 * (src/evolution.jl:62-128, 238-275).  A Julia `HIPExp <: EvolutionSolver`
 * (a thin ccall layer) forwards the three EvolutionSolver methods
 * (src/evolution.jl:5-33) to the entry points below:
 *   update_solver!(solver, H, dt, force)  -> lmh_set_field / lmh_set_onsite +
 *                                            lmh_update_solver (parametric H), or
 *                                            lmh_set_csc_pattern / lmh_set_csc_values +
 *                                            lmh_update_solver (the reference's own
 *                                            SparseMatrixCSC)
 *   step!(solver, P, cache)               -> lmh_step
 *   evolution_cache(solver, P)            -> lmh_state_create (library-owned)
 *
 * Conventions
 *   - Complex numbers are interleaved (re, im) doubles; dense matrices are
 *     column-major with a leading dimension, as Julia's Matrix{ComplexF64}.
 *   - Site and bond indices are 0-based.
 *   - Every function returns 0 on success or one of the LMH_E* codes; the
 *     message of the last failure on the calling thread is lmh_last_error().
 *     LMH_EINVAL corresponds to the reference's ArgumentError.
 *   - A context owns one HIP stream on one device and is not thread-safe;
 *     separate contexts are independent.  Host pointers are borrowed only for
 *     the duration of a call.
 */
#ifndef LMHIP_H
#define LMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LMH_OK 0
#define LMH_EINVAL 1     /* invalid argument (ArgumentError)                */
#define LMH_ENOMEM 2     /* device or host allocation failed                */
#define LMH_EDEVICE 3    /* HIP runtime / library error                     */
#define LMH_ESTATE 4     /* call out of order (e.g. step before update)     */
#define LMH_ERANGE 5     /* a fixed-capacity limit was exceeded             */

/* Field component kinds for lmh_set_field. params holds 3 doubles each:
 *   LMH_FIELD_LANDAU        {B, 0, 0}    src/zoo/magneticfields.jl:11-15
 *   LMH_FIELD_SYMMETRIC     {B, 0, 0}    src/zoo/magneticfields.jl:27-31
 *   LMH_FIELD_POINT_AXIAL   {flux, x, y} src/zoo/magneticfields.jl:66-85
 *   LMH_FIELD_POINT_SINGULAR{flux, x, y} src/zoo/magneticfields.jl:89-104
 * A list of components is a FieldSum (src/operators/magneticfield.jl:100-104),
 * summed left to right; PointFluxes are one component per flux.  Periodic
 * image expansion (src/zoo/magneticfields.jl:237-266) is done by the caller. */
#define LMH_FIELD_LANDAU 1
#define LMH_FIELD_SYMMETRIC 2
#define LMH_FIELD_POINT_AXIAL 3
#define LMH_FIELD_POINT_SINGULAR 4

typedef struct lmh_context lmh_context;
typedef struct lmh_state lmh_state;

/* Library version (major*10000 + minor*100 + patch). */
int lmh_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char *lmh_last_error(void);

/* Number of visible HIP devices. */
int lmh_device_count(int *count);

/* Context on `device` with its own stream.  Replaces the CachedExp solver
 * object (src/evolution.jl:62-68). */
int lmh_create(lmh_context **ctx, int device);
int lmh_destroy(lmh_context *ctx);

/*
 * Lattice geometry and hopping terms: one entry per bond produced by the
 * reference's bond iteration (src/core/bonds.jl:415-422 through
 * add_term!, src/operators/constructoperator.jl:14-38), in that order.
 * Bond b contributes H[i,j] += amp_b * twist_b * exp(-2πi ∫_{r_i}^{r_j} A·dl)
 * and, when i != j, the conjugate to H[j,i] (src/operators/builder.jl:282-309).
 *   r_i, r_j : 2*nbonds doubles, r_j UNWRAPPED across periodic boundaries
 *              (ResolvedSite.old_site, src/core/boundaries.jl:280-289)
 *   amp      : 2*nbonds doubles (complex hopping amplitude)
 *   twist    : 2*nbonds doubles (boundary twist factor; 1+0i if NULL)
 * Replaces the host-side H assembly of src/operators/builder.jl:86-151,
 * 296-309: the phases are recomputed on the device from the field.
 * Internal degrees of freedom (N states per site, src/operators/system.jl:16):
 * nsites counts matrix indices (site i, state a -> i N + a); a hopping block
 * B is one bond per nonzero B[a,b] (row-major, builder.jl:58-63) with the
 * lattice bond's coordinates, and the off-diagonal part of a Hermitian
 * on-site block one bond per element a < b with r_i == r_j (phase 1).
 */
int lmh_set_geometry(lmh_context *ctx, int64_t nsites, int64_t nbonds,
                     const int64_t *bond_i, const int64_t *bond_j,
                     const double *r_i, const double *r_j,
                     const double *amp, const double *twist);

/* Optional: site coordinates xy[2*nsites] (the lattice's site coordinates,
 * src/lattices/bravais/unitcell.jl:119-122).  Used only to order the columns
 * of the per-term right pass by locality (Z-order); results do not depend
 * on it. */
int lmh_set_site_coords(lmh_context *ctx, const double *xy);

/* On-site energies: diag[i] = sum of the on-site terms of site i
 * (src/operators/constructoperator.jl:21-30).  NULL means none. */
int lmh_set_onsite(lmh_context *ctx, const double *diag);

/* Gauge field as a list of components (see LMH_FIELD_*).  ncomp = 0 is
 * NoField.  Cheap: only the parameters travel; H(t) is rebuilt on the device
 * by the next lmh_update_solver. */
int lmh_set_field(lmh_context *ctx, int64_t ncomp, const int32_t *kind, const double *params);

/*
 * The Hamiltonian as a sparse matrix instead of lattice + terms + field: the
 * SparseMatrixCSC{ComplexF64} that the reference's own builders produce
 * (to_operator, src/operators/builder.jl:317-318, held in Hamiltonian.data,
 * src/operators/system.jl:380-386) and that Evolution hands to
 * update_solver! as _data(H) (src/evolution.jl:244).
 *
 * lmh_set_csc_pattern fixes the structure: n x n, colptr[n+1] and
 * rowval[nnz] as in Julia's SparseMatrixCSC (rows strictly increasing in each
 * column), index base 1 (Julia's arrays, passed as they are) or 0.  The
 * pattern must be symmetric; its entries above the diagonal become the bonds
 * of the context's geometry, in column order (bond b = the b-th stored entry
 * H[i,j] with i < j, j ascending, then i ascending; that is the order of the
 * bond currents lmh_observables returns).  It replaces any geometry, field and
 * on-site terms of the context and must be followed by lmh_set_csc_values
 * before lmh_update_solver.  Calling it again with another pattern is
 * allowed (existing states must keep their size); lmh_set_geometry leaves
 * the CSC mode.
 *
 * lmh_set_csc_values uploads the nnz stored values (interleaved complex, in
 * the pattern's order: Julia's nonzeros(H)); H must be Hermitian exactly as
 * the reference's builders make it (H[j,i] stored as the conjugate of
 * H[i,j], real diagonal), else LMH_EINVAL and lmh_update_solver refuses to run
 * until valid values arrive.  The values replace the bond amplitudes and
 * on-site energies on the device (one copy and one gather kernel); the next
 * lmh_update_solver rebuilds H and its Taylor order.  lmh_set_field (with
 * components) and lmh_set_onsite return LMH_ESTATE in this mode, and so does
 * lmh_tape_begin: the adjoint differentiates the parametric Hamiltonian.
 * The reference's cache test (src/evolution.jl:83-92: same matrix object or
 * an equal one, and dt ≈ the previous dt) is the caller's: an unchanged
 * matrix needs no lmh_set_csc_values call.
 */
int lmh_set_csc_pattern(lmh_context *ctx, int64_t n, const int64_t *colptr, const int64_t *rowval, int32_t base);
int lmh_set_csc_values(lmh_context *ctx, int64_t nnz, const double *nzval);

/*
 * update_solver!(solver, H, dt, force) (src/evolution.jl:83-92): rebuilds
 * H(t) on the device when the field, on-site terms or geometry changed since
 * the last call, and selects the Taylor order K and scaling s with the
 * reference's myexp! rule (src/evolution.jl:93-118: s = nextpow(2, max|H|dt),
 * K = first n with max|T_n| <= threshold/s, T_n droptol'd at nztol).
 * Skipped (cache hit) when nothing changed and dt is ≈ the previous dt,
 * unless force != 0.  K_out / s_out may be NULL.
 *
 * With K_out and s_out both NULL the call need not wait for the GPU
 * (LMH_OPT_ASYNC_ORDER): when the steps run on the strip kernel with
 * deferred passes (LMH_OPT_DEFER, grid-stencil H, N <= 10240) the order
 * search's result stays on the device and lmh_step's launch reads K, s and
 * the Horner coefficients from there; the host takes the result up at the
 * next call that needs it (the next update, a read, lmh_sync,
 * lmh_solver_order).  If the new H no longer suits the launch chosen (not
 * strip-shaped, a change in the x bonds' form, K = 0, or a search that
 * needs larger tables) that launch does nothing and the step runs again,
 * on the right path, when the result is taken up -- with the same results
 * as a blocking update.  An order-search error then comes back from that
 * later call.
 */
int lmh_update_solver(lmh_context *ctx, double dt, int force, double threshold, double nztol,
                      int32_t *K_out, double *s_out);
/* K and s of the last lmh_update_solver (waits for them if they are still on
 * the device).  No separate counterpart in the reference: CachedExp keeps
 * U, not K (src/evolution.jl:49-68). */
int lmh_solver_order(lmh_context *ctx, int32_t *K, double *s);

/* A density matrix owned by the library (P plus its two evolution caches,
 * src/evolution.jl:80,188-196).  Contents undefined until uploaded. */
int lmh_state_create(lmh_context *ctx, lmh_state **st);
int lmh_state_destroy(lmh_state *st);
/* Host <-> device copies of P (column-major, leading dimension ld >= N). */
int lmh_state_upload(lmh_state *st, const double *P, int64_t ld);
int lmh_state_download(lmh_state *st, double *P, int64_t ld);
/* Device-side copy (the Evolution constructor copies each state,
 * src/evolution.jl:193). */
int lmh_state_copy(lmh_state *dst, const lmh_state *src);
/* Sparse access to P without moving N^2 values: set_entries makes P the
 * matrix with the given entries (repeated (row, col) pairs add, all others
 * 0), get_entries reads out[k] = P[rows[k], cols[k]] (complex, interleaved).
 * For losses of a few entries of P (localdensity, DensityCurrents on a cut)
 * and their cotangents -- the reference reads such entries one by one
 * (src/operators/latticeutils.jl:41-45, src/zoo/currents.jl:81-102). */
int lmh_state_set_entries(lmh_state *st, int64_t count, const int64_t *rows, const int64_t *cols, const double *vals);
int lmh_state_get_entries(lmh_state *st, int64_t count, const int64_t *rows, const int64_t *cols, double *out);
/* max |P P - P| (idempotency error of a pure-state projector), through a
 * dense rocBLAS product on the device.  An accuracy diagnostic with no
 * counterpart in the reference. */
int lmh_state_idempotency(lmh_state *st, double *max_err);
/* Current device pointer and leading dimension of P (changes after a step;
 * valid until the next call on this state). */
int lmh_state_device_ptr(lmh_state *st, void **dptr, int64_t *ld);

/* step!(::CachedExp, P, cache) (src/evolution.jl:73-78): P <- U P U' with
 * U the order-K Taylor polynomial of exp(-i dt H) selected by the last
 * lmh_update_solver, applied as Horner passes on the device (part of the
 * work may be deferred to the next step or read: LMH_OPT_DEFER). */
int lmh_step(lmh_context *ctx, lmh_state *st);

/* localdensity(P) (src/operators/latticeutils.jl:41-45) into rho[N] and
 * DensityCurrents(H, P)[i, j] = 2 Im(H[i,j] P[j,i]) (src/zoo/currents.jl:81-102)
 * for every geometry bond into J[nbonds], using the H of the current
 * parameters -- after lmh_update_solver + lmh_step that is the H of the step
 * (the H yielded with P by the iterator, src/evolution.jl:266-275).
 * Either output may be NULL.  Raw values: no |J| < 1e-10 dropping. */
int lmh_observables(lmh_context *ctx, lmh_state *st, double *rho, double *J);

/* Download the device-built H: diag[N], and for every bond b the entry
 * H[i_b, j_b] into hij[2*nbonds].  For parity checks of the in-kernel
 * Peierls phases against src/operators/builder.jl:282-309. */
int lmh_get_hamiltonian(lmh_context *ctx, double *diag, double *hij);

/*
 * densitymatrix(H, mu = mu; T = T) of a one-particle system with dense
 * diagonalisation (src/spectrum.jl:226-232, 253-257, 317-333), written into
 * `st`: P0 = V diag(f(E)) V' with f(E) = 1 if E - mu == 0 at T == 0 and the
 * Fermi-Dirac occupation otherwise.  H is the device-built Hamiltonian of the
 * current geometry / on-site / field parameters.  The eigensystem comes from
 * rocSOLVER zheevd (ascending E); the projector product is an FP64 MFMA
 * kernel.  Optional outputs: the number of columns with f > 0 and the
 * eigenvalues E[N] (ascending).
 */
int lmh_densitymatrix(lmh_context *ctx, lmh_state *st, double mu, double T, int64_t *nocc, double *E);

/*
 * The rest of the densitymatrix family (src/spectrum.jl:226-288) splits the
 * same way the reference does: `diagonalize(dense(H))` (src/spectrum.jl:48-55,
 * 108-120) on the device, the occupation function on the host, the projector
 * `projector(f, eig) = V * (f .* V')` (src/spectrum.jl:226-228) on the device.
 *   lmh_diagonalize       eigenvectors of the device-built H kept in the
 *                         context (one N x N buffer) until the next call or
 *                         lmh_eigensystem_free; eigenvalues (ascending) into
 *                         E[N].
 *   lmh_projector         st <- V diag(f) V' for any real weights f[N]
 *                         (negative ones too: Bose-Einstein below mu); ncols
 *                         (may be NULL): the columns between the first and the
 *                         last nonzero weight, the only ones in the product.
 * The host side computes f for the Fermi-Dirac / Bose-Einstein ensembles,
 * the Gibbs distribution, the ground state, the Fermi sphere and the fixed-N
 * chemical potential (Newton's iteration of src/spectrum.jl:266-285).
 */
int lmh_diagonalize(lmh_context *ctx, double *E);
int lmh_projector(lmh_context *ctx, lmh_state *st, const double *f, int64_t *ncols);
int lmh_eigensystem_free(lmh_context *ctx);

/*
 * Kets: step!(::CachedExp, psi::AbstractVector, cache) (src/evolution.jl:69-72)
 * -- psi <- U psi with the same U as lmh_step -- for a block of m kets stored
 * as an N x m column-major matrix (each column one state of an Evolution).
 * lmh_kets_observables gives localdensity (|psi_i|^2) and DensityCurrents
 * (2 Im(H_ij psi_j conj(psi_i)) per bond) of column `col` (either output may
 * be NULL), with the H of the current parameters.
 */
typedef struct lmh_kets lmh_kets;
int lmh_kets_create(lmh_context *ctx, int64_t m, lmh_kets **kets);
int lmh_kets_destroy(lmh_kets *kets);
int lmh_kets_upload(lmh_kets *kets, const double *psi, int64_t ld);
int lmh_kets_download(lmh_kets *kets, double *psi, int64_t ld);
int lmh_kets_step(lmh_context *ctx, lmh_kets *kets);
int lmh_kets_observables(lmh_context *ctx, lmh_kets *kets, int64_t col, double *rho, double *J);

/*
 * Krylov ket propagation: the counterpart of KrylovKitExp
 * (src/evolution.jl:130-154), whose step! calls KrylovKit.exponentiate(H,
 * -im dt, psi; kw...) (KrylovKit 0.4 - 0.9, Project.toml:26).
 * lmh_krylov_step replaces every ket psi of `kets` by exp(-i dt H) psi, with H
 * the Hamiltonian of the context's current parameters (lmh_set_field /
 * lmh_set_onsite; no lmh_update_solver needed): a Lanczos basis of at most
 * krylovdim vectors per ket (all kets advance together on the device), the
 * exponential of the small tridiagonal matrix on the host, and sub-steps of
 * the time when the basis does not reach the tolerance.  A (sub-)step of
 * length tau is accepted when the estimated error beta_j |e_j' exp(-i tau
 * T_j) e_1| |psi| is at most tol tau |psi| (tol is per unit time, as
 * KrylovKit's); LMH_ERANGE ("did not converge", the reference's
 * ArgumentError) after maxiter Lanczos restarts.  KrylovKit's defaults:
 * krylovdim 30, tol 1e-12, maxiter 100.  *matvecs (may be NULL) receives the
 * number of H applications per ket.  The call returns with the last
 * combination queued on the context's stream; every later call on the
 * context (observables, download, the next step) sees its result.
 */
int lmh_krylov_step(lmh_context *ctx, lmh_kets *kets, double dt, int krylovdim, double tol, int maxiter,
                    int64_t *matvecs);

/*
 * Currents between regions of the lattice for DensityCurrents(H, P)
 * (src/zoo/currents.jl:92-102; J_ij = 2 Im(H_ij P_ji), H of the current
 * parameters): currentsfrom(curr, src) (src/currents.jl:76-88) into per_site[N]
 * (0 on src) when dst is NULL, and currentsfromto(curr, src, dst)
 * (src/currents.jl:90-110; dst NULL = the complement of src) into *total.
 * src / dst: N bytes, nonzero = in the region.  Either output may be NULL.
 */
int lmh_currents_region(lmh_context *ctx, lmh_state *st, const uint8_t *src, const uint8_t *dst, double *per_site,
                        double *total);

/*
 * lmh_currents_region without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): src,
 * dst (N bytes, nonzero = in the region; dst may be NULL), per_site (N
 * doubles) and total (1 double) are DEVICE pointers the caller owns, either
 * output may be NULL.  The per-workgroup partial sums live in a static
 * device array of the library (8 KB per GPU, held until the process ends),
 * so N is limited to 262144 sites (LMH_ERANGE beyond).  Asynchronous on the
 * context's stream (lmh_sync waits); it does not synchronise or allocate.
 * Calls that overlap on one GPU share the array and race.
 */
int lmh_currents_region_noalloc(lmh_context *ctx, lmh_state *st, const uint8_t *src, const uint8_t *dst,
                                double *per_site, double *total);

/*
 * TimeSequence recording on the device (src/timesequence.jl:5-120): a series
 * keeps one entry per lmh_series_record call in device memory (no host
 * transfer per step), times strictly increasing.  Entry width:
 *   LMH_SERIES_DENSITY   localdensity, N values
 *   LMH_SERIES_CURRENTS  DensityCurrents on every bond of the geometry, nbonds values
 *   LMH_SERIES_REGION    currentsfromto(src, dst), 1 value (dst NULL: complement)
 * lmh_series_differentiate / lmh_series_integrate apply differentiate! (the
 * symmetric difference, times moved to the midpoints, one entry fewer) and
 * integrate! (trapezoids, cumulative, the first entry 0) in place
 * (src/timesequence.jl:213-225, 255-268).  lmh_series_download copies the
 * times[len] and values[len * width] (entry-major) to the host.
 */
#define LMH_SERIES_DENSITY 1
#define LMH_SERIES_CURRENTS 2
#define LMH_SERIES_REGION 3
typedef struct lmh_series lmh_series;
int lmh_series_create(lmh_context *ctx, int kind, int64_t capacity, const uint8_t *src, const uint8_t *dst,
                      lmh_series **series);
int lmh_series_destroy(lmh_series *series);
int lmh_series_record(lmh_series *series, lmh_state *st, double t);
int lmh_series_length(lmh_series *series, int64_t *len, int64_t *width);
int lmh_series_differentiate(lmh_series *series);
int lmh_series_integrate(lmh_series *series);
int lmh_series_download(lmh_series *series, double *times, double *values);

/*
 * Adjoint gradients.  The reference has no derivative of its evolution; this
 * is the reverse-mode counterpart of the driver loop (src/evolution.jl:238-275,
 * step! at :73-78): for a loss L(P(T)) it returns dL/d(field strengths) of
 * every step, dL/d(on-site energies) and dL/dP(0), the pullback a
 * ChainRulesCore rrule of the evolution needs.
 *
 * lmh_tape_begin(ctx, st, C): from now on every lmh_step of `st` records the
 *   step's field parameters, dt and Taylor order, and keeps a device copy of P
 *   every C steps (P_0, P_C, ...).  Geometry and on-site energies must stay
 *   fixed while recording; field strengths may change every step (the number
 *   and kinds of field components may not).
 * lmh_tape_backward(ctx, lam, grad_field, grad_onsite): `lam` holds
 *   dL/dP(T) (any complex matrix; its Hermitian part is used, since P stays
 *   Hermitian) and receives dL/dP(0).  grad_field[k * ncomp + c] = dL/da_c at
 *   step k (a_c = strength of field component c as passed to lmh_set_field);
 *   grad_onsite[i] = dL/dV_i summed over all steps.  P_k is recomputed by
 *   exact reverse unitary steps P_k = U_k^H P_{k+1} U_k, restarted from the
 *   checkpoints.  Needs s = 1 (max|H_ij| |dt| <= 1); otherwise LMH_ERANGE.
 *   For K N <= 10240 (N <= 2048) a column's K Horner iterates stay in LDS
 *   (fused reverse kernels); larger systems use per-term passes over all
 *   columns with the iterates recomputed (device memory: the checkpoints plus
 *   four N x N buffers besides the two states).  The recorded state's caches
 *   are overwritten.  Call lmh_update_solver before the next lmh_step.
 * lmh_hamiltonian_derivative(ctx, c, dh): dh[b] = dH[i_b, j_b]/da_c for every
 *   bond (complex, interleaved) -- for losses that depend on H directly, such
 *   as DensityCurrents (src/zoo/currents.jl:81-102).
 */
int lmh_tape_begin(lmh_context *ctx, lmh_state *st, int64_t checkpoint_every);
int lmh_tape_end(lmh_context *ctx);
int lmh_tape_length(lmh_context *ctx, int64_t *nsteps, int64_t *ncomp);
/* Checkpoints of the recording so far held in device and in pinned host
 * memory (LMH_OPT_CHECKPOINTS); either pointer may be NULL. */
int lmh_tape_checkpoints(lmh_context *ctx, int64_t *on_device, int64_t *on_host);
/* Progress of lmh_tape_backward, which runs every recorded step in one call:
 * fn(user, done, total) is called on the calling thread each time the
 * reverse steps of a checkpoint segment have completed on the device (done
 * and total count reverse steps; the last call has done == total).  fn NULL
 * stops the calls.  The reference has no counterpart: its Evolution is a
 * Julia iterator (src/evolution.jl:277-298), so a caller sees every step. */
typedef void (*lmh_progress_fn)(void *user, int64_t done, int64_t total);
int lmh_set_progress(lmh_context *ctx, lmh_progress_fn fn, void *user);
int lmh_tape_backward(lmh_context *ctx, lmh_state *lam, double *grad_field, double *grad_onsite);
int lmh_hamiltonian_derivative(lmh_context *ctx, int64_t comp, double *dh);

/*
 * Batches: many independent evolutions -- e.g. disorder realisations of
 * different lattice sizes, each with its own context (geometry, on-site
 * energies, field, solver K/s/dt from lmh_update_solver) and state -- advanced
 * together.  The reference would run one Evolution per realisation
 * (src/evolution.jl:183-275) one after the other; here one step of all
 * members whose H is a grid stencil with uniform strips (square lattices in a
 * Landau gauge or without field) is one persistent launch per side over all
 * their columns, whatever their sizes (LMH_OPT_BATCH_STRIP), and one step of
 * the others one launch per side and size class; lmh_batch_step with
 * use_graph != 0 replays the launches of three steps (one rotation of the
 * three state buffers) as a captured HIP graph.  Members need N <= 10240.
 * A member's H may change between lmh_batch_step calls (a time-dependent H:
 * set its parameters and call lmh_update_solver, then lmh_batch_step with
 * the steps that H holds for, as the reference evaluates h(t) at every step,
 * src/evolution.jl:238-250): the batch's tables take up its new K, s and dt.
 * The member's H must keep the form its kernel was chosen for (strip form
 * for the strip launch, value classes for a value-class variant), else
 * LMH_ESTATE (create the batch again); so is a parameter change without
 * lmh_update_solver, a geometry change, or stepping a member directly.
 * States stay readable through lmh_state_download.
 */
typedef struct lmh_batch lmh_batch;
int lmh_batch_create(lmh_batch **b, int64_t nmembers, lmh_context *const *ctxs, lmh_state *const *states);
int lmh_batch_destroy(lmh_batch *b);
int lmh_batch_step(lmh_batch *b, int64_t nsteps, int use_graph);
/* mean(localdensity(P_r)) = Tr(P_r)/N_r per member into out[nmembers]. */
int lmh_batch_mean_density(lmh_batch *b, double *out);
/* The kernel member `member` runs in lmh_batch_step: 32 for the persistent
 * strip launch (LMH_OPT_BATCH_STRIP), else the fused variant of its size class
 * (0..31, as LMH_OPT_FUSED_VARIANT).  No counterpart in the reference (it has
 * no batches); for tests and schedule reports. */
int lmh_batch_member_kernel(lmh_batch *b, int64_t member, int32_t *kernel);

/* Per-kernel device timing with HIP events recorded around every launch on
 * the context's stream (used by the benchmark to report the dominant
 * kernel's average duration).  kind: LMH_KERNEL_*.  lmh_get_profile
 * synchronises the stream and returns the number of launches and their
 * summed duration since profiling was last enabled. */
#define LMH_KERNEL_HORNER_LEFT 0
#define LMH_KERNEL_HORNER_RIGHT 1
#define LMH_KERNEL_BUILD_H 2
#define LMH_KERNEL_TAYLOR_ORDER 3
#define LMH_KERNEL_OBSERVABLES 4
#define LMH_KERNEL_PROJECTOR 5
#define LMH_KERNEL_FUSED 6
#define LMH_KERNEL_TRANSPOSE 7
#define LMH_KERNEL_ADJOINT 8        /* reverse-step side kernels (lmh_tape_backward) */
#define LMH_KERNEL_ADJOINT_REDUCE 9 /* gradient reduction per reverse step */
#define LMH_KERNEL_KRYLOV 10        /* lmh_krylov_step: Lanczos iterations and combinations */
#define LMH_NUM_KERNEL_KINDS 11
int lmh_set_profiling(lmh_context *ctx, int enable);
int lmh_get_profile(lmh_context *ctx, int kind, int64_t *launches, double *total_ms);

/* How lmh_step applies U P U' (same polynomial, same results to rounding):
 *   LMH_STEP_PER_PASS  one streaming kernel per Horner term and side
 *                      (2K launches, any N);
 *   LMH_STEP_FUSED_T   all K terms of a column on chip, two column passes
 *                      whose results are stored conjugate-transposed
 *                      ((U P) U' = (U (U P)')'), N <= 10240;
 *   LMH_STEP_FUSED_N   as FUSED_T with plain stores plus a tiled conjugate
 *                      transpose after each pass;
 *   LMH_STEP_AUTO      FUSED_T when it applies, else PER_PASS (default). */
#define LMH_STEP_AUTO 0
#define LMH_STEP_PER_PASS 1
#define LMH_STEP_FUSED_T 2
#define LMH_STEP_FUSED_N 3
int lmh_set_step_mode(lmh_context *ctx, int mode);

/* Tuning options (performance only; results are the same to rounding).  A
 * fused variant or banded left pass that no longer fits after
 * lmh_set_geometry reverts to its default:
 *   LMH_OPT_STEP_MODE   one of LMH_STEP_*
 *   LMH_OPT_RIGHT_PASS  per-term right side: 1 = left passes on the transpose
 *                       (each side's last term written conjugate-transposed:
 *                       (U (U P)')' = U P U'), 2 = column group staged in LDS,
 *                       3 = unstaged column sweep in Z-order (default),
 *                       4 = as 3 with column groups innermost per XCD,
 *                       5 = as 1 with each side's result transposed by a
 *                       tiled transpose pass instead of the last term
 *   LMH_OPT_LEFT_PASS   per-term left pass (and the adjoint's reverse terms):
 *                       0 = banded when every entry of H lies within 1024
 *                       (reverse terms: 512) rows of the diagonal, else
 *                       gather (default), 1 = gather through L1/L2, 2 = banded (a
 *                       block's rows plus the band staged in LDS per column),
 *                       3 = banded with two Horner terms per launch (band <=
 *                       512 rows; the halo rows' H held in LDS) */
#define LMH_OPT_STEP_MODE 1
#define LMH_OPT_RIGHT_PASS 2
#define LMH_OPT_LEFT_COLS 3      /* columns swept per left-pass thread (32)  */
#define LMH_OPT_RIGHT_COLS 4     /* Z-ordered columns per right-pass block (16) */
#define LMH_OPT_FUSED_MAX_N 5    /* LMH_STEP_AUTO uses fused passes up to this N (10240) */
#define LMH_OPT_EIGENSOLVER 6    /* lmh_densitymatrix / lmh_diagonalize, first solver tried:
                                    0 divide and conquer (zheevd; dsyevd for real H), 1 QR
                                    (zheev / dsyev), 2 Jacobi (zheevdj / dsyevdj); H whose
                                    hoppings are all real goes to the real solvers first */
#define LMH_OPT_FUSED_VARIANT 7  /* fused column pass shape (threads, sites/thread, columns,
                                    H in registers, x in registers, site H layout): -1 = by
                                    N, else a variant index (0..32).  Variant 32 is the strip
                                    kernel for grid-stencil H (square lattices): a thread
                                    owns a run of consecutive sites with its bond values in
                                    registers.  It runs (and is the default) whenever the
                                    geometry is a grid and the last H build gave every
                                    strip uniform bonds (Landau gauge, no field); else the
                                    by-size default runs */
#define LMH_OPT_ADJOINT_MODE 8   /* lmh_tape_backward: 0 = fused when K N <= 10240 (else
                                    per-term), 1 = fused only, 2 = per-term passes (the
                                    K-1 Horner iterates of a side kept in HBM when it has
                                    room, else recomputed), 3 = per-term, always
                                    recomputing the iterates */
#define LMH_OPT_LEFT_PASS 9
#define LMH_OPT_CHECKPOINTS 10   /* where lmh_tape_begin's checkpoints live: 0 = device
                                    memory while the backward sweep's work buffers and
                                    Horner iterates still fit beside them, else pinned host
                                    memory (default); 1 = device only; 2 = host only */
#define LMH_OPT_KRYLOV_GRAPHS 11 /* lmh_krylov_step: 1 = the first Lanczos chunk of a step
                                    (start, as many iterations as the previous step
                                    needed, the next beta) replayed from a captured HIP
                                    graph per source buffer and length (default); 0 =
                                    plain launches */
#define LMH_OPT_VALUE_CLASSES 12 /* fused passes (N <= 10240, at most 4 neighbours per
                                    site): H's slots whose values agree, directly or
                                    conjugated, at this many probes of the field
                                    amplitudes (0..3, default 3; probe 0 is the zero
                                    field) share a value class, and fused variants 29-31
                                    read the class values from LDS.  Every H build
                                    checks each slot against its class; on a mismatch
                                    those variants run as their plain counterparts
                                    (20-22).  0 = no classes */
#define LMH_OPT_FUSED_LAUNCHED 13 /* read only (lmh_get_option): the fused variant the last
                                     fused launch of lmh_step / lmh_kets_step ran (-1: none
                                     yet); it differs from LMH_OPT_FUSED_VARIANT when that
                                     is -1 (by size) or its kernel does not apply to the
                                     current H */
#define LMH_OPT_EIGEN_USED 14     /* read only: the solver of the last lmh_densitymatrix /
                                     lmh_diagonalize, method (0..2 as above) + 4 if it was
                                     the real one (-1: none yet) */
#define LMH_OPT_BATCH_STRIP 15     /* lmh_batch_create (option of the first member's context):
                                     1 = members whose H is a grid stencil with uniform strips
                                     (the strip kernel of LMH_OPT_FUSED_VARIANT 32) advance in
                                     one persistent launch per side over all of them (default);
                                     0 = every member runs its size class's fused kernel */
#define LMH_OPT_DEFER 16           /* fused steps on the strip kernel (LMH_OPT_FUSED_VARIANT 32):
                                     1 = lmh_step launches the step's first pass only and
                                     defers the second, which then runs in the same launch
                                     as the next step's first pass (one read and one write of
                                     the matrix per step instead of two) or, before anything
                                     reads the state (downloads, entries, observables, series,
                                     copies, device pointers, tapes, batches, lmh_sync), on
                                     its own (default); 0 = both passes at every lmh_step.
                                     Results are the same either way */
#define LMH_OPT_ASYNC_ORDER 17     /* 1 = lmh_update_solver without K_out / s_out leaves K on
                                     the device where the strip steps can take it from there
                                     (default); 0 = every update waits for K.  Results are the
                                     same either way */
#define LMH_OPT_STEP_COUNT 18      /* read only: lmh_step calls since lmh_set_profiling or
                                     context creation */
#define LMH_OPT_STEP_TERMS 19      /* read only: the sum of their K */
#define LMH_OPT_ASYNC_STEPS 20     /* read only: of those, the steps launched with K on the
                                     device (LMH_OPT_ASYNC_ORDER) */
#define LMH_OPT_ASYNC_REDONE 21    /* read only: of those, the steps whose launch found another
                                     result than the one assumed and ran again */
int lmh_set_option(lmh_context *ctx, int key, int64_t value);
int lmh_get_option(lmh_context *ctx, int key, int64_t *value);
/* H's value classes as of the last lmh_update_solver (LMH_OPT_VALUE_CLASSES):
 * the number of classes (0: none for this geometry and field) and whether the
 * last H build matched every slot to its class (1) or not (0).  No
 * counterpart in the reference: it reports which fused kernel the steps use.
 * Either pointer may be NULL. */
int lmh_value_classes(lmh_context *ctx, int64_t *nclasses, int64_t *matched);

/* Autotuner: after lmh_update_solver, times `reps` steps of each candidate
 * schedule (step mode x left/right tile shapes) on a scratch copy of `st`
 * (st itself is not advanced) and keeps the fastest in the context's
 * options.  K and s stay the reference rule's; only the schedule changes.
 * best_ms (may be NULL) receives the winner's time per step. */
int lmh_autotune(lmh_context *ctx, lmh_state *st, int32_t reps, double *best_ms);

/* The context's HIP stream (hipStream_t) for callers that order their own
 * work against the library's. */
int lmh_get_stream(lmh_context *ctx, void **stream);

/* Run the deferred passes of the context's states (LMH_OPT_DEFER), then block
 * until all work queued on the context's stream is done. */
int lmh_sync(lmh_context *ctx);

#ifdef __cplusplus
}
#endif
#endif /* LMHIP_H */
