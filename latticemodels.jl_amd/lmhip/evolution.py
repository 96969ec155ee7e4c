"""`Evolution` iterator and the `HIPExp` solver of the host mirror.

Mirrors the reference's EvolutionSolver plugin interface (src/evolution.jl:5-33)
and driver (src/evolution.jl:183-298):

  update_solver!(solver, H, dt, force)  -> HIPExp.update_solver
  step!(solver, state, cache)           -> HIPExp.step
  evolution_cache(solver, state)        -> the library owns P and its caches

  Evolution([solver,] h, states...; timedomain)  copies every state
  iterate: dt_i = t_i - t_{i-1} (first: t_1 - time); |dt| < 1e-15 is no step;
           H = h(time) at the left end of the step; time += dt;
           yields EvolutionTimestamp(H_used, states, time)
  step! with dt < -1e-15 raises (ArgumentError -> ValueError)

Every numeric step runs in liblmhip.so on the device; this module only moves
parameters.
"""
import ctypes
import math
import warnings
import weakref

import numpy as np

from . import _lib
from ._lib import check, lib, ptr
from .operators import Hamiltonian, SparseHamiltonian


_DEFAULT_CONTEXTS = {}


def default_context(device, n):
    """Shared context per (device, lattice size) used when none is given."""
    key = (int(device), int(n))
    if key not in _DEFAULT_CONTEXTS:
        _DEFAULT_CONTEXTS[key] = Context(device)
    return _DEFAULT_CONTEXTS[key]


class Context:
    """One liblmhip context (device + stream + device-side H)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib().lmh_create(ctypes.byref(h), int(device)))
        self.h = h
        self.device = device
        self.geom_key = None
        self.geom_n = None
        self.geom_h = None                 # the Hamiltonian whose geometry is on the device
        self.internal = 1
        self._csc_current = None           # the SparseHamiltonian whose values the device holds
        self._states = weakref.WeakSet()

    def close(self):
        # lmh_destroy frees the context's states too: detach them first so
        # their own finalizers do not free them a second time
        if self.h:
            for st in list(self._states):
                st.close()
            lib().lmh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_hamiltonian(self, H):
        if isinstance(H, SparseHamiltonian):
            return self._set_sparse(H)
        self._csc_current = None
        key = H.geometry_key()
        if key != self.geom_key:
            bonds, amp = H.geometry()
            nb = len(bonds)
            bi = np.ascontiguousarray(bonds.i, dtype=np.int64)
            bj = np.ascontiguousarray(bonds.j, dtype=np.int64)
            ri = np.ascontiguousarray(bonds.ri, dtype=np.float64)
            rj = np.ascontiguousarray(bonds.rj, dtype=np.float64)
            am = np.ascontiguousarray(amp, dtype=np.complex128)
            tw = np.ascontiguousarray(bonds.twist, dtype=np.complex128)
            check(lib().lmh_set_geometry(self.h, H.n, nb, ptr(bi), ptr(bj), ptr(ri), ptr(rj), ptr(am), ptr(tw)))
            xy = np.ascontiguousarray(H.coords, dtype=np.float64)
            check(lib().lmh_set_site_coords(self.h, ptr(xy)))
            self.geom_key = key
            self.geom_n = H.n
            self.geom_h = H
            self.internal = H.internal
            self._geom_refs = (bi, bj, ri, rj, am, tw, bonds)
        diag = None if H.onsite is None else np.ascontiguousarray(H.onsite, dtype=np.float64)
        check(lib().lmh_set_onsite(self.h, ptr(diag)))
        kinds, params = H.field_params()
        kinds = np.ascontiguousarray(kinds)
        params = np.ascontiguousarray(params)
        check(lib().lmh_set_field(self.h, len(kinds), ptr(kinds), ptr(params)))

    def _set_sparse(self, H, upload=False):
        """A CSC Hamiltonian: the pattern once (again when it changes), the
        values unless the device already holds this very object's (upload:
        always)."""
        pat = getattr(self, "_csc_pattern", None)
        same_pattern = self.geom_key == ("csc",) and pat is not None and (
            pat[0].shape == H.colptr.shape and np.array_equal(pat[0], H.colptr) and np.array_equal(pat[1], H.rowval))
        if same_pattern and self._csc_current is H and not upload:
            return
        if not same_pattern:
            check(lib().lmh_set_csc_pattern(self.h, H.n, ptr(H.colptr), ptr(H.rowval), 0))
            self.geom_key = ("csc",)
            self.geom_n = H.n
            self.geom_h = H
            self.internal = 1
            self._csc_pattern = (H.colptr.copy(), H.rowval.copy())
            self._geom_refs = H.bond_ends()
        self._csc_current = None
        check(lib().lmh_set_csc_values(self.h, H.nnz, ptr(H.nzval)))
        self._csc_current = H

    def hamiltonian_entries(self, H):
        """Device-built diag[N] and H[i_b, j_b] for every bond of the device
        geometry (matrix indices; with internal states one per nonzero block
        element, `H.geometry()`), for parity checks."""
        self.set_hamiltonian(H)
        diag = np.zeros(H.n)
        hij = np.zeros(len(H.geometry()[0]), dtype=np.complex128)
        check(lib().lmh_get_hamiltonian(self.h, ptr(diag), ptr(hij)))
        return diag, hij


class DensityMatrix:
    """A density matrix held on the device (handle on an lmh_state)."""

    def __init__(self, ctx, n):
        s = ctypes.c_void_p()
        check(lib().lmh_state_create(ctx.h, ctypes.byref(s)))
        self.h = s
        self.ctx = ctx
        self.n = n
        ctx._states.add(self)

    @classmethod
    def from_array(cls, ctx, P):
        P = np.asfortranarray(P, dtype=np.complex128)
        if P.shape != (ctx.geom_n, ctx.geom_n):
            raise ValueError("density matrix of shape %s for %d sites" % (P.shape, ctx.geom_n))
        d = cls(ctx, ctx.geom_n)
        check(lib().lmh_state_upload(d.h, ptr(P), P.shape[0]))
        return d

    def copy(self):
        d = DensityMatrix(self.ctx, self.n)
        check(lib().lmh_state_copy(d.h, self.h))
        return d

    def to_numpy(self):
        out = np.zeros((self.n, self.n), dtype=np.complex128, order="F")
        check(lib().lmh_state_download(self.h, ptr(out), self.n))
        return out

    def idempotency_error(self):
        """max |P P - P| (a dense product on the device)."""
        out = ctypes.c_double()
        check(lib().lmh_state_idempotency(self.h, ctypes.byref(out)))
        return out.value

    def get_entries(self, rows, cols):
        """P[rows[k], cols[k]] without downloading P."""
        r = np.ascontiguousarray(rows, dtype=np.int64)
        c = np.ascontiguousarray(cols, dtype=np.int64)
        out = np.zeros(len(r), dtype=np.complex128)
        check(lib().lmh_state_get_entries(self.h, len(r), ptr(r), ptr(c), ptr(out)))
        return out

    def set_entries(self, rows, cols, vals):
        """P <- the matrix with these entries (repeats add; the rest 0)."""
        r = np.ascontiguousarray(rows, dtype=np.int64)
        c = np.ascontiguousarray(cols, dtype=np.int64)
        v = np.ascontiguousarray(vals, dtype=np.complex128)
        check(lib().lmh_state_set_entries(self.h, len(r), ptr(r), ptr(c), ptr(v)))
        return self

    def device_ptr(self):
        p = ctypes.c_void_p()
        ld = ctypes.c_int64()
        check(lib().lmh_state_device_ptr(self.h, ctypes.byref(p), ctypes.byref(ld)))
        return p.value, ld.value

    def close(self):
        if self.h:
            if self.ctx.h:                   # else freed with its context
                lib().lmh_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Kets:
    """Wavefunctions held on the device: an N x m block of kets (handle on
    lmh_kets); a 1-D array is one Ket.  step!(::CachedExp, ::AbstractVector)
    (src/evolution.jl:69-72) evolves them as psi <- U psi."""

    def __init__(self, ctx, n, m=1):
        k = ctypes.c_void_p()
        check(lib().lmh_kets_create(ctx.h, int(m), ctypes.byref(k)))
        self.h = k
        self.ctx = ctx
        self.n = n
        self.m = int(m)
        self.vector = True
        ctx._states.add(self)

    @classmethod
    def from_array(cls, ctx, psi):
        psi = np.asarray(psi, dtype=np.complex128)
        vector = psi.ndim == 1
        block = np.asfortranarray(psi.reshape(psi.shape[0], -1))
        if block.shape[0] != ctx.geom_n:
            raise ValueError("ket of length %d for %d sites" % (block.shape[0], ctx.geom_n))
        k = cls(ctx, ctx.geom_n, block.shape[1])
        k.vector = vector
        check(lib().lmh_kets_upload(k.h, ptr(block), block.shape[0]))
        return k

    def copy(self):
        return Kets.from_array(self.ctx, self.to_numpy())

    def to_numpy(self):
        out = np.zeros((self.n, self.m), dtype=np.complex128, order="F")
        check(lib().lmh_kets_download(self.h, ptr(out), self.n))
        return out[:, 0].copy() if self.vector and self.m == 1 else out

    def observables(self, col=0, rho=True, J=True):
        r = np.zeros(self.n) if rho else None
        nb = len(self.ctx._geom_refs[0])
        j = np.zeros(nb) if J else None
        check(lib().lmh_kets_observables(self.ctx.h, self.h, int(col), ptr(r), ptr(j)))
        return r, j

    def close(self):
        if self.h:
            if self.ctx.h:
                lib().lmh_kets_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EvolutionSolver:
    """Abstract solver (src/evolution.jl:25-33)."""

    def update_solver(self, H, dt, force=False):
        raise NotImplementedError

    def step(self, state):
        raise NotImplementedError


class HIPExp(EvolutionSolver):
    """Device counterpart of CachedExp(threshold=1e-10, nztol=1e-14)
    (src/evolution.jl:49-68): same Taylor order rule, applied as Horner passes."""

    def __init__(self, threshold=1e-10, nztol=1e-14, device=0, ctx=None, step_mode=_lib.STEP_AUTO):
        self.threshold = float(threshold)
        self.nztol = float(nztol)
        self.device = device
        self.ctx = ctx
        self.step_mode = int(step_mode)
        self._order = (0, 1.0)
        self.H = None

    # K and s of the last update_solver.  The update does not wait for them
    # (the library keeps them on the device for the steps); reading either
    # one waits.
    @property
    def K(self):
        return self._read_order()[0]

    @property
    def s(self):
        return self._read_order()[1]

    def _read_order(self):
        if self._order is None:
            K = ctypes.c_int32()
            s = ctypes.c_double()
            check(lib().lmh_solver_order(self.ctx.h, ctypes.byref(K), ctypes.byref(s)))
            self._order = (K.value, s.value)
        return self._order

    def _context(self, H):
        if self.ctx is None:
            self.ctx = default_context(self.device, H.n)
        return self.ctx

    def _apply_mode(self):
        check(lib().lmh_set_step_mode(self.ctx.h, self.step_mode))

    def update_solver(self, H, dt, force=False):
        if isinstance(H, SparseHamiltonian):
            # the reference's cache test for a matrix (src/evolution.jl:83-92):
            # the same object or an equal matrix is not uploaded again, and
            # the library then keeps U when -i dt ≈ the previous factor
            ctx = self._context(H)
            cur = ctx._csc_current
            if force or cur is None or not (cur is H or cur == H):
                ctx._set_sparse(H, upload=True)
        elif isinstance(H, Hamiltonian):
            self._context(H).set_hamiltonian(H)
        else:
            raise TypeError("HIPExp needs a parametric Hamiltonian (lmhip.Hamiltonian) or a SparseHamiltonian")
        self._order = None
        check(lib().lmh_update_solver(self.ctx.h, float(dt), int(bool(force)), self.threshold, self.nztol,
                                      None, None))
        self.H = H

    def step(self, state):
        if self.step_mode != _lib.STEP_AUTO or not getattr(self, "_tuned", False):
            self._apply_mode()
        if isinstance(state, Kets):
            check(lib().lmh_kets_step(self.ctx.h, state.h))
        else:
            check(lib().lmh_step(self.ctx.h, state.h))

    def autotune(self, state, reps=3):
        """Run the library's schedule autotuner on a scratch copy of `state`
        (after an update_solver); returns the chosen options and time/step."""
        best = ctypes.c_double()
        check(lib().lmh_autotune(self.ctx.h, state.h, int(reps), ctypes.byref(best)))
        self._tuned = True
        out = {"best_ms": best.value}
        for key, name in ((_lib.OPT_STEP_MODE, "step_mode"), (_lib.OPT_LEFT_COLS, "left_cols"),
                          (_lib.OPT_RIGHT_COLS, "right_cols"), (_lib.OPT_RIGHT_PASS, "right_pass"),
                          (_lib.OPT_FUSED_VARIANT, "fused_variant"), (_lib.OPT_LEFT_PASS, "left_pass")):
            v = ctypes.c_int64()
            check(lib().lmh_get_option(self.ctx.h, key, ctypes.byref(v)))
            out[name] = v.value
        return out

    def make_state(self, P, H0):
        """evolution_cache + copy(state): the Evolution owns a copy of P (a
        density matrix, or kets: a Kets or a 1-D array)."""
        ctx = self._context(H0)
        if isinstance(P, Kets) or (isinstance(P, np.ndarray) and P.ndim == 1):
            psi = P.to_numpy() if isinstance(P, Kets) else P
            ctx.set_hamiltonian(H0)
            return Kets.from_array(ctx, psi)
        if isinstance(P, DensityMatrix):
            if P.ctx is ctx:
                return P.copy()
            P = P.to_numpy()
        ctx.set_hamiltonian(H0)
        if hasattr(P, "rows") and hasattr(P, "cols") and hasattr(P, "vals"):
            # sparse entries (e.g. a SparseCotangent or a product state):
            # scattered on the device, no dense host matrix
            return DensityMatrix(ctx, H0.n).set_entries(P.rows, P.cols, P.vals)
        return DensityMatrix.from_array(ctx, P)


class HIPKrylovExp(EvolutionSolver):
    """Device counterpart of KrylovKitExp(; krylovdim, tol, maxiter)
    (src/evolution.jl:130-154): update_solver! keeps H and dt, step! replaces
    each ket psi by exp(-i dt H) psi (KrylovKit.exponentiate semantics: a
    Lanczos basis of at most krylovdim vectors, tol per unit time, time
    sub-steps, at most maxiter restarts) with the Lanczos recurrences of all
    kets of a block run together on the device (lmh_krylov_step).  Evolves
    kets only, as the reference's step!(::KrylovKitExp, ::AbstractVector)."""

    def __init__(self, krylovdim=30, tol=1e-12, maxiter=100, device=0, ctx=None):
        self.krylovdim = int(krylovdim)
        self.tol = float(tol)
        self.maxiter = int(maxiter)
        self.device = device
        self.ctx = ctx
        self.dt = 0.0
        self.H = None
        self.matvecs = 0          # H applications per ket in the last step

    def _context(self, H):
        if self.ctx is None:
            self.ctx = default_context(self.device, H.n)
        return self.ctx

    def update_solver(self, H, dt, force=False):
        if not isinstance(H, Hamiltonian):
            raise TypeError("HIPKrylovExp needs a parametric Hamiltonian (lmhip.operators.Hamiltonian)")
        self._context(H).set_hamiltonian(H)
        self.dt, self.H = float(dt), H

    def step(self, state):
        if not isinstance(state, Kets):
            raise TypeError("HIPKrylovExp evolves kets (step!(::KrylovKitExp, ::AbstractVector))")
        mv = ctypes.c_int64()
        check(lib().lmh_krylov_step(self.ctx.h, state.h, self.dt, self.krylovdim, self.tol, self.maxiter,
                                    ctypes.byref(mv)))
        self.matvecs = mv.value

    def make_state(self, psi, H0):
        ctx = self._context(H0)
        ctx.set_hamiltonian(H0)
        if isinstance(psi, Kets):
            return psi.copy() if psi.ctx is ctx else Kets.from_array(ctx, psi.to_numpy())
        psi = np.asarray(psi)
        if psi.ndim not in (1, 2) or (psi.ndim == 2 and psi.shape[0] == psi.shape[1] and psi.shape[1] > 1):
            raise TypeError("HIPKrylovExp evolves kets: a length-N vector or an N x m block")
        return Kets.from_array(ctx, psi)


class EvolutionTimestamp:
    """(states..., H, t) when iterated (src/evolution.jl:277-298)."""

    def __init__(self, H, states, t):
        self.H = H
        self.states = states
        self.t = t

    def __iter__(self):
        yield from self.states
        yield self.H
        yield self.t

    @property
    def state(self):
        (s,) = self.states
        return s

    def __getitem__(self, i):
        return self.states[i]


class Evolution:
    """Evolution([solver,] h, states...; timedomain=None) (src/evolution.jl:183-236)."""

    def __init__(self, *args, timedomain=None, **solver_kw):
        if args and isinstance(args[0], EvolutionSolver):
            solver, args = args[0], args[1:]
        elif args and isinstance(args[0], type) and issubclass(args[0], EvolutionSolver):
            solver, args = args[0](**solver_kw), args[1:]
        else:
            solver = HIPExp(**solver_kw)
        if len(args) < 2:
            raise ValueError("No states provided")
        self.solver = solver
        self.hamiltonian = args[0]
        H0 = self._eval(0.0)
        self.states = [solver.make_state(P, H0) for P in args[1:]]
        self.time = 0.0
        self.timedomain = timedomain

    def _eval(self, t):
        h = self.hamiltonian
        return h(t) if callable(h) else h

    def step(self, dt):
        if dt < -1e-15:
            raise ValueError("negative time step")
        H = self._eval(self.time)
        if abs(dt) < 1e-15:
            return H
        self.solver.update_solver(H, dt)
        for st in self.states:
            self.solver.step(st)
        self.time += dt
        return H

    def __call__(self, times):
        times = list(times)
        prev = None
        for i, t in enumerate(times):
            dt = (t - prev) if i > 0 else (t - self.time)
            prev = t
            H = self.step(dt)
            yield EvolutionTimestamp(H, self.states, float(self.time))

    def __iter__(self):
        if self.timedomain is None:
            raise TypeError("Evolution without a time domain; call it with the times")
        return self(self.timedomain)


class BatchedEvolution:
    """Many independent density-matrix evolutions advanced together, e.g. the
    disorder realisations of an ensemble average.  Each member is
    (Hamiltonian, P0), the Hamiltonian either fixed or a function t -> H(t);
    the reference runs one `Evolution` per member (src/evolution.jl:183-275),
    evaluating h(t) at the start of every step (:238-250).  Here each member
    gets its own context and state, and one step of all members is one
    persistent strip launch (or one launch per side and size class);
    `step(n, graph=True)` replays three-step HIP graphs.

    With time-dependent members every step updates each of them with H(t)
    at the step's start and then advances the batch one step; a member whose
    H leaves the form its kernel was chosen for (strip form, value classes)
    stops the batch with an error."""

    def __init__(self, members, dt, threshold=1e-10, nztol=1e-14, device=0, strip=True):
        self.dt = float(dt)
        self.solvers, self.states, self.hfuncs = [], [], []
        for H, P0 in members:
            f = H if callable(H) else None
            H0 = f(0.0) if f else H
            sol = HIPExp(threshold=threshold, nztol=nztol, device=device, ctx=Context(device))
            st = sol.make_state(P0, H0)
            sol.update_solver(H0, self.dt)
            self.solvers.append(sol)
            self.states.append(st)
            self.hfuncs.append(f)
        m = len(self.states)
        # members with uniform-strip grid stencils: one persistent strip launch
        # per side (LMH_OPT_BATCH_STRIP, read from the first member's context)
        check(lib().lmh_set_option(self.solvers[0].ctx.h, _lib.OPT_BATCH_STRIP, int(bool(strip))))
        ctxs = (ctypes.c_void_p * m)(*[s.ctx.h.value for s in self.solvers])
        sts = (ctypes.c_void_p * m)(*[s.h.value for s in self.states])
        b = ctypes.c_void_p()
        check(lib().lmh_batch_create(ctypes.byref(b), m, ctxs, sts))
        self.h = b
        self.time = 0.0

    def step(self, nsteps=1, graph=True):
        if not any(self.hfuncs):
            check(lib().lmh_batch_step(self.h, int(nsteps), int(bool(graph))))
            self.time += nsteps * self.dt
            return
        for _ in range(int(nsteps)):
            for sol, f in zip(self.solvers, self.hfuncs):
                if f is not None:
                    sol.update_solver(f(self.time), self.dt)
            check(lib().lmh_batch_step(self.h, 1, int(bool(graph))))
            self.time += self.dt

    def mean_density(self):
        out = np.zeros(len(self.states))
        check(lib().lmh_batch_mean_density(self.h, ptr(out)))
        return out

    def member_kernels(self):
        """Per member: 32 (the persistent strip launch) or its size class's
        fused variant."""
        out = []
        for r in range(len(self.states)):
            v = ctypes.c_int32()
            check(lib().lmh_batch_member_kernel(self.h, r, ctypes.byref(v)))
            out.append(v.value)
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().lmh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


FermiDirac, BoseEinstein = 1, -1        # ParticleStatistics (src/operators/system.jl:60-63)
_TEMP_TOL = 1e-10                        # src/spectrum.jl:235


def _div(a, b):
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        return float(np.float64(a) / np.float64(b))


def _exp(x):
    with np.errstate(over="ignore", invalid="ignore"):
        return float(np.exp(np.float64(x)))


def densfun(T, mu, statistics=FermiDirac):
    """The occupation E -> (E - mu == T == 0) ? 1 : 1 / (exp((E - mu) / T) +
    statistics) (src/spectrum.jl:230-232), IEEE division as in Julia."""
    def f(E):
        if E - mu == T and T == 0:
            return 1.0
        return _div(1.0, _exp(_div(E - mu, T)) + statistics)
    return f


def ddensfun(T, mu, statistics=FermiDirac):
    """d densfun / dE (src/spectrum.jl:233-235)."""
    def f(E):
        ex = np.float64(_exp(_div(E - mu, T)))
        with np.errstate(over="ignore", invalid="ignore"):
            den = (ex + np.float64(statistics)) ** 2
        return _div(_div(-ex, T), den)
    return f


class Eigensystem:
    """diagonalize(dense(H)) (src/spectrum.jl:48-55, 108-120) kept on the
    device: `values` (ascending) on the host, the eigenvectors in the
    context until `close()` or the next diagonalisation there."""

    def __init__(self, ctx, values):
        self.ctx = ctx
        self.values = values

    def __len__(self):
        return len(self.values)

    def projector(self, f):
        """projector(f, eig) = V * (f .* V') (src/spectrum.jl:226-228) for the
        weights f (a callable of E or an array of N values): a DensityMatrix."""
        w = np.array([f(e) for e in self.values] if callable(f) else f, dtype=np.float64)
        if w.shape != self.values.shape:
            raise ValueError("need one weight per eigenvalue")
        P = DensityMatrix(self.ctx, len(self.values))
        ncols = ctypes.c_int64()
        check(lib().lmh_projector(self.ctx.h, P.h, ptr(w), ctypes.byref(ncols)))
        P.nocc = ncols.value
        return P

    def groundstate_densitymatrix(self):
        """psi psi' of the first (lowest) eigenvector (src/spectrum.jl:237-241, 195)."""
        w = np.zeros(len(self.values))
        w[0] = 1.0
        return self.projector(w)

    def gibbs_densitymatrix(self, T=0.0):
        """src/spectrum.jl:242-250."""
        if abs(T) < _TEMP_TOL:
            return self.groundstate_densitymatrix()
        w = np.array([_exp(_div(-e, T)) for e in self.values])
        return self.projector(w / float(np.sum(w)))

    def ensemble_densitymatrix(self, mu=0.0, T=0.0, statistics=FermiDirac):
        """src/spectrum.jl:251-257."""
        return self.projector(densfun(T, mu, statistics))

    def fermisphere_densitymatrix(self, N):
        """src/spectrum.jl:258-265."""
        n = len(self.values)
        if n < N:
            raise ValueError("cannot build Fermi sphere with %d particles: only %d bands present" % (N, n))
        if N < n and np.isclose(self.values[N - 1], self.values[N]):
            warnings.warn("degenerate levels on the Fermi sphere")
        w = np.zeros(n)
        w[:N] = 1.0
        return self.projector(w)

    def fixn_densitymatrix(self, N, T=0.0, statistics=FermiDirac, maxiter=1000, atol=math.sqrt(np.finfo(float).eps)):
        """src/spectrum.jl:266-285: at |T| < 1e-10 the Fermi sphere (or N times
        the ground state for bosons), else Newton's iteration for mu from 0;
        the P gets `.mu` (None at T = 0)."""
        if abs(T) < _TEMP_TOL:
            if statistics == BoseEinstein:
                w = np.zeros(len(self.values))
                w[0] = float(N)
                P = self.projector(w)
            else:
                P = self.fermisphere_densitymatrix(N)
            P.mu = None
            return P
        newmu = 0.0
        for _ in range(maxiter):
            f, df = densfun(T, newmu, statistics), ddensfun(T, newmu, statistics)
            n_ev = sum(f(e) for e in self.values)
            dn_ev = sum(df(e) for e in self.values)
            dmu = _div(n_ev - N, dn_ev)
            if abs(dmu) < atol:
                P = self.ensemble_densitymatrix(mu=newmu, T=T, statistics=statistics)
                P.mu = newmu
                return P
            newmu += dmu
        raise ValueError("did not converge")

    def densitymatrix(self, T=0.0, mu=None, N=None, statistics=None):
        """densitymatrix(eig; kw...) of a one-particle system
        (src/spectrum.jl:317-329): N -> fixed particle number (mu found), mu or
        statistics -> ensemble, neither -> Gibbs distribution."""
        if N is not None:
            if mu is not None:
                raise ValueError("cannot specify both N and μ")
            return self.fixn_densitymatrix(N, T=T, statistics=FermiDirac if statistics is None else statistics)
        if mu is not None or statistics is not None:
            return self.ensemble_densitymatrix(mu=0.0 if mu is None else mu, T=T,
                                               statistics=FermiDirac if statistics is None else statistics)
        return self.gibbs_densitymatrix(T)

    def close(self):
        if self.ctx is not None and self.ctx.h:
            check(lib().lmh_eigensystem_free(self.ctx.h))


def diagonalize(H, device=0, ctx=None):
    """diagonalize(dense(H)) of the device-built H: an Eigensystem."""
    if not isinstance(H, Hamiltonian):
        raise TypeError("diagonalize needs a parametric Hamiltonian")
    ctx = ctx if ctx is not None else default_context(device, H.n)
    ctx.set_hamiltonian(H)
    E = np.zeros(H.n)
    check(lib().lmh_diagonalize(ctx.h, ptr(E)))
    return Eigensystem(ctx, E)


def densitymatrix(H, mu=None, T=0.0, N=None, statistics=None, device=0, ctx=None, return_energies=False):
    """densitymatrix(H; T, mu, N, statistics) of a one-particle system
    (src/spectrum.jl:317-333 with dense diagonalisation), computed on the device
    from the device-built H (or from an Eigensystem).  Dispatch as the
    reference's: N -> fixed particle number with mu found by Newton's
    iteration; mu or statistics -> Fermi-Dirac / Bose-Einstein ensemble;
    neither -> Gibbs distribution (the ground state at T = 0).  Returns a
    DensityMatrix (device state)."""
    if T < 0:
        raise ValueError("negative temperature")
    if isinstance(H, Eigensystem):
        P = H.densitymatrix(T=T, mu=mu, N=N, statistics=statistics)
        return (P, H.values) if return_energies else P
    if not isinstance(H, Hamiltonian):
        raise TypeError("densitymatrix needs a parametric Hamiltonian")
    ctx = ctx if ctx is not None else default_context(device, H.n)
    ctx.set_hamiltonian(H)
    if N is None and mu is not None and statistics in (None, FermiDirac):
        # the Fermi-Dirac ensemble in one library call
        P = DensityMatrix(ctx, H.n)
        nocc = ctypes.c_int64()
        E = np.zeros(H.n)
        check(lib().lmh_densitymatrix(ctx.h, P.h, float(mu), float(T), ctypes.byref(nocc), ptr(E)))
        P.nocc = nocc.value
        return (P, E) if return_energies else P
    eig = diagonalize(H, ctx=ctx)
    try:
        P = eig.densitymatrix(T=T, mu=mu, N=N, statistics=statistics)
    finally:
        eig.close()
    return (P, eig.values) if return_energies else P


def site_density(ctx, rho):
    """Per-site density: the sum over a site's internal states
    (src/operators/latticeutils.jl:41-45)."""
    N = ctx.internal
    return rho if N == 1 else rho.reshape(-1, N).sum(axis=1)


def _current_blocks(H):
    """For every lattice bond k = (i, j) and internal pair (a, b): the scalar
    bond carrying H[iN+a, jN+b] (sign +1) or its partner (sign -1), so that
    DensityCurrents[i, j] = sum_{a,b} 2 Im(H[iN+a, jN+b] P[jN+b, iN+a])
    (src/zoo/currents.jl:92-102) is a signed sum of the per-bond currents,
    each stored entry counted once."""
    agg = getattr(H, "_agg", None)
    if agg is not None:
        return agg
    N = H.internal
    bonds, _amp = H.geometry()
    first = {}
    for b, (r, c) in enumerate(zip(bonds.i.tolist(), bonds.j.tolist())):
        if r // N == c // N or (r, c) in first or (c, r) in first:
            continue
        first[(r, c)] = b
    ks, bs, sg = [], [], []
    sb = H.bonds
    for k, (i, j) in enumerate(zip(sb.i.tolist(), sb.j.tolist())):
        for a in range(N):
            for bb in range(N):
                r, c = i * N + a, j * N + bb
                if (r, c) in first:
                    ks.append(k); bs.append(first[(r, c)]); sg.append(1.0)
                elif (c, r) in first:
                    ks.append(k); bs.append(first[(c, r)]); sg.append(-1.0)
    H._agg = (np.array(ks, dtype=np.int64), np.array(bs, dtype=np.int64), np.array(sg))
    return H._agg


def site_currents(ctx, J):
    """Per-lattice-bond currents from the per-scalar-bond ones of the device."""
    if ctx.internal == 1:
        return J
    H = ctx.geom_h
    ks, bs, sg = _current_blocks(H)
    out = np.zeros(len(H.bonds))
    np.add.at(out, ks, sg * J[bs])
    return out


def localdensity(P, solver=None):
    """localdensity(P) (src/operators/latticeutils.jl:41-45) of a device state
    (for kets |psi_i|^2 of the first ket), summed over internal states."""
    if isinstance(P, Kets):
        return site_density(P.ctx, P.observables(0, J=False)[0])
    rho = np.zeros(P.n)
    check(lib().lmh_observables(P.ctx.h, P.h, ptr(rho), None))
    return site_density(P.ctx, rho)


class DensityCurrents:
    """DensityCurrents(H, P) (src/zoo/currents.jl:81-102): 2 Im(H_ij P_ji) per
    bond of H, evaluated on the device with the H the solver last built."""

    def __init__(self, H, P):
        self.H = H
        self.P = P

    def values(self):
        self.P.ctx.set_hamiltonian(self.H)
        if isinstance(self.P, Kets):
            return site_currents(self.P.ctx, self.P.observables(0, rho=False)[1])
        J = np.zeros(len(self.P.ctx._geom_refs[0]))
        check(lib().lmh_observables(self.P.ctx.h, self.P.h, None, ptr(J)))
        return site_currents(self.P.ctx, J)

    def __array__(self, dtype=None):
        return self.values() if dtype is None else self.values().astype(dtype)


def observables(P, H=None):
    """Both observables in one fused launch: (rho[N], J[nbonds]); J with the
    given H (default: the Hamiltonian of the context's current parameters,
    i.e. the one of the last step)."""
    if H is not None:
        P.ctx.set_hamiltonian(H)
    if isinstance(P, Kets):
        r, j = P.observables(0)
        return site_density(P.ctx, r), site_currents(P.ctx, j)
    nb = len(P.ctx._geom_refs[0])
    rho = np.zeros(P.n)
    J = np.zeros(nb)
    check(lib().lmh_observables(P.ctx.h, P.h, ptr(rho), ptr(J)))
    return site_density(P.ctx, rho), site_currents(P.ctx, J)
