"""Oracle: numerics (ctypes over lmoracle.c), spectrum, observables, Evolution.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates:
  src/evolution.jl:62-92      CachedExp, update_solver! (cache hit when the factor is
                              ≈ the old one and H equals the old H)
  src/evolution.jl:93-128     myexp!                       -> lmoracle.c
  src/evolution.jl:73-78      step! on a density matrix    -> lmoracle.c
  src/evolution.jl:238-275    step!(::Evolution) / iterate: dt_i = t_i - t_{i-1}, the
                              first dt = t_1 - time; |dt| < 1e-15 is no step; H is
                              evaluated at the time BEFORE the step; time += dt;
                              yields (H used, states, time)
  src/spectrum.jl:226-232,253-257,317-333   projector(densfun(T, mu), eig) with
                              f(E) = 1 if E - mu == 0 at T == 0, Fermi-Dirac otherwise
  src/currents.jl:76-110      currentsfrom / currentsfromto of DensityCurrents
  src/timesequence.jl:213-268 differentiate! / integrate!
  src/spectrum.jl:233-288     the densitymatrix family: ddensfun, ensemble (Fermi-Dirac /
                              Bose-Einstein), Gibbs, ground state, Fermi sphere, fixed N
                              (Newton's iteration for mu), the keyword dispatch
  src/operators/latticeutils.jl:41-45      localdensity = real(diag P)
  src/zoo/currents.jl:81-102  DensityCurrents[i, j] = 2 Im(H[i,j] P[j,i])

Pins (tests/test_oracle_numerics.py): U vs scipy.linalg.expm at 1e-10; K vs a
brute-force dense evaluation of the stopping rule; P0 idempotent with
Tr P0 = #(E <= mu) and [H, P0] = 0; Gibbs T=0 equals the ground-state
projector (test/test_operators.jl:232-236); current antisymmetry and the
Heisenberg sum rule (test/test_currents.jl:14-26); the von Neumann density
(test/test_operators.jl:35-42); stationarity of P0 under constant H.
The droptol boundary of myexp! and LAPACK's basis inside degenerate
eigenspaces are parity unpinned (no reference run is possible here).
"""
import ctypes
import math
import os
from fractions import Fraction

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    """Load oracle/_build/liblmoracle.so (build with `make -C oracle`)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "liblmoracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", _HERE])
        L = ctypes.CDLL(path)
        vp, i64, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
        L.lmo_csc_new.restype = vp
        L.lmo_csc_new.argtypes = [i64, vp, vp, vp]
        L.lmo_csc_free.argtypes = [vp]
        L.lmo_csc_nnz.restype = i64
        L.lmo_csc_nnz.argtypes = [vp]
        L.lmo_csc_export.argtypes = [vp, vp, vp, vp]
        L.lmo_myexp.restype = vp
        L.lmo_myexp.argtypes = [vp, dbl, dbl, dbl, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(dbl), vp, ctypes.c_int]
        L.lmo_step.argtypes = [vp, i64, vp, vp, ctypes.c_int]
        L.lmo_left_cols.argtypes = [vp, i64, vp, vp, i64, i64, ctypes.c_int]
        L.lmo_right_rows.argtypes = [vp, i64, vp, vp, i64, i64, ctypes.c_int]
        L.lmo_horner_step.argtypes = [vp, i64, dbl, ctypes.c_int, dbl, vp, vp, vp, ctypes.c_int]
        L.lmo_spmm.argtypes = [vp, i64, i64, vp, vp, ctypes.c_int]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class SparseCSC:
    """Owning handle on an lmoracle CSC matrix."""

    def __init__(self, handle, n):
        self.h = handle
        self.n = n

    @classmethod
    def from_csc(cls, csc):
        colptr = np.ascontiguousarray(csc.colptr, dtype=np.int64)
        rowval = np.ascontiguousarray(csc.rowval, dtype=np.int64)
        nzval = np.ascontiguousarray(csc.nzval, dtype=np.complex128)
        h = lib().lmo_csc_new(csc.n, _ptr(colptr), _ptr(rowval), _ptr(nzval))
        return cls(h, csc.n)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.lmo_csc_free(self.h)
            self.h = None

    @property
    def nnz(self):
        return lib().lmo_csc_nnz(self.h)

    def to_scipy(self):
        """The same matrix as scipy.sparse.csc_matrix (0-based)."""
        import scipy.sparse as sp
        nnz = self.nnz
        colptr = np.zeros(self.n + 1, dtype=np.int64)
        rowval = np.zeros(nnz, dtype=np.int64)
        nzval = np.zeros(nnz, dtype=np.complex128)
        lib().lmo_csc_export(self.h, _ptr(colptr), _ptr(rowval), _ptr(nzval))
        return sp.csc_matrix((nzval, rowval, colptr), shape=(self.n, self.n))

    def dense(self):
        nnz = self.nnz
        colptr = np.zeros(self.n + 1, dtype=np.int64)
        rowval = np.zeros(nnz, dtype=np.int64)
        nzval = np.zeros(nnz, dtype=np.complex128)
        lib().lmo_csc_export(self.h, _ptr(colptr), _ptr(rowval), _ptr(nzval))
        d = np.zeros((self.n, self.n), dtype=np.complex128)
        for j in range(self.n):
            s, e = colptr[j], colptr[j + 1]
            d[rowval[s:e], j] = nzval[s:e]
        return d


def myexp(H, dt, threshold=1e-10, nztol=1e-14, droptol=True):
    """myexp!(U, H, -im*dt) (src/evolution.jl:93-128). Returns (U, K, s, deltas)."""
    Hh = SparseCSC.from_csc(H.csc)
    K = ctypes.c_int(0)
    s = ctypes.c_double(1.0)
    deltas = np.zeros(256)
    h = lib().lmo_myexp(Hh.h, float(dt), float(threshold), float(nztol), int(droptol),
                        ctypes.byref(K), ctypes.byref(s), _ptr(deltas), 256)
    return SparseCSC(h, H.n), K.value, s.value, deltas[:K.value].copy()


def step(U, P, nthreads=0):
    """step!(::CachedExp, P, cache): P <- (U*P)*U' in place (column-major P)."""
    n = U.n
    assert P.shape == (n, n) and P.dtype == np.complex128
    Pf = np.asfortranarray(P)
    work = np.zeros((n, n), dtype=np.complex128, order="F")
    lib().lmo_step(U.h, n, _ptr(Pf), _ptr(work), nthreads)
    if Pf is not P:
        P[...] = Pf
    return P


def apply(U, X, nthreads=0):
    """U @ X for a thin column-major X (n x m)."""
    X = np.asfortranarray(X, dtype=np.complex128)
    n, m = X.shape
    Y = np.zeros((n, m), dtype=np.complex128, order="F")
    lib().lmo_spmm(U.h, n, m, _ptr(X), _ptr(Y), nthreads)
    return Y


def horner_step(H, dt, K, s, P, nthreads=0):
    """Same truncated series as myexp! but without droptol, as Horner passes."""
    n = H.n
    Hh = SparseCSC.from_csc(H.csc)
    Pf = np.asfortranarray(P)
    w1 = np.zeros((n, n), dtype=np.complex128, order="F")
    w2 = np.zeros((n, n), dtype=np.complex128, order="F")
    lib().lmo_horner_step(Hh.h, n, float(dt), int(K), float(s), _ptr(Pf), _ptr(w1), _ptr(w2), nthreads)
    if Pf is not P:
        P[...] = Pf
    return P


FERMI_DIRAC, BOSE_EINSTEIN = 1, -1        # ParticleStatistics (src/operators/system.jl:60-63)
TEMP_TOL = 1e-10                           # src/spectrum.jl:235


def _fdiv(a, b):
    """IEEE a / b as Julia computes it (x / 0 = +-Inf, 0 / 0 = NaN)."""
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        return float(np.float64(a) / np.float64(b))


def _exp(x):
    with np.errstate(over="ignore", invalid="ignore"):
        return float(np.exp(np.float64(x)))


def densfun(T, mu, statistics=FERMI_DIRAC):
    """densfun(T, mu, statistics) (src/spectrum.jl:230-232):
    E -> (E - mu == T == 0) ? 1 : 1 / (exp((E - mu) / T) + Int(statistics))."""
    def f(E):
        if E - mu == T and T == 0:
            return 1.0
        return _fdiv(1.0, _exp(_fdiv(E - mu, T)) + statistics)
    return f


def ddensfun(T, mu, statistics=FERMI_DIRAC):
    """ddensfun (src/spectrum.jl:233-235): -exp(x) / T / (exp(x) + s)^2, x = (E - mu) / T."""
    def f(E):
        ex = np.float64(_exp(_fdiv(E - mu, T)))
        with np.errstate(over="ignore", invalid="ignore"):
            den = (ex + np.float64(statistics)) ** 2
        return _fdiv(_fdiv(-ex, T), den)
    return f


def projector_weights(V, f):
    """projector(f, eig) = V * (f .* V') (src/spectrum.jl:226-228) for weights f."""
    f = np.asarray(f, dtype=np.float64)
    return (V * f[None, :]) @ V.conj().T


def ensemble_densitymatrix(E, V, mu=0.0, T=0.0, statistics=FERMI_DIRAC):
    """src/spectrum.jl:253-257."""
    fn = densfun(T, mu, statistics)
    return projector_weights(V, [fn(e) for e in E])


def groundstate_densitymatrix(E, V):
    """src/spectrum.jl:237-241: psi psi' with psi the eigenvector of the lowest
    eigenvalue (the first one findgroundstate meets, src/spectrum.jl:180-194)."""
    k = int(np.argmin(E))
    psi = V[:, k]
    return np.outer(psi, psi.conj())


def gibbs_densitymatrix(E, V, T=0.0):
    """src/spectrum.jl:242-250: ground state for |T| < 1e-10, else exp(-E/T)/Z."""
    if abs(T) < TEMP_TOL:
        return groundstate_densitymatrix(E, V)
    w = np.array([_exp(_fdiv(-e, T)) for e in E])
    Z = float(np.sum(w))
    return projector_weights(V, w / Z)


def fermisphere_densitymatrix(E, V, N):
    """src/spectrum.jl:258-265: the N lowest levels (ArgumentError if N exceeds
    the number of levels)."""
    if len(E) < N:
        raise ValueError("cannot build Fermi sphere with %d particles: only %d bands present" % (N, len(E)))
    return projector_weights(V[:, :N], np.ones(N))


def fixn_densitymatrix(E, V, N, T=0.0, statistics=FERMI_DIRAC, maxiter=1000, atol=math.sqrt(np.finfo(float).eps)):
    """src/spectrum.jl:266-285: Fermi sphere (or N x ground state for bosons) at
    |T| < 1e-10, else Newton's iteration for mu from 0 until |dmu| < atol;
    returns (P, mu)."""
    if abs(T) < TEMP_TOL:
        if statistics == BOSE_EINSTEIN:
            return groundstate_densitymatrix(E, V) * N, None
        return fermisphere_densitymatrix(E, V, N), None
    newmu = 0.0
    for _ in range(maxiter):
        f, df = densfun(T, newmu, statistics), ddensfun(T, newmu, statistics)
        n_ev = sum(f(e) for e in E)
        dn_ev = sum(df(e) for e in E)
        dmu = _fdiv(n_ev - N, dn_ev)
        if abs(dmu) < atol:
            return ensemble_densitymatrix(E, V, mu=newmu, T=T, statistics=statistics), newmu
        newmu += dmu
    raise ValueError("did not converge")


def densitymatrix_kw(Hdense, T=0.0, mu=None, N=None, statistics=None):
    """densitymatrix(h; kw...) of a one-particle system with dense
    diagonalisation (src/spectrum.jl:317-333, OneParticleSystem dispatch):
    N -> fixed particle number; mu or statistics -> ensemble; else Gibbs."""
    E, V = np.linalg.eigh(Hdense)
    if N is not None:
        if mu is not None:
            raise ValueError("cannot specify both N and μ")
        return fixn_densitymatrix(E, V, N, T=T, statistics=FERMI_DIRAC if statistics is None else statistics)[0]
    if mu is not None or statistics is not None:
        return ensemble_densitymatrix(E, V, mu=0.0 if mu is None else mu, T=T,
                                      statistics=FERMI_DIRAC if statistics is None else statistics)
    return gibbs_densitymatrix(E, V, T)


def densitymatrix(Hdense, mu=0.0, T=0.0):
    """densitymatrix(h, mu=mu) for a OneParticleSystem with dense diagonalisation:
    projector(densfun(T, mu), eigen) = V * (f.(E) .* V')  (src/spectrum.jl:226-228)."""
    E, V = np.linalg.eigh(Hdense)
    f = np.array([densfun(T, mu)(e) for e in E])
    return (V * f[None, :]) @ V.conj().T, E


def localdensity(P, internal=1):
    """localdensity (src/operators/latticeutils.jl:41-45): real(P_ii) summed
    over the `internal` states of each site."""
    d = np.real(np.diag(P))
    if internal == 1:
        return d.copy()
    return d.reshape(-1, internal).sum(axis=1)


def site_current(H, P, i, j):
    """DensityCurrents(H, P)[i, j] (src/zoo/currents.jl:92-102): the sum over
    the internal states a, b of 2 Im(H[iN+a, jN+b] P[jN+b, iN+a])."""
    N = getattr(H, "internal", 1)
    if N == 1:
        return 2 * (H[(i, j)] * P[j, i]).imag
    acc = 0.0
    for a in range(N):
        for b in range(N):
            acc += 2 * (H[(i * N + a, j * N + b)] * P[j * N + b, i * N + a]).imag
    return acc


def density_currents(H, P, bonds=None):
    """DensityCurrents(H, P)[i, j] for each bond (i, j) of sites."""
    bonds = H.bonds if bonds is None else bonds
    out = np.empty(len(bonds))
    for b, (i, j) in enumerate(bonds):
        out[b] = site_current(H, P, i, j)
    return out


def _site_pairs(H):
    N = getattr(H, "internal", 1)
    return sorted({(r // N, c // N) for (r, c) in H.entries if r // N != c // N})


def currentsfrom(H, P, src):
    """currentsfrom(DensityCurrents(H, P), src) (src/currents.jl:76-88) with
    curr[i, j] = 2 Im(H[i,j] P[j,i]) over the stored entries of H:
    out[j] = 0 for j in src, else sum over i in src of curr[i, j]."""
    n = H.n // getattr(H, "internal", 1)
    src = set(int(i) for i in src)
    out = np.zeros(n)
    for (i, j) in _site_pairs(H):
        if i in src and j not in src:
            out[j] += site_current(H, P, i, j)
    return out


def currentsfromto(H, P, src, dst=None):
    """currentsfromto(DensityCurrents(H, P), src, dst) (src/currents.jl:90-110):
    sum of curr[i, j] over i in src, j in dst (dst = the other sites if None)."""
    n = H.n // getattr(H, "internal", 1)
    src = set(int(i) for i in src)
    dst = set(range(n)) - src if dst is None else set(int(j) for j in dst)
    tot = 0.0
    for (i, j) in _site_pairs(H):
        if i in src and j in dst:
            tot += site_current(H, P, i, j)
    return tot


def differentiate(times, values):
    """differentiate!(TimeSequence) (src/timesequence.jl:213-225):
    v[i-1] <- v[i] / dt - v[i-1] / dt at t[i-1] + dt / 2, the last entry dropped."""
    td = [float(t) for t in times]
    vs = [np.array(v, dtype=float) for v in values]
    if len(td) < 2:
        raise RuntimeError("Cannot differentiate TimeSequence of length %d" % len(td))
    for i in range(1, len(td)):
        dt = td[i] - td[i - 1]
        vs[i - 1] = (1 / dt) * vs[i] + (-1 / dt) * vs[i - 1]
        td[i - 1] += dt / 2
    return td[:-1], vs[:-1]


def integrate(times, values):
    """integrate!(TimeSequence) (src/timesequence.jl:255-268): trapezoids
    (dt/2) (v[i] + v[i-1]), shifted behind a zero, then the running sum."""
    td = [float(t) for t in times]
    vs = [np.array(v, dtype=float) for v in values]
    if len(td) < 2:
        raise RuntimeError("Cannot integrate TimeSequence of length %d" % len(td))
    for i in range(1, len(td)):
        dt = td[i] - td[i - 1]
        vs[i - 1] = (dt / 2) * vs[i] + (dt / 2) * vs[i - 1]
    last = vs.pop()
    vs.insert(0, np.zeros_like(last))
    for i in range(1, len(vs)):
        vs[i] = 1 * vs[i - 1] + 1 * vs[i]
    return td, vs


def julia_range(start, stp, stop):
    """Elements of Julia's `start:step:stop` for decimal literals: k*step rounded
    once from the exact decimal values (StepRangeLen with TwicePrecision)."""
    a = Fraction(repr(float(start))) if not isinstance(start, Fraction) else start
    s = Fraction(repr(float(stp))) if not isinstance(stp, Fraction) else stp
    b = Fraction(repr(float(stop))) if not isinstance(stop, Fraction) else stop
    n = int(math.floor((b - a) / s)) + 1
    return [float(a + k * s) for k in range(n)]


class Evolution:
    """Evolution(h, P0) with the default CachedExp solver (src/evolution.jl:183-298).

    `hfun(t)` returns an OracleHamiltonian.  Iterating `ev(times)` yields
    (H_used, P, t) after each step, P being the evolved copy (not a snapshot).
    """

    def __init__(self, hfun, P0, threshold=1e-10, nztol=1e-14, droptol=True, nthreads=0):
        self.hfun = hfun
        self.P = np.array(P0, dtype=np.complex128, order="F", copy=True)
        self.time = 0.0
        self.threshold = threshold
        self.nztol = nztol
        self.droptol = droptol
        self.nthreads = nthreads
        self._mat = None
        self._factor = 0j
        self.U = None
        self.K = 0
        self.s = 1.0
        self.K_history = []

    def _same_h(self, H):
        if self._mat is None:
            return False
        if H is self._mat:
            return True
        a, b = self._mat.csc, H.csc
        return (a.n == b.n and np.array_equal(a.colptr, b.colptr) and
                np.array_equal(a.rowval, b.rowval) and np.array_equal(a.nzval, b.nzval))

    def update_solver(self, H, dt, force=False):
        factor = complex(0.0, -dt)
        if not force:
            same_f = factor == self._factor or abs(factor - self._factor) <= math.sqrt(2.220446049250313e-16) * max(abs(factor), abs(self._factor))
            if same_f and self._same_h(H):
                return False
        self._mat = H
        self.U, self.K, self.s, _ = myexp(H, dt, self.threshold, self.nztol, self.droptol)
        self._factor = factor
        return True

    def step(self, dt):
        if dt < -1e-15:
            raise ValueError("negative time step")
        H = self.hfun(self.time)
        if abs(dt) < 1e-15:
            return H
        self.update_solver(H, dt)
        self.K_history.append(self.K)
        step(self.U, self.P, self.nthreads)
        self.time += dt
        return H

    def __call__(self, times):
        prev = None
        for i, t in enumerate(times):
            dt = (t - prev) if i > 0 else (t - self.time)
            prev = t
            H = self.step(dt)
            yield H, self.P, float(self.time)
