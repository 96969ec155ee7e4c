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


def densfun(T, mu):
    """densfun(T, mu, FermiDirac) (src/spectrum.jl:230-232)."""
    def f(E):
        if E - mu == 0 and T == 0:
            return 1.0
        if T == 0:
            return 1.0 if E - mu < 0 else 0.0
        return 1.0 / (math.exp((E - mu) / T) + 1)
    return f


def densitymatrix(Hdense, mu=0.0, T=0.0):
    """densitymatrix(h, mu=mu) for a OneParticleSystem with dense diagonalisation:
    projector(densfun(T, mu), eigen) = V * (f.(E) .* V')  (src/spectrum.jl:226-228)."""
    E, V = np.linalg.eigh(Hdense)
    f = np.array([densfun(T, mu)(e) for e in E])
    return (V * f[None, :]) @ V.conj().T, E


def localdensity(P):
    return np.real(np.diag(P)).copy()


def density_currents(H, P, bonds=None):
    """DensityCurrents(H, P)[i, j] = 2 Im(H[i,j] * P[j,i]) for each bond (i, j)."""
    bonds = H.bonds if bonds is None else bonds
    out = np.empty(len(bonds))
    for b, (i, j) in enumerate(bonds):
        out[b] = 2 * (H[(i, j)] * P[j, i]).imag
    return out


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
