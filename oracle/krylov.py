"""Krylov ket propagation: the oracle of ``KrylovKitExp``.

TEST INFRASTRUCTURE ONLY (see the package docstring).

The reference's ``KrylovKitExp`` (src/evolution.jl:130-154) stores
``factor = -im * dt`` in ``update_solver!`` and calls
``KrylovKit.exponentiate(H, factor, psi; kw...)`` in ``step!``, throwing
``ArgumentError("`exponentiate` did not converge")`` when KrylovKit reports
no convergence.  KrylovKit is a third-party package (compat "0.4 - 0.9",
Project.toml:26) that is not part of the reference tree.  Its published
algorithm for a Hermitian operator: a Lanczos basis of at most ``krylovdim``
vectors (default 30) per restart, the exponential of the small tridiagonal
matrix, and an adaptive time step so that the error per unit time stays below
``tol`` (default 1e-12), at most ``maxiter`` (default 100) restarts.  This
module restates that algorithm with the acceptance rule the device uses: a
(sub-)step of length tau is accepted when beta_j |e_j' exp(-i tau T_j) e_1|
|psi| <= tol tau |psi|, and a basis that does not reach it halves the step.

Pins (tests/test_oracle_krylov.py): scipy.linalg.expm on small Hermitian
matrices and lattice Hamiltonians (a textbook routine: the exact propagator,
which KrylovKit approximates to tol per unit time); the reference's own test
(test/test_timedeps.jl:42-68: QWZ on SquareLattice(10, 10), its ground state,
``0:0.1:10``, component 2 against the dense ``exp`` at atol 1e-10); norm
conservation; the sub-step path (tiny krylovdim) against the same exact
result; and the non-convergence error for maxiter too small.  KrylovKit's own
arithmetic (its orthogonalisation and step-size heuristics) is parity
unpinned: only its result, exp(-i dt H) psi to tol per unit time, is.
"""
import numpy as np


class KrylovError(ArithmeticError):
    """exponentiate did not converge (the reference's ArgumentError)."""


def _apply(H, v):
    return H @ v


def lanczos_expv(H, dt, psi, krylovdim=30, tol=1e-12, maxiter=100):
    """exp(-i dt H) psi for a Hermitian H (dense or scipy.sparse), as
    KrylovKit.exponentiate(H, -im*dt, psi) is used by step!(::KrylovKitExp)
    (src/evolution.jl:150-153).  Returns (psi_new, matvecs)."""
    psi = np.asarray(psi, dtype=np.complex128).copy()
    T = abs(float(dt))
    sgn = -1.0 if dt < 0 else 1.0
    rem = T
    matvecs = 0
    restarts = 0
    while rem > 1e-15 * T:
        if restarts >= maxiter:
            raise KrylovError("`exponentiate` did not converge")
        restarts += 1
        b0 = np.linalg.norm(psi)
        if b0 == 0.0:
            return psi, matvecs
        V = [psi / b0]
        alpha, beta = [], []
        accepted = None
        for j in range(1, krylovdim + 1):
            w = _apply(H, V[-1])
            matvecs += 1
            if j > 1:
                w = w - beta[-1] * V[-2]
            a = np.vdot(V[-1], w)
            w = w - a * V[-1]
            alpha.append(a.real)
            b = np.linalg.norm(w)
            beta.append(b)
            V.append(w / b if b > 0 else w)
            Tj = np.diag(alpha) + np.diag(beta[:-1], 1) + np.diag(beta[:-1], -1)
            lam, Z = np.linalg.eigh(Tj)
            y = b0 * (Z @ (np.exp(-1j * sgn * rem * lam) * Z[0, :]))
            if b * abs(y[-1]) <= tol * rem * b0:
                accepted = (rem, y)
                break
        if accepted is None:
            tau = rem
            for _ in range(200):
                y = b0 * (Z @ (np.exp(-1j * sgn * tau * lam) * Z[0, :]))
                if beta[-1] * abs(y[-1]) <= tol * tau * b0:
                    accepted = (tau, y)
                    break
                tau *= 0.5
            if accepted is None:
                raise KrylovError("`exponentiate` did not converge")
        tau, y = accepted
        psi = sum(y[k] * V[k] for k in range(len(y)))
        rem -= tau
    return psi, matvecs
