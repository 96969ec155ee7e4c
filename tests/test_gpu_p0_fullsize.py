"""GPU: the Fermi projector P0 = densitymatrix(H, mu = 0) of the N = 10^4
benchmark lattice (SquareLattice 100 x 100, periodic, on-site disorder
uniform in [-0.5, 0.5], no field), checked at full size through what the
mathematics fixes for an orthogonal projector onto the eigenvectors with
E <= mu (src/spectrum.jl:226-232, 317-333):

    max |P^2 - P| <= 1e-10        (idempotent; dense product on the device)
    Tr P = #(E <= mu)             (rank; and the library's own count)
    max |H P - P H| <= 1e-10      (commutes with H; sparse H on the host)
    P = P^H, E ascending and inside H's Gershgorin interval

With no field every hopping is real, so H is real symmetric and the library
takes the real eigensolver; the test asserts that it did.  The eigenvalues
are pinned against LAPACK on a 2500-site slice of the same construction
(SquareLattice 50 x 50, the first 2500 disorder values): every eigenvalue to
1e-12 and the projector at a mu inside a spectral gap to 1e-10.
"""
import ctypes
import sys

import numpy as np
import pytest
import scipy.sparse

import lmhip
from helpers import Model
from lmhip import _lib
from oracle import numerics as ON

pytestmark = pytest.mark.gpu

SEED, L, W = 20240601, 100, 1.0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _disorder(n):
    return np.random.default_rng(SEED).uniform(-W / 2, W / 2, n)


def _eigen_used(ctx):
    """LMH_OPT_EIGEN_USED: solver method, + 4 when it was the real one."""
    v = ctypes.c_int64()
    _lib.check(_lib.lib().lmh_get_option(ctx.h, _lib.OPT_EIGEN_USED, ctypes.byref(v)))
    return v.value


def test_fermi_projector_n10000_full_size():
    V = _disorder(L * L)
    m = Model("square", (L, L), (True, True), onsite=V)
    ctx = lmhip.Context(0)
    say("case P0 n10000 mu=0 (real H)")
    P, E = lmhip.densitymatrix(m.device_h([]), mu=0.0, ctx=ctx, return_energies=True)
    used = _eigen_used(ctx)
    say("eigensolver used: method %d, %s" % (used % 4, "real" if used >= 4 else "complex"))
    assert used >= 4, "real symmetric H did not take the real eigensolver"
    n = L * L
    nocc = int(np.sum(E <= 0.0))
    assert P.nocc == nocc
    assert np.all(np.diff(E) >= 0.0)
    assert E[0] >= -4.0 - W / 2 - 1e-12 and E[-1] <= 4.0 + W / 2 + 1e-12
    idem = P.idempotency_error()
    say("max |P^2 - P| = %.3e, nocc = %d" % (idem, nocc))
    assert idem <= 1e-10
    Pg = P.to_numpy()
    P.close()
    tr = float(np.trace(Pg).real)
    assert abs(tr - nocc) <= 1e-8, (tr, nocc)
    assert np.max(np.abs(Pg - Pg.conj().T)) <= 1e-12
    csc = m.oracle_h([]).csc
    Hc = scipy.sparse.csc_matrix((csc.nzval, csc.rowval, csc.colptr), shape=(n, n))
    Hr = Hc.tocsr()
    comm = 0.0
    for c0 in range(0, n, 2500):
        HP = Hr @ Pg[:, c0:c0 + 2500]                   # (H P)[:, cols]
        PH = Pg @ Hc[:, c0:c0 + 2500]                   # (P H)[:, cols]
        comm = max(comm, float(np.max(np.abs(HP - PH))))
    say("max |HP - PH| = %.3e" % comm)
    assert comm <= 1e-10


def test_fermi_projector_slice_n2500_against_lapack():
    Ls = 50
    V = _disorder(L * L)[: Ls * Ls]
    m = Model("square", (Ls, Ls), (True, True), onsite=V)
    Hd = m.oracle_h([]).dense()
    assert np.max(np.abs(Hd.imag)) == 0.0
    E_ref = np.linalg.eigvalsh(Hd.real)
    # mu inside the widest of the gaps around 0 (the projector is unique there)
    k = int(np.searchsorted(E_ref, 0.0))
    g, i = max((E_ref[j + 1] - E_ref[j], j) for j in range(k - 4, k + 4))
    mu = 0.5 * (E_ref[i] + E_ref[i + 1])
    ctx = lmhip.Context(0)
    say("case P0 slice n2500 mu=%.6f gap=%.2e" % (mu, g))
    P, E = lmhip.densitymatrix(m.device_h([]), mu=mu, ctx=ctx, return_energies=True)
    assert _eigen_used(ctx) >= 4
    assert np.max(np.abs(E - E_ref)) <= 1e-12
    P_ref, _ = ON.densitymatrix(Hd, mu=mu)
    Pg = P.to_numpy()
    assert np.max(np.abs(Pg - P_ref)) <= 1e-10
    assert P.nocc == int(np.sum(E_ref <= mu))
    # a field makes H complex: the complex solver, same checks
    fs = [("landau", 0.01, 0, 0)]
    Hc = m.oracle_h(fs).dense()
    Ec = np.linalg.eigvalsh(Hc)
    Pc, E2 = lmhip.densitymatrix(m.device_h(fs), mu=mu, ctx=ctx, return_energies=True)
    assert _eigen_used(ctx) < 4
    assert np.max(np.abs(E2 - Ec)) <= 1e-12
