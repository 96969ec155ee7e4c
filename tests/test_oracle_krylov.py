"""Pins of the Krylov ket oracle (CPU only): the exact propagator
(scipy.linalg.expm, a textbook routine) on random Hermitian matrices and
lattice Hamiltonians, the reference's own KrylovKitExp test
(test/test_timedeps.jl:42-68), norm conservation, the sub-step path and the
non-convergence error."""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import fields as F
from oracle import geometry as G
from oracle import hamiltonian as H
from oracle import krylov as KR


def _herm(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return (A + A.conj().T) / 2


def _ket(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("dt", [0.05, 0.3, -0.2])
def test_random_hermitian_against_expm(dt):
    Hd = _herm(40, 1)
    psi = _ket(40, 2)
    y, mv = KR.lanczos_expv(Hd, dt, psi)
    ref = sl.expm(-1j * dt * Hd) @ psi
    assert np.max(np.abs(y - ref)) <= 1e-11
    assert 1 <= mv <= 30


@pytest.mark.parametrize("make", [
    lambda: H.tightbinding_hamiltonian(G.SquareLattice(8, 7, periodic=(True, False)), field=F.LandauGauge(0.03)),
    lambda: H.haldane(G.HoneycombLattice(5, 4), 1.0, 0.1, 0.2, field=F.PointFlux(0.2, (2.7, 1.9))),
])
def test_lattice_hamiltonians_against_expm(make):
    Hd = make().dense()
    psi = _ket(len(Hd), 3)
    for dt in (0.1, 0.5):
        y, _ = KR.lanczos_expv(Hd, dt, psi)
        assert np.max(np.abs(y - sl.expm(-1j * dt * Hd) @ psi)) <= 1e-11


def test_reference_qwz_groundstate_evolution():
    """test/test_timedeps.jl:42-68: QWZ on SquareLattice(10, 10) (m = 1, no
    field), its ground state, Evolution over 0:0.1:10 with KrylovKitExp;
    component 2 (1-based) after every step against the dense exp at atol
    1e-10 -- here the whole ket, at the same tolerance."""
    Hd = H.qwz(G.SquareLattice(10, 10)).dense()
    E, Vv = np.linalg.eigh(Hd)
    psi = Vv[:, int(np.argmin(E))]
    U = sl.expm(-1j * 0.1 * Hd)
    exact = psi.copy()
    worst = 0.0
    for _ in range(100):
        psi, _ = KR.lanczos_expv(Hd, 0.1, psi)
        exact = U @ exact
        worst = max(worst, np.max(np.abs(psi - exact)))
        assert abs(psi[1] - exact[1]) <= 1e-10
    assert worst <= 1e-10
    assert abs(np.linalg.norm(psi) - 1.0) <= 1e-12


def test_substeps_when_the_basis_is_short():
    Hd = H.tightbinding_hamiltonian(G.SquareLattice(9, 9), field=F.LandauGauge(0.04)).dense()
    psi = _ket(81, 5)
    y, mv = KR.lanczos_expv(Hd, 1.5, psi, krylovdim=10)
    assert mv > 20                     # restarted with sub-steps
    assert np.max(np.abs(y - sl.expm(-1.5j * Hd) @ psi)) <= 1e-10


def test_non_convergence_raises():
    Hd = _herm(30, 7)
    with pytest.raises(KR.KrylovError):
        KR.lanczos_expv(Hd, 5.0, _ket(30, 8), krylovdim=2, maxiter=1)
