"""CPU pins of the discrete adjoint (dense numpy model of the algorithm the
device kernels implement) against central finite differences of the forward
evolution: the oracle's H builder for H(theta), the same order-K polynomial
for U.  Tolerances: FD with h = 1e-6 is good to ~1e-8 relative here."""
import numpy as np
import pytest

import adjoint_model as AM
from oracle import fields as OF
from oracle import geometry as OG
from oracle import hamiltonian as OH
from oracle import numerics as ON


def _h(lat, V, flux, point):
    return OH.tightbinding_hamiltonian(lat, onsite=V, field=OF.PointFlux(flux, point)).dense()


def _setup(seed=0, L=(4, 3), nsteps=7, dt=0.1):
    rng = np.random.default_rng(seed)
    lat = OG.SquareLattice(*L)
    n = L[0] * L[1]
    V = rng.normal(0, 0.3, n)
    B = rng.uniform(0, 0.3, nsteps)
    point = (2.3, 1.7)
    H0 = _h(lat, V, 0.0, point)
    P0, _ = ON.densitymatrix(H0, mu=0.1)
    Ks = [ON.myexp(OH.tightbinding_hamiltonian(lat, onsite=V, field=OF.PointFlux(b, point)), dt)[1] for b in B]
    return lat, V, B, point, P0, Ks, dt


def _loss_and_cot(P, w):
    # L = sum_i w_i rho_i + (Im P_01)^2 : a diagonal part and a non-Hermitian-looking cotangent
    L = float(np.sum(w * np.real(np.diag(P)))) + float(np.imag(P[0, 1])) ** 2
    Lam = np.diag(w).astype(np.complex128)
    # d(Im P01)^2 = 2 Im P01 * Im dP01 = Re(conj(Lam01) dP01) with conj(Lam01) = -2i Im P01
    Lam[0, 1] += np.conj(-2j * np.imag(P[0, 1]))
    return L, Lam


def _run(lat, V, B, point, P0, Ks, dt):
    Hs = [_h(lat, V, b, point) for b in B]
    Ps = AM.forward(Hs, [dt] * len(B), Ks, P0)
    return Hs, Ps


@pytest.mark.parametrize("C", [1, 3, 10])
def test_adjoint_model_matches_finite_differences(C):
    lat, V, B, point, P0, Ks, dt = _setup()
    n = len(V)
    w = np.linspace(0.5, 1.5, n)
    Hs, Ps = _run(lat, V, B, point, P0, Ks, dt)
    L0, Lam = _loss_and_cot(Ps[-1], w)
    cps = {k: Ps[k] for k in range(0, len(B), C)}
    Lam0, Fs = AM.backward(Hs, [dt] * len(B), Ks, Ps[-1], Lam, cps, C)
    hstep = 1e-6
    gV = sum(np.real(np.diag(F)) for F in Fs)
    gB = []
    for k, F in enumerate(Fs):
        dH = (_h(lat, V, B[k] + hstep, point) - _h(lat, V, B[k] - hstep, point)) / (2 * hstep)
        gB.append(float(np.real(np.sum(F * dH))))

    def loss(Vx, Bx, P0x=P0):
        _, Ps_ = _run(lat, Vx, Bx, point, P0x, Ks, dt)
        return _loss_and_cot(Ps_[-1], w)[0]

    for i in (0, 3, n - 1):
        e = np.zeros(n); e[i] = hstep
        fd = (loss(V + e, B) - loss(V - e, B)) / (2 * hstep)
        assert abs(fd - gV[i]) <= 1e-7 * max(1.0, abs(fd)), (i, fd, gV[i])
    for k in (0, 2, len(B) - 1):
        e = np.zeros(len(B)); e[k] = hstep
        fd = (loss(V, B + e) - loss(V, B - e)) / (2 * hstep)
        assert abs(fd - gB[k]) <= 1e-7 * max(1.0, abs(fd)), (k, fd, gB[k])
    # dL/dP0 along a Hermitian direction
    rng = np.random.default_rng(5)
    D = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    D = D + D.conj().T
    fd = (loss(V, B, P0 + hstep * D) - loss(V, B, P0 - hstep * D)) / (2 * hstep)
    assert abs(fd - float(np.real(np.sum(Lam0.conj() * D)))) <= 1e-7 * max(1.0, abs(fd))


def test_adjoint_model_reverse_recompute_is_exact_to_unitarity():
    """P_k recomputed by reverse unitary steps (no checkpoints but P_0) stays
    within the polynomial's unitarity defect of the forward P_k."""
    lat, V, B, point, P0, Ks, dt = _setup(nsteps=9)
    Hs, Ps = _run(lat, V, B, point, P0, Ks, dt)
    P = Ps[-1]
    for k in range(len(B) - 1, -1, -1):
        Y = AM.apply_Udag(Hs[k], P, dt, Ks[k])
        P = AM.apply_Udag(Hs[k], Y.conj().T, dt, Ks[k])
        assert np.max(np.abs(P - Ps[k])) <= 1e-9
