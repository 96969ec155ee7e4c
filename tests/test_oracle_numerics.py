"""Pins of the oracle's numerics (CPU only): the sparse Taylor exponential
against scipy's expm and a brute-force evaluation of the stopping rule, the
two-sided step, the Fermi projector, the observables' sum rules and the
Evolution iterator semantics."""
import math
import os

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import fields as F
from oracle import geometry as G
from oracle import hamiltonian as H
from oracle import numerics as N


def _models():
    rng = np.random.default_rng(0)
    yield "open_flux", H.tightbinding_hamiltonian(G.SquareLattice(7, 6), field=F.PointFlux(0.13, (3.5, 3.5)))
    yield "pbc_landau_disorder", H.tightbinding_hamiltonian(
        G.SquareLattice(6, 6, periodic=(True, True)), onsite=rng.uniform(-0.5, 0.5, 36), field=F.LandauGauge(0.05))
    yield "haldane", H.haldane(G.HoneycombLattice(4, 4), 1.0, 0.1, 0.2, field=F.PointFlux(0.2, (2.2, 1.9)))


def _brute_K(Hd, dt, thr=1e-10):
    """Stopping rule of src/evolution.jl:105-118 on dense powers (no droptol)."""
    s = 1.0
    mx = np.max(np.abs(Hd))
    while s < mx * dt:
        s *= 2.0
    A = Hd / s
    T = np.eye(len(Hd), dtype=complex)
    n = 1
    while True:
        T = (A @ T) * complex(0.0, -dt) / n
        if np.max(np.abs(T)) <= thr / s:
            return n, s
        n += 1


@pytest.mark.parametrize("dt", [0.02, 0.05, 0.1, 0.7])
def test_myexp_vs_expm_and_brute_force_order(dt):
    for name, h in _models():
        U, K, s, deltas = N.myexp(h, dt)
        Hd = h.dense()
        ref = sl.expm(-1j * dt * Hd)
        assert np.max(np.abs(U.dense() - ref)) < 1e-10, name
        Kb, sb = _brute_K(Hd, dt)
        assert (K, s) == (Kb, sb), name
        assert deltas[-1] <= 1e-10 / s < deltas[-2]


def test_myexp_zero_operator_and_zero_dt():
    h = H.tightbinding_hamiltonian(G.SquareLattice(3, 3), t1=0.0)
    U, K, s, _ = N.myexp(h, 0.1)
    assert K == 0 and np.array_equal(U.dense(), np.eye(9))
    h2 = H.tightbinding_hamiltonian(G.SquareLattice(3, 3))
    U2, K2, _, _ = N.myexp(h2, 0.0)
    assert K2 == 0 and np.array_equal(U2.dense(), np.eye(9))


def test_step_matches_dense_product():
    _name, h = next(_models())
    U, K, s, _ = N.myexp(h, 0.1)
    Ud = U.dense()
    rng = np.random.default_rng(5)
    P = rng.normal(size=(h.n, h.n)) + 1j * rng.normal(size=(h.n, h.n))
    ref = Ud @ P @ Ud.conj().T
    out = N.step(U, P.copy())
    assert np.max(np.abs(out - ref)) < 1e-13


def test_droptol_vs_exact_polynomial_is_below_parity_bar():
    """The device applies the same order-K polynomial without droptol; the
    difference after 200 cfg1-like steps stays far below 1e-10."""
    l = G.SquareLattice(10, 10)
    tau = 10.0
    hf = lambda t: H.tightbinding_hamiltonian(l, field=F.PointFlux(0.2 * min(t, tau) / tau, (5.5, 5.5)))
    P0, _ = N.densitymatrix(hf(0).dense(), 0.05)
    ts = N.julia_range(0, 0.1, 20)
    ev = N.Evolution(hf, P0)
    Ph = np.array(P0, order="F")
    worst = 0.0
    for k, (Hk, P, t) in enumerate(ev(ts)):
        if k == 0:
            continue
        N.horner_step(Hk, ts[k] - ts[k - 1], ev.K, ev.s, Ph)
        worst = max(worst, float(np.max(np.abs(P - Ph))))
    assert worst < 1e-12


def test_densitymatrix_projector_properties():
    _name, h = list(_models())[1]
    Hd = h.dense()
    P, E = N.densitymatrix(Hd, mu=0.0)
    assert np.max(np.abs(P @ P - P)) < 1e-12
    assert abs(np.trace(P).real - np.sum(E <= 0.0)) < 1e-12
    assert np.max(np.abs(Hd @ P - P @ Hd)) < 1e-12
    # ground-state projector (test/test_operators.jl:232-236): mu between E0 and E1
    P1, E1 = N.densitymatrix(Hd, mu=0.5 * (E[0] + E[1]))
    w, v = np.linalg.eigh(Hd)
    assert np.max(np.abs(P1 - np.outer(v[:, 0], v[:, 0].conj()))) < 1e-12
    # tie rule: E - mu == 0 counts as occupied at T = 0
    assert N.densfun(0.0, 1.5)(1.5) == 1.0 and N.densfun(0.0, 1.5)(1.5000001) == 0.0
    # finite temperature: trace = sum of Fermi factors
    PT, ET = N.densitymatrix(Hd, mu=0.1, T=0.3)
    assert abs(np.trace(PT).real - np.sum(1 / (np.exp((ET - 0.1) / 0.3) + 1))) < 1e-12


def test_currents_antisymmetry_and_heisenberg():
    # test/test_currents.jl:14-26 and test/test_operators.jl:35-42
    l = G.SquareLattice(5, 5)
    h = H.tightbinding_hamiltonian(l, field=F.LandauGauge(0.1))
    h0 = H.tightbinding_hamiltonian(l)
    P, _ = N.densitymatrix(h0.dense(), mu=0.2)
    Hd = h.dense()
    J = dict(zip(h.bonds, N.density_currents(h, P)))
    Jrev = N.density_currents(h, P, [(j, i) for (i, j) in h.bonds])
    assert np.max(np.abs(np.array(list(J.values())) + Jrev)) < 1e-14
    dPdt = -1j * (Hd @ P - P @ Hd)
    for s in range(l.__len__()):
        out = sum(v if i == s else -v for (i, j), v in J.items() if s in (i, j))
        heis = (1j * (P @ Hd - Hd @ P))[s, s]
        assert abs(out - heis.real) < 1e-12 and abs(heis.imag) < 1e-12
        assert abs(N.localdensity(dPdt)[s] - heis.real) < 1e-12


def test_ket_evolution_matches_dense_exp():
    # test/test_timedeps.jl:42-68 (pure state as a rank-1 P), atol 1e-10
    l = G.SquareLattice(6, 6)
    h = H.tightbinding_hamiltonian(l, field=F.LandauGauge(0.07))
    Hd = h.dense()
    psi = np.linalg.eigh(H.tightbinding_hamiltonian(l).dense())[1][:, 0]
    ts = N.julia_range(0, 0.1, 10)
    ev = N.Evolution(lambda t: h, np.outer(psi, psi.conj()))
    Ue = sl.expm(-1j * 0.1 * Hd)
    for k, (_H, P, t) in enumerate(ev(ts)):
        ref = np.outer(psi, psi.conj())
        assert np.max(np.abs(P - ref)) < 1e-10
        psi = Ue @ psi


def test_evolution_iterator_semantics():
    calls = []
    l = G.SquareLattice(3, 3)

    def hf(t):
        calls.append(t)
        return H.tightbinding_hamiltonian(l, field=F.LandauGauge(0.1 * t))

    ev = N.Evolution(hf, np.eye(9) / 9)
    ts = N.julia_range(0.0, 0.1, 0.3)
    out = [(t, ev.K) for (_H, _P, t) in ev(ts)]
    assert [o[0] for o in out] == pytest.approx([0.0, 0.1, 0.2, 0.3], abs=1e-15)
    # first timestamp is no step; H evaluated at the left end of each step
    assert calls[0] == 0.0 and calls[1] == 0.0 and calls[2] == pytest.approx(0.1) and calls[3] == pytest.approx(0.2)
    with pytest.raises(ValueError):
        ev.step(-0.5)
    # julia float range: correctly rounded k * step
    assert N.julia_range(0, 0.1, 0.3)[3] == 0.3 and (0.1 * 3) != 0.3


def test_golden_fixture_regression():
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg1.npz"))
    l = G.SquareLattice(10, 10)
    tau = 10.0
    hf = lambda t: H.tightbinding_hamiltonian(l, field=F.PointFlux(0.2 * min(t, tau) / tau, (5.5, 5.5)))
    ev = N.Evolution(hf, g["P0"])
    ts = list(g["ts"])
    for k, (Hk, P, t) in enumerate(ev(ts)):
        if k == 0:
            continue
        assert ev.K == g["K"][k - 1]
        if k % 20 == 0:
            assert np.max(np.abs(N.localdensity(P) - g["rho"][k - 1])) < 1e-13
            assert np.max(np.abs(N.density_currents(Hk, P) - g["J"][k - 1])) < 1e-13
    assert np.max(np.abs(ev.P - g["P_final"])) < 1e-13
    assert [tuple(b) for b in g["bonds"]] == hf(0.0).bonds


def test_densitymatrix_family_reference_pins():
    """test/test_operators.jl:226-236 on SquareLattice(5, 5): fixed N = 3 at
    T = 1 (Fermi-Dirac, mu found) has Tr P = 3; the Gibbs distribution at T = 0
    is the ground-state projector; the BoseEinstein ensemble at T = 0, mu = 0
    (test/test_workflows.jl:22) is finite."""
    Hd = H.tightbinding_hamiltonian(G.SquareLattice(5, 5)).dense()
    E, V = np.linalg.eigh(Hd)
    P, mu = N.fixn_densitymatrix(E, V, 3, T=1.0)
    assert abs(np.trace(P).real - 3.0) < 1e-7
    from scipy.optimize import brentq
    mu_ref = brentq(lambda m: np.sum(1.0 / (np.exp(E - m) + 1.0)) - 3.0, -10.0, 10.0, xtol=1e-15)
    assert abs(mu - mu_ref) < 1e-7
    Pg = N.densitymatrix_kw(Hd, T=0.0)
    assert np.max(np.abs(Pg - np.outer(V[:, 0], V[:, 0].conj()))) < 1e-12
    Pb = N.densitymatrix_kw(Hd, statistics=N.BOSE_EINSTEIN)
    assert np.all(np.isfinite(Pb))
    # its weights: -1 below mu = 0, 1 at 0 (the E - mu == T == 0 rule), 0 above
    f = N.densfun(0.0, 0.0, N.BOSE_EINSTEIN)
    assert (f(-0.5), f(0.0), f(0.5)) == (-1.0, 1.0, 0.0)


def test_densitymatrix_family_closed_forms():
    """Each distribution against the matrix function of H it defines (scipy
    expm / inverse), on a disordered lattice with a nondegenerate spectrum."""
    _name, h = list(_models())[1]
    Hd = h.dense()
    n = len(Hd)
    E, V = np.linalg.eigh(Hd)
    I = np.eye(n)
    T = 0.7
    Ez = sl.expm(-Hd / T)
    assert np.max(np.abs(N.gibbs_densitymatrix(E, V, T) - Ez / np.trace(Ez))) < 1e-12
    mu = 0.3
    fd = np.linalg.inv(sl.expm((Hd - mu * I) / T) + I)
    assert np.max(np.abs(N.ensemble_densitymatrix(E, V, mu, T, N.FERMI_DIRAC) - fd)) < 1e-12
    mub = E[0] - 0.4
    be = np.linalg.inv(sl.expm((Hd - mub * I) / T) - I)
    assert np.max(np.abs(N.ensemble_densitymatrix(E, V, mub, T, N.BOSE_EINSTEIN) - be)) < 1e-10
    # Fermi sphere: the T = 0 Fermi-Dirac projector at mu between levels N-1 and N
    for k in (1, 7, n - 1, n):
        Pk = N.fermisphere_densitymatrix(E, V, k)
        assert abs(np.trace(Pk).real - k) < 1e-10
        assert np.max(np.abs(Pk @ Pk - Pk)) < 1e-12
        if k < n:
            Pm = N.ensemble_densitymatrix(E, V, 0.5 * (E[k - 1] + E[k]), 0.0)
            assert np.max(np.abs(Pk - Pm)) < 1e-12
    with pytest.raises(ValueError):
        N.fermisphere_densitymatrix(E, V, n + 1)
    # fixed N at T = 0: Fermi sphere (fermions), N x ground state (bosons)
    assert np.max(np.abs(N.fixn_densitymatrix(E, V, 5)[0] - N.fermisphere_densitymatrix(E, V, 5))) < 1e-15
    Pb0 = N.fixn_densitymatrix(E, V, 4, statistics=N.BOSE_EINSTEIN)[0]
    assert np.max(np.abs(Pb0 - 4 * N.groundstate_densitymatrix(E, V))) < 1e-15
    # fixed N at T > 0 for bosons: Newton's iteration from mu = 0 (as the
    # reference's; it may settle inside the spectrum) ends with Tr P = N
    Pbn, mun = N.fixn_densitymatrix(E, V, 3, T=0.5, statistics=N.BOSE_EINSTEIN)
    assert abs(np.trace(Pbn).real - 3.0) < 1e-6
    # keyword dispatch (src/spectrum.jl:317-329)
    with pytest.raises(ValueError):
        N.densitymatrix_kw(Hd, N=3, mu=0.0)
    assert np.max(np.abs(N.densitymatrix_kw(Hd, statistics=N.FERMI_DIRAC) -
                         N.ensemble_densitymatrix(E, V, 0.0, 0.0))) < 1e-12


def test_timesequence_doctests():
    """The reference's doctests (src/timesequence.jl:178-212, 227-254) on
    f(t) = t over 0:0.1:10: the derivative is 1 at the midpoints 0.05 ..
    9.95; the integral at 0, 0.1, ..., 0.8 shows as 0.0, 0.005, 0.02, 0.045,
    0.08, 0.125, 0.18, 0.245, 0.32 (rounded to 5 digits as the doctest prints)."""
    ts = N.julia_range(0, 0.1, 10)
    td, dv = N.differentiate(ts, ts)
    assert len(td) == 100
    assert abs(td[0] - 0.05) < 1e-12 and abs(td[-1] - 9.95) < 1e-12
    assert np.max(np.abs(np.array(dv) - 1.0)) < 1e-10
    ti, iv = N.integrate(ts, ts)
    shown = [round(float(v), 5) for v in iv[:9]]
    assert shown == [0.0, 0.005, 0.02, 0.045, 0.08, 0.125, 0.18, 0.245, 0.32]
    # closed forms: trapezoids are exact for a linear f, the symmetric
    # difference is exact at the midpoint for a quadratic one
    assert np.max(np.abs(np.array(iv) - np.array(ts) ** 2 / 2)) < 1e-10
    tq, dq = N.differentiate(ts, [t * t for t in ts])
    assert np.max(np.abs(np.array(dq) - 2 * np.array(tq))) < 1e-10
    with pytest.raises(RuntimeError):
        N.differentiate([0.0], [1.0])


def test_region_currents_sum_rules():
    """currentsfrom / currentsfromto against each other, antisymmetry, and the
    continuity equation from the Heisenberg equation (test/test_currents.jl:
    14-26): d/dt of a region's density is currentsfromto(region, rest)."""
    _name, h = list(_models())[0]
    Hd = h.dense()
    n = h.n
    rng = np.random.default_rng(3)
    Q = np.linalg.qr(rng.normal(size=(n, 9)) + 1j * rng.normal(size=(n, 9)))[0]
    P = Q @ Q.conj().T
    src = [0, 1, 2, 7, 8, 9]
    out = N.currentsfrom(h, P, src)
    tot = N.currentsfromto(h, P, src)
    assert abs(np.sum(out) - tot) < 1e-13 and np.all(out[src] == 0)
    # dP/dt = -i [H, P] gives d rho_i / dt = sum_j J_ij, so with this sign
    # convention d/dt sum_{i in src} rho_i = currentsfromto(src, rest)
    dP = -1j * (Hd @ P - P @ Hd)
    drho = np.real(np.diag(dP))[src].sum()
    assert abs(drho - tot) < 1e-12
    # to a region and back: antisymmetry
    dst = [3, 4, 10, 11]
    assert abs(N.currentsfromto(h, P, src, dst) + N.currentsfromto(h, P, dst, src)) < 1e-13


def test_qwz_internal_spin_pins():
    """Internal degrees of freedom (SpinBasis(1//2), index i*2 + a):
    the QWZ torus spectrum in closed form, +-sqrt((m + cos kx + cos ky)^2 +
    sin^2 kx + sin^2 ky); and test/test_currents.jl:1-26 on SquareLattice(4, 4):
    P = densitymatrix(diagonalize(qwz(l)), statistics=FermiDirac), dc =
    DensityCurrents(qwz(l, field=LandauGauge(0.1)), P): antisymmetry, the four
    bond currents of site 6 sum to d/dt of its density, which equals
    currentsfromto(dc, 6)."""
    m = 0.7
    lt = G.SquareLattice(5, 4, periodic=(True, True))
    E = np.sort(np.linalg.eigvalsh(H.qwz(lt, m=m).dense()))
    ex = []
    for kx in 2 * np.pi * np.arange(5) / 5:
        for ky in 2 * np.pi * np.arange(4) / 4:
            r = math.sqrt((m + math.cos(kx) + math.cos(ky)) ** 2 + math.sin(kx) ** 2 + math.sin(ky) ** 2)
            ex += [r, -r]
    assert np.max(np.abs(E - np.sort(ex))) < 1e-12
    l = G.SquareLattice(4, 4)
    H0, H1 = H.qwz(l), H.qwz(l, field=F.LandauGauge(0.1))
    P = N.densitymatrix_kw(H0.dense(), statistics=N.FERMI_DIRAC)
    s1, s2 = 5, 10                                  # l[6], l[11]
    assert N.site_current(H1, P, s1, s2) == -N.site_current(H1, P, s2, s1)
    assert abs(N.site_current(H1, P, s1, s1)) <= np.finfo(float).eps
    rho = np.zeros((32, 32))
    rho[2 * s1, 2 * s1] = rho[2 * s1 + 1, 2 * s1 + 1] = 1.0
    Hd = H1.dense()
    dens_dt = np.trace(1j * (Hd @ rho - rho @ Hd) @ P)
    nbrs = [9, 6, 1, 4]                             # s1 + axis1, + axis2, - axis1, - axis2
    total = sum(N.site_current(H1, P, s1, t) for t in nbrs)
    assert abs(total - dens_dt) < 1e-12 and abs(dens_dt.imag) < 1e-12
    assert abs(N.currentsfromto(H1, P, [s1]) - dens_dt.real) < 1e-12
    # localdensity sums the two spin states of a site
    assert np.max(np.abs(N.localdensity(P, 2) - np.real(np.diag(P)).reshape(-1, 2).sum(1))) == 0
