"""GPU parity of the Fermi projector P0 = densitymatrix(H, mu; T) and of the
evolution at the full N = 10^4 workload size.

P0: device eigensolver + MFMA projector against numpy's LAPACK eigh at
1e-10 (chemical potentials chosen inside spectral gaps, where the projector is
unique; exactly degenerate levels at mu make the reference's own P0 depend on
LAPACK rounding and are not comparable).
"""
import sys

import numpy as np
import pytest

import lmhip
from helpers import Model
from oracle import numerics as ON

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gap_mu(E, target=0.0):
    """A mu in the spectral gap nearest to target (midpoint of two levels)."""
    E = np.sort(E)
    k = int(np.searchsorted(E, target))
    k = min(max(k, 1), len(E) - 1)
    gaps = [(E[i + 1] - E[i], i) for i in range(max(0, k - 3), min(len(E) - 1, k + 3))]
    g, i = max(gaps)
    return 0.5 * (E[i] + E[i + 1])


CASES = [
    ("square_open_flux", lambda: Model("square", (10, 10)), [("axial", 0.1, 5.5, 5.5)], 0.0),
    ("square_pbc_disorder", lambda: Model("square", (12, 12), (True, True),
                                          onsite=np.random.default_rng(1).uniform(-0.5, 0.5, 144)),
     [("landau", 0.01, 0, 0)], 0.0),
    ("haldane", lambda: Model("honeycomb", (6, 7), t1=1.0, t2=0.1, m=0.2), [("axial", 0.05, 3.1, 2.6)], 0.0),
    ("square_T", lambda: Model("square", (9, 8), onsite=np.random.default_rng(2).uniform(-1, 1, 72)),
     [("landau", 0.03, 0, 0)], 0.35),
    # real symmetric H (no field) above the size where rocsolver_zheevd has
    # returned NaN eigenvectors: the library must detect it and use another solver
    ("square_real_n400", lambda: Model("square", (20, 20), onsite=np.random.default_rng(3).normal(0, 0.1, 400)),
     [], 0.0),
]


@pytest.mark.parametrize("name,make,fspec,T", CASES, ids=[c[0] for c in CASES])
def test_densitymatrix_matches_lapack(name, make, fspec, T):
    say("case %s" % name)
    m = make()
    Hd = m.oracle_h(fspec).dense()
    E_ref = np.linalg.eigvalsh(Hd)
    mu = _gap_mu(E_ref) if T == 0 else 0.1
    P_ref, _ = ON.densitymatrix(Hd, mu=mu, T=T)
    P, E = lmhip.densitymatrix(m.device_h(fspec), mu=mu, T=T, return_energies=True)
    Pg = P.to_numpy()
    assert np.max(np.abs(E - E_ref)) <= 1e-12
    assert np.max(np.abs(Pg - P_ref)) <= TOL
    if T == 0:
        assert P.nocc == int(np.sum(E_ref <= mu))
        assert np.max(np.abs(Pg @ Pg - Pg)) <= 1e-10
    assert abs(np.trace(Pg).real - np.trace(P_ref).real) <= 1e-10


def test_densitymatrix_then_evolution():
    """P0 made on the device feeds Evolution directly (state copy on device)."""
    m = Model("square", (8, 8), onsite=np.random.default_rng(4).uniform(-0.5, 0.5, 64))
    tau = 1.0
    fs = lambda t: [("axial", 0.2 * min(t, tau) / tau, 4.5, 4.5)]
    E = np.linalg.eigvalsh(m.oracle_h(fs(0)).dense())
    mu = _gap_mu(E)
    P0 = lmhip.densitymatrix(m.device_h(fs(0)), mu=mu)
    P0_ref, _ = ON.densitymatrix(m.oracle_h(fs(0)).dense(), mu=mu)
    ts = [0.1 * k for k in range(31)]
    ev_d = lmhip.Evolution(lambda t: m.device_h(fs(t)), P0)
    ev_o = ON.Evolution(lambda t: m.oracle_h(fs(t)), P0_ref)
    worst = 0.0
    for st, (_H, Po, _t) in zip(ev_d(ts), ev_o(ts)):
        worst = max(worst, float(np.max(np.abs(st.states[0].to_numpy() - Po))))
    assert worst <= TOL
    # the Evolution owns a copy: the original P0 is untouched
    assert np.max(np.abs(P0.to_numpy() - P0_ref)) <= TOL


def test_full_size_n10000_against_factored_reference():
    """SquareLattice(100,100) PBC, Landau ramp + W=1 disorder, dt = 0.05: the
    device P after 6 ramp steps equals V V' with V = U_6...U_1 Q evolved by the
    oracle's own sparse U (myexp! with droptol), Q a rank-8 orthonormal basis.
    Size-independent check at the benchmark's full size."""
    L, tau, bmax, dt = 100, 10.0, 0.01, 0.05
    rng = np.random.default_rng(123)
    V = rng.uniform(-0.5, 0.5, L * L)
    m = Model("square", (L, L), (True, True), onsite=V)
    Q = np.linalg.qr(rng.normal(size=(L * L, 8)) + 1j * rng.normal(size=(L * L, 8)))[0]
    fs = lambda t: [("landau", bmax * min(t, tau) / tau, 0, 0)]
    t0 = 4.0
    ts = [t0 + dt * k for k in range(7)]
    say("case n10000 steps=%d" % (len(ts) - 1))
    ev = lmhip.Evolution(lambda t: m.device_h(fs(t)), Q @ Q.conj().T)
    ev.time = t0
    Vm = np.asfortranarray(Q)
    t = t0
    for k, st in enumerate(ev(ts)):
        if k == 0:
            continue
        U, K, s, _ = ON.myexp(m.oracle_h(fs(t)), ts[k] - ts[k - 1])
        assert ev.solver.K == K and ev.solver.s == s
        Vm = ON.apply(U, Vm)
        t = ts[k]
    Pg = ev.states[0].to_numpy()
    worst = 0.0
    for c0 in range(0, L * L, 1000):
        ref = Vm @ Vm[c0:c0 + 1000].conj().T
        worst = max(worst, float(np.max(np.abs(Pg[:, c0:c0 + 1000] - ref))))
    say("n10000 worst |dP| = %.3e" % worst)
    assert worst <= TOL
    rho, J = lmhip.observables(ev.states[0])
    assert abs(rho.sum() - 8.0) <= 1e-8


FAMILY = [
    # (id, keyword arguments of densitymatrix)
    ("gibbs_T0", dict(T=0.0)),                                    # the ground state
    ("gibbs_T", dict(T=0.6)),
    ("fermi_dirac_T", dict(mu=0.2, T=0.4)),
    ("bose_einstein_T", dict(mu=-4.5, T=0.5, statistics=lmhip.BoseEinstein)),
    ("bose_einstein_T0_mu0", dict(statistics=lmhip.BoseEinstein)),  # weights -1 / 1 / 0: signed projector
    ("fixn_fermi_T", dict(N=7, T=1.0)),
    ("fixn_fermi_T0", dict(N=7)),                                 # Fermi sphere
    ("fixn_bose_T0", dict(N=4, statistics=lmhip.BoseEinstein)),  # N x ground state
]


@pytest.mark.parametrize("kw", [f[1] for f in FAMILY], ids=[f[0] for f in FAMILY])
def test_densitymatrix_family_matches_oracle(kw):
    """The densitymatrix family (device diagonalisation + signed-weight MFMA
    projector, host occupations) against the oracle's restatement of
    src/spectrum.jl:226-333, on a disordered lattice with a nondegenerate
    spectrum (so the ground state and the Fermi sphere are unique)."""
    say("case densitymatrix %s" % kw)
    m = Model("square", (9, 8), onsite=np.random.default_rng(4).uniform(-1, 1, 72))
    spec = [("landau", 0.03, 0, 0)]
    okw = dict(kw)
    if "statistics" in okw:
        okw["statistics"] = ON.BOSE_EINSTEIN if okw["statistics"] == lmhip.BoseEinstein else ON.FERMI_DIRAC
    P_ref = ON.densitymatrix_kw(m.oracle_h(spec).dense(), **okw)
    P = lmhip.densitymatrix(m.device_h(spec), **kw)
    d = float(np.max(np.abs(P.to_numpy() - P_ref)))
    assert d <= TOL, d
    if "N" in kw and kw.get("statistics") != lmhip.BoseEinstein:
        assert abs(np.trace(P.to_numpy()).real - kw["N"]) < 1e-7


def test_eigensystem_reuse_and_release():
    """One diagonalisation, several projectors; the eigenvectors stay in the
    context until released, after which a projector is refused."""
    m = Model("square", (8, 8), onsite=np.random.default_rng(6).uniform(-1, 1, 64))
    spec = [("axial", 0.2, 4.5, 4.5)]
    Hd = m.oracle_h(spec).dense()
    E_ref, V_ref = np.linalg.eigh(Hd)
    eig = lmhip.diagonalize(m.device_h(spec))
    assert np.max(np.abs(eig.values - E_ref)) < 1e-12
    for N in (1, 10, 63):
        P = eig.fermisphere_densitymatrix(N)
        assert P.nocc == N
        d = float(np.max(np.abs(P.to_numpy() - ON.fermisphere_densitymatrix(E_ref, V_ref, N))))
        assert d <= TOL, (N, d)
    w = np.linspace(-1.0, 2.0, 64)
    Pw = eig.projector(w)
    assert float(np.max(np.abs(Pw.to_numpy() - ON.projector_weights(V_ref, w)))) <= TOL
    with pytest.raises(ValueError):
        eig.fermisphere_densitymatrix(65)
    eig.close()
    with pytest.raises(lmhip._lib.LMHError):
        eig.projector(w)
