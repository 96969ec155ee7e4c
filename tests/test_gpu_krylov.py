"""GPU parity of the Krylov ket propagator (lmh_krylov_step, the counterpart
of KrylovKitExp, src/evolution.jl:130-154) against the exact propagator
exp(-i dt H) of the oracle's Hamiltonian (scipy.linalg.expm, or
scipy.sparse.linalg.expm_multiply for a large lattice) after every step,
1e-10 -- the bar of the reference's own KrylovKitExp test
(test/test_timedeps.jl:42-68), which is reproduced here."""
import ctypes
import sys

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import lmhip
from helpers import Model
from oracle import geometry as OG
from oracle import hamiltonian as OH

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = {
    "square_open_axial": (lambda: Model("square", (10, 10)), [("axial", 0.2, 5.5, 5.5)], 0.1, 30),
    "square_pbc_landau_disorder": (
        lambda: Model("square", (12, 10), (True, True), onsite=np.random.default_rng(7).uniform(-0.5, 0.5, 120)),
        [("landau", 0.02, 0, 0)], 0.05, 30),
    "haldane_pointflux": (lambda: Model("honeycomb", (5, 6), t1=1.0, t2=0.1, m=0.2), [("axial", 0.15, 2.75, 2.2)],
                          0.2, 20),
}


@pytest.mark.parametrize("graphs", [1, 0])
@pytest.mark.parametrize("nkets", [1, 3])
@pytest.mark.parametrize("name", sorted(CASES))
def test_krylov_every_step(name, nkets, graphs):
    """A ramped field (H changes every step) and a block of kets: every ket
    after every step against exp(-i dt H(t_{k-1})) applied to the previous
    exact ket; the first Lanczos chunk of a step replayed from a captured
    graph or launched plainly."""
    make, fspec0, dt, nsteps = CASES[name]
    m = make()
    say("case krylov %s kets=%d graphs=%d" % (name, nkets, graphs))
    tau = nsteps * dt / 2
    ramp = lambda t: [(k, a * min(t, tau) / tau, px, py) for (k, a, px, py) in fspec0]
    rng = np.random.default_rng(5)
    psi = rng.normal(size=(m.n, nkets)) + 1j * rng.normal(size=(m.n, nkets))
    psi /= np.linalg.norm(psi, axis=0)
    solver = lmhip.HIPKrylovExp(ctx=lmhip.Context(0))
    lmhip._lib.check(lmhip.lib().lmh_set_option(solver.ctx.h, lmhip._lib.OPT_KRYLOV_GRAPHS, graphs))
    ev = lmhip.Evolution(solver, lambda t: m.device_h(ramp(t)), psi[:, 0] if nkets == 1 else psi)
    ts = [k * dt for k in range(nsteps + 1)]
    exact = psi.copy()
    worst = 0.0
    for k, st in enumerate(ev(ts)):
        if k == 0:
            continue
        Hd = m.oracle_h(ramp(ts[k - 1])).dense()
        exact = sl.expm(-1j * (ts[k] - ts[k - 1]) * Hd) @ exact
        got = st.states[0].to_numpy().reshape(m.n, nkets)
        worst = max(worst, float(np.max(np.abs(got - exact))))
        assert worst <= TOL, (k, worst)
        assert 1 <= solver.matvecs <= 30
    say("  worst %.2e" % worst)


def test_reference_qwz_groundstate_case():
    """test/test_timedeps.jl:42-68 on the device: QWZ (m = 1) on
    SquareLattice(10, 10), its ground state, 0:0.1:10 with the Krylov
    solver; the whole ket (the test checks component 2) against the dense
    exponential after every step, atol 1e-10."""
    lat = lmhip.SquareLattice(10, 10)
    Hdev = lmhip.qwz(lat)
    Hd = OH.qwz(OG.SquareLattice(10, 10)).dense()
    E, V = np.linalg.eigh(Hd)
    psi = V[:, int(np.argmin(E))]
    U = sl.expm(-1j * 0.1 * Hd)
    ev = lmhip.Evolution(lmhip.HIPKrylovExp, Hdev, psi)
    from fractions import Fraction
    ts = [float(k * Fraction("0.1")) for k in range(101)]
    exact = psi.copy()
    say("case krylov reference qwz N=%d" % len(psi))
    for k, st in enumerate(ev(ts)):
        if k:
            exact = U @ exact
        got = st.states[0].to_numpy()
        assert abs(got[1] - exact[1]) <= TOL
        assert np.max(np.abs(got - exact)) <= TOL, k


def test_large_single_ket_against_expm_multiply():
    """One ket on SquareLattice(150, 100) with a Landau field (N = 15000,
    above the fused kernel's range; the benchmark's regime): five steps
    against scipy.sparse.linalg.expm_multiply of the oracle's H."""
    m = Model("square", (150, 100), onsite=np.random.default_rng(3).uniform(-0.5, 0.5, 15000))
    spec = [("landau", 0.01, 0, 0)]
    Ho = m.oracle_h(spec)
    c = Ho.csc
    Hs = sp.csc_matrix((c.nzval, c.rowval, c.colptr), shape=(c.n, c.n))
    rng = np.random.default_rng(9)
    psi = rng.normal(size=m.n) + 1j * rng.normal(size=m.n)
    psi /= np.linalg.norm(psi)
    solver = lmhip.HIPKrylovExp(ctx=lmhip.Context(0))
    ev = lmhip.Evolution(solver, m.device_h(spec), psi)
    exact = psi.copy()
    say("case krylov large N=%d" % m.n)
    for k, st in enumerate(ev([0.1 * q for q in range(6)])):
        if k == 0:
            continue
        exact = spl.expm_multiply(-0.1j * Hs, exact)
        d = float(np.max(np.abs(st.states[0].to_numpy() - exact)))
        assert d <= TOL, (k, d)


def _direct(m, spec, psi, dt, krylovdim=30, tol=1e-12, maxiter=100):
    ctx = lmhip.Context(0)
    ctx.set_hamiltonian(m.device_h(spec))
    k = lmhip.Kets.from_array(ctx, psi)
    mv = ctypes.c_int64()
    rc = lmhip.lib().lmh_krylov_step(ctx.h, k.h, dt, krylovdim, tol, maxiter, ctypes.byref(mv))
    return rc, k.to_numpy(), mv.value


def test_substeps_negative_time_zero_ket_and_failure():
    """A short basis takes time sub-steps (still exact to 1e-10); a negative
    dt gives exp(+i|dt| H); a zero ket stays zero beside nonzero ones; too
    few restarts fail with LMH_ERANGE (the reference's ArgumentError)."""
    m = Model("square", (9, 9))
    spec = [("landau", 0.04, 0, 0)]
    Hd = m.oracle_h(spec).dense()
    rng = np.random.default_rng(5)
    psi = rng.normal(size=(81, 3)) + 1j * rng.normal(size=(81, 3))
    psi[:, 1] = 0.0
    say("case krylov substeps")
    rc, got, mv = _direct(m, spec, psi, 1.5, krylovdim=10)
    assert rc == 0 and mv > 20
    assert np.max(np.abs(got - sl.expm(-1.5j * Hd) @ psi)) <= TOL
    assert np.all(got[:, 1] == 0)
    say("case krylov negative dt")
    rc, got, _ = _direct(m, spec, psi, -0.3)
    assert rc == 0 and np.max(np.abs(got - sl.expm(0.3j * Hd) @ psi)) <= TOL
    say("case krylov non-convergence")
    rc, _, _ = _direct(m, spec, psi, 5.0, krylovdim=2, maxiter=1)
    assert rc == lmhip._lib.LMH_ERANGE
    assert b"did not converge" in lmhip.lib().lmh_last_error()


def test_internal_states_qwz_block():
    """QWZ with a field (internal states, matrix-valued hoppings) and two kets."""
    lat = lmhip.SquareLattice(8, 7)
    field = lmhip.LandauGauge(0.03)
    Hdev = lmhip.qwz(lat, 1.3, field=field)
    from oracle import fields as OF
    Hd = OH.qwz(OG.SquareLattice(8, 7), 1.3, field=OF.LandauGauge(0.03)).dense()
    rng = np.random.default_rng(2)
    psi = rng.normal(size=(len(Hd), 2)) + 1j * rng.normal(size=(len(Hd), 2))
    ev = lmhip.Evolution(lmhip.HIPKrylovExp(), Hdev, psi)
    U = sl.expm(-0.25j * Hd)
    exact = psi.copy()
    say("case krylov qwz block N=%d" % len(Hd))
    for k, st in enumerate(ev([0.25 * q for q in range(9)])):
        if k:
            exact = U @ exact
        assert np.max(np.abs(st.states[0].to_numpy() - exact)) <= TOL, k


def test_graphs_follow_geometry_changes():
    """One context and ket block evolved with an open lattice, then a periodic
    one of the same size (new geometry: the captured chunks are rebuilt), then
    the open one again, against the exact propagators."""
    ma = Model("square", (8, 8))
    mb = Model("square", (8, 8), (True, True))
    spec = [("landau", 0.03, 0, 0)]
    rng = np.random.default_rng(4)
    psi = rng.normal(size=(64, 2)) + 1j * rng.normal(size=(64, 2))
    solver = lmhip.HIPKrylovExp(ctx=lmhip.Context(0))
    kets = solver.make_state(psi, ma.device_h(spec))
    exact = psi.copy()
    say("case krylov graphs across geometries")
    for mdl, nsteps in ((ma, 3), (mb, 3), (ma, 2)):
        H = mdl.device_h(spec)
        U = sl.expm(-0.2j * mdl.oracle_h(spec).dense())
        for _ in range(nsteps):
            solver.update_solver(H, 0.2)
            solver.step(kets)
            exact = U @ exact
            assert np.max(np.abs(kets.to_numpy() - exact)) <= TOL
