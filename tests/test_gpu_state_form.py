"""GPU: a fused step stores (U P U')^H, so between reads a state alternates
between holding P and holding P^H (one transposition per step instead of
two).  Everything that reads a state must see P: observables and region
currents read the stored form directly, downloads, entry reads, device
pointers, tapes and batches bring the state back to P first.

Every case steps several times without downloading in between and checks
what it reads against the oracle's evolution (1e-10): densities and bond
currents after every step, entries after odd and even step counts, the
final matrix, a non-Hermitian P, both fused step modes, a recording series,
a batch built from states left in the P^H form, and an adjoint tape begun
on one.
"""
import ctypes
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


CASES = {
    "strip_landau": (lambda: Model("square", (10, 12), (True, True),
                                   onsite=np.random.default_rng(1).uniform(-0.5, 0.5, 120)),
                     [("landau", 0.02, 0, 0)]),
    "general_pointflux": (lambda: Model("square", (9, 7)), [("axial", 0.3, 4.5, 3.5)]),
    "haldane": (lambda: Model("honeycomb", (5, 5), t1=1.0, t2=0.1, m=0.2), [("axial", 0.1, 2.75, 2.2)]),
}


@pytest.mark.parametrize("mode", ["fusedT", "fusedN"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_reads_between_unread_steps(name, mode):
    make, fspec0 = CASES[name]
    m = make()
    n = m.n
    dt, nsteps = 0.05, 9
    tau = nsteps * dt / 2
    ramp = lambda t: [(k, a * min(t, tau) / tau, px, py) for (k, a, px, py) in fspec0]
    rng = np.random.default_rng(3)
    P0 = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))       # not Hermitian
    ts = [k * dt for k in range(nsteps + 1)]
    refs = [(Ho, Po.copy()) for (Ho, Po, _t) in ON.Evolution(lambda t: m.oracle_h(ramp(t)), P0)(ts)]
    md = lmhip._lib.STEP_FUSED_T if mode == "fusedT" else lmhip._lib.STEP_FUSED_N
    solver = lmhip.HIPExp(step_mode=md)
    st = solver.make_state(P0, m.device_h(ramp(0.0)))
    rows = rng.integers(0, n, 40)
    cols = rng.integers(0, n, 40)
    say("case %s %s" % (name, mode))
    for k in range(1, nsteps + 1):
        solver.update_solver(m.device_h(ramp(ts[k - 1])), dt)
        solver.step(st)
        Ho, Po = refs[k]
        rho, J = lmhip.observables(st, solver.H)
        assert np.max(np.abs(rho - ON.localdensity(Po))) <= TOL, k
        assert np.max(np.abs(J - ON.density_currents(Ho, Po))) <= TOL, k
        if k in (3, 6):
            got = st.get_entries(rows, cols)
            assert np.max(np.abs(got - Po[rows, cols])) <= TOL, k
    assert np.max(np.abs(st.to_numpy() - refs[-1][1])) <= TOL


def test_series_and_region_currents_in_both_forms():
    """A density, current and region series recorded after every step (the
    state alternating forms) against the oracle's evolution."""
    make, fspec = CASES["strip_landau"]
    m = make()
    n = m.n
    rng = np.random.default_rng(4)
    Q = np.linalg.qr(rng.normal(size=(n, 30)) + 1j * rng.normal(size=(n, 30)))[0]
    P0 = Q @ Q.conj().T
    dt, nsteps = 0.05, 7
    ts = [k * dt for k in range(nsteps + 1)]
    refs = [(Ho, Po.copy()) for (Ho, Po, _t) in ON.Evolution(lambda t: m.oracle_h(fspec), P0)(ts)]
    solver = lmhip.HIPExp()
    H = m.device_h(fspec)
    st = solver.make_state(P0, H)
    src = np.arange(0, 24)
    sd = lmhip.DeviceSeries(solver.ctx, "density", nsteps + 1)
    sc = lmhip.DeviceSeries(solver.ctx, "currents", nsteps + 1)
    sr = lmhip.DeviceSeries(solver.ctx, "region", nsteps + 1, src=src)
    solver.update_solver(H, dt)
    for k in range(nsteps + 1):
        if k:
            solver.update_solver(H, dt)
            solver.step(st)
        for s in (sd, sc, sr):
            s.record(st, ts[k])
    Ho = m.oracle_h(fspec)
    td, tc, tr = sd.to_timesequence(), sc.to_timesequence(), sr.to_timesequence()
    for k in range(nsteps + 1):
        Po = refs[k][1]
        assert np.max(np.abs(td.values[k] - ON.localdensity(Po))) <= TOL
        assert np.max(np.abs(tc.values[k] - ON.density_currents(Ho, Po))) <= TOL
        assert abs(tr.values[k] - ON.currentsfromto(Ho, Po, src)) <= TOL
    for s in (sd, sc, sr):
        s.close()


def _hip_to_host(ptr, nbytes):
    hip = ctypes.CDLL("libamdhip64.so")
    out = np.zeros(nbytes // 16, dtype=np.complex128)
    rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(nbytes), 2)
    assert rc == 0
    return out


def test_device_pointer_and_copy_see_P():
    """lmh_state_device_ptr hands out P (the state is brought back first);
    a copy of a state in the P^H form reads as the same P."""
    make, fspec = CASES["general_pointflux"]
    m = make()
    n = m.n
    rng = np.random.default_rng(5)
    P0 = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    solver = lmhip.HIPExp()
    H = m.device_h(fspec)
    st = solver.make_state(P0, H)
    solver.update_solver(H, 0.05)
    U = ON.myexp(m.oracle_h(fspec), 0.05)[0]
    step = lambda P: ON.apply(U, ON.apply(U, P).conj().T).conj().T     # U P U'
    solver.step(st)                                   # P^H form now
    cp = st.copy()
    ref = step(P0)
    assert np.max(np.abs(cp.to_numpy() - ref)) <= TOL
    solver.step(st)
    solver.step(st)                                   # P^H form again
    ptr, ld = st.device_ptr()
    lmhip._lib.check(lmhip._lib.lib().lmh_sync(solver.ctx.h))
    raw = _hip_to_host(ptr, n * ld * 16).reshape(n, ld)[:, :n].T      # column-major
    assert np.max(np.abs(raw - step(step(ref)))) <= TOL


def test_batch_from_states_in_both_forms():
    """Batch members left in the P^H form (one unread step) and in the P form
    (two) step together; each matches its own oracle evolution."""
    L_ = lmhip._lib.lib()
    rng = np.random.default_rng(6)
    dt = 0.05
    solvers, states, refs = [], [], []
    for r, L in enumerate((9, 10)):
        m = Model("square", (L, L), onsite=rng.uniform(-1, 1, L * L))
        spec = [("landau", 0.005, 0, 0)]
        Q = np.linalg.qr(rng.normal(size=(L * L, 12)) + 1j * rng.normal(size=(L * L, 12)))[0]
        P0 = Q @ Q.conj().T
        solver = lmhip.HIPExp(ctx=lmhip.Context(0))
        H = m.device_h(spec)
        st = solver.make_state(P0, H)
        solver.update_solver(H, dt)
        nfirst = r + 1
        for _ in range(nfirst):
            solver.step(st)
        solvers.append(solver)
        states.append(st)
        ts = [k * dt for k in range(nfirst + 4)]
        refs.append([Po.copy() for (_H, Po, _t) in ON.Evolution(lambda t, m=m, s=spec: m.oracle_h(s), P0)(ts)][-1])
    ctxs = (ctypes.c_void_p * 2)(*[s.ctx.h.value for s in solvers])
    sts = (ctypes.c_void_p * 2)(*[s.h.value for s in states])
    b = ctypes.c_void_p()
    lmhip._lib.check(L_.lmh_batch_create(ctypes.byref(b), 2, ctxs, sts))
    try:
        lmhip._lib.check(L_.lmh_batch_step(b, 3, 0))
        for st, ref in zip(states, refs):
            assert np.max(np.abs(st.to_numpy() - ref)) <= TOL
    finally:
        L_.lmh_batch_destroy(b)


def test_tape_begun_on_a_state_in_the_conjugate_form():
    """A tape begun on a state in the P^H form (one unread step) gives the
    same gradients as one begun on that step's P uploaded afresh."""
    m = Model("square", (6, 6), onsite=np.random.default_rng(7).uniform(-0.5, 0.5, 36))
    fs = lambda t: [("axial", 0.1 + 0.3 * t, 2.5, 3.5)]
    rng = np.random.default_rng(8)
    Q = np.linalg.qr(rng.normal(size=(36, 10)) + 1j * rng.normal(size=(36, 10)))[0]
    P0 = Q @ Q.conj().T
    dt = 0.1
    cot = lmhip.localdensity_cotangent(36, 1.0, sites=[0, 1, 6, 7])

    def taped(ev):
        tape = lmhip.EvolutionTape(ev, checkpoint_every=2)
        for _ in range(6):
            ev.step(dt)
        g = tape.backward(cot)
        out = (g.field.copy(), g.onsite.copy(), g.P0.to_numpy())
        tape.close()
        return out

    say("case tape on the conjugate form")
    evA = lmhip.Evolution(lmhip.HIPExp(ctx=lmhip.Context(0)), lambda t: m.device_h(fs(t)), P0)
    evA.step(dt)                                     # unread: P^H form
    gA = taped(evA)
    evB0 = lmhip.Evolution(lmhip.HIPExp(ctx=lmhip.Context(0)), lambda t: m.device_h(fs(t)), P0)
    evB0.step(dt)
    P1 = evB0.states[0].to_numpy()
    evB = lmhip.Evolution(lmhip.HIPExp(ctx=lmhip.Context(0)), lambda t: m.device_h(fs(t)), P1)
    evB.time = dt
    gB = taped(evB)
    for a, b in zip(gA, gB):
        assert np.max(np.abs(a - b)) <= 1e-13
