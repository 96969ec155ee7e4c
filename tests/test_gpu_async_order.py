"""GPU: K on the device (LMH_OPT_ASYNC_ORDER).  An H update whose caller
does not ask for K leaves the Taylor order on the device; the strip step's
launch reads K, s and its coefficients from there and does nothing unless
the result is the one the host assumed; the host takes the result up at the
next call that needs it and runs again any step whose launch did nothing.

Every case is checked against the oracle's evolution (1e-10) and bit for bit
against the same evolution with blocking updates: a ramped field (no step
redone), a point flux switched on from zero strength (the H leaves the strip
form: the launch does nothing and the step runs again on the general
kernel), a K = 0 update, an update whose order search needs the larger
tables, two states stepped by one update, reads of every kind between
updates, and the K / s the host reports.
Reference: src/evolution.jl:93-118 (the order rule), 238-275 (the iterator).
"""
import ctypes
import sys

import numpy as np
import pytest

import lmhip
from helpers import Model
from lmhip import _lib
from oracle import numerics as ON

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _option(ctx, key):
    v = ctypes.c_int64()
    _lib.check(_lib.lib().lmh_get_option(ctx.h, key, ctypes.byref(v)))
    return v.value


def _model():
    return Model("square", (10, 12), (True, True), onsite=np.random.default_rng(1).uniform(-0.5, 0.5, 120))


def _p0(n, seed=2):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))        # not Hermitian


def _solver():
    return lmhip.HIPExp(ctx=lmhip.Context(0), step_mode=_lib.STEP_FUSED_T)


def _run(m, specs, dts, P0s, async_order, reads=()):
    """specs[k], dts[k]: the H and dt of step k + 1.  States from P0s, all
    stepped after each update.  Returns the matrices read after the steps in
    `reads`, the final ones, the K's (read after the run) and the solver."""
    solver = _solver()
    sts = [solver.make_state(P, m.device_h(specs[0])) for P in P0s]
    _lib.check(_lib.lib().lmh_set_option(solver.ctx.h, _lib.OPT_ASYNC_ORDER, int(async_order)))
    _lib.check(_lib.lib().lmh_set_profiling(solver.ctx.h, 1))
    got, Ks = {}, []
    for k, (spec, dt) in enumerate(zip(specs, dts), start=1):
        solver.update_solver(m.device_h(spec), dt)
        for st in sts:
            solver.step(st)
        if k in reads:
            got[k] = [st.to_numpy() for st in sts]
            Ks.append((k, solver.K, solver.s))
    final = [st.to_numpy() for st in sts]
    return got, final, Ks, solver


def _oracle(m, specs, dts, P0):
    P = P0
    out = []
    for spec, dt in zip(specs, dts):
        if dt != 0.0:
            U = ON.myexp(m.oracle_h(spec), dt)[0]
            P = ON.apply(U, ON.apply(U, P).conj().T).conj().T
        out.append(P)
    return out


def _compare(m, specs, dts, P0s, reads):
    ga, fa, Ka, sa = _run(m, specs, dts, P0s, True, reads)
    gb, fb, Kb, sb = _run(m, specs, dts, P0s, False, reads)
    for q, P0 in enumerate(P0s):
        ref = _oracle(m, specs, dts, P0)
        for k in reads:
            assert np.max(np.abs(ga[k][q] - ref[k - 1])) <= TOL, (q, k)
            assert np.array_equal(ga[k][q], gb[k][q]), (q, k)
        assert np.max(np.abs(fa[q] - ref[-1])) <= TOL, q
        assert np.array_equal(fa[q], fb[q]), q
    assert Ka == Kb
    stats = lambda s: [_option(s.ctx, key) for key in (_lib.OPT_STEP_COUNT, _lib.OPT_STEP_TERMS,
                                                       _lib.OPT_ASYNC_STEPS, _lib.OPT_ASYNC_REDONE)]
    return stats(sa), stats(sb)


def test_ramp_keeps_k_on_device():
    m = _model()
    tau = 0.6
    specs = [[("landau", 0.02 * min(0.05 * k, tau) / tau, 0, 0)] for k in range(24)]
    dts = [0.05] * 24
    say("case landau ramp, async vs blocking")
    (n, terms, nasync, redone), (nb, termsb, nasyncb, redoneb) = _compare(m, specs, dts, [_p0(m.n)], {1, 2, 5, 9, 17})
    assert (n, terms) == (nb, termsb) and n == 24
    # the first update has no earlier result to go by; every later one that
    # changes H leaves K on the device (an unchanged H and dt keep the known
    # K), and none of their launches is redone
    changes = sum(1 for k in range(1, 24) if specs[k] != specs[k - 1])
    assert changes == 12
    assert nasync == changes and redone == 0, (nasync, redone)
    assert nasyncb == 0 and redoneb == 0


def test_point_flux_from_zero_leaves_the_strip_form():
    m = _model()
    specs = [[("axial", 0.3 * min(0.05 * k, 0.4) / 0.4, 4.5, 5.5)] for k in range(10)]
    dts = [0.05] * 10
    say("case point flux from zero: strip launch does nothing, step redone")
    (n, terms, nasync, redone), (nb, termsb, _, _) = _compare(m, specs, dts, [_p0(m.n, 3)], {1, 2, 3, 6})
    assert (n, terms) == (nb, termsb)
    assert nasync >= 1 and redone >= 1, (nasync, redone)


def test_k_zero_and_larger_tables():
    """dt = 0 (K = 0: the launch does nothing and the step is a no-op), then
    dt = 0.9 (the order bound exceeds the grid window, which fits at
    dt = 0.05 on this 24 x 24 torus: the order search needs its larger
    tables, run when the result is taken up; the later searches start
    there)."""
    m = Model("square", (24, 24), (True, True), onsite=np.random.default_rng(1).uniform(-0.5, 0.5, 576))
    lan = [("landau", 0.02, 0, 0)]
    specs = [lan] * 9
    dts = [0.05, 0.05, 0.0, 0.05, 0.9, 0.05, 0.9, 0.9, 0.05]
    say("case K = 0 and larger order tables")
    (n, terms, nasync, redone), _ = _compare(m, specs, dts, [_p0(m.n, 4)], {2, 3, 4, 5, 6, 8})
    assert nasync >= 1 and redone >= 2, (nasync, redone)


def test_two_states_one_update():
    m = _model()
    specs = [[("landau", 0.01 + 0.001 * k, 0, 0)] for k in range(6)] + \
            [[("axial", 0.2, 4.5, 5.5)]] + [[("landau", 0.015, 0, 0)]] * 3
    dts = [0.05] * 10
    say("case two states per update")
    (n, terms, nasync, redone), _ = _compare(m, specs, dts, [_p0(m.n, 5), _p0(m.n, 6)], {3, 7, 8})
    assert n == 20 and nasync >= 8 and redone >= 2, (n, nasync, redone)


def test_reads_and_host_order_between_updates():
    m = _model()
    n = m.n
    rng = np.random.default_rng(3)
    Q = np.linalg.qr(rng.normal(size=(n, 20)) + 1j * rng.normal(size=(n, 20)))[0]
    P0 = Q @ Q.conj().T
    solver = _solver()
    spec_of = lambda k: [("landau", 0.004 * k, 0, 0)]
    st = solver.make_state(P0, m.device_h(spec_of(0)))
    series = lmhip.DeviceSeries(solver.ctx, "density", 16)
    ref = P0
    try:
        for k in range(1, 13):
            Ho = m.oracle_h(spec_of(k))
            solver.update_solver(m.device_h(spec_of(k)), 0.05)
            solver.step(st)
            U = ON.myexp(Ho, 0.05)[0]
            ref = ON.apply(U, ON.apply(U, ref).conj().T).conj().T
            kind = k % 4
            if kind == 0:
                rho, J = lmhip.observables(st, solver.H)
                assert np.max(np.abs(rho - ON.localdensity(ref))) <= TOL, k
                assert np.max(np.abs(J - ON.density_currents(Ho, ref))) <= TOL, k
            elif kind == 1:
                cp = st.copy()
                assert np.max(np.abs(cp.to_numpy() - ref)) <= TOL, k
                cp.close()
            elif kind == 2:
                series.record(st, 0.05 * k)
                assert np.max(np.abs(series.to_timesequence().values[-1] - ON.localdensity(ref))) <= TOL, k
            else:
                # the host's K and s are the blocking update's
                K = ctypes.c_int32()
                s = ctypes.c_double()
                _lib.check(_lib.lib().lmh_update_solver(solver.ctx.h, 0.05, 0, 1e-10, 1e-14, ctypes.byref(K),
                                                        ctypes.byref(s)))
                assert (solver.K, solver.s) == (K.value, s.value), k
                assert np.max(np.abs(st.to_numpy() - ref)) <= TOL, k
        assert np.max(np.abs(st.to_numpy() - ref)) <= TOL
        assert _option(solver.ctx, _lib.OPT_ASYNC_STEPS) >= 8
    finally:
        series.close()
