"""GPU: deferred second passes (LMH_OPT_DEFER).  A fused step on the strip
kernel launches its first pass only; the deferred second pass runs in the
same launch as the next step's first pass, or alone before the state is read.

Checked against the oracle's evolution (1e-10) and against the same
evolution with deferral off (bit for bit): long runs of unread steps, reads
of every kind between them (observables, entries, copies, device pointers,
series, allocation-free region currents, idempotency), a step whose H no
longer qualifies for the strip kernel while a pass is pending, a K = 0 step,
the option switched off with a pass pending, a geometry change, and the
launch count (one per step plus one for the last read).
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


def _model():
    return Model("square", (10, 12), (True, True), onsite=np.random.default_rng(1).uniform(-0.5, 0.5, 120))


def _p0(n, seed=2):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))        # not Hermitian


def _option(ctx, key):
    v = ctypes.c_int64()
    _lib.check(_lib.lib().lmh_get_option(ctx.h, key, ctypes.byref(v)))
    return v.value


def _launches(ctx):
    n, ms = ctypes.c_int64(), ctypes.c_double()
    _lib.check(_lib.lib().lmh_get_profile(ctx.h, _lib.KERNEL_FUSED, ctypes.byref(n), ctypes.byref(ms)))
    return n.value


def _run(m, spec_of, P0, ts, defer, reads=()):
    """Device evolution; reads: step indices after which the state is read
    (downloaded).  Returns the matrices read and the final one."""
    solver = lmhip.HIPExp(ctx=lmhip.Context(0), step_mode=_lib.STEP_FUSED_T)
    st = solver.make_state(P0, m.device_h(spec_of(ts[0])))
    _lib.check(_lib.lib().lmh_set_option(solver.ctx.h, _lib.OPT_DEFER, int(defer)))
    got = {}
    for k in range(1, len(ts)):
        solver.update_solver(m.device_h(spec_of(ts[k - 1])), ts[k] - ts[k - 1])
        solver.step(st)
        if k in reads:
            got[k] = st.to_numpy()
    return got, st.to_numpy(), solver, st


def test_unread_steps_match_oracle_and_undeferred():
    m = _model()
    tau = 0.6
    spec_of = lambda t: [("landau", 0.02 * min(t, tau) / tau, 0, 0)]
    P0 = _p0(m.n)
    ts = [0.05 * k for k in range(25)]
    refs = [Po.copy() for (_H, Po, _t) in ON.Evolution(lambda t: m.oracle_h(spec_of(t)), P0)(ts)]
    reads = {1, 2, 5, 6, 11, 17}
    say("case deferred steps, reads after %s" % sorted(reads))
    gd, fd, sol, _st = _run(m, spec_of, P0, ts, True, reads)
    assert _option(sol.ctx, _lib.OPT_FUSED_LAUNCHED) == _lib.FUSED_STRIP
    gn, fn, _s, _t = _run(m, spec_of, P0, ts, False, reads)
    for k in reads:
        assert np.max(np.abs(gd[k] - refs[k])) <= TOL, k
        assert np.array_equal(gd[k], gn[k]), k
    assert np.max(np.abs(fd - refs[-1])) <= TOL
    assert np.array_equal(fd, fn)


def test_launch_count_one_per_step():
    m = _model()
    spec = [("landau", 0.02, 0, 0)]
    P0 = _p0(m.n)
    solver = lmhip.HIPExp(ctx=lmhip.Context(0), step_mode=_lib.STEP_FUSED_T)
    H = m.device_h(spec)
    st = solver.make_state(P0, H)
    solver.update_solver(H, 0.05)
    _lib.check(_lib.lib().lmh_set_profiling(solver.ctx.h, 1))
    for _ in range(9):
        solver.step(st)
    assert _launches(solver.ctx) == 9
    _lib.check(_lib.lib().lmh_sync(solver.ctx.h))          # the last second pass
    assert _launches(solver.ctx) == 10
    _lib.check(_lib.lib().lmh_sync(solver.ctx.h))          # nothing left
    assert _launches(solver.ctx) == 10
    U = ON.myexp(m.oracle_h(spec), 0.05)[0]
    ref = P0
    for _ in range(9):
        ref = ON.apply(U, ON.apply(U, ref).conj().T).conj().T
    assert np.max(np.abs(st.to_numpy() - ref)) <= TOL


def test_reads_of_every_kind_between_deferred_steps():
    m = _model()
    spec = [("landau", 0.015, 0, 0)]
    Ho = m.oracle_h(spec)
    n = m.n
    rng = np.random.default_rng(3)
    Q = np.linalg.qr(rng.normal(size=(n, 20)) + 1j * rng.normal(size=(n, 20)))[0]
    P0 = Q @ Q.conj().T
    U = ON.myexp(Ho, 0.05)[0]
    step = lambda P: ON.apply(U, ON.apply(U, P).conj().T).conj().T
    solver = lmhip.HIPExp(ctx=lmhip.Context(0), step_mode=_lib.STEP_FUSED_T)
    H = m.device_h(spec)
    st = solver.make_state(P0, H)
    solver.update_solver(H, 0.05)
    series = lmhip.DeviceSeries(solver.ctx, "density", 16)
    ref = P0
    rows = rng.integers(0, n, 30)
    cols = rng.integers(0, n, 30)
    src = np.zeros(n, dtype=np.uint8)
    src[:37] = 1
    hip = ctypes.CDLL("libamdhip64.so")
    dsrc, dout = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dsrc), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(dout), ctypes.c_size_t(8)) == 0
    assert hip.hipMemcpy(dsrc, ctypes.c_void_p(src.ctypes.data), ctypes.c_size_t(n), 1) == 0
    try:
        for k in range(1, 13):
            solver.step(st)
            ref = step(ref)
            kind = k % 6
            if kind == 0:
                rho, J = lmhip.observables(st, solver.H)
                assert np.max(np.abs(rho - ON.localdensity(ref))) <= TOL, k
                assert np.max(np.abs(J - ON.density_currents(Ho, ref))) <= TOL, k
            elif kind == 1:
                got = st.get_entries(rows, cols)
                assert np.max(np.abs(got - ref[rows, cols])) <= TOL, k
            elif kind == 2:
                cp = st.copy()
                assert np.max(np.abs(cp.to_numpy() - ref)) <= TOL, k
                cp.close()
            elif kind == 3:
                series.record(st, 0.05 * k)
                series_ref = ON.localdensity(ref)
                assert np.max(np.abs(series.to_timesequence().values[-1] - series_ref)) <= TOL, k
            elif kind == 4:
                _lib.check(_lib.lib().lmh_currents_region_noalloc(solver.ctx.h, st.h, dsrc, None, None, dout))
                _lib.check(_lib.lib().lmh_sync(solver.ctx.h))
                tot = np.zeros(1)
                assert hip.hipMemcpy(ctypes.c_void_p(tot.ctypes.data), dout, ctypes.c_size_t(8), 2) == 0
                assert abs(tot[0] - ON.currentsfromto(Ho, ref, np.nonzero(src)[0])) <= TOL, k
            else:
                err = st.idempotency_error()
                assert err <= 1e-9, (k, err)
        assert np.max(np.abs(st.to_numpy() - ref)) <= TOL
    finally:
        hip.hipFree(dsrc)
        hip.hipFree(dout)
        series.close()


def test_pending_pass_then_other_kernels_option_and_geometry():
    """A pass pending when: the next H no longer has uniform strips (a point
    flux: the general kernels run), the step has K = 0 (dt = 0: nothing to
    do), the option is switched off, and the geometry changes."""
    m = _model()
    n = m.n
    P0 = _p0(n, 5)
    landau = [("landau", 0.02, 0, 0)]
    flux = [("axial", 0.3, 4.5, 5.5)]
    dt = 0.05
    solver = lmhip.HIPExp(ctx=lmhip.Context(0), step_mode=_lib.STEP_FUSED_T)
    st = solver.make_state(P0, m.device_h(landau))
    U_l = ON.myexp(m.oracle_h(landau), dt)[0]
    U_f = ON.myexp(m.oracle_h(flux), dt)[0]
    step = lambda U, P: ON.apply(U, ON.apply(U, P).conj().T).conj().T
    ref = P0
    say("case pending pass across kernel changes")
    # Landau (strip, deferred), then the point flux (general kernel)
    solver.update_solver(m.device_h(landau), dt)
    solver.step(st)
    ref = step(U_l, ref)
    solver.update_solver(m.device_h(flux), dt)
    solver.step(st)
    assert _option(solver.ctx, _lib.OPT_FUSED_LAUNCHED) != _lib.FUSED_STRIP
    ref = step(U_f, ref)
    assert np.max(np.abs(st.to_numpy() - ref)) <= TOL
    # deferred, then a K = 0 step (dt = 0 leaves U = I), then one more
    solver.update_solver(m.device_h(landau), dt)
    solver.step(st)
    ref = step(U_l, ref)
    solver.update_solver(m.device_h(landau), 0.0)
    assert solver.K == 0
    solver.step(st)
    solver.update_solver(m.device_h(landau), dt)
    solver.step(st)
    ref = step(U_l, ref)
    assert np.max(np.abs(st.to_numpy() - ref)) <= TOL
    # a pass pending when the option goes off
    solver.step(st)
    ref = step(U_l, ref)
    _lib.check(_lib.lib().lmh_set_option(solver.ctx.h, _lib.OPT_DEFER, 0))
    solver.step(st)
    ref = step(U_l, ref)
    assert np.max(np.abs(st.to_numpy() - ref)) <= TOL
    _lib.check(_lib.lib().lmh_set_option(solver.ctx.h, _lib.OPT_DEFER, 1))
    # a pass pending when the geometry changes (same size, open lattice):
    # it runs with the geometry it was made for
    solver.step(st)
    ref = step(U_l, ref)
    m2 = Model("square", (10, 12), onsite=np.random.default_rng(1).uniform(-0.5, 0.5, 120))
    solver.update_solver(m2.device_h(landau), dt)
    assert np.max(np.abs(st.to_numpy() - ref)) <= TOL
    solver.step(st)
    ref = step(ON.myexp(m2.oracle_h(landau), dt)[0], ref)
    assert np.max(np.abs(st.to_numpy() - ref)) <= TOL
