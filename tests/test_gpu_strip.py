"""GPU: the strip kernel (fused variant 32) for grid-stencil Hamiltonians.

A square lattice's H has the diagonal and the four grid neighbours per row;
under a Landau gauge (or no field) every bond of one direction along a run
of consecutive sites of one x column has the same value, and the strip
kernel keeps those values and the run's sites in registers.  The library
detects the grid from the bond pattern and checks the strip values after
every H build; when they are not uniform (a point flux) the steps run the
general fused kernels.

Each case is a ramped evolution checked against the oracle after every step
(max |P_gpu - P_oracle| <= 1e-10), with the fused variant that actually ran
(LMH_OPT_FUSED_LAUNCHED) asserted: strips of every length, open and periodic
edges, the overlapping last strip of an x column, several columns per
workgroup, s > 1 (squaring), non-Hermitian P, and kets (plain output).
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


def launched(solver):
    v = ctypes.c_int64(-2)
    lmhip._lib.check(lmhip._lib.lib().lmh_get_option(solver.ctx.h, lmhip._lib.OPT_FUSED_LAUNCHED, ctypes.byref(v)))
    return v.value


def disorder(n, seed, w=0.5):
    return np.random.default_rng(seed).uniform(-w, w, n)


# (dims, periodic, onsite seed, field spec at full ramp, expect strip)
CASES = {
    "pbc_landau_disorder_10x12": ((10, 12), (True, True), 1, [("landau", 0.013, 0, 0)], True),
    "open_landau_9x14": ((9, 14), (False, False), 2, [("landau", 0.02, 0, 0)], True),
    "ypbc_nofield_11x13": ((11, 13), (False, True), 3, [], True),
    "xpbc_landau_7x17": ((7, 17), (True, False), 4, [("landau", 0.03, 0, 0)], True),
    "pbc_landau_5x23": ((5, 23), (True, True), 5, [("landau", 0.011, 0, 0)], True),
    "pbc_landau_24x31": ((24, 31), (True, True), 6, [("landau", 0.007, 0, 0)], True),
    "open_landau_30x41": ((30, 41), (False, False), 7, [("landau", 0.004, 0, 0)], True),
    "pbc_pointflux_fallback": ((10, 12), (True, True), 8, [("axial", 0.2, 4.5, 5.5)], False),
    "open_symmetric_fallback": ((9, 11), (False, False), 9, [("symmetric", 0.03, 0, 0)], False),
}


def _ramp(fspec0, tau):
    return lambda t: [(k, a * min(t, tau) / tau, px, py) for (k, a, px, py) in fspec0]


def _evolve_both(m, ramp, P0, dt, nsteps, mode=lmhip._lib.STEP_AUTO):
    ts = [k * dt for k in range(nsteps + 1)]
    # the oracle's Evolution yields its own state, updated in place: copy each
    refs = [Po.copy() for (_Ho, Po, _t) in ON.Evolution(lambda t: m.oracle_h(ramp(t)), P0)(ts)]
    solver = lmhip.HIPExp(step_mode=mode)
    st = solver.make_state(P0, m.device_h(ramp(0.0)))
    seen = set()
    worst = 0.0
    for k in range(1, nsteps + 1):
        solver.update_solver(m.device_h(ramp(ts[k - 1])), dt)
        solver.step(st)
        seen.add(launched(solver))
        d = float(np.max(np.abs(st.to_numpy() - refs[k])))
        worst = max(worst, d)
        assert d <= TOL, (k, d)
    return seen, worst, solver


@pytest.mark.parametrize("name", sorted(CASES))
def test_strip_every_step(name):
    dims, periodic, seed, fspec0, strip = CASES[name]
    m = Model("square", dims, periodic, onsite=disorder(dims[0] * dims[1], seed))
    dt, nsteps = 0.05, 16
    ramp = _ramp(fspec0, nsteps * dt / 2)
    P0, _ = ON.densitymatrix(m.oracle_h(ramp(0.0)).dense(), mu=0.0)
    say("case %s N=%d" % (name, m.n))
    seen, worst, _ = _evolve_both(m, ramp, P0, dt, nsteps)
    say("  launched %s worst %.2e" % (sorted(seen), worst))
    if strip:
        assert seen == {lmhip._lib.FUSED_STRIP}, seen
    else:
        # the first step's H (zero field at t = 0) has uniform strips; every
        # later one does not, and those steps run a general fused kernel
        assert seen - {lmhip._lib.FUSED_STRIP}, seen


@pytest.mark.parametrize("axis,strip", [(0, True), (1, False)])
def test_strip_twisted_boundary(axis, strip):
    """Twisted periodic edges (TwistedBoundary, src/core/boundaries.jl:46-59):
    a twist across x makes the seam strips' x bonds complex but still uniform
    (the kernel's complex x-bond path); a twist across y makes one bond of a
    strip differ from the others, and the steps fall back."""
    from oracle import geometry as OG
    twists = (0.7, 0.0) if axis == 0 else (0.0, 0.7)
    m = Model("square", (10, 12), (True, True), onsite=disorder(120, 21))
    m.olat = OG.SquareLattice(10, 12, periodic=(True, True), twists=twists)
    m.llat = lmhip.SquareLattice(10, 12, boundaries=(("axis1", twists[0] or True), ("axis2", twists[1] or True)))
    ramp = _ramp([("landau", 0.02, 0, 0)], 0.4)
    P0, _ = ON.densitymatrix(m.oracle_h(ramp(0.0)).dense(), mu=0.0)
    say("case twist axis %d" % axis)
    seen, worst, _ = _evolve_both(m, ramp, P0, 0.05, 12)
    if strip:
        assert seen == {lmhip._lib.FUSED_STRIP}, seen
    else:
        assert lmhip._lib.FUSED_STRIP not in seen, seen


@pytest.mark.parametrize("mode", ["fusedT", "fusedN"])
def test_strip_non_hermitian_state(mode):
    """P <- U P U' for any P: a random non-Hermitian complex matrix."""
    m = Model("square", (10, 12), (True, True), onsite=disorder(120, 11))
    ramp = _ramp([("landau", 0.02, 0, 0)], 0.4)
    rng = np.random.default_rng(12)
    P0 = rng.normal(size=(120, 120)) + 1j * rng.normal(size=(120, 120))
    md = lmhip._lib.STEP_FUSED_T if mode == "fusedT" else lmhip._lib.STEP_FUSED_N
    say("case non-Hermitian %s" % mode)
    seen, worst, _ = _evolve_both(m, ramp, P0, 0.05, 12, mode=md)
    assert seen == {lmhip._lib.FUSED_STRIP}, seen


def test_strip_scaling_and_squaring():
    """max|H| dt > 1 selects s > 1: p(H/s) applied s times within one launch."""
    m = Model("square", (8, 12), (True, True), onsite=disorder(96, 13, w=2.0))
    spec = [("landau", 0.05, 0, 0)]
    rng = np.random.default_rng(14)
    Q = np.linalg.qr(rng.normal(size=(96, 9)) + 1j * rng.normal(size=(96, 9)))[0]
    P0 = Q @ Q.conj().T
    dt = 0.6
    assert ON.myexp(m.oracle_h(spec), dt)[2] > 1
    say("case squaring")
    seen, worst, solver = _evolve_both(m, lambda t: spec, P0, dt, 8)
    assert solver.s > 1
    assert seen == {lmhip._lib.FUSED_STRIP}, seen


def test_strip_odd_and_even_term_counts():
    """The kernel sweeps a strip's sites upwards and downwards on alternate
    Horner terms, so odd and even K take different paths through it.  Steps
    with dt giving K = 4, 5, 7, 8, 9 (the oracle's orders are asserted): one
    polynomial (the first step, and the plain pass when the state is read)
    and two per launch (the previous step's second side with K_prev, then
    this step's first side with K), both of odd and even total, checked
    against the oracle at three reads; then kets (one polynomial, plain
    stores) at odd and even K, checked after every step."""
    m = Model("square", (10, 12), (True, True), onsite=disorder(120, 1))
    spec = [("landau", 0.013, 0, 0)]
    Ho = m.oracle_h(spec)
    dts = [0.1, 0.05, 0.03, 0.03, 0.01, 0.001, 0.1, 0.05, 0.1, 0.01]
    Us = {dt: ON.myexp(Ho, dt) for dt in set(dts)}
    assert sorted({Us[dt][1] for dt in dts}) == [4, 5, 7, 8, 9], {dt: Us[dt][1] for dt in dts}
    rng = np.random.default_rng(22)
    P0 = rng.normal(size=(120, 120)) + 1j * rng.normal(size=(120, 120))
    solver = lmhip.HIPExp(step_mode=lmhip._lib.STEP_FUSED_T)
    H = m.device_h(spec)
    st = solver.make_state(P0, H)
    ref = P0
    say("case odd and even K, density matrix")
    for k, dt in enumerate(dts):
        solver.update_solver(H, dt)
        solver.step(st)
        assert launched(solver) == lmhip._lib.FUSED_STRIP
        assert solver.K == Us[dt][1], (k, dt, solver.K)
        U = Us[dt][0]
        ref = ON.apply(U, ON.apply(U, ref).conj().T).conj().T
        # (a read applies the deferred pass: between reads the launches
        # carry two polynomials)
        if k in (3, 6, len(dts) - 1):
            d = float(np.max(np.abs(st.to_numpy() - ref)))
            assert d <= TOL, (k, dt, d)
    say("case odd and even K, kets")
    psi0 = rng.normal(size=(120, 5)) + 1j * rng.normal(size=(120, 5))
    kets = lmhip.Kets.from_array(solver.ctx, psi0)
    ref = psi0.copy()
    for k, dt in enumerate(dts[:5]):
        solver.update_solver(H, dt)
        solver.step(kets)
        assert launched(solver) == lmhip._lib.FUSED_STRIP
        ref = ON.apply(Us[dt][0], ref)
        d = float(np.max(np.abs(kets.to_numpy() - ref)))
        assert d <= TOL, (k, dt, d)


def test_strip_kets_plain_output():
    """Kets psi <- U psi through the strip kernel's plain (untransposed) stores,
    several kets per workgroup, against the oracle's sparse U."""
    m = Model("square", (12, 16), (False, True), onsite=disorder(192, 15))
    spec = [("landau", 0.015, 0, 0)]
    rng = np.random.default_rng(16)
    psi0 = rng.normal(size=(192, 7)) + 1j * rng.normal(size=(192, 7))
    dt = 0.05
    U = ON.myexp(m.oracle_h(spec), dt)[0]
    solver = lmhip.HIPExp()
    H = m.device_h(spec)
    ctx = solver._context(H)
    ctx.set_hamiltonian(H)
    kets = lmhip.Kets.from_array(ctx, psi0)
    ref = psi0.copy()
    say("case kets")
    for k in range(6):
        solver.update_solver(H, dt)
        solver.step(kets)
        assert launched(solver) == lmhip._lib.FUSED_STRIP
        ref = ON.apply(U, ref)
        d = float(np.max(np.abs(kets.to_numpy() - ref)))
        assert d <= TOL, (k, d)


def test_strip_pinned_variant_matches_general_kernels():
    """The same evolution with the strip kernel and with a pinned general
    variant (20) agree to 1e-12 (device against device) and both against the
    oracle."""
    m = Model("square", (20, 36), (True, True), onsite=disorder(720, 17))
    ramp = _ramp([("landau", 0.01, 0, 0)], 0.3)
    P0, _ = ON.densitymatrix(m.oracle_h(ramp(0.0)).dense(), mu=0.0)
    dt, nsteps = 0.05, 10
    ts = [k * dt for k in range(nsteps + 1)]
    outs = {}
    for v in (-1, 20):
        solver = lmhip.HIPExp(step_mode=lmhip._lib.STEP_FUSED_T)
        st = solver.make_state(P0, m.device_h(ramp(0.0)))
        lmhip._lib.check(lmhip._lib.lib().lmh_set_option(solver.ctx.h, lmhip._lib.OPT_FUSED_VARIANT, v))
        try:
            for k in range(1, nsteps + 1):
                solver.update_solver(m.device_h(ramp(ts[k - 1])), dt)
                solver.step(st)
            outs[v] = (st.to_numpy(), launched(solver))
        finally:
            lmhip._lib.check(lmhip._lib.lib().lmh_set_option(solver.ctx.h, lmhip._lib.OPT_FUSED_VARIANT, -1))
    assert outs[-1][1] == lmhip._lib.FUSED_STRIP and outs[20][1] == 20
    assert np.max(np.abs(outs[-1][0] - outs[20][0])) <= 1e-12
    ref = [Po.copy() for (_H, Po, _t) in ON.Evolution(lambda t: m.oracle_h(ramp(t)), P0)(ts)][-1]
    assert np.max(np.abs(outs[-1][0] - ref)) <= TOL
