"""GPU: value classes of H for the fused passes (variants 29-31).

Slots of H whose values agree, directly or conjugated, at every probe of the
field amplitudes share a class, and those variants read the class values
from LDS instead of H's values from L2.  Each case is a ramped evolution
checked against the oracle after every step (max|P_gpu - P_oracle| <= 1e-10):

* with the classes in use (three probes): the check after every H build
  must accept them;
* with classes from the zero-field probe alone, which puts every Landau or
  point-flux phase in the class of the bare hopping: every H build with a
  nonzero field must reject them, and the steps run the plain variant;
* with no classes.
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


def classes(solver):
    n, ok = ctypes.c_int64(-1), ctypes.c_int64(-1)
    lmhip._lib.check(lmhip._lib.lib().lmh_value_classes(solver.ctx.h, ctypes.byref(n), ctypes.byref(ok)))
    return n.value, ok.value


CASES = {
    "square_pbc_landau_disorder": (
        lambda: Model("square", (12, 10), (True, True), onsite=np.random.default_rng(7).uniform(-0.5, 0.5, 120)),
        [("landau", 0.013, 0, 0)]),
    "square_open_axial": (lambda: Model("square", (10, 10)), [("axial", 0.1, 5.5, 5.5)]),
    "square_y_periodic_sum": (lambda: Model("square", (6, 9), (False, True)),
                              [("landau", 0.05, 0, 0), ("singular", 0.3, 2.5, 4.5)]),
}


@pytest.mark.parametrize("probes", [3, 1, 0])
@pytest.mark.parametrize("variant", [29, 30, 31])
@pytest.mark.parametrize("name", sorted(CASES))
def test_value_class_variants_every_step(name, variant, probes):
    make, fspec0 = CASES[name]
    m = make()
    dt, nsteps = 0.05, 24
    tau = nsteps * dt / 2
    ramp = lambda t: [(k, a * min(t, tau) / tau, px, py) for (k, a, px, py) in fspec0]
    rng = np.random.default_rng(5)
    Q = np.linalg.qr(rng.normal(size=(m.n, m.n // 3)) + 1j * rng.normal(size=(m.n, m.n // 3)))[0]
    P0 = Q @ Q.conj().T
    ts = [k * dt for k in range(nsteps + 1)]
    # the oracle's Evolution yields its own state, updated in place: copy each
    refs = [Po.copy() for (_Ho, Po, _t) in ON.Evolution(lambda t: m.oracle_h(ramp(t)), P0)(ts)]
    say("case %s variant %d probes %d" % (name, variant, probes))
    solver = lmhip.HIPExp(step_mode=lmhip._lib.STEP_FUSED_T)
    st = solver.make_state(P0, m.device_h(ramp(0.0)))
    solver.update_solver(m.device_h(ramp(0.0)), dt)
    L = lmhip._lib.lib()
    lmhip._lib.check(L.lmh_set_option(solver.ctx.h, lmhip._lib.OPT_VALUE_CLASSES, probes))
    lmhip._lib.check(L.lmh_set_option(solver.ctx.h, lmhip._lib.OPT_FUSED_VARIANT, variant))
    seen = set()
    try:
        for k in range(1, nsteps + 1):
            solver.update_solver(m.device_h(ramp(ts[k - 1])), dt)
            n, ok = classes(solver)
            seen.add((n > 0, ok))
            solver.step(st)
            d = float(np.max(np.abs(st.to_numpy() - refs[k])))
            assert d <= TOL, (k, d)
    finally:
        # contexts may outlive the solver: leave the defaults behind
        lmhip._lib.check(L.lmh_set_option(solver.ctx.h, lmhip._lib.OPT_VALUE_CLASSES, 3))
        lmhip._lib.check(L.lmh_set_option(solver.ctx.h, lmhip._lib.OPT_FUSED_VARIANT, -1))
    if probes == 3:
        assert seen == {(True, 1)}, seen
    elif probes == 1:
        # the zero field at t = 0 matches the merged classes; every later step rejects them
        assert (True, 0) in seen and (False, 1) not in seen, seen
    else:
        assert seen == {(False, 0)}, seen


def test_landau_classes_are_few():
    """SquareLattice 12x10, periodic, Landau gauge: the x bonds carry the bare
    hopping, a y bond's phase depends on its x coordinate only, and the two
    directions of a bond are conjugate -- so a handful of classes per column
    of sites, however many sites."""
    make, fspec = CASES["square_pbc_landau_disorder"]
    m = make()
    solver = lmhip.HIPExp(step_mode=lmhip._lib.STEP_FUSED_T)
    H = m.device_h(fspec)
    solver.make_state(np.eye(m.n, dtype=complex), H)
    lmhip._lib.check(lmhip._lib.lib().lmh_set_option(solver.ctx.h, lmhip._lib.OPT_VALUE_CLASSES, 3))
    solver.update_solver(H, 0.05)
    n, ok = classes(solver)
    say("case landau classes %d of %d slots" % (n, 4 * m.n))
    assert ok == 1
    assert 2 <= n <= 2 + 2 * 12, n
