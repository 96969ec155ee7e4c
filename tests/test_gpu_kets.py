"""GPU parity of ket evolution (step!(::CachedExp, ::AbstractVector),
src/evolution.jl:69-72: psi <- U psi) against the oracle's sparse U (myexp!
with droptol) after every step, 1e-10; fused column passes and per-term
passes; observables |psi_i|^2 and 2 Im(H_ij psi_j conj(psi_i))."""
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
    "square_open_axial": (lambda: Model("square", (10, 10)), [("axial", 0.2, 5.5, 5.5)], 0.1, 40),
    "square_pbc_landau_disorder": (
        lambda: Model("square", (12, 10), (True, True), onsite=np.random.default_rng(7).uniform(-0.5, 0.5, 120)),
        [("landau", 0.02, 0, 0)], 0.05, 40),
    "haldane_pointflux": (lambda: Model("honeycomb", (5, 6), t1=1.0, t2=0.1, m=0.2), [("axial", 0.15, 2.75, 2.2)],
                          0.05, 30),
}


@pytest.mark.parametrize("path", ["fused", "perterm"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_ket_evolution_every_step(name, path):
    make, fspec0, dt, nsteps = CASES[name]
    m = make()
    say("case kets %s %s" % (name, path))
    tau = nsteps * dt / 2
    ramp = lambda t: [(k, a * min(t, tau) / tau, px, py) for (k, a, px, py) in fspec0]
    rng = np.random.default_rng(5)
    psi0 = rng.normal(size=m.n) + 1j * rng.normal(size=m.n)
    psi0 /= np.linalg.norm(psi0)
    ctx = lmhip.Context(0)
    solver = lmhip.HIPExp(ctx=ctx)
    ev = lmhip.Evolution(solver, lambda t: m.device_h(ramp(t)), psi0)
    if path == "perterm":
        lmhip._lib.check(lmhip.lib().lmh_set_option(ctx.h, lmhip._lib.OPT_FUSED_MAX_N, 0))
    ts = [k * dt for k in range(nsteps + 1)]
    psi = psi0.copy()
    worst = 0.0
    for k, st in enumerate(ev(ts)):
        if k == 0:
            continue
        Ho = m.oracle_h(ramp(ts[k - 1]))
        U, K, s, _ = ON.myexp(Ho, ts[k] - ts[k - 1])
        assert solver.K == K
        psi = ON.apply(U, psi.reshape(-1, 1))[:, 0]
        got = st.states[0].to_numpy()
        d = float(np.max(np.abs(got - psi)))
        worst = max(worst, d)
        assert d <= TOL, (k, d)
        rho, J = lmhip.observables(st.states[0], st.H)
        assert np.max(np.abs(rho - np.abs(psi) ** 2)) <= TOL
        b = st.H.bonds
        hij = np.array([Ho[(int(i), int(j))] for i, j in zip(b.i, b.j)])
        Jref = 2 * np.imag(hij * psi[b.j] * np.conj(psi[b.i]))
        assert np.max(np.abs(J - Jref)) <= TOL
    assert abs(np.linalg.norm(ev.states[0].to_numpy()) - 1.0) <= 1e-9
    say("kets %s %s worst %.2e" % (name, path, worst))


def test_ket_block_matches_columns():
    """m kets in one block evolve like m separate kets."""
    m = Model("square", (9, 8), (True, False), onsite=np.random.default_rng(2).uniform(-1, 1, 72))
    spec = [("landau", 0.03, 0, 0)]
    rng = np.random.default_rng(8)
    block = rng.normal(size=(m.n, 5)) + 1j * rng.normal(size=(m.n, 5))
    ctx = lmhip.Context(0)
    H = m.device_h(spec)
    solver = lmhip.HIPExp(ctx=ctx)
    ctx.set_hamiltonian(H)
    kets = lmhip.Kets.from_array(ctx, block)
    solver.update_solver(H, 0.07)
    for _ in range(6):
        solver.step(kets)
    U, K, s, _ = ON.myexp(m.oracle_h(spec), 0.07)
    ref = block
    for _ in range(6):
        ref = ON.apply(U, ref)
    assert np.max(np.abs(kets.to_numpy() - ref)) <= TOL
