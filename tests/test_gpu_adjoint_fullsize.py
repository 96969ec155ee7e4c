"""GPU: the adjoint gradient of the large workload at its full size --
HoneycombLattice 120 x 120 (N = 28800), Haldane t1 = 1, t2 = 0.1, M = 0.2, a
point flux at the central hexagon ramped 0.5 / 10 per unit time (one flux
value per step), dt = 0.02, on-site potentials V (the gradient is taken at
V = 0 and at a seeded V ~ N(0, 0.05^2)), P0 a half-filled product state.

Four steps with a checkpoint every two; the loss is the density on the edge
sites (fewer neighbours than the bulk).  The gradient with respect to every
on-site potential and the flux of every step, contracted with two seeded
random unit directions, against central finite differences (h = 1e-6) of the
device forward evolution: |fd - <g, v>| <= 2e-7 + 1e-5 |fd|, with the Taylor
orders of the perturbed runs equal to the taped run's.
Reference: the evolution is src/evolution.jl:183-250 (myexp!,
src/evolution.jl:93-128); the loss is localdensity (src/currents.jl).
"""
import sys

import numpy as np
import pytest

import lmhip

pytestmark = pytest.mark.gpu

L, DT, NSTEPS, EVERY, H_FD = 120, 0.02, 4, 2, 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("vsig", [0.0, 0.05])
def test_haldane_n28800_edge_density_gradient_against_device_fd(vsig):
    c = L // 2
    pt = (c * 1.5 - 0.5, c * np.sqrt(3.0) / 2.0 - np.sqrt(3.0) / 6.0)     # central hexagon
    lat = lmhip.HoneycombLattice(L, L)
    n = len(lat)
    rng = np.random.default_rng(2805)
    V = rng.normal(0.0, vsig, n) if vsig else np.zeros(n)
    B = 0.5 * np.minimum(np.arange(NSTEPS) * DT, 10.0) / 10.0
    times = [DT * k for k in range(NSTEPS + 1)]
    idx = lambda t: min(int(round(t / DT)), NSTEPS - 1)
    hfun = lambda Vx, Bx: (lambda t: lmhip.haldane(lat, 1.0, 0.1, 0.2, Vx, field=lmhip.PointFlux(Bx[idx(t)], pt)))
    H0 = lmhip.haldane(lat, 1.0, 0.1, 0.2, V, field=lmhip.PointFlux(0.0, pt))
    deg = np.zeros(n, dtype=int)
    np.add.at(deg, H0.bonds.i, 1)
    np.add.at(deg, H0.bonds.j, 1)
    edge = np.nonzero(deg < np.max(deg))[0]
    occ = np.sort(np.random.default_rng(77).permutation(n)[: n // 2])
    ctx = lmhip.Context(0)
    ctx.set_hamiltonian(H0)
    P0 = lmhip.DensityMatrix(ctx, n).set_entries(occ, occ, np.ones(len(occ)))

    def forward(Vx, Bx, every=None):
        ev = lmhip.Evolution(lmhip.HIPExp(ctx=ctx), hfun(Vx, Bx), P0)
        tape = lmhip.EvolutionTape(ev, every) if every else None
        Ks = [ev.solver.K for _ in ev(times)][1:]
        loss = float(np.sum(lmhip.localdensity(ev.states[0])[edge]))
        return loss, ev, tape, Ks

    say("case adjoint n28800 vsig=%g steps=%d checkpoint every %d (%d edge sites)" % (vsig, NSTEPS, EVERY, len(edge)))
    loss, ev, tape, Ks = forward(V, B, every=EVERY)
    g = tape.backward(lmhip.localdensity_cotangent(n, 1.0, sites=edge))
    gV, gB = g.onsite.copy(), g.field[:, 0].copy()
    tape.close()
    ev.states[0].close()
    say("loss %.12f K=%s |dL/dV| %.4e |dL/dB| %.4e" % (loss, Ks, np.linalg.norm(gV), np.linalg.norm(gB)))
    assert gV.shape == (n,) and gB.shape == (NSTEPS,)
    dirs = np.random.default_rng(2806)
    try:
        for d in range(2):
            vV, vB = dirs.normal(size=n), dirs.normal(size=NSTEPS)
            nrm = np.sqrt(np.sum(vV ** 2) + np.sum(vB ** 2))
            vV, vB = vV / nrm, vB / nrm
            out = []
            for sgn in (1.0, -1.0):
                lx, evx, _t, Kx = forward(V + sgn * H_FD * vV, B + sgn * H_FD * vB)
                evx.states[0].close()
                assert Kx == Ks, "Taylor order changed under the perturbation"
                out.append(lx)
            fd = (out[0] - out[1]) / (2 * H_FD)
            pred = float(np.dot(gV, vV) + np.dot(gB, vB))
            say("direction %d: fd %.10e  adjoint %.10e  diff %.2e" % (d, fd, pred, abs(fd - pred)))
            assert abs(fd - pred) <= 2e-7 + 1e-5 * abs(fd), (d, fd, pred)
    finally:
        P0.close()
