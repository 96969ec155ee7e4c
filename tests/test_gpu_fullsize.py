"""Full-size checks of the large workloads against the oracle, through
size-independent constructions (no dense N x N work on the host).

HoneycombLattice 120x120 (N = 28800), Haldane t1 = 1, t2 = 0.1, M = 0.2, point
flux at the central hexagon, dt = 0.02: from a half-filled product state
P0 = sum_{i in occ} e_i e_i^T, two device steps give P2 = W P0 W' with
W = U2 U1, so column j of P2 is W[:, occ] conj(W[j, occ])^T -- computed from
the oracle's own sparse U (myexp! with droptol) for a few columns and compared
entry by entry (1e-10), with the Taylor orders equal to the oracle's.
"""
import sys

import numpy as np
import pytest

import lmhip
from oracle import fields as OF
from oracle import geometry as OG
from oracle import hamiltonian as OH
from oracle import numerics as ON

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_haldane_n28800_two_steps_against_oracle_columns():
    L, dt, t0, flux_max, tau = 120, 0.02, 5.0, 0.5, 10.0
    c = L // 2
    pt = (c * 1.5 - 0.5, c * np.sqrt(3.0) / 2.0 - np.sqrt(3.0) / 6.0)     # central hexagon
    flux = lambda t: flux_max * min(t, tau) / tau
    lat = lmhip.HoneycombLattice(L, L)
    olat = OG.HoneycombLattice(L, L)
    n = len(lat)
    h = lambda t: lmhip.haldane(lat, 1.0, 0.1, 0.2, field=lmhip.PointFlux(flux(t), pt))
    occ = np.sort(np.random.default_rng(77).permutation(n)[: n // 2])
    ts = [t0, t0 + dt, t0 + 2 * dt]
    say("case haldane n28800 steps=2")
    ctx = lmhip.Context(0)
    ctx.set_hamiltonian(h(t0))
    P0 = lmhip.DensityMatrix(ctx, n).set_entries(occ, occ, np.ones(len(occ)))
    ev = lmhip.Evolution(lmhip.HIPExp(ctx=ctx), h, P0)
    P0.close()
    ev.time = t0
    Ks = [ev.solver.K for _ in ev(ts)][1:]
    say("device steps done, K = %s" % Ks)
    W = None
    Kref = []
    for k in range(2):
        Ho = OH.haldane(olat, 1.0, 0.1, 0.2, field=OF.PointFlux(flux(ts[k]), pt))
        U, K, s, _ = ON.myexp(Ho, ts[k + 1] - ts[k])
        Kref.append(K)
        assert s == 1.0
        Us = U.to_scipy()
        W = Us if W is None else Us @ W
    assert Ks == Kref, (Ks, Kref)
    W = W.tocsc()
    Wocc = W[:, occ]
    cols = np.array([0, 1, n // 3, n // 2 + 7, n - 2, n - 1, 14399, 20000])
    worst = 0.0
    rows = np.arange(n)
    for j in cols:
        ref = np.asarray((Wocc @ Wocc[j, :].conj().T).todense()).ravel()
        got = ev.states[0].get_entries(rows, np.full(n, j))
        worst = max(worst, float(np.max(np.abs(got - ref))))
    say("n28800 worst |dP| over %d columns = %.3e" % (len(cols), worst))
    assert worst <= TOL
    rho = lmhip.localdensity(ev.states[0])
    assert abs(rho.sum() - n // 2) <= 1e-8
