"""GPU: the batched workload at its own configuration -- 64 disorder
realisations on open L x L squares with L drawn uniformly from 24..96
(N = 576 .. 9216), on-site disorder uniform in [-1, 1] (W = 2) seeded per
realisation, a constant Landau field of 0.005 flux quanta per plaquette,
dt = 0.05 -- advanced as one BatchedEvolution for 7 steps (eager steps and
a graph replay), every member checked against the oracle; once with the
persistent strip launch (one per side over all 64 members) and once with the
per-size-class launches.

The check is size independent: each member starts from a rank-8 projector
P0 = Q Q^H with Q the unit vectors of 8 seeded sites (set on the device from
its 8 entries), and the oracle advances Q alone with its sparse U (myexp!
with droptol, src/evolution.jl:93-128), so P_k = Q_k Q_k^H.  After every
batch call the whole diagonal and 4000 seeded off-diagonal entries of every
member agree with Q_k Q_k^H to 1e-10, Tr P agrees with the oracle's to 1e-10
and stays 8 to 1e-8, and the per-member mean density equals Tr P / N.
Reference: one Evolution per realisation, src/evolution.jl:183-250.
"""
import sys

import numpy as np
import pytest

import lmhip
from helpers import Model
from oracle import numerics as ON

pytestmark = pytest.mark.gpu

TOL = 1e-10
MEMBERS, LMIN, LMAX, W, B, DT, RANK, SEED = 64, 24, 96, 2.0, 0.005, 0.05, 8, 20240602


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Entries:
    def __init__(self, rows, cols, vals):
        self.rows, self.cols, self.vals = rows, cols, vals


@pytest.mark.parametrize("strip", [True, False], ids=["strip", "classes"])
def test_batched_configuration_against_oracle(strip):
    rng = np.random.default_rng(SEED)
    sizes = rng.integers(LMIN, LMAX + 1, MEMBERS)
    members, oracles = [], []
    for r, L in enumerate(sizes):
        L = int(L)
        n = L * L
        rr = np.random.default_rng(SEED + 1 + r)
        V = rr.uniform(-W / 2, W / 2, n)
        m = Model("square", (L, L), onsite=V)
        spec = [("landau", B, 0, 0)]
        occ = np.sort(rr.choice(n, RANK, replace=False))
        members.append((m.device_h(spec), _Entries(occ, occ, np.ones(RANK))))
        U, K, s, _ = ON.myexp(m.oracle_h(spec), DT)
        Q = np.zeros((n, RANK), dtype=np.complex128, order="F")
        Q[occ, np.arange(RANK)] = 1.0
        pick_i = rr.integers(0, n, 4000)
        pick_j = rr.integers(0, n, 4000)
        oracles.append({"U": U, "K": K, "Q": Q, "i": pick_i, "j": pick_j})
        if r % 16 == 15:
            say("oracle members %d/%d (L up to %d)" % (r + 1, MEMBERS, int(max(sizes[: r + 1]))))
    say("case batched configuration: 64 members, L %d..%d, sum N = %d, strip=%s"
        % (sizes.min(), sizes.max(), int(np.sum(sizes.astype(np.int64) ** 2)), strip))
    batch = lmhip.BatchedEvolution(members, DT, strip=strip)
    del members
    kernels = batch.member_kernels()
    assert (kernels == [lmhip._lib.FUSED_STRIP] * MEMBERS) if strip else (lmhip._lib.FUSED_STRIP not in kernels)
    assert [s.K for s in batch.solvers] == [o["K"] for o in oracles]
    done = 0
    try:
        # eager steps up to a multiple of three, then one replay of the
        # captured three-step graph, then an eager step
        for nsteps, graph in [(2, False), (1, False), (3, True), (1, False)]:
            batch.step(nsteps, graph=graph)
            done += nsteps
            worst = 0.0
            md = batch.mean_density()
            for r, (st, o) in enumerate(zip(batch.states, oracles)):
                for _ in range(nsteps):
                    o["Q"] = ON.apply(o["U"], o["Q"])
                Q = o["Q"]
                n = st.n
                diag = np.arange(n)
                got_d = st.get_entries(diag, diag)
                ref_d = np.sum(np.abs(Q) ** 2, axis=1)
                got_o = st.get_entries(o["i"], o["j"])
                ref_o = np.sum(Q[o["i"]] * np.conj(Q[o["j"]]), axis=1)
                d = max(float(np.max(np.abs(got_d - ref_d))), float(np.max(np.abs(got_o - ref_o))))
                worst = max(worst, d)
                assert d <= TOL, (r, done, d)
                tr = float(np.sum(got_d.real))
                assert abs(tr - float(np.sum(ref_d))) <= TOL, (r, done)
                assert abs(tr - RANK) <= 1e-8, (r, done, tr)
                assert abs(md[r] - tr / n) <= 1e-13, (r, done)
            say("after %d steps (%s): worst |dP| over 64 members %.2e" % (done, "graph" if graph else "eager", worst))
    finally:
        batch.close()
