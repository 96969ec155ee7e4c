"""GPU checks of the adjoint gradients (lmh_tape_backward through the C ABI).

* dH/da from the library against central differences of the oracle's H
  builder (every field kind, open and periodic lattices);
* device gradients against the dense numpy model of the same discrete
  adjoint (tests/adjoint_model.py) -- same math, 1e-9 relative;
* the differentiation workload (SquareLattice 32x32, 100-step piecewise
  constant PointFlux schedule, 1024 on-site potentials, loss = corner density
  + squared current through a column cut, checkpoint every 10 steps) against
  central finite differences (h = 1e-6) of the device forward evolution;
* a smaller case against central finite differences of the oracle's CPU
  evolution (the reference algorithm: sparse U with droptol).
"""
import sys
import time

import numpy as np
import pytest

import adjoint_model as AM
import lmhip
from oracle import fields as OF
from oracle import geometry as OG
from oracle import hamiltonian as OH
from oracle import numerics as ON

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Problem:
    """Square lattice, open, PointFlux schedule B_k (one per step), on-site V."""

    def __init__(self, L, nsteps, dt, seed, point=None, bmax=0.2, vsig=0.1, corner=5, cut=None):
        rng = np.random.default_rng(seed)
        self.L, self.n, self.nsteps, self.dt = L, L * L, nsteps, dt
        self.B = rng.uniform(0.0, bmax, nsteps)
        self.V = rng.normal(0.0, vsig, L * L)
        self.point = point if point is not None else (L / 2 + 0.5, L / 2 + 0.5)
        self.lat = lmhip.SquareLattice(L, L)
        self.olat = OG.SquareLattice(L, L)
        xy = self.lat.coords
        self.corner = np.nonzero((xy[:, 0] <= corner) & (xy[:, 1] <= corner))[0]
        self.cutx = cut if cut is not None else L // 2
        self.times = [dt * k for k in range(nsteps + 1)]

    def hdev(self, V, b):
        return lmhip.tightbinding_hamiltonian(self.lat, V, field=lmhip.PointFlux(b, self.point))

    def horc(self, V, b):
        return OH.tightbinding_hamiltonian(self.olat, onsite=V, field=OF.PointFlux(b, self.point))

    def index(self, t):
        return min(int(round(t / self.dt)), self.nsteps - 1)

    def cut_mask(self, H):
        xy = self.lat.coords
        b = H.bonds
        xi, xj = xy[b.i, 0], xy[b.j, 0]
        return ((xi == self.cutx) & (xj == self.cutx + 1)) | ((xj == self.cutx) & (xi == self.cutx + 1))

    def p0(self):
        H = self.hdev(self.V, 0.0)
        return lmhip.densitymatrix(H, mu=0.0, ctx=lmhip.Context(0)).to_numpy()

    def loss_from(self, rho, J, mask):
        S = float(np.sum(J[mask]))
        return float(np.sum(rho[self.corner])) + S * S, S

    def forward(self, V, B, P0, every=None):
        ctx = lmhip.Context(0)
        ev = lmhip.Evolution(lmhip.HIPExp(ctx=ctx), lambda t: self.hdev(V, B[self.index(t)]), P0)
        tape = lmhip.EvolutionTape(ev, every) if every else None
        H_last, Ks = None, []
        for ts in ev(self.times):
            H_last = ts.H
            Ks.append(ev.solver.K)
        rho, J = lmhip.observables(ev.states[0], H_last)
        L, S = self.loss_from(rho, J, self.cut_mask(H_last))
        return L, S, ev, tape, H_last, Ks[1:]

    def gradient(self, P0, every, mode=0, right_pass=None, left_pass=None):
        L, S, ev, tape, H_last, Ks = self.forward(self.V, self.B, P0, every)
        lmhip._lib.check(lmhip.lib().lmh_set_option(ev.solver.ctx.h, lmhip._lib.OPT_ADJOINT_MODE, mode))
        if right_pass is not None:
            lmhip._lib.check(lmhip.lib().lmh_set_option(ev.solver.ctx.h, lmhip._lib.OPT_RIGHT_PASS, right_pass))
        if left_pass is not None:
            lmhip._lib.check(lmhip.lib().lmh_set_option(ev.solver.ctx.h, lmhip._lib.OPT_LEFT_PASS, left_pass))
        n = self.n
        mask = self.cut_mask(H_last)
        lam = lmhip.localdensity_cotangent(n, 1.0, sites=self.corner)
        lamJ, gdirect = lmhip.currents_cotangent(H_last, ev.states[0], 2.0 * S * mask.astype(float))
        g = tape.backward(lam + lamJ)
        gB = g.field[:, 0].copy()
        gB[-1] += gdirect[0]
        return L, gB, g.onsite.copy(), g, Ks, ev, tape

    def loss_oracle(self, V, B, P0):
        ev = ON.Evolution(lambda t: self.horc(V, B[self.index(t)]), P0)
        Hl = None
        for Hl, _P, _t in ev(self.times):
            pass
        P = ev.P
        rho = np.real(np.diag(P))
        Hd = self.hdev(V, 0.0)          # bond list (geometry only)
        b = Hd.bonds
        hij = np.array([Hl[(int(i), int(j))] for i, j in zip(b.i, b.j)])
        J = 2.0 * np.imag(hij * P[b.j, b.i])
        return self.loss_from(rho, J, self.cut_mask(Hd))[0]


def test_hamiltonian_derivative_matches_builder_differences():
    """dH/d(strength) per user-level field strength (sum over its lowered
    components, e.g. the periodic images of a point flux) against central
    differences of the oracle's builder."""
    from helpers import fields_pair
    h = 1e-6
    cases = [
        (OG.SquareLattice(6, 5), lmhip.SquareLattice(6, 5), [("landau", 0.07, 0, 0)]),
        (OG.SquareLattice(6, 5), lmhip.SquareLattice(6, 5), [("axial", 0.3, 2.5, 2.5), ("symmetric", 0.05, 0, 0)]),
        (OG.SquareLattice(5, 5, periodic=(True, True)),
         lmhip.SquareLattice(5, 5, boundaries=(("axis1", True), ("axis2", True))), [("singular", 0.4, 2.5, 2.5)]),
        (OG.SquareLattice(6, 4, periodic=(True, False)),
         lmhip.SquareLattice(6, 4, boundaries=(("axis1", True),)), [("landau", 0.03, 0, 0), ("singular", 0.2, 3.5, 2.5)]),
    ]
    for olat, llat, spec in cases:
        _, lf = fields_pair(spec)
        H = lmhip.tightbinding_hamiltonian(llat, t2=0.3, field=lf)
        ctx = lmhip.Context(0)
        kinds, _ = H.field_params()
        owner = lmhip.strength_index(lf, llat)
        b = H.bonds
        dcomp = [lmhip.hamiltonian_derivative(ctx, H, c) for c in range(len(kinds))]
        for s_i in range(len(spec)):
            d = sum(dcomp[c] for c in range(len(kinds)) if owner[c] == s_i)

            def build(delta):
                sp = [(k, a + (delta if q == s_i else 0.0), px, py) for q, (k, a, px, py) in enumerate(spec)]
                of, _ = fields_pair(sp)
                Ho = OH.tightbinding_hamiltonian(olat, t2=0.3, field=of)
                return np.array([Ho[(int(i), int(j))] for i, j in zip(b.i, b.j)])

            fd = (build(h) - build(-h)) / (2 * h)
            assert np.max(np.abs(fd - d)) <= 1e-8, (spec, s_i, np.max(np.abs(fd - d)))


def test_device_adjoint_matches_dense_model():
    pb = Problem(L=5, nsteps=12, dt=0.1, seed=3, corner=2, cut=2)
    pb.V = np.random.default_rng(4).normal(0, 0.3, pb.n)
    P0 = pb.p0()
    C = 4
    say("case adjoint vs dense model N=%d steps=%d C=%d" % (pb.n, pb.nsteps, C))
    L, gB, gV, g, Ks, ev, tape = pb.gradient(P0, C)
    # the same discrete adjoint, dense, with the device's K per step
    Hs = [pb.horc(pb.V, b).dense() for b in pb.B]
    Ps = AM.forward(Hs, [pb.dt] * pb.nsteps, Ks, P0)
    Pd = ev.states[0].to_numpy()
    assert np.max(np.abs(Pd - Ps[-1])) <= 1e-10
    H_last = pb.hdev(pb.V, pb.B[-1])
    mask = pb.cut_mask(H_last)
    n = pb.n
    lam = lmhip.localdensity_cotangent(n, 1.0, sites=pb.corner)
    rho, J = lmhip.observables(ev.states[0], H_last)
    S = float(np.sum(J[mask]))
    lamJ, gdirect = lmhip.currents_cotangent(H_last, ev.states[0], 2.0 * S * mask.astype(float))
    cps = {k: Ps[k] for k in range(0, pb.nsteps, C)}
    Lam0, Fs = AM.backward(Hs, [pb.dt] * pb.nsteps, Ks, Ps[-1], (lam + lamJ).dense(n), cps, C)
    gV_model = sum(np.real(np.diag(F)) for F in Fs)
    hstep = 1e-7
    gB_model = []
    for k, F in enumerate(Fs):
        dH = (pb.horc(pb.V, pb.B[k] + hstep).dense() - pb.horc(pb.V, pb.B[k] - hstep).dense()) / (2 * hstep)
        gB_model.append(float(np.real(np.sum(F * dH))))
    gB_model[-1] += gdirect[0]
    scale = max(1.0, float(np.max(np.abs(gV_model))))
    assert np.max(np.abs(gV - gV_model)) <= 1e-9 * scale, np.max(np.abs(gV - gV_model))
    assert np.max(np.abs(gB - np.array(gB_model))) <= 1e-7 * max(1.0, np.max(np.abs(gB_model)))
    assert np.max(np.abs(g.P0.to_numpy() - Lam0)) <= 1e-9 * max(1.0, float(np.max(np.abs(Lam0))))


def test_differentiation_workload_against_device_finite_differences():
    """SquareLattice 32x32, 100 steps of dt=0.1, B_k ~ U[0,0.2], V ~ N(0, 0.1^2),
    checkpoint every 10 steps; adjoint vs central differences h = 1e-6."""
    pb = Problem(L=32, nsteps=100, dt=0.1, seed=2024)
    P0 = pb.p0()
    say("case differentiation N=1024 steps=100 C=10")
    t0 = time.perf_counter()
    L, gB, gV, g, Ks, ev, tape = pb.gradient(P0, 10)
    t_grad = time.perf_counter() - t0
    t0 = time.perf_counter()
    L2, *_ = pb.forward(pb.V, pb.B, P0)
    t_fwd = time.perf_counter() - t0
    assert abs(L - L2) <= 1e-12 * max(1.0, abs(L))
    say("loss %.12f  K=%s  host wall: forward %.3f s, forward+tape+backward %.3f s" % (L, sorted(set(Ks)), t_fwd, t_grad))
    h = 1e-6

    def fd(dV=None, dB=None):
        Vp = pb.V + (dV if dV is not None else 0.0)
        Vm = pb.V - (dV if dV is not None else 0.0)
        Bp = pb.B + (dB if dB is not None else 0.0)
        Bm = pb.B - (dB if dB is not None else 0.0)
        lp, _, evp, _, _, Kp = pb.forward(Vp, Bp, P0)
        lm, _, evm, _, _, Km = pb.forward(Vm, Bm, P0)
        assert Kp == Ks and Km == Ks, "Taylor order changed under the perturbation"
        return (lp - lm) / (2 * h)

    worst = 0.0
    for k in (0, 37, 99):
        e = np.zeros(pb.nsteps); e[k] = h
        f = fd(dB=e)
        worst = max(worst, abs(f - gB[k]) / max(1e-3, abs(f)))
        assert abs(f - gB[k]) <= 2e-7 + 1e-5 * abs(f), (k, f, gB[k])
    for i in (0, int(pb.corner[-1]), 500, 1023):
        e = np.zeros(pb.n); e[i] = h
        f = fd(dV=e)
        assert abs(f - gV[i]) <= 2e-7 + 1e-5 * abs(f), (i, f, gV[i])
    rng = np.random.default_rng(77)
    for _ in range(2):
        vB = rng.normal(size=pb.nsteps)
        vV = rng.normal(size=pb.n)
        nrm = np.sqrt(np.sum(vB ** 2) + np.sum(vV ** 2))
        vB, vV = vB / nrm, vV / nrm
        f = fd(dV=h * vV, dB=h * vB)
        pred = float(np.dot(vB, gB) + np.dot(vV, gV))
        assert abs(f - pred) <= 2e-7 + 1e-5 * abs(f), (f, pred)
    say("differentiation: gradient norm |dL/dB| %.4e |dL/dV| %.4e" % (np.linalg.norm(gB), np.linalg.norm(gV)))


def test_adjoint_against_oracle_cpu_finite_differences():
    """Directional derivative of the oracle's CPU evolution (reference algorithm
    with droptol) against the device adjoint."""
    pb = Problem(L=8, nsteps=20, dt=0.1, seed=11, corner=3, cut=4)
    P0 = pb.p0()
    say("case adjoint vs oracle FD N=%d steps=%d" % (pb.n, pb.nsteps))
    L, gB, gV, g, Ks, ev, tape = pb.gradient(P0, 5)
    Lo = pb.loss_oracle(pb.V, pb.B, P0)
    assert abs(L - Lo) <= 1e-10 * max(1.0, abs(L))
    rng = np.random.default_rng(8)
    h = 1e-6
    for _ in range(3):
        vB = rng.normal(size=pb.nsteps)
        vV = rng.normal(size=pb.n)
        f = (pb.loss_oracle(pb.V + h * vV, pb.B + h * vB, P0) - pb.loss_oracle(pb.V - h * vV, pb.B - h * vB, P0)) / (2 * h)
        pred = float(np.dot(vB, gB) + np.dot(vV, gV))
        assert abs(f - pred) <= 1e-7 + 1e-5 * abs(f), (f, pred)


@pytest.mark.parametrize("mode,right_pass,left_pass", [(2, 3, 0), (3, 3, 0), (2, 5, 0), (3, 1, 0), (2, 3, 1),
                                                        (3, 5, 1), (2, 5, 3), (3, 1, 3)])
@pytest.mark.parametrize("L,nsteps,C", [(8, 14, 5), (20, 10, 4)])
def test_per_term_adjoint_matches_fused(L, nsteps, C, mode, right_pass, left_pass):
    """The per-term reverse passes (any N) against the fused reverse kernels
    (same discrete adjoint, different summation order): Horner iterates kept
    (mode 2) or recomputed per reverse term (mode 3), P_k by right passes
    (right_pass 3) or by left passes between conjugate transposes (1, 5),
    banded (left_pass 0), gathering (1) or banded two-term (3) left passes,
    banded or gathering reverse terms."""
    pb = Problem(L=L, nsteps=nsteps, dt=0.1, seed=21, corner=3, cut=L // 2)
    P0 = pb.p0()
    say("case per-term vs fused adjoint N=%d steps=%d C=%d mode=%d right_pass=%d left_pass=%d"
        % (pb.n, nsteps, C, mode, right_pass, left_pass))
    out = {}
    for md in (1, mode):
        Lv, gB, gV, g, Ks, ev, tape = pb.gradient(P0, C, mode=md, right_pass=right_pass, left_pass=left_pass)
        out[md] = (Lv, gB, gV, g.P0.to_numpy())
    (L1, gB1, gV1, P1), (L2, gB2, gV2, P2) = out[1], out[mode]
    assert L1 == L2
    assert np.max(np.abs(gB1 - gB2)) <= 1e-10 * max(1.0, np.max(np.abs(gB1)))
    assert np.max(np.abs(gV1 - gV2)) <= 1e-10 * max(1.0, np.max(np.abs(gV1)))
    assert np.max(np.abs(P1 - P2)) <= 1e-10 * max(1.0, np.max(np.abs(P1)))


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_haldane_edge_density_gradient(mode):
    """Honeycomb Haldane model (NN + complex NNN, zmax 9) with a ramped point
    flux at a hexagon centre; loss = density on the edge sites; gradient with
    respect to every on-site potential and the flux of every step, fused and
    per-term reverse passes, against central differences of the device
    forward evolution."""
    L1, L2, nsteps, dt = 5, 5, 12, 0.05
    lat = lmhip.HoneycombLattice(L1, L2)
    n = len(lat)
    rng = np.random.default_rng(31)
    V = rng.normal(0.0, 0.1, n)
    B = 0.3 * np.minimum(np.arange(nsteps) * dt, 0.4) / 0.4
    pt = (3 * 1.0 + 3 * 0.5 - 0.5, 3 * np.sqrt(3) / 2 - np.sqrt(3) / 6)
    xy = lat.coords
    deg = np.zeros(n, dtype=int)
    H0 = lmhip.haldane(lat, 1.0, 0.1, 0.2, V, field=lmhip.PointFlux(0.0, pt))
    np.add.at(deg, H0.bonds.i, 1)
    np.add.at(deg, H0.bonds.j, 1)
    edge = np.nonzero(deg < np.max(deg))[0]
    P0 = lmhip.densitymatrix(H0, mu=0.0, ctx=lmhip.Context(0)).to_numpy()
    times = [dt * k for k in range(nsteps + 1)]
    idx = lambda t: min(int(round(t / dt)), nsteps - 1)

    def forward(Vx, Bx, every=None):
        ctx = lmhip.Context(0)
        ev = lmhip.Evolution(lmhip.HIPExp(ctx=ctx),
                             lambda t: lmhip.haldane(lat, 1.0, 0.1, 0.2, Vx, field=lmhip.PointFlux(Bx[idx(t)], pt)),
                             P0)
        tape = lmhip.EvolutionTape(ev, every) if every else None
        Ks = [ev.solver.K for _ in ev(times)]
        rho = lmhip.localdensity(ev.states[0])
        return float(np.sum(rho[edge])), ev, tape, Ks

    say("case haldane edge density mode=%d N=%d" % (mode, n))
    Lv, ev, tape, Ks = forward(V, B, every=4)
    lmhip._lib.check(lmhip.lib().lmh_set_option(ev.solver.ctx.h, lmhip._lib.OPT_ADJOINT_MODE, mode))
    calls = []
    g = tape.backward(lmhip.localdensity_cotangent(n, 1.0, sites=edge), progress=lambda d, t: calls.append((d, t)))
    # one call per checkpoint segment (checkpoints at steps 0, 4, 8), after the device ran it
    assert calls == [(4, nsteps), (8, nsteps), (12, nsteps)], calls
    gV, gB = g.onsite, g.field[:, 0]
    h = 1e-6
    vV, vB = rng.normal(size=n), rng.normal(size=nsteps)
    lp, _, _, Kp = forward(V + h * vV, B + h * vB)
    lm_, _, _, Km = forward(V - h * vV, B - h * vB)
    assert Kp == Ks and Km == Ks
    fd = (lp - lm_) / (2 * h)
    pred = float(np.dot(gV, vV) + np.dot(gB, vB))
    assert abs(fd - pred) <= 2e-7 + 1e-5 * abs(fd), (fd, pred)
    for i in (int(edge[0]), n // 2):
        e = np.zeros(n); e[i] = h
        f = (forward(V + e, B)[0] - forward(V - e, B)[0]) / (2 * h)
        assert abs(f - gV[i]) <= 2e-7 + 1e-5 * abs(f), (i, f, gV[i])


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_host_checkpoints_give_the_same_gradient(mode):
    """Checkpoints kept in pinned host memory (LMH_OPT_CHECKPOINTS 2) give
    exactly the gradients of device checkpoints, for the fused and both
    per-term reverse sweeps."""
    L1, L2, nsteps, dt = 4, 4, 9, 0.05
    lat = lmhip.HoneycombLattice(L1, L2)
    n = len(lat)
    V = np.random.default_rng(8).normal(0.0, 0.1, n)
    B = np.linspace(0.0, 0.3, nsteps)
    pt = (2 * 1.0 + 2 * 0.5 - 0.5, 2 * np.sqrt(3) / 2 - np.sqrt(3) / 6)
    H0 = lmhip.haldane(lat, 1.0, 0.1, 0.2, V, field=lmhip.PointFlux(0.0, pt))
    P0 = lmhip.densitymatrix(H0, mu=0.0, ctx=lmhip.Context(0)).to_numpy()
    times = [dt * k for k in range(nsteps + 1)]
    idx = lambda t: min(int(round(t / dt)), nsteps - 1)
    grads = []
    for place in (1, 2):
        ctx = lmhip.Context(0)
        lmhip._lib.check(lmhip.lib().lmh_set_option(ctx.h, lmhip._lib.OPT_CHECKPOINTS, place))
        lmhip._lib.check(lmhip.lib().lmh_set_option(ctx.h, lmhip._lib.OPT_ADJOINT_MODE, mode))
        ev = lmhip.Evolution(lmhip.HIPExp(ctx=ctx),
                             lambda t: lmhip.haldane(lat, 1.0, 0.1, 0.2, V, field=lmhip.PointFlux(B[idx(t)], pt)), P0)
        tape = lmhip.EvolutionTape(ev, 3)
        for _ in ev(times):
            pass
        say("case host checkpoints mode=%d place=%d N=%d" % (mode, place, n))
        assert tape.checkpoints == ((3, 0) if place == 1 else (0, 3))
        g = tape.backward(lmhip.localdensity_cotangent(n, 1.0))
        grads.append((g.onsite.copy(), g.field.copy(), g.P0.to_numpy()))
        tape.close()
    for a, b in zip(grads[0], grads[1]):
        assert np.array_equal(a, b)


def test_state_entries_roundtrip():
    """lmh_state_set_entries / lmh_state_get_entries against the dense copy."""
    pb = Problem(L=6, nsteps=2, dt=0.1, seed=5, corner=2, cut=2)
    P0 = pb.p0()
    ctx = lmhip.Context(0)
    ctx.set_hamiltonian(pb.hdev(pb.V, 0.0))
    st = lmhip.DensityMatrix.from_array(ctx, P0)
    rng = np.random.default_rng(3)
    r = rng.integers(0, pb.n, 200)
    c = rng.integers(0, pb.n, 200)
    assert np.array_equal(st.get_entries(r, c), P0[r, c])
    v = rng.normal(size=200) + 1j * rng.normal(size=200)
    r[:20] = r[20:40]                       # repeated pairs add
    c[:20] = c[20:40]
    st.set_entries(r, c, v)
    want = np.zeros((pb.n, pb.n), dtype=np.complex128)
    np.add.at(want, (r, c), v)
    assert np.max(np.abs(st.to_numpy() - want)) <= 1e-15
    with pytest.raises(ValueError):
        st.get_entries([pb.n], [0])


def test_tape_rejects_changes_and_unsupported_sizes():
    pb = Problem(L=4, nsteps=3, dt=0.1, seed=1, corner=1, cut=1)
    P0 = pb.p0()
    L, S, ev, tape, H_last, Ks = pb.forward(pb.V, pb.B, P0, every=2)
    assert tape.nsteps == 3
    # on-site energies changed after recording -> backward refuses
    ev.solver.ctx.set_hamiltonian(pb.hdev(pb.V + 0.1, 0.0))
    with pytest.raises(lmhip.LMHError):
        tape.backward(np.eye(pb.n))
    tape.close()
