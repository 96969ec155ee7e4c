"""Internal degrees of freedom (SpinBasis(1//2) per site, matrix index 2 i + a)
on the device against the oracle: the QWZ model (src/zoo/models.jl:130-136)
and a construct_hamiltonian with a Hermitian on-site matrix -- H entries,
evolution after every step, site densities and bond currents (block sums,
src/zoo/currents.jl:92-102), region currents, the Fermi projector and kets."""
import sys

import numpy as np
import pytest

import lmhip
from helpers import fields_pair, lattices
from oracle import hamiltonian as OH
from oracle import numerics as ON

pytestmark = pytest.mark.gpu

TOL = 1e-10
SX = [[0, 1], [1, 0]]
SY = [[0, -1j], [1j, 0]]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class SpinModel:
    """The same spinful model on both sides."""

    def __init__(self, kind, dims, periodic=(False, False), m=1.0, seed=0):
        self.kind = kind
        self.olat, self.llat = lattices("square", dims, periodic)
        rng = np.random.default_rng(seed)
        self.V = rng.uniform(-0.4, 0.4, len(self.llat))
        self.W = rng.uniform(-0.3, 0.3, len(self.llat))
        self.m = m

    @property
    def n(self):
        return 2 * len(self.llat)

    def oracle_h(self, spec):
        of, _ = fields_pair(spec)
        if self.kind == "qwz":
            h = OH.qwz(self.olat, self.m, field=of)
            return h
        terms = [OH.Onsite(self.V, SX), OH.Onsite(self.W, SY), OH.Onsite(self.V, 1.0),
                 OH.Hopping(1.0, self.olat.nn_translations(1)),
                 OH.Hopping([[0.2, 0.1j], [0.1j, -0.2]], self.olat.nn_translations(2))]
        return OH.OracleHamiltonian(self.olat, terms, of, internal=2)

    def device_h(self, spec):
        _, lf = fields_pair(spec)
        if self.kind == "qwz":
            return lmhip.qwz(self.llat, self.m, field=lf)
        return lmhip.construct_hamiltonian(
            self.llat, (np.array(SX), self.V), (np.array(SY), self.W), (1.0, self.V),
            (1.0, lmhip.NearestNeighbor(1)), (np.array([[0.2, 0.1j], [0.1j, -0.2]]), lmhip.NearestNeighbor(2)),
            field=lf, internal=lmhip.SpinBasis(0.5))


CASES = {
    "qwz_open_pointflux": (lambda: SpinModel("qwz", (7, 6), m=0.8), [("axial", 0.2, 3.5, 2.5)]),
    "qwz_pbc_landau": (lambda: SpinModel("qwz", (6, 6), (True, True), m=-1.2), [("landau", 0.05, 0, 0)]),
    "onsite_matrices_nnn": (lambda: SpinModel("custom", (6, 5), seed=3), [("axial", 0.15, 2.5, 2.5)]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_hamiltonian_entries(name):
    say("case %s entries" % name)
    make, spec = CASES[name]
    m = make()
    Ho, Hd = m.oracle_h(spec), m.device_h(spec)
    D = Ho.dense()
    ctx = lmhip.Context(0)
    diag, hij = ctx.hamiltonian_entries(Hd)
    b, _amp = Hd.geometry()
    assert np.max(np.abs(diag - np.real(np.diag(D)))) <= 1e-14
    assert np.max(np.abs(hij - D[b.i, b.j])) <= 1e-12
    # every stored entry of the oracle's H is reached by a bond or the diagonal
    cover = set(zip(b.i.tolist(), b.j.tolist())) | set(zip(b.j.tolist(), b.i.tolist()))
    cover |= {(k, k) for k in range(m.n)}
    assert set(Ho.entries) <= cover
    ctx.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_evolution_and_observables(name):
    """Ramped field, P after every step, site densities and bond currents."""
    say("case %s evolution" % name)
    make, spec0 = CASES[name]
    m = make()
    dt, nsteps = 0.05, 30
    tau = nsteps * dt / 2
    ramp = lambda t: [(k, a * min(t, tau) / tau, px, py) for (k, a, px, py) in spec0]
    P0, _ = ON.densitymatrix(m.oracle_h(ramp(0.0)).dense(), mu=0.05)
    rng = np.random.default_rng(4)
    Q = np.linalg.qr(rng.normal(size=(m.n, 7)) + 1j * rng.normal(size=(m.n, 7)))[0]
    P0 = 0.5 * P0 + 0.5 * (Q @ Q.conj().T)                # not stationary
    ts = [k * dt for k in range(nsteps + 1)]
    ev_o = ON.Evolution(lambda t: m.oracle_h(ramp(t)), P0)
    ev_d = lmhip.Evolution(lambda t: m.device_h(ramp(t)), P0)
    worst = dict(P=0.0, rho=0.0, J=0.0)
    for (Ho, Po, to), tsd in zip(ev_o(ts), ev_d(ts)):
        P_d, H_d, t_d = tsd
        assert t_d == to
        worst["P"] = max(worst["P"], float(np.max(np.abs(P_d.to_numpy() - Po))))
        rho, J = lmhip.observables(P_d, H_d)
        worst["rho"] = max(worst["rho"], float(np.max(np.abs(rho - ON.localdensity(Po, 2)))))
        Jo = ON.density_currents(Ho, Po)
        worst["J"] = max(worst["J"], float(np.max(np.abs(J - Jo))))
        assert ev_d.solver.K == ev_o.K or to == 0.0
    assert worst["P"] <= TOL and worst["rho"] <= TOL and worst["J"] <= TOL, worst


def test_region_currents_projector_and_kets():
    """QWZ with a point flux: the Fermi projector, region currents of site
    regions, and kets with two spin components per site."""
    say("case qwz projector / regions / kets")
    m = SpinModel("qwz", (6, 6), m=0.6)
    spec = [("axial", 0.3, 2.5, 3.5)]
    Ho, Hd = m.oracle_h(spec), m.device_h(spec)
    D = Ho.dense()
    E = np.linalg.eigvalsh(D)
    k = int(np.searchsorted(E, 0.0))
    mu = 0.5 * (E[k - 1] + E[k])
    P_ref, _ = ON.densitymatrix(D, mu=mu)
    P = lmhip.densitymatrix(Hd, mu=mu)
    assert np.max(np.abs(P.to_numpy() - P_ref)) <= TOL
    rng = np.random.default_rng(7)
    Q = np.linalg.qr(rng.normal(size=(m.n, 9)) + 1j * rng.normal(size=(m.n, 9)))[0]
    Pn = Q @ Q.conj().T
    Pd = lmhip.DensityMatrix.from_array(P.ctx, Pn)
    curr = lmhip.DensityCurrents(Hd, Pd)
    src = [0, 1, 6, 7, 13]
    dst = [2, 8, 14, 20]
    assert np.max(np.abs(lmhip.currentsfrom(curr, src) - ON.currentsfrom(Ho, Pn, src))) <= TOL
    assert abs(lmhip.currentsfromto(curr, src) - ON.currentsfromto(Ho, Pn, src)) <= TOL
    assert abs(lmhip.currentsfromto(curr, src, dst) - ON.currentsfromto(Ho, Pn, src, dst)) <= TOL
    assert np.max(np.abs(curr.values() - ON.density_currents(Ho, Pn))) <= TOL
    # kets: psi <- U psi, site densities |psi_{2i}|^2 + |psi_{2i+1}|^2
    psi0 = Q[:, 0].copy()
    U, _K, _s, _d = ON.myexp(Ho, 0.05)
    ev = lmhip.Evolution(lambda t: Hd, psi0)
    ts = [0.05 * j for j in range(6)]
    psi = psi0.copy()
    for j, tsd in enumerate(ev(ts)):
        if j > 0:
            psi = U.to_scipy() @ psi
        d = tsd.states[0].to_numpy().ravel()
        assert np.max(np.abs(d - psi)) <= TOL
        rho_k = lmhip.localdensity(tsd.states[0])
        assert np.max(np.abs(rho_k - (np.abs(psi) ** 2).reshape(-1, 2).sum(1))) <= TOL


def test_site_level_loss_gradient():
    """Adjoint gradient of a loss on site densities and block-summed site
    currents (two internal states per site: localdensity_cotangent with
    internal=2, currents_cotangent on the lattice bonds) against central
    differences (h = 1e-6) of the device forward: QWZ on a 5 x 4 open
    lattice, point-flux strength B_k per step (6 steps, dt = 0.1),
    L = sum of three sites' densities + S + S^2, S = w . J over the lattice
    bonds, checkpoint every 2 steps."""
    say("case qwz site-level loss gradient")
    _olat, llat = lattices("square", (5, 4))
    pt = (2.5, 1.5)
    nsteps, dt = 6, 0.1
    rng = np.random.default_rng(31)
    B = rng.uniform(0.05, 0.3, nsteps)
    times = [dt * k for k in range(nsteps + 1)]
    n = 2 * len(llat)
    Q = np.linalg.qr(rng.normal(size=(n, 11)) + 1j * rng.normal(size=(n, 11)))[0]
    P0 = Q @ Q.conj().T
    corner = [0, 1, 4]

    def hfun(Bs):
        return lambda t: lmhip.qwz(llat, 0.7, field=lmhip.PointFlux(Bs[min(int(round(t / dt)), nsteps - 1)], pt))

    nb = len(hfun(B)(0.0).bonds.i)
    wJ = rng.normal(size=nb)

    def forward(Bs, every=None):
        ev = lmhip.Evolution(lmhip.HIPExp(ctx=lmhip.Context(0)), hfun(Bs), P0)
        tape = lmhip.EvolutionTape(ev, every) if every else None
        H_last = None
        for ts in ev(times):
            H_last = ts.H
        rho, J = lmhip.observables(ev.states[0], H_last)
        assert len(rho) == len(llat) and len(J) == nb
        S = float(np.dot(wJ, J))
        return float(np.sum(rho[corner])) + S + S * S, S, ev, tape, H_last

    L, S, ev, tape, H_last = forward(B, every=2)
    lam = lmhip.localdensity_cotangent(n, 1.0, sites=corner, internal=2)
    lamJ, gdirect = lmhip.currents_cotangent(H_last, ev.states[0], (1.0 + 2.0 * S) * wJ)
    g = tape.backward(lam + lamJ)
    assert g.field.shape == (nsteps, 1), g.field.shape
    gB = g.field[:, 0].copy()
    gB[-1] += gdirect[0]
    h = 1e-6
    fd = []
    for k in range(nsteps):
        Bp, Bm = B.copy(), B.copy()
        Bp[k] += h
        Bm[k] -= h
        fd.append((forward(Bp)[0] - forward(Bm)[0]) / (2 * h))
    fd = np.array(fd)
    say("  L %.12f  dL/dB adjoint %s  central differences %s" % (L, np.array2string(gB, precision=9),
                                                                np.array2string(fd, precision=9)))
    assert np.max(np.abs(gB - fd)) <= 1e-7 * max(1.0, float(np.max(np.abs(fd))))
    tape.close()
