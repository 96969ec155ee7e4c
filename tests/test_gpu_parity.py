"""GPU parity: liblmhip.so (through the C ABI via the lmhip mirror) against the
oracle on the same inputs.

Bars: Hamiltonian entries within 1e-12 (a few ulp of the phase), the Taylor
order K identical to the reference rule's, and after EVERY step
max|P_gpu - P_oracle| <= 1e-10, densities and currents within 1e-10
(complex FP64 throughout).
"""
import ctypes
import sys

import numpy as np
import pytest

import lmhip
from helpers import Model, bond_dict
from oracle import numerics as ON

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


MODELS = {
    "square_open_axial": (lambda: Model("square", (10, 10)), [("axial", 0.1, 5.5, 5.5)]),
    "square_open_axial_offgrid": (lambda: Model("square", (9, 7)), [("axial", 0.37, 3.3, 2.9)]),
    "square_pbc_landau_disorder": (
        lambda: Model("square", (12, 10), (True, True), onsite=np.random.default_rng(7).uniform(-0.5, 0.5, 120)),
        [("landau", 0.013, 0, 0)]),
    "square_pbc_pointflux": (lambda: Model("square", (8, 8), (True, True)), [("axial", 0.21, 4.5, 3.5)]),
    "square_pbc_y_only_sum": (lambda: Model("square", (6, 9), (False, True), t2=0.3),
                              [("landau", 0.05, 0, 0), ("singular", 0.3, 2.5, 4.5)]),
    "square_symmetric_nnn": (lambda: Model("square", (7, 7), t2=-0.25), [("symmetric", 0.07, 0, 0)]),
    "haldane_pointflux": (lambda: Model("honeycomb", (5, 6), t1=1.0, t2=0.1, m=0.2,
                                        onsite=np.random.default_rng(3).normal(0, 0.1, 60)),
                          [("axial", 0.15, 2.75, 2.2)]),
    "haldane_pbc_landau": (lambda: Model("honeycomb", (6, 6), (True, True), t1=1.0, t2=0.1, m=0.2),
                           [("landau", 0.02, 0, 0)]),
}


@pytest.mark.parametrize("name", sorted(MODELS))
def test_hamiltonian_entries(gpu_ctx, name):
    make, fspec = MODELS[name]
    say("case %s" % name)
    m = make()
    Ho = m.oracle_h(fspec)
    Hd = m.device_h(fspec)
    diag, hij = gpu_ctx.hamiltonian_entries(Hd)
    bonds = Hd.bonds
    # same bond list (order and indices) as the reference's iteration
    assert [(int(i), int(j)) for i, j in zip(bonds.i, bonds.j)] == Ho.bonds
    ref = np.array([Ho[(int(i), int(j))] for i, j in zip(bonds.i, bonds.j)])
    assert np.max(np.abs(hij - ref)) <= 1e-12
    dref = np.array([Ho[(i, i)].real for i in range(m.n)])
    assert np.max(np.abs(diag - dref)) <= 1e-15


@pytest.mark.parametrize("name", sorted(MODELS))
@pytest.mark.parametrize("dt", [0.02, 0.05, 0.1, 0.3])
def test_taylor_order(name, dt):
    make, fspec = MODELS[name]
    say("case %s dt=%g" % (name, dt))
    m = make()
    solver = lmhip.HIPExp()
    solver.update_solver(m.device_h(fspec), dt)
    _U, K, s, _d = ON.myexp(m.oracle_h(fspec), dt)
    assert (solver.K, solver.s) == (K, s)


def _run_both(m, fspec_of_t, P0, ts, check_every=1, mode=lmhip._lib.STEP_AUTO):
    """Evolve on both sides; return max deviations of P, rho, J over the checks."""
    ho = lambda t: m.oracle_h(fspec_of_t(t))
    hd = lambda t: m.device_h(fspec_of_t(t))
    ev_o = ON.Evolution(ho, P0)
    ev_d = lmhip.Evolution(lmhip.HIPExp(step_mode=mode), hd, P0)
    worst = dict(P=0.0, rho=0.0, J=0.0, Kmismatch=0)
    for k, ((Ho, Po, to), tsd) in enumerate(zip(ev_o(ts), ev_d(ts))):
        P_d, H_d, t_d = tsd
        assert t_d == to
        if k % check_every and k != len(ts) - 1:
            continue
        Pg = P_d.to_numpy()
        worst["P"] = max(worst["P"], float(np.max(np.abs(Pg - Po))))
        rho, J = lmhip.observables(P_d, H_d)
        worst["rho"] = max(worst["rho"], float(np.max(np.abs(rho - ON.localdensity(Po)))))
        Jo = ON.density_currents(Ho, Po)
        worst["J"] = max(worst["J"], float(np.max(np.abs(J - Jo))) if len(J) else 0.0)
        if k > 0 and ev_d.solver.K != ev_o.K:
            worst["Kmismatch"] += 1
    return worst


def test_cfg1_every_step_against_golden_and_oracle():
    """SquareLattice(10,10), PointFlux ramp through (5.5,5.5), 200 steps of 0.1."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg1.npz"))
    m = Model("square", (10, 10))
    tau = 10.0
    fspec = lambda t: [("axial", 0.2 * min(t, tau) / tau, 5.5, 5.5)]
    ts = list(g["ts"])
    say("case cfg1 N=100 steps=%d" % (len(ts) - 1))
    ev_d = lmhip.Evolution(lambda t: m.device_h(fspec(t)), g["P0"])
    ev_o = ON.Evolution(lambda t: m.oracle_h(fspec(t)), g["P0"])
    worst = 0.0
    for k, (tsd, (Ho, Po, to)) in enumerate(zip(ev_d(ts), ev_o(ts))):
        P_d, H_d, t_d = tsd
        if k == 0:
            continue
        rho, J = lmhip.observables(P_d, H_d)
        assert ev_d.solver.K == g["K"][k - 1]
        assert np.max(np.abs(rho - g["rho"][k - 1])) <= TOL, k
        assert np.max(np.abs(J - g["J"][k - 1])) <= TOL, k
        Pg = P_d.to_numpy()
        d = float(np.max(np.abs(Pg - Po)))
        worst = max(worst, d)
        assert d <= TOL, (k, d)
    Pg = ev_d.states[0].to_numpy()
    assert np.max(np.abs(Pg - g["P_final"])) <= TOL
    say("cfg1 worst |dP| = %.3e" % worst)


MODES = {"auto": lmhip._lib.STEP_AUTO, "perpass": lmhip._lib.STEP_PER_PASS,
         "fusedT": lmhip._lib.STEP_FUSED_T, "fusedN": lmhip._lib.STEP_FUSED_N}


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("name,dt,nsteps", [
    ("square_pbc_landau_disorder", 0.05, 120),
    ("haldane_pointflux", 0.02, 100),
    ("square_pbc_pointflux", 0.1, 60),
    ("square_pbc_y_only_sum", 0.05, 60),
    ("haldane_pbc_landau", 0.05, 60),
])
def test_ramped_evolution_every_step(name, dt, nsteps, mode):
    make, fspec0 = MODELS[name]
    m = make()
    say("case %s dt=%g steps=%d mode=%s" % (name, dt, nsteps, mode))
    tau = nsteps * dt / 2
    ramp = lambda t: [(k, a * min(t, tau) / tau, px, py) for (k, a, px, py) in fspec0]
    P0, _ = ON.densitymatrix(m.oracle_h(ramp(0.0)).dense(), mu=0.0)
    ts = [k * dt for k in range(nsteps + 1)]
    w = _run_both(m, ramp, P0, ts, mode=MODES[mode])
    assert w["Kmismatch"] == 0, w
    assert w["P"] <= TOL and w["rho"] <= TOL and w["J"] <= TOL, w


def test_constant_h_quench_and_stationarity():
    """Constant H: U is cached (no rebuild); P0 of H itself must stay put."""
    m = Model("square", (10, 10), (True, True), onsite=np.random.default_rng(11).uniform(-0.5, 0.5, 100))
    spec = [("landau", 0.01, 0, 0)]
    Ho = m.oracle_h(spec)
    P0, _ = ON.densitymatrix(Ho.dense(), mu=0.0)
    ts = [k * 0.05 for k in range(101)]
    w = _run_both(m, lambda t: spec, P0, ts, check_every=10)
    assert w["P"] <= TOL, w
    ev = lmhip.Evolution(lambda t: m.device_h(spec), P0)
    for _ in ev(ts):
        pass
    assert np.max(np.abs(ev.states[0].to_numpy() - P0)) <= 1e-9


@pytest.mark.parametrize("mode", sorted(MODES))
def test_scaling_and_squaring_branch(mode):
    """max|H| * dt > 1 selects s > 1: U = p(H/s)^s, the reference's squaring."""
    m = Model("square", (6, 6), onsite=np.random.default_rng(5).uniform(-2.0, 2.0, 36))
    spec = [("axial", 0.1, 3.5, 3.5)]
    P0, _ = ON.densitymatrix(m.oracle_h(spec).dense(), mu=0.3)
    rng = np.random.default_rng(9)
    Q = np.linalg.qr(rng.normal(size=(36, 5)) + 1j * rng.normal(size=(36, 5)))[0]
    P0 = Q @ Q.conj().T                               # a non-stationary pure projector
    ts = [k * 0.6 for k in range(11)]
    w = _run_both(m, lambda t: spec, P0, ts, mode=MODES[mode])
    solver_s = ON.myexp(m.oracle_h(spec), 0.6)[2]
    assert solver_s > 1
    assert w["P"] <= TOL and w["Kmismatch"] == 0, w


def test_autotuned_schedule_keeps_parity():
    """The autotuner only changes the schedule (mode, tile shapes); the result
    after every step still matches the oracle."""
    m = Model("square", (12, 12), (True, True), onsite=np.random.default_rng(21).uniform(-0.5, 0.5, 144))
    tau = 2.0
    fs = lambda t: [("landau", 0.02 * min(t, tau) / tau, 0, 0)]
    P0, _ = ON.densitymatrix(m.oracle_h(fs(0.0)).dense(), mu=0.0)
    ts = [0.05 * k for k in range(61)]
    solver = lmhip.HIPExp()
    ev_d = lmhip.Evolution(solver, lambda t: m.device_h(fs(t)), P0)
    ev_o = ON.Evolution(lambda t: m.oracle_h(fs(t)), P0)
    worst = 0.0
    for k, (st, (_H, Po, _t)) in enumerate(zip(ev_d(ts), ev_o(ts))):
        if k == 1:
            cfg = solver.autotune(ev_d.states[0], reps=2)
            say("autotuned: %s" % cfg)
        worst = max(worst, float(np.max(np.abs(st.states[0].to_numpy() - Po))))
    assert worst <= TOL


class _Shuffled(lmhip.operators.Hamiltonian):
    """The same Hamiltonian with its sites renumbered by perm (new index of
    site i: perm[i]); positions, so Peierls phases, are unchanged."""

    def __init__(self, H, perm):
        self.H, self.perm = H, np.asarray(perm, dtype=np.int64)
        self.internal = 1
        inv = np.argsort(self.perm)
        self.onsite = None if H.onsite is None else np.asarray(H.onsite)[inv]
        self._coords = np.asarray(H.coords)[inv]

    @property
    def n(self):
        return self.H.n

    @property
    def coords(self):
        return self._coords

    def geometry(self):
        b, amp = self.H.geometry()
        return lmhip.lattice.Bonds(self.perm[b.i], self.perm[b.j], b.ri, b.rj, b.twist), amp

    def geometry_key(self):
        return ("shuffled", id(self))

    def field_params(self):
        return self.H.field_params()


SHUFFLE_MODELS = {
    "square_pbc_24x24_landau_disorder": (
        lambda: Model("square", (24, 24), (True, True), onsite=np.random.default_rng(5).uniform(-0.5, 0.5, 576)),
        [("landau", 0.011, 0, 0)]),
    "haldane_open_10x10_pointflux": (lambda: Model("honeycomb", (10, 10), t1=1.0, t2=0.1, m=0.2),
                                     [("axial", 0.2, 4.5, 4.0)]),
}


@pytest.mark.parametrize("mode,left_pass,right_pass", [("fusedT", None, None), ("fusedN", None, None),
                                                       ("perpass", 0, 1), ("perpass", 1, 3), ("perpass", 0, 4),
                                                       ("perpass", 1, 5)])
@pytest.mark.parametrize("name", sorted(SHUFFLE_MODELS))
def test_shuffled_site_order(name, mode, left_pass, right_pass):
    """The sites of a lattice renumbered at random (so H's band is of the
    order of N and no kernel can rely on lattice order): the device evolution
    of the renumbered problem is the renumbered oracle evolution, after every
    step, for the fused and the per-term schedules."""
    make, spec = SHUFFLE_MODELS[name]
    m = make()
    n = m.n
    rng = np.random.default_rng(29)
    Q = np.linalg.qr(rng.normal(size=(n, 40)) + 1j * rng.normal(size=(n, 40)))[0]
    P0 = Q @ Q.conj().T
    dt, nsteps = 0.05, 3
    U, _K, _s, _d = ON.myexp(m.oracle_h(spec), dt)
    refs, P = [], P0
    for _ in range(nsteps):
        P = ON.step(U, P.copy())
        refs.append(P)
    perm = np.random.default_rng(3).permutation(n)
    inv = np.argsort(perm)
    H = _Shuffled(m.device_h(spec), perm)
    modes = {"fusedT": lmhip._lib.STEP_FUSED_T, "fusedN": lmhip._lib.STEP_FUSED_N,
             "perpass": lmhip._lib.STEP_PER_PASS}
    solver = lmhip.HIPExp(step_mode=modes[mode])
    st = solver.make_state(P0[np.ix_(inv, inv)], H)
    solver.update_solver(H, dt)
    L = lmhip._lib.lib()
    if left_pass is not None:
        lmhip._lib.check(L.lmh_set_option(solver.ctx.h, lmhip._lib.OPT_LEFT_PASS, left_pass))
        lmhip._lib.check(L.lmh_set_option(solver.ctx.h, lmhip._lib.OPT_RIGHT_PASS, right_pass))
    say("case %s mode %s left_pass %s right_pass %s" % (name, mode, left_pass, right_pass))
    for k in range(nsteps):
        solver.step(st)
        d = float(np.max(np.abs(st.to_numpy() - refs[k][np.ix_(inv, inv)])))
        assert d <= TOL, (k, d)


@pytest.mark.parametrize("name", ["square_pbc_landau_disorder", "haldane_pbc_landau"])
def test_every_fused_variant(name):
    """Each entry of the fused kernel's variant table that fits the lattice
    (threads x sites, columns per workgroup, H slots in registers, x in
    registers, compile-time slot chunks) gives the oracle's P after every
    step, from a non-stationary pure state."""
    make, spec = MODELS[name]
    m = make()
    n = m.n
    rng = np.random.default_rng(17)
    Q = np.linalg.qr(rng.normal(size=(n, n // 3)) + 1j * rng.normal(size=(n, n // 3)))[0]
    P0 = Q @ Q.conj().T
    dt, nsteps = 0.05, 5
    Ho = m.oracle_h(spec)
    U, _K, _s, _d = ON.myexp(Ho, dt)
    refs, P = [], P0
    for _ in range(nsteps):
        P = ON.step(U, P.copy())
        refs.append(P)
    solver = lmhip.HIPExp(step_mode=lmhip._lib.STEP_FUSED_T)
    H = m.device_h(spec)
    st = solver.make_state(P0, H)
    solver.update_solver(H, dt)
    L = lmhip._lib.lib()
    tried = 0
    for v in range(64):
        rc = L.lmh_set_option(solver.ctx.h, lmhip._lib.OPT_FUSED_VARIANT, v)
        if rc == lmhip._lib.LMH_EINVAL:
            break                                  # past the end of the table
        if rc == lmhip._lib.LMH_ERANGE:
            continue                               # does not fit this lattice
        lmhip._lib.check(rc)
        say("case %s variant %d" % (name, v))
        x = st.copy()
        for k in range(nsteps):
            solver.step(x)
            d = float(np.max(np.abs(x.to_numpy() - refs[k])))
            assert d <= TOL, (v, k, d)
        tried += 1
    lmhip._lib.check(L.lmh_set_option(solver.ctx.h, lmhip._lib.OPT_FUSED_VARIANT, -1))
    assert tried >= (20 if m.kind == "square" else 10), tried


SCHEDULE_MODELS = {
    # several 1024-row blocks of the banded pass, the last one partial
    "square_open_48x48": (lambda: Model("square", (48, 48), onsite=np.random.default_rng(2).uniform(-0.5, 0.5, 2304)),
                          [("axial", 0.3, 23.5, 20.5)]),
    # zmax = 9 (Haldane), band 2 * 20 + 2 rows
    "haldane_open_20x20": (lambda: Model("honeycomb", (20, 20), t1=1.0, t2=0.1, m=0.2), [("landau", 0.01, 0, 0)]),
    # periodic: the wrap-around bonds put H's band above 1024 rows (gather only)
    "square_pbc_40x40": (lambda: Model("square", (40, 40), (True, True)), [("landau", 0.02, 0, 0)]),
}


@pytest.mark.parametrize("left_pass,right_pass", [(0, 1), (1, 1), (2, 1), (0, 3), (1, 4), (0, 5), (1, 5),
                                                   (3, 1), (3, 3), (3, 5)])
@pytest.mark.parametrize("name", sorted(SCHEDULE_MODELS))
def test_per_term_schedules(name, left_pass, right_pass):
    """Per-term steps with each left-pass kernel (gather / banded / banded
    with two terms per launch) and right side (left passes on the transpose,
    with the last term transposed in place or by the transpose kernel; column
    sweeps) give the oracle's U P U' for a non-Hermitian P (so the transposed
    schedules are exact, not relying on P = P'), at two time steps whose
    orders K differ (odd and even term counts for the pairing)."""
    make, spec = SCHEDULE_MODELS[name]
    m = make()
    n = m.n
    rng = np.random.default_rng(5)
    P0 = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / n
    dt, nsteps = 0.04, 2
    solver = lmhip.HIPExp(step_mode=lmhip._lib.STEP_PER_PASS)
    H = m.device_h(spec)
    st = solver.make_state(P0, H)
    solver.update_solver(H, dt)
    L = lmhip._lib.lib()
    rc = L.lmh_set_option(solver.ctx.h, lmhip._lib.OPT_LEFT_PASS, left_pass)
    if rc == lmhip._lib.LMH_ERANGE:
        assert name == "square_pbc_40x40" and left_pass in (2, 3)
        pytest.skip("band above 1024 rows: the banded passes do not apply (rejected as they should be)")
    lmhip._lib.check(rc)
    lmhip._lib.check(L.lmh_set_option(solver.ctx.h, lmhip._lib.OPT_RIGHT_PASS, right_pass))
    say("case %s left_pass %d right_pass %d" % (name, left_pass, right_pass))
    P = P0.copy()
    Ks = []
    for dt_k in (dt, 0.07):
        U, _K, _s, _d = ON.myexp(m.oracle_h(spec), dt_k)
        solver.update_solver(H, dt_k)
        Ks.append(solver.K)
        for k in range(nsteps):
            solver.step(st)
            P = ON.step(U, P)
            d = float(np.max(np.abs(st.to_numpy() - P)))
            assert d <= TOL, (dt_k, k, d)
    assert Ks[0] % 2 != Ks[1] % 2, Ks


def test_adjacency_matrix_hamiltonian(gpu_ctx):
    """Hopping over an AdjacencyMatrix (src/core/bonds.jl:164-345): a periodic
    12 x 12 square lattice with the SiteDistance(d <= 1.5) bonds (nearest and
    diagonal neighbours, nearest images) and 20 of them removed, on-site
    disorder and a Landau ramp.  Device H entries against the oracle's
    (1e-12), then P after every step of an evolution (1e-10)."""
    from oracle import fields as OF
    from oracle import geometry as OG
    from oracle import hamiltonian as OH
    lo = OG.SquareLattice(12, 12, periodic=(True, True))
    lm = lmhip.SquareLattice(12, 12, boundaries=(("axis1", True), ("axis2", True)))
    am_o = OG.AdjacencyMatrix(lo, OG.SiteDistance(lambda d: d <= 1.5))
    am_m = lmhip.AdjacencyMatrix(lm, lmhip.SiteDistance(lambda d: d <= 1.5))
    rng = np.random.default_rng(11)
    pairs = sorted(am_o.pairs)
    for q in rng.choice(len(pairs), 20, replace=False):
        i, j = pairs[q]
        am_o.set(i, j, False)
        am_m[i, j] = False
    V = rng.uniform(-0.5, 0.5, 144)
    bt = lambda t: 0.03 * min(t, 0.2) / 0.2
    ho = lambda t: OH.OracleHamiltonian(lo, [OH.Onsite(V), OH.Hopping(0.9, [am_o])], OF.LandauGauge(bt(t)))
    hd = lambda t: lmhip.construct_hamiltonian(lm, V, (0.9, am_m), field=lmhip.LandauGauge(bt(t)))
    say("case adjacency entries")
    Ho, Hd = ho(0.3), hd(0.3)
    diag, hij = gpu_ctx.hamiltonian_entries(Hd)
    b = Hd.bonds
    ref = np.array([Ho.entries[(int(i), int(j))] for i, j in zip(b.i, b.j)])
    assert len(b) == len(am_o.pairs)
    assert np.max(np.abs(hij - ref)) <= 1e-12
    assert np.max(np.abs(diag - V)) <= 1e-15
    say("case adjacency evolution")
    P0, _ = ON.densitymatrix(ho(0.0).dense(), mu=0.0)
    ts = [0.05 * k for k in range(9)]
    ev_d = lmhip.Evolution(hd, P0)
    ev_o = ON.Evolution(ho, P0)
    for k, (st, (_H, Po, _t)) in enumerate(zip(ev_d(ts), ev_o(ts))):
        d = float(np.max(np.abs(st.states[0].to_numpy() - Po)))
        assert d <= TOL, (k, d)


def test_schedule_options_follow_geometry_changes():
    """A fused variant pinned for a nearest-neighbour lattice (packed site
    records, zmax <= 4) reverts to the by-size default when a geometry with
    next-nearest bonds (zmax 8) replaces it in the same context -- schedule
    options choose kernels, never results -- and the evolution still matches
    the oracle after every step."""
    ctx = lmhip.Context(0)
    spec = [("landau", 0.02, 0, 0)]
    ma = Model("square", (12, 12), (True, True))
    mb = Model("square", (12, 12), (True, True), t2=0.3)
    solver = lmhip.HIPExp(ctx=ctx, step_mode=lmhip._lib.STEP_FUSED_T)
    solver.update_solver(ma.device_h(spec), 0.05)
    L = lmhip.lib()
    lmhip._lib.check(L.lmh_set_option(ctx.h, lmhip._lib.OPT_FUSED_VARIANT, 16))
    say("case options across geometries")
    rng = np.random.default_rng(3)
    Q = np.linalg.qr(rng.normal(size=(144, 30)) + 1j * rng.normal(size=(144, 30)))[0]
    P0 = Q @ Q.conj().T
    st = solver.make_state(P0, mb.device_h(spec))
    solver.update_solver(mb.device_h(spec), 0.05)
    v = ctypes.c_int64()
    lmhip._lib.check(L.lmh_get_option(ctx.h, lmhip._lib.OPT_FUSED_VARIANT, ctypes.byref(v)))
    assert v.value == -1
    U, _K, _s, _d = ON.myexp(mb.oracle_h(spec), 0.05)
    P = P0
    for k in range(3):
        solver.step(st)
        P = ON.step(U, P.copy())
        assert float(np.max(np.abs(st.to_numpy() - P))) <= TOL, k
