"""GPU parity of the region current sums (currentsfrom / currentsfromto,
src/currents.jl:76-110) and of TimeSequence recording on the device
(src/timesequence.jl: differentiate!, integrate!) against the oracle."""
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


def _random_projector(n, k, seed):
    rng = np.random.default_rng(seed)
    Q = np.linalg.qr(rng.normal(size=(n, k)) + 1j * rng.normal(size=(n, k)))[0]
    return Q @ Q.conj().T


@pytest.mark.parametrize("name", ["square_open_flux", "haldane_pbc"])
def test_region_currents_match_oracle(name):
    if name == "square_open_flux":
        m, spec = Model("square", (9, 7)), [("axial", 0.23, 4.5, 3.5)]
    else:
        m, spec = Model("honeycomb", (6, 6), (True, True), t1=1.0, t2=0.1, m=0.2), [("landau", 0.02, 0, 0)]
    say("case region currents %s" % name)
    n = m.n
    P0 = _random_projector(n, n // 4, 1)
    Ho, Hd = m.oracle_h(spec), m.device_h(spec)
    ctx = lmhip.Context(0)
    ctx.set_hamiltonian(Hd)
    P = lmhip.DensityMatrix.from_array(ctx, P0)
    curr = lmhip.DensityCurrents(Hd, P)
    rng = np.random.default_rng(2)
    src = np.sort(rng.choice(n, n // 5, replace=False))
    dst = np.sort(rng.choice(np.setdiff1d(np.arange(n), src), n // 4, replace=False))
    out = lmhip.currentsfrom(curr, src)
    assert np.max(np.abs(out - ON.currentsfrom(Ho, P0, src))) <= TOL
    mask = np.zeros(n, dtype=bool)
    mask[src] = True
    assert np.max(np.abs(lmhip.currentsfrom(curr, mask) - out)) == 0.0
    assert abs(lmhip.currentsfromto(curr, src) - ON.currentsfromto(Ho, P0, src)) <= TOL
    assert abs(lmhip.currentsfromto(curr, src, dst) - ON.currentsfromto(Ho, P0, src, dst)) <= TOL
    # overlapping regions: pairs inside both cancel as in the reference's sum
    both = np.union1d(src, dst[:3])
    assert abs(lmhip.currentsfromto(curr, src, both) - ON.currentsfromto(Ho, P0, src, both)) <= TOL
    ctx.close()


def test_series_recording_differentiate_integrate():
    """Density, bond currents and a region current recorded on the device at
    every step of an evolution agree with the oracle's evolution; the device
    differentiate! / integrate! agree with the oracle's restatement and with
    the host TimeSequence; the continuity equation holds to O(dt^2)."""
    m = Model("square", (8, 8), onsite=np.random.default_rng(5).uniform(-0.5, 0.5, 64))
    tau = 1.0
    fs = lambda t: [("axial", 0.3 * min(t, tau) / tau, 3.5, 4.5)]
    P0 = _random_projector(64, 20, 3)
    dt, nsteps = 0.05, 40
    ts = [dt * k for k in range(nsteps + 1)]
    src = np.array([0, 1, 8, 9, 16, 17, 24, 25])
    say("case series recording")
    ev_d = lmhip.Evolution(lambda t: m.device_h(fs(t)), P0)
    ev_o = ON.Evolution(lambda t: m.oracle_h(fs(t)), P0)
    it_d, it_o = iter(ev_d(ts)), iter(ev_o(ts))
    first = next(it_d)
    next(it_o)
    ctx = ev_d.solver.ctx
    sd = lmhip.DeviceSeries(ctx, "density", nsteps + 1)
    sc = lmhip.DeviceSeries(ctx, "currents", nsteps + 1)
    sr = lmhip.DeviceSeries(ctx, "region", nsteps + 1, src=src)
    for s in (sd, sc, sr):
        s.record(first.states[0], ts[0])
    rho_o, J_o, R_o = [np.real(np.diag(P0)).copy()], [], []
    Ho0 = m.oracle_h(fs(0.0))
    J_o.append(ON.density_currents(Ho0, P0))
    R_o.append(ON.currentsfromto(Ho0, P0, src))
    for k in range(1, nsteps + 1):
        tsd = next(it_d)
        Ho, Po, _t = next(it_o)
        for s in (sd, sc, sr):
            s.record(tsd.states[0], ts[k])
        rho_o.append(np.real(np.diag(Po)).copy())
        J_o.append(ON.density_currents(Ho, Po))
        R_o.append(ON.currentsfromto(Ho, Po, src))
    td = sd.to_timesequence()
    assert len(td) == nsteps + 1 and td.times == ts
    assert max(np.max(np.abs(v - r)) for v, r in zip(td.values, rho_o)) <= TOL
    tc = sc.to_timesequence()
    assert max(np.max(np.abs(v - r)) for v, r in zip(tc.values, J_o)) <= TOL
    tr = sr.to_timesequence()
    assert max(abs(v - r) for v, r in zip(tr.values, R_o)) <= TOL
    # continuity: d/dt of the region's density = the current into it (checked
    # once the field is constant: a recorded current uses the H of the step
    # that led to it, which lags by one step while the field ramps)
    dens = [float(np.sum(v[src])) for v in td.values]
    _tm, dd = ON.differentiate(ts, dens)
    mid = [0.5 * (a + b) for a, b in zip(tr.values[:-1], tr.values[1:])]
    late = [i for i in range(nsteps) if ts[i] > tau + dt]
    assert len(late) > 10
    assert np.max(np.abs(np.array(dd)[late] - np.array(mid)[late])) <= 5 * dt * dt
    # device differentiate! / integrate! against the oracle and the host TimeSequence
    ref_t, ref_v = ON.differentiate(ts, td.values)
    host = td.differentiate()
    sd.differentiate_()
    dev = sd.to_timesequence()
    assert dev.times == host.times
    assert np.max(np.abs(np.array(dev.times) - np.array(ref_t))) == 0.0
    scale = max(1.0, max(np.max(np.abs(v)) for v in ref_v))
    assert max(np.max(np.abs(a - b)) for a, b in zip(dev.values, ref_v)) <= 1e-12 * scale
    assert max(np.max(np.abs(a - b)) for a, b in zip(dev.values, host.values)) <= 1e-12 * scale
    ri_t, ri_v = ON.integrate(tr.times, tr.values)
    sr.integrate_()
    dev_i = sr.to_timesequence()
    assert dev_i.times == ri_t
    assert max(abs(a - float(b)) for a, b in zip(dev_i.values, ri_v)) <= 1e-12
    assert max(abs(a - b) for a, b in zip(dev_i.values, tr.integrate().values)) <= 1e-12
    # errors: times must increase, the capacity is fixed, short series
    with pytest.raises(ValueError):
        sc.record(ev_d.states[0], ts[-1])
    with pytest.raises(lmhip._lib.LMHError):
        sc.record(ev_d.states[0], ts[-1] + 1.0)
    short = lmhip.DeviceSeries(ctx, "region", 4, src=src)
    short.record(ev_d.states[0], 0.0)
    with pytest.raises(lmhip._lib.LMHError):
        short.differentiate_()
    for s in (sd, sc, sr, short):
        s.close()
