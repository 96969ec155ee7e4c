"""GPU parity of the batched evolution (many disorder realisations of uneven
lattice sizes advanced together, three-step HIP graphs replayed): the
persistent strip launch (one per side over all members) and the per-size-class
fused launches (LMH_OPT_BATCH_STRIP 0).

Every member is checked against (a) the same member evolved on its own with
lmh_step and (b) for small members, the oracle's Evolution; both after every
graph replay, with eager device work (downloads, per-member steps, trace
kernels) between replays.  Bars: 1e-10 against the oracle, 1e-12 against the
single-member device path (same arithmetic, same kernels family).
"""
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


def _members(sizes, W=2.0, B=0.005, seed=100):
    out = []
    for r, L in enumerate(sizes):
        rng = np.random.default_rng(seed + r)
        m = Model("square", (L, L), onsite=rng.uniform(-W / 2, W / 2, L * L))
        out.append(m)
    return out, [("landau", B, 0, 0)]


def _device_p0(m, spec):
    H = m.device_h(spec)
    return lmhip.densitymatrix(H, mu=0.0, ctx=lmhip.Context(0)).to_numpy()


@pytest.mark.parametrize("strip", [True, False], ids=["strip", "classes"])
@pytest.mark.parametrize("sizes", [
    [6, 9, 7, 12, 5],                 # one size class (N <= 1024), ragged
    [8, 33, 5, 40, 20, 47, 11],       # N = 64 .. 2209: three classes
])
def test_batched_matches_single_and_oracle(sizes, strip):
    say("case batched sizes=%s strip=%s" % (sizes, strip))
    dt = 0.05
    models, spec = _members(sizes)
    P0s = [_device_p0(m, spec) for m in models]
    batch = lmhip.BatchedEvolution([(m.device_h(spec), P0) for m, P0 in zip(models, P0s)], dt, strip=strip)
    kernels = batch.member_kernels()
    if strip:
        assert kernels == [lmhip._lib.FUSED_STRIP] * len(sizes), kernels
    else:
        assert lmhip._lib.FUSED_STRIP not in kernels, kernels
    singles = []
    for m, P0 in zip(models, P0s):
        ev = lmhip.Evolution(lambda t, m=m: m.device_h(spec), P0)
        singles.append(ev)
    oracles = {r: ON.Evolution(lambda t, m=m: m.oracle_h(spec), P0)
               for r, (m, P0) in enumerate(zip(models, P0s)) if m.n <= 144}
    # schedule: eager single steps, graph replays of 3, remainders; each call
    # is followed by eager checks (downloads + single-member steps)
    plan = [(1, False), (3, True), (3, True), (5, True), (2, False), (6, True), (3, True)]
    done = 0
    worst_single = worst_oracle = 0.0
    for nsteps, graph in plan:
        batch.step(nsteps, graph=graph)
        ts = [dt * k for k in range(done, done + nsteps + 1)]
        done += nsteps
        for r, ev in enumerate(singles):
            for _ in ev(ts):
                pass
            Pb = batch.states[r].to_numpy()
            Ps = ev.states[0].to_numpy()
            d = float(np.max(np.abs(Pb - Ps)))
            worst_single = max(worst_single, d)
            assert d <= 1e-12, (r, done, d)
            if r in oracles:
                for _ in oracles[r](ts):
                    pass
                do = float(np.max(np.abs(Pb - oracles[r].P)))
                worst_oracle = max(worst_oracle, do)
                assert do <= TOL, (r, done, do)
        md = batch.mean_density()
        want = np.array([np.trace(s.to_numpy()).real / s.n for s in batch.states])
        assert np.max(np.abs(md - want)) <= 1e-13
    say("batched worst vs single %.2e, vs oracle %.2e" % (worst_single, worst_oracle))
    batch.close()


def test_batch_rejects_changed_member():
    models, spec = _members([5, 6])
    P0s = [_device_p0(m, spec) for m in models]
    batch = lmhip.BatchedEvolution([(m.device_h(spec), P0) for m, P0 in zip(models, P0s)], 0.05)
    batch.step(2)
    sol = batch.solvers[0]
    sol.step(batch.states[0])          # stepping a member outside the batch
    with pytest.raises(lmhip.LMHError):
        batch.step(1)
    batch.close()


@pytest.mark.parametrize("strip", [True, False], ids=["strip", "classes"])
def test_batched_value_class_variants(strip):
    """With the per-class launches, a member large enough for the value-class
    variant (N = 5184: ten sites per thread, one column per workgroup, variant
    29) batched with two of the two-column class (N = 3249, 3600, plain
    records); with the strip launch, all three in it.  Each matches its
    single-matrix evolution after eager steps and graph replays."""
    sizes = [57, 60, 72]
    say("case batched value classes sizes=%s strip=%s" % (sizes, strip))
    dt = 0.05
    models, spec = _members(sizes)
    rng = np.random.default_rng(9)
    P0s = []
    for m in models:
        Q = np.linalg.qr(rng.normal(size=(m.n, 16)) + 1j * rng.normal(size=(m.n, 16)))[0]
        P0s.append(Q @ Q.conj().T)
    batch = lmhip.BatchedEvolution([(m.device_h(spec), P0) for m, P0 in zip(models, P0s)], dt, strip=strip)
    kernels = batch.member_kernels()
    if strip:
        assert kernels == [lmhip._lib.FUSED_STRIP] * 3, kernels
    else:
        assert kernels[2] == 29 and 29 not in kernels[:2], kernels
    singles = [lmhip.Evolution(lambda t, m=m: m.device_h(spec), P0) for m, P0 in zip(models, P0s)]
    done = 0
    for nsteps, graph in [(1, False), (3, True), (2, False)]:
        batch.step(nsteps, graph=graph)
        ts = [dt * k for k in range(done, done + nsteps + 1)]
        done += nsteps
        for r, ev in enumerate(singles):
            for _ in ev(ts):
                pass
            d = float(np.max(np.abs(batch.states[r].to_numpy() - ev.states[0].to_numpy())))
            assert d <= 1e-12, (r, done, d)
    batch.close()


@pytest.mark.parametrize("strip", [True, False], ids=["strip", "classes"])
def test_batch_time_dependent_members(strip):
    """Members whose H depends on time (a Landau field ramped at a rate of
    its own per member, with the member's disorder), beside one with a fixed
    H: every step evaluates H at the step's start (src/evolution.jl:238-250),
    as the oracle's Evolution does; eager steps and graph replays."""
    sizes = [6, 9, 7, 12]
    say("case batched time-dependent sizes=%s strip=%s" % (sizes, strip))
    dt = 0.05
    models, _spec = _members(sizes)
    spec_of = lambda r, t: [("landau", 0.004 + 0.01 * (r + 1) * t, 0, 0)]
    fixed = 2                                   # this member's H does not change
    rng = np.random.default_rng(31)
    P0s = []
    for m in models:
        Q = np.linalg.qr(rng.normal(size=(m.n, 6)) + 1j * rng.normal(size=(m.n, 6)))[0]
        P0s.append(Q @ Q.conj().T)
    members = []
    for r, (m, P0) in enumerate(zip(models, P0s)):
        if r == fixed:
            members.append((m.device_h(spec_of(r, 0.0)), P0))
        else:
            members.append((lambda t, m=m, r=r: m.device_h(spec_of(r, t)), P0))
    batch = lmhip.BatchedEvolution(members, dt, strip=strip)
    kernels = batch.member_kernels()
    assert (lmhip._lib.FUSED_STRIP in kernels) == strip, kernels
    refs = [ON.Evolution((lambda t, m=m, r=r: m.oracle_h(spec_of(r, 0.0 if r == fixed else t))), P0)
            for r, (m, P0) in enumerate(zip(models, P0s))]
    done = 0
    for nsteps, graph in [(1, False), (3, True), (2, False), (3, True)]:
        batch.step(nsteps, graph=graph)
        ts = [dt * k for k in range(done, done + nsteps + 1)]
        done += nsteps
        for r, ev in enumerate(refs):
            for (_H, Po, _t) in ev(ts):
                pass
            err = np.max(np.abs(batch.states[r].to_numpy() - Po))
            assert err <= TOL, (r, done, err)
    assert abs(batch.time - done * dt) < 1e-12
    batch.close()


def test_batch_member_leaving_strip_form_is_refused():
    """A member whose H leaves the strip form (a point flux switched on from
    zero) after the batch chose the strip launch for it: LMH_ESTATE, with the
    states left as they were after the last step."""
    models, _spec = _members([6, 8])
    dt = 0.05
    specs = [lambda t: [("landau", 0.01, 0, 0)], lambda t: [("axial", 0.5 * t, 3.5, 3.5)]]
    rng = np.random.default_rng(5)
    P0s = [np.eye(m.n)[:, :3] @ np.eye(m.n)[:3, :] + 0.0j for m in models]
    members = [(lambda t, m=m, f=f: m.device_h(f(t)), P0) for m, f, P0 in zip(models, specs, P0s)]
    batch = lmhip.BatchedEvolution(members, dt, strip=True)
    assert batch.member_kernels() == [lmhip._lib.FUSED_STRIP] * 2
    batch.step(1)
    refs = [ON.Evolution(lambda t, m=m, f=f: m.oracle_h(f(t)), P0) for m, f, P0 in zip(models, specs, P0s)]
    after = [list(ev([0.0, dt]))[-1][1] for ev in refs]
    with pytest.raises(lmhip.LMHError):
        batch.step(1)
    for r in range(2):
        assert np.max(np.abs(batch.states[r].to_numpy() - after[r])) <= TOL, r
    batch.close()


def test_batch_refuses_taped_member():
    """A member whose state is being recorded by an adjoint tape: refused by
    lmh_batch_create, and by lmh_batch_step when the tape begins after the
    batch was made (the tape records only lmh_step)."""
    import ctypes
    from lmhip import _lib
    L_ = _lib.lib()
    models, spec = _members([5, 6])
    P0s = [_device_p0(m, spec) for m in models]
    sols, sts = [], []
    for m, P0 in zip(models, P0s):
        sol = lmhip.HIPExp(ctx=lmhip.Context(0))
        st = sol.make_state(P0, m.device_h(spec))
        sol.update_solver(m.device_h(spec), 0.05)
        sols.append(sol)
        sts.append(st)
    ctxs = (ctypes.c_void_p * 2)(*[s.ctx.h.value for s in sols])
    hs = (ctypes.c_void_p * 2)(*[s.h.value for s in sts])
    b = ctypes.c_void_p()
    _lib.check(L_.lmh_tape_begin(sols[1].ctx.h, sts[1].h, 1))
    assert L_.lmh_batch_create(ctypes.byref(b), 2, ctxs, hs) == _lib.LMH_ESTATE
    _lib.check(L_.lmh_tape_end(sols[1].ctx.h))
    _lib.check(L_.lmh_batch_create(ctypes.byref(b), 2, ctxs, hs))
    try:
        _lib.check(L_.lmh_batch_step(b, 1, 0))
        _lib.check(L_.lmh_tape_begin(sols[0].ctx.h, sts[0].h, 1))
        assert L_.lmh_batch_step(b, 1, 0) == _lib.LMH_ESTATE
        _lib.check(L_.lmh_tape_end(sols[0].ctx.h))
        _lib.check(L_.lmh_batch_step(b, 1, 0))
    finally:
        L_.lmh_batch_destroy(b)
