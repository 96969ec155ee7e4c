"""GPU: the Hamiltonian handed over as a sparse matrix -- the
SparseMatrixCSC{ComplexF64} that the reference's own builders produce and
that its Evolution passes to update_solver! (src/evolution.jl:244,
src/operators/builder.jl:317-318, src/operators/system.jl:380-386) -- through
lmh_set_csc_pattern / lmh_set_csc_values.

* The small workload (SquareLattice 10x10, open, PointFlux ramp through
  (5.5, 5.5), P0 the golden Fermi projector, 200 steps of 0.1) with h(t)
  returning the oracle builder's CSC matrix: P against the oracle's
  evolution after every step (1e-10), K and densities against the golden
  file, bond currents 2 Im(H_ij P_ji) on the pattern's bonds (1e-10).
* The reference's cache rule (src/evolution.jl:83-92): the same or an equal
  matrix with dt ≈ the previous dt builds nothing; new values, a new dt or
  force do.
* Index base 1 (Julia's arrays as they are) gives the same device H as base 0.
* The N = 10^4 lattice (periodic, Landau ramp, disorder) through CSC matrices
  runs the strip kernel and matches a factored rank-8 reference evolved by the
  oracle's sparse U.
* Error paths: non-Hermitian values, an unsymmetric pattern, complex
  diagonal, size mismatches, calls out of order.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import lmhip
from helpers import Model
from lmhip import _lib
from oracle import numerics as ON

pytestmark = pytest.mark.gpu

TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sparse_of(Ho):
    c = Ho.csc
    return lmhip.SparseHamiltonian(Ho.n, c.colptr, c.rowval, c.nzval)


def currents_on(Hs, Hd, P):
    i, j = Hs.bond_ends()
    return 2.0 * np.imag(Hd[i, j] * P[j, i])


def _profile(ctx, kind):
    nl, ms = ctypes.c_int64(), ctypes.c_double()
    _lib.check(_lib.lib().lmh_get_profile(ctx.h, kind, ctypes.byref(nl), ctypes.byref(ms)))
    return nl.value


def test_small_workload_through_csc_every_step():
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cfg1.npz"))
    m = Model("square", (10, 10))
    tau = 10.0
    fspec = lambda t: [("axial", 0.2 * min(t, tau) / tau, 5.5, 5.5)]
    ts = list(g["ts"])
    say("case csc small workload N=100 steps=%d" % (len(ts) - 1))
    ev_d = lmhip.Evolution(lambda t: sparse_of(m.oracle_h(fspec(t))), g["P0"])
    ev_o = ON.Evolution(lambda t: m.oracle_h(fspec(t)), g["P0"])
    worst = 0.0
    for k, (tsd, (Ho, Po, to)) in enumerate(zip(ev_d(ts), ev_o(ts))):
        P_d, H_d, t_d = tsd
        if k == 0:
            continue
        assert isinstance(H_d, lmhip.SparseHamiltonian)
        rho, J = lmhip.observables(P_d, H_d)
        assert ev_d.solver.K == g["K"][k - 1], k
        assert np.max(np.abs(rho - g["rho"][k - 1])) <= TOL, k
        Pg = P_d.to_numpy()
        d = float(np.max(np.abs(Pg - Po)))
        worst = max(worst, d)
        assert d <= TOL, (k, d)
        assert np.max(np.abs(J - currents_on(H_d, Ho.dense(), Po))) <= TOL, k
    assert np.max(np.abs(ev_d.states[0].to_numpy() - g["P_final"])) <= TOL
    say("csc small workload worst |dP| = %.3e" % worst)


def test_cache_rule_of_the_reference():
    m = Model("square", (8, 8), (True, True), onsite=np.random.default_rng(5).uniform(-0.5, 0.5, 64))
    Ho = m.oracle_h([("landau", 0.02, 0, 0)])
    H = sparse_of(Ho)
    rng = np.random.default_rng(6)
    Q = np.linalg.qr(rng.normal(size=(64, 9)) + 1j * rng.normal(size=(64, 9)))[0]
    P0 = Q @ Q.conj().T
    solver = lmhip.HIPExp(ctx=lmhip.Context(0))
    st = solver.make_state(P0, H)
    ctx = solver.ctx
    _lib.check(_lib.lib().lmh_set_profiling(ctx.h, 1))
    builds = lambda: _profile(ctx, _lib.KERNEL_BUILD_H)
    orders = lambda: _profile(ctx, _lib.KERNEL_TAYLOR_ORDER)
    solver.update_solver(H, 0.05)
    b0, o0 = builds(), orders()
    assert b0 >= 1 and o0 >= 1
    solver.update_solver(H, 0.05)                                   # same object
    solver.update_solver(sparse_of(Ho), 0.05 * (1 + 1e-12))        # equal matrix, dt ≈
    assert (builds(), orders()) == (b0, o0)
    solver.update_solver(H, 0.05, force=True)                      # force
    assert builds() == b0 + 1 and orders() > o0
    b1, o1 = builds(), orders()
    solver.update_solver(H, 0.1)                                    # new dt: new U, same H
    assert builds() == b1 and orders() > o1
    H2 = sparse_of(m.oracle_h([("landau", 0.03, 0, 0)]))           # new values
    solver.update_solver(H2, 0.1)
    assert builds() == b1 + 1
    # the step uses H2: compare with the oracle
    solver.step(st)
    U = ON.myexp(m.oracle_h([("landau", 0.03, 0, 0)]), 0.1)[0]
    ref = ON.apply(U, ON.apply(U, P0).conj().T).conj().T
    assert np.max(np.abs(st.to_numpy() - ref)) <= TOL


def test_index_base_one_matches_base_zero():
    m = Model("honeycomb", (4, 5), t1=1.0, t2=0.1, m=0.2, onsite=np.random.default_rng(8).normal(0, 0.1, 40))
    Hs = sparse_of(m.oracle_h([("axial", 0.13, 2.75, 2.2)]))
    L_ = _lib.lib()
    out = []
    for base in (0, 1):
        ctx = lmhip.Context(0)
        cp = np.ascontiguousarray(Hs.colptr + base)
        rv = np.ascontiguousarray(Hs.rowval + base)
        _lib.check(L_.lmh_set_csc_pattern(ctx.h, Hs.n, _lib.ptr(cp), _lib.ptr(rv), base))
        _lib.check(L_.lmh_set_csc_values(ctx.h, Hs.nnz, _lib.ptr(Hs.nzval)))
        K = ctypes.c_int32()
        s = ctypes.c_double()
        _lib.check(L_.lmh_update_solver(ctx.h, 0.05, 0, 1e-10, 1e-14, ctypes.byref(K), ctypes.byref(s)))
        i, j = Hs.bond_ends()
        diag = np.zeros(Hs.n)
        hij = np.zeros(len(i), dtype=np.complex128)
        _lib.check(L_.lmh_get_hamiltonian(ctx.h, _lib.ptr(diag), _lib.ptr(hij)))
        out.append((K.value, diag, hij))
        ctx.close()
    Hd = m.oracle_h([("axial", 0.13, 2.75, 2.2)]).dense()
    i, j = Hs.bond_ends()
    for K, diag, hij in out:
        assert K == out[0][0]
        assert np.max(np.abs(diag - np.real(np.diag(Hd)))) <= 1e-14
        assert np.max(np.abs(hij - Hd[i, j])) <= 1e-14


def test_n10000_through_csc_runs_the_strip_kernel():
    L, tau, bmax, dt = 100, 10.0, 0.01, 0.05
    rng = np.random.default_rng(123)
    V = rng.uniform(-0.5, 0.5, L * L)
    m = Model("square", (L, L), (True, True), onsite=V)
    Q = np.linalg.qr(rng.normal(size=(L * L, 8)) + 1j * rng.normal(size=(L * L, 8)))[0]
    fs = lambda t: [("landau", bmax * min(t, tau) / tau, 0, 0)]
    t0 = 4.0
    ts = [t0 + dt * k for k in range(4)]
    hs = {}
    for t in ts[:-1]:
        hs[t] = m.oracle_h(fs(t))
    say("case csc n10000 steps=%d" % (len(ts) - 1))
    ev = lmhip.Evolution(lambda t: sparse_of(hs[min(hs, key=lambda u: abs(u - t))]), Q @ Q.conj().T)
    ev.time = t0
    Vm = np.asfortranarray(Q)
    launched = []
    for k, st in enumerate(ev(ts)):
        if k == 0:
            continue
        v = ctypes.c_int64()
        _lib.check(_lib.lib().lmh_get_option(ev.solver.ctx.h, _lib.OPT_FUSED_LAUNCHED, ctypes.byref(v)))
        launched.append(v.value)
        U, K, s, _ = ON.myexp(hs[ts[k - 1]], ts[k] - ts[k - 1])
        assert ev.solver.K == K and ev.solver.s == s
        Vm = ON.apply(U, Vm)
    assert launched == [_lib.FUSED_STRIP] * 3, launched
    n = L * L
    pick = np.random.default_rng(9).integers(0, n, (2, 20000))
    got = ev.states[0].get_entries(pick[0], pick[1])
    ref = np.sum(Vm[pick[0]] * np.conj(Vm[pick[1]]), axis=1)
    d = float(np.max(np.abs(got - ref)))
    diag = ev.states[0].get_entries(np.arange(n), np.arange(n))
    d = max(d, float(np.max(np.abs(diag - np.sum(np.abs(Vm) ** 2, axis=1)))))
    say("csc n10000 worst |dP| = %.3e" % d)
    assert d <= TOL


def test_error_paths():
    L_ = _lib.lib()
    m = Model("square", (4, 4))
    Hs = sparse_of(m.oracle_h([("axial", 0.1, 1.5, 1.5)]))
    ctx = lmhip.Context(0)
    K = ctypes.c_int32()
    s = ctypes.c_double()
    upd = lambda: L_.lmh_update_solver(ctx.h, 0.05, 0, 1e-10, 1e-14, ctypes.byref(K), ctypes.byref(s))
    vals = lambda v: L_.lmh_set_csc_values(ctx.h, len(v), _lib.ptr(np.ascontiguousarray(v)))
    assert vals(Hs.nzval) == _lib.LMH_ESTATE                     # values before a pattern
    cp, rv = Hs.colptr, Hs.rowval
    _lib.check(L_.lmh_set_csc_pattern(ctx.h, Hs.n, _lib.ptr(cp), _lib.ptr(rv), 0))
    assert upd() == _lib.LMH_ESTATE                                # update before values
    assert vals(Hs.nzval[:-1]) == _lib.LMH_EINVAL                 # wrong count
    bad = Hs.nzval.copy()
    k = int(np.nonzero(Hs.rowval != np.repeat(np.arange(Hs.n), np.diff(Hs.colptr)))[0][0])
    bad[k] *= 1.0 + 1e-15j                                         # H[j,i] no longer conj(H[i,j])
    assert vals(bad) == _lib.LMH_EINVAL
    assert b"Hermitian" in L_.lmh_last_error()
    assert upd() == _lib.LMH_ESTATE                                # refused until valid values
    bad = Hs.nzval.copy()
    bad[0] = np.nan
    assert vals(bad) == _lib.LMH_EINVAL
    _lib.check(vals(Hs.nzval))
    _lib.check(upd())
    kinds = np.array([1], dtype=np.int32)
    params = np.array([0.1, 0.0, 0.0])
    assert L_.lmh_set_field(ctx.h, 1, _lib.ptr(kinds), _lib.ptr(params)) == _lib.LMH_ESTATE
    diag = np.zeros(Hs.n)
    assert L_.lmh_set_onsite(ctx.h, _lib.ptr(diag)) == _lib.LMH_ESTATE
    st = lmhip.DensityMatrix(ctx, Hs.n)
    assert L_.lmh_tape_begin(ctx.h, st.h, 2) == _lib.LMH_ESTATE
    # an unsymmetric pattern: drop one stored entry below the diagonal
    cols = np.repeat(np.arange(Hs.n), np.diff(Hs.colptr))
    drop = int(np.nonzero(Hs.rowval > cols)[0][0])
    rv2 = np.delete(Hs.rowval, drop)
    cp2 = Hs.colptr.copy()
    cp2[cols[drop] + 1:] -= 1
    assert L_.lmh_set_csc_pattern(ctx.h, Hs.n, _lib.ptr(cp2), _lib.ptr(rv2), 0) == _lib.LMH_EINVAL
    rv3 = Hs.rowval.copy()
    c0 = int(np.nonzero(np.diff(Hs.colptr) >= 2)[0][0])
    a, b = Hs.colptr[c0], Hs.colptr[c0] + 1
    rv3[a], rv3[b] = rv3[b], rv3[a]                                # rows not increasing
    assert L_.lmh_set_csc_pattern(ctx.h, Hs.n, _lib.ptr(cp), _lib.ptr(rv3), 0) == _lib.LMH_EINVAL
    assert L_.lmh_set_csc_pattern(ctx.h, Hs.n, _lib.ptr(cp), _lib.ptr(rv), 2) == _lib.LMH_EINVAL
    # a complex diagonal entry
    m2 = Model("square", (4, 4), onsite=np.linspace(-1, 1, 16))
    Hs2 = sparse_of(m2.oracle_h([]))
    _lib.check(L_.lmh_set_csc_pattern(ctx.h, Hs2.n, _lib.ptr(Hs2.colptr), _lib.ptr(Hs2.rowval), 0))
    bad = Hs2.nzval.copy()
    kd = int(np.nonzero(Hs2.rowval == np.repeat(np.arange(16), np.diff(Hs2.colptr)))[0][0])
    bad[kd] += 1e-3j
    assert vals(bad) == _lib.LMH_EINVAL
    st.close()
    ctx.close()
