"""GPU: lmh_currents_region_noalloc (region currents with caller-owned device
buffers and the library's static device scratch, no allocation) against
lmh_currents_region on the same states: currentsfrom per site and
currentsfromto totals, two calls in a row (the second on a smaller lattice,
so the shared scratch is reused with fewer partial sums), a state left in the
P^H form by an unread fused step, and the size limit."""
import ctypes
import sys

import numpy as np
import pytest

import lmhip
from helpers import Model

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _state(m, spec, seed, steps):
    rng = np.random.default_rng(seed)
    n = m.n
    Q = np.linalg.qr(rng.normal(size=(n, n // 3)) + 1j * rng.normal(size=(n, n // 3)))[0]
    solver = lmhip.HIPExp(ctx=lmhip.Context(0))
    H = m.device_h(spec)
    st = solver.make_state(Q @ Q.conj().T, H)
    solver.update_solver(H, 0.05)
    for _ in range(steps):
        solver.step(st)
    return solver, st


def _existing(L, solver, st, src, dst):
    n = st.n
    ps = np.zeros(n)
    tot = np.zeros(1)
    lmhip._lib.check(L.lmh_currents_region(solver.ctx.h, st.h, src.ctypes.data_as(ctypes.c_void_p),
                                           None if dst is None else dst.ctypes.data_as(ctypes.c_void_p),
                                           ps.ctypes.data_as(ctypes.c_void_p), tot.ctypes.data_as(ctypes.c_void_p)))
    return ps, tot[0]


class _Dev:
    """Caller-owned device memory through the HIP runtime the library uses."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.ptrs.append(p)
        return p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0
        return p

    def download(self, p, arr):
        assert self.hip.hipMemcpy(arr.ctypes.data, p, arr.nbytes, 2) == 0
        return arr

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(L, solver, st, src, dst):
    dev = _Dev()
    n = st.n
    d_src = dev.upload(src)
    d_dst = None if dst is None else dev.upload(dst)
    d_ps = dev.upload(np.full(n, -1.0))
    d_tot = dev.upload(np.full(1, -1.0))
    lmhip._lib.check(L.lmh_currents_region_noalloc(solver.ctx.h, st.h, d_src, d_dst, d_ps, d_tot))
    lmhip._lib.check(L.lmh_sync(solver.ctx.h))
    ps = dev.download(d_ps, np.zeros(n))
    tot = dev.download(d_tot, np.zeros(1))[0]
    dev.free()
    return ps, tot


def test_noalloc_region_currents_match_the_allocating_entry_point():
    L = lmhip._lib.lib()
    cases = [
        # (lattice, field, unread fused steps): larger first, then smaller
        (Model("square", (23, 19), onsite=np.random.default_rng(1).uniform(-0.5, 0.5, 437)),
         [("axial", 0.2, 11.5, 9.5)], 1),
        (Model("square", (8, 9)), [("landau", 0.03, 0, 0)], 2),
    ]
    for k, (m, spec, steps) in enumerate(cases):
        say("case no-alloc region currents N=%d steps=%d" % (m.n, steps))
        solver, st = _state(m, spec, 10 + k, steps)
        rng = np.random.default_rng(20 + k)
        src = (rng.random(m.n) < 0.3).astype(np.uint8) * 3       # any nonzero byte is "in"
        dst = (rng.random(m.n) < 0.4).astype(np.uint8)
        for d in (None, dst):
            ps_ref, tot_ref = _existing(L, solver, st, src, d)
            ps, tot = _noalloc(L, solver, st, src, d)
            assert np.max(np.abs(ps - ps_ref)) <= TOL
            assert abs(tot - tot_ref) <= TOL * max(1.0, abs(tot_ref))
        solver.ctx.close()


def test_noalloc_region_currents_size_limit():
    """More sites than the static scratch covers: LMH_ERANGE before anything
    runs (no state is needed for the check)."""
    L = lmhip._lib.lib()
    lat = lmhip.SquareLattice(520, 520)
    ctx = lmhip.Context(0)
    ctx.set_hamiltonian(lmhip.tightbinding_hamiltonian(lat))
    rc = L.lmh_currents_region_noalloc(ctx.h, None, None, None, None, None)
    assert rc == lmhip._lib.LMH_ERANGE
    assert b"262144" in L.lmh_last_error()
    ctx.close()
