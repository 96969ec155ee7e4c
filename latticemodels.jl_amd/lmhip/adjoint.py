"""Reverse-mode gradients of the density-matrix evolution (device adjoint).

The reference has no derivative of `Evolution` (src/evolution.jl:183-275);
this module is the host side of what a ChainRulesCore `rrule` for it needs:

    ev = Evolution(h, P0)
    tape = EvolutionTape(ev, checkpoint_every=10)   # record the steps of ev.states[0]
    for (P, H, t) in ev(times): ...
    L, dL_dP = my_loss(P)                             # cotangent of P(T)
    g = tape.backward(dL_dP)
    g.field[k, c]   dL/d(strength of lowered field component c) at step k
    g.onsite[i]     dL/dV_i (summed over steps)
    g.P0            dL/dP(0) (Hermitian, N x N)

`rrule_evolution` packages the same as (P_T, pullback) in the ChainRulesCore
shape.  Cotangent helpers build dL/dP for losses made of `localdensity`
(src/operators/latticeutils.jl:41-45) and `DensityCurrents`
(src/zoo/currents.jl:81-102) values, including the currents' direct
dependence on the last step's Hamiltonian.

Every numeric step runs in liblmhip.so; this module only moves parameters.
"""
import ctypes

import numpy as np

from ._lib import PROGRESS_FN, check, lib, ptr
from .evolution import DensityMatrix, Evolution, _current_blocks
from .field import strength_gradient


class SparseCotangent:
    """dL/dP with few nonzero entries (rows, cols, vals); `+` concatenates
    (repeated entries add when the library scatters them)."""

    def __init__(self, rows, cols, vals):
        self.rows = np.asarray(rows, dtype=np.int64)
        self.cols = np.asarray(cols, dtype=np.int64)
        self.vals = np.asarray(vals, dtype=np.complex128)

    def __add__(self, other):
        return SparseCotangent(np.concatenate([self.rows, other.rows]), np.concatenate([self.cols, other.cols]),
                               np.concatenate([self.vals, other.vals]))

    def dense(self, n):
        out = np.zeros((n, n), dtype=np.complex128)
        np.add.at(out, (self.rows, self.cols), self.vals)
        return out


class Gradients:
    def __init__(self, field, onsite, P0):
        self.field = field
        self.onsite = onsite
        self.P0 = P0

    def field_strengths(self, fields, lattice):
        """dL/d(user-level strengths) per step, given the field object of each
        step (maps periodic images and FieldSum members back)."""
        return np.array([strength_gradient(f, lattice, self.field[k]) for k, f in enumerate(fields)])


class EvolutionTape:
    """Records every step of one state of an Evolution (lmh_tape_begin)."""

    def __init__(self, evolution, checkpoint_every=10, index=0):
        self.state = evolution.states[index]
        self.ctx = self.state.ctx
        check(lib().lmh_tape_begin(self.ctx.h, self.state.h, int(checkpoint_every)))
        self.checkpoint_every = int(checkpoint_every)

    def _length(self):
        ns = ctypes.c_int64()
        nc = ctypes.c_int64()
        check(lib().lmh_tape_length(self.ctx.h, ctypes.byref(ns), ctypes.byref(nc)))
        return ns.value, nc.value

    @property
    def nsteps(self):
        return self._length()[0]

    @property
    def checkpoints(self):
        """(in device memory, in pinned host memory) so far."""
        d = ctypes.c_int64()
        h = ctypes.c_int64()
        check(lib().lmh_tape_checkpoints(self.ctx.h, ctypes.byref(d), ctypes.byref(h)))
        return d.value, h.value

    def backward(self, dL_dP, progress=None):
        """dL_dP: cotangent of the recorded state's current P -- a
        SparseCotangent, an N x N numpy array or a DensityMatrix of the same
        context.  Returns Gradients (g.P0 is a DensityMatrix).  progress:
        optional callable(done, total), called as each checkpoint segment of
        reverse steps completes on the device."""
        nsteps, ncomp = self._length()
        if isinstance(dL_dP, DensityMatrix):
            lam = dL_dP.copy()
        elif isinstance(dL_dP, SparseCotangent):
            lam = DensityMatrix(self.ctx, self.state.n).set_entries(dL_dP.rows, dL_dP.cols, dL_dP.vals)
        else:
            lam = DensityMatrix.from_array(self.ctx, np.asarray(dL_dP, dtype=np.complex128))
        gf = np.zeros((nsteps, ncomp))
        gv = np.zeros(self.state.n)
        if progress is None:
            check(lib().lmh_tape_backward(self.ctx.h, lam.h, ptr(gf), ptr(gv)))
            return Gradients(gf, gv, lam)
        cb = PROGRESS_FN(lambda _user, done, total: progress(int(done), int(total)))
        check(lib().lmh_set_progress(self.ctx.h, cb, None))
        try:
            check(lib().lmh_tape_backward(self.ctx.h, lam.h, ptr(gf), ptr(gv)))
        finally:
            check(lib().lmh_set_progress(self.ctx.h, PROGRESS_FN(), None))
        return Gradients(gf, gv, lam)

    def close(self):
        if self.ctx is not None and self.ctx.h:
            lib().lmh_tape_end(self.ctx.h)
        self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hamiltonian_derivative(ctx, H, comp):
    """dH[i_b, j_b]/d(strength of lowered component comp) for every bond of
    the device geometry of H (`H.geometry()`)."""
    ctx.set_hamiltonian(H)
    out = np.zeros(len(H.geometry()[0]), dtype=np.complex128)
    check(lib().lmh_hamiltonian_derivative(ctx.h, int(comp), ptr(out)))
    return out


def localdensity_cotangent(n, weights, sites=None, internal=1):
    """dL/dP for L = sum_i w_i rho_i, rho = localdensity(P): a
    SparseCotangent on the diagonal (all sites, or the given ones).  n is the
    matrix size; with internal states (`internal` = N per site) i is a
    lattice site and rho_i = sum_a P[i N + a, i N + a]
    (src/operators/latticeutils.jl:41-45), so w_i goes to each of the N
    diagonal entries of its block."""
    N = int(internal)
    w = np.asarray(weights, dtype=np.float64)
    idx = np.arange(n // N) if sites is None else np.asarray(sites, dtype=np.int64)
    w = np.broadcast_to(w, idx.shape)
    rows = (idx[:, None] * N + np.arange(N)[None, :]).ravel()
    return SparseCotangent(rows, rows, np.repeat(w, N).astype(np.complex128))


def currents_cotangent(H, P, dL_dJ, ctx=None):
    """For J_b = DensityCurrents(H, P)[i_b, j_b] = 2 Im(H[i_b,j_b] P[j_b,i_b])
    and per-bond dL/dJ_b: returns (Lambda, g_direct) with Lambda the cotangent
    of P (a SparseCotangent on the bonds with nonzero dL/dJ_b) and
    g_direct[c] = sum_b dL/dJ_b * 2 Im(dH[i_b,j_b]/da_c P[j_b,i_b]) (H's own
    dependence on the field strengths of the step that produced P).  P is a
    DensityMatrix (only the needed entries are read) or an N x N array.

    The bonds b are those of `H.bonds`, the ones `observables` reports
    currents for.  With internal states (N per site) these are lattice bonds
    and J_b is the block sum sum_{a,c} 2 Im(H[i N + a, j N + c] P[j N + c, i N + a])
    (src/zoo/currents.jl:92-102): dL/dJ_b is spread over the matrix-index
    bonds of the block with the signs that `site_currents` sums them with."""
    if isinstance(P, DensityMatrix):
        ctx = ctx if ctx is not None else P.ctx
    elif ctx is None:
        raise ValueError("currents_cotangent needs ctx= when P is a host array")
    w_in = np.asarray(dL_dJ, dtype=np.float64)
    if len(w_in) != len(H.bonds.i):
        raise ValueError("dL_dJ has %d entries for %d bonds" % (len(w_in), len(H.bonds.i)))
    if getattr(H, "internal", 1) == 1:
        bonds, w = H.bonds, w_in
    else:
        # per matrix-index bond of the device geometry: the signed sum of
        # the lattice bonds' weights whose blocks it carries
        bonds = H.geometry()[0]
        ks, bs, sg = _current_blocks(H)
        w = np.zeros(len(bonds.i))
        np.add.at(w, bs, sg * w_in[ks])
    sel = np.nonzero(w)[0]
    bi, bj, ws = bonds.i[sel], bonds.j[sel], w[sel]
    diag, hij = ctx.hamiltonian_entries(H)
    hs = hij[sel]
    pji = P.get_entries(bj, bi) if isinstance(P, DensityMatrix) else np.asarray(P)[bj, bi]
    # dJ_b = 2 Re(-i H_ij dP_ji) = Re(conj(Lam_ji) dP_ji)  ->  Lam_ji = 2i conj(H_ij) w_b
    lam = SparseCotangent(bj, bi, 2j * np.conj(hs) * ws)
    kinds, _ = H.field_params()
    g = np.zeros(len(kinds))
    for c in range(len(kinds)):
        dh = hamiltonian_derivative(ctx, H, c)[sel]
        g[c] = float(np.sum(ws * 2.0 * np.imag(dh * pji)))
    return lam, g


def rrule_evolution(h, P0, times, checkpoint_every=10, **solver_kw):
    """ChainRulesCore-shaped rrule of P(T) = last state of Evolution(h, P0)(times).

    Returns (P_T, H_last, pullback): P_T the device state after the last time,
    H_last the Hamiltonian of the last step (the one DensityCurrents pairs with
    P_T) and pullback(dL_dP_T) -> Gradients."""
    ev = Evolution(h, P0, **solver_kw)
    tape = EvolutionTape(ev, checkpoint_every=checkpoint_every)
    H_last = None
    for ts in ev(times):
        H_last = ts.H
    P_T = ev.states[0]

    def pullback(dL_dP_T):
        return tape.backward(dL_dP_T)

    pullback.tape = tape
    pullback.evolution = ev
    return P_T, H_last, pullback
