"""TimeSequence and region currents (src/timesequence.jl, src/currents.jl:76-110).

`TimeSequence` is the host container of the reference (times + values, exact
lookups within sqrt(eps), `differentiate` / `integrate`); `DeviceSeries`
records per-step observables of an evolution into device memory through
`lmh_series_*` (no host transfer per step) and applies the same
differentiate! / integrate! there.  `currentsfrom` / `currentsfromto` sum
`DensityCurrents` over regions on the device (`lmh_currents_region`).
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import check, lib, ptr

_ATOL = math.sqrt(np.finfo(float).eps)


def _axpby(a, x, b, y):
    """a x + b y, products and sum rounded separately (BLAS axpby)."""
    if isinstance(x, np.ndarray):
        return np.add(np.multiply(a, x), np.multiply(b, y))
    return a * x + b * y


class TimeSequence:
    """A time-ordered sequence of values (src/timesequence.jl:5-160)."""

    def __init__(self, times=(), values=()):
        times, values = list(times), list(values)
        if len(times) != len(values):
            raise ValueError("Keys/values length mismatch:\n%d timestamps, %d snapshots" % (len(times), len(values)))
        self.times = [float(t) for t in times]
        self.values = values

    def __len__(self):
        return len(self.times)

    def _find(self, t):
        for k, tk in enumerate(self.times):
            if abs(tk - t) <= _ATOL:
                return k
        return None

    def __getitem__(self, t):
        k = self._find(t)
        if k is None:
            raise KeyError(t)
        return self.values[k]

    def __setitem__(self, t, v):
        k = self._find(t)
        if k is None:
            k = int(np.searchsorted(self.times, t))
            self.times.insert(k, float(t))
            self.values.insert(k, v)
        else:
            self.values[k] = v

    def __iter__(self):
        return iter(zip(self.times, self.values))

    def copy(self):
        return TimeSequence(list(self.times), [v.copy() if isinstance(v, np.ndarray) else v for v in self.values])

    def differentiate_(self):
        """differentiate! (src/timesequence.jl:213-225): symmetric differences
        at the midpoints, in place."""
        if len(self) < 2:
            raise RuntimeError("Cannot differentiate TimeSequence of length %d" % len(self))
        td = self.times
        for i in range(1, len(self)):
            dt = td[i] - td[i - 1]
            self.values[i - 1] = _axpby(1 / dt, self.values[i], -1 / dt, self.values[i - 1])
            td[i - 1] += dt / 2
        td.pop()
        self.values.pop()
        return self

    def integrate_(self):
        """integrate! (src/timesequence.jl:255-268): cumulative trapezoids, the
        first value 0, in place."""
        if len(self) < 2:
            raise RuntimeError("Cannot integrate TimeSequence of length %d" % len(self))
        td = self.times
        for i in range(1, len(self)):
            dt = td[i] - td[i - 1]
            self.values[i - 1] = _axpby(dt / 2, self.values[i], dt / 2, self.values[i - 1])
        last = self.values.pop()
        self.values.insert(0, np.zeros_like(last) if isinstance(last, np.ndarray) else 0.0 * last)
        for i in range(1, len(self)):
            self.values[i] = _axpby(1, self.values[i - 1], 1, self.values[i])
        return self

    def differentiate(self):
        return self.copy().differentiate_()

    def integrate(self):
        return self.copy().integrate_()


def _mask(n, region):
    """A region as n bytes: a boolean mask of length n or site indices."""
    if region is None:
        return None
    a = np.asarray(region)
    m = np.zeros(n, dtype=np.uint8)
    if a.dtype == bool:
        if a.shape != (n,):
            raise ValueError("a region mask needs one entry per site")
        m[a] = 1
    else:
        idx = np.atleast_1d(a).astype(np.int64)
        if idx.size and (idx.min() < 0 or idx.max() >= n):
            raise ValueError("site index out of range")
        m[idx] = 1
    return m


def _curr_state(curr):
    P, H = curr.P, curr.H
    P.ctx.set_hamiltonian(H)
    return P


def _site_mask(ctx, nsites, region):
    """A region of sites as a mask over matrix indices (all internal states)."""
    m = _mask(nsites, region)
    if m is None or ctx.internal == 1:
        return m
    return np.repeat(m, ctx.internal)


def currentsfrom(curr, src):
    """currentsfrom(DensityCurrents(H, P), src) (src/currents.jl:76-88): for
    every site j outside src the current from src into j (0 on src)."""
    P = _curr_state(curr)
    N = P.ctx.internal
    out = np.zeros(P.n)
    m = _site_mask(P.ctx, P.n // N, src)
    check(lib().lmh_currents_region(P.ctx.h, P.h, ptr(m), None, ptr(out), None))
    return out if N == 1 else out.reshape(-1, N).sum(axis=1)


def currentsfromto(curr, src, dst=None):
    """currentsfromto(DensityCurrents(H, P), src[, dst]) (src/currents.jl:90-110):
    the total current from src to dst (the rest of the lattice if None)."""
    P = _curr_state(curr)
    ns = P.n // P.ctx.internal
    tot = ctypes.c_double()
    ms, md = _site_mask(P.ctx, ns, src), _site_mask(P.ctx, ns, dst)
    check(lib().lmh_currents_region(P.ctx.h, P.h, ptr(ms), None if md is None else ptr(md), None, ctypes.byref(tot)))
    return tot.value


class DeviceSeries:
    """Per-step observables of an evolution recorded in device memory:
    kind "density" (localdensity), "currents" (DensityCurrents per bond) or
    "region" (currentsfromto(src, dst) per entry)."""

    KINDS = {"density": _lib.SERIES_DENSITY, "currents": _lib.SERIES_CURRENTS, "region": _lib.SERIES_REGION}

    def __init__(self, ctx, kind, capacity, src=None, dst=None, n=None):
        if kind not in self.KINDS:
            raise ValueError("kind must be one of %s" % sorted(self.KINDS))
        self.ctx = ctx
        self.kind = kind
        n = n if n is not None else ctx.geom_n // ctx.internal
        ms, md = _site_mask(ctx, n, src), _site_mask(ctx, n, dst)
        h = ctypes.c_void_p()
        check(lib().lmh_series_create(ctx.h, self.KINDS[kind], int(capacity), None if ms is None else ptr(ms),
                                      None if md is None else ptr(md), ctypes.byref(h)))
        self.h = h
        ctx._states.add(self)

    def record(self, state, t):
        check(lib().lmh_series_record(self.h, state.h, float(t)))

    def __len__(self):
        ln, w = ctypes.c_int64(), ctypes.c_int64()
        check(lib().lmh_series_length(self.h, ctypes.byref(ln), ctypes.byref(w)))
        return ln.value

    def differentiate_(self):
        check(lib().lmh_series_differentiate(self.h))
        return self

    def integrate_(self):
        check(lib().lmh_series_integrate(self.h))
        return self

    def to_timesequence(self):
        ln, w = ctypes.c_int64(), ctypes.c_int64()
        check(lib().lmh_series_length(self.h, ctypes.byref(ln), ctypes.byref(w)))
        times = np.zeros(ln.value)
        vals = np.zeros((ln.value, w.value))
        check(lib().lmh_series_download(self.h, ptr(times), ptr(vals)))
        if self.kind == "region":
            return TimeSequence(times, [float(v[0]) for v in vals])
        from .evolution import site_currents, site_density
        agg = site_density if self.kind == "density" else site_currents
        return TimeSequence(times, [agg(self.ctx, v.copy()) for v in vals])

    def close(self):
        if self.h and self.ctx is not None and self.ctx.h:
            lib().lmh_series_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
