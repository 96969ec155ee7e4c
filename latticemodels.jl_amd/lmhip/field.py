"""Gauge fields of the host mirror, lowered to device field components.

  NoField                   src/operators/magneticfield.jl:43-44
  LandauGauge(B)            src/zoo/magneticfields.jl:11-15
  SymmetricGauge(B)         src/zoo/magneticfields.jl:27-31
  PointFlux(flux, point; gauge)   src/zoo/magneticfields.jl:46-104
  PointFluxes(fluxes, points; gauge)   src/zoo/magneticfields.jl:119-184
  f1 + f2 (FieldSum)        src/operators/magneticfield.jl:100-121
On a lattice with boundaries, point fluxes switch to the singular gauge and
are replicated over the 3^m boundary images (src/zoo/magneticfields.jl:237-266);
that bookkeeping happens here, the line integrals happen on the device.
"""
import warnings

import numpy as np

from . import _lib


class AbstractField:
    def __add__(self, other):
        return FieldSum(self, other)

    def components(self, lat):
        """[(kind, a, px, py)] after adapting to the lattice boundaries."""
        raise NotImplementedError


class NoField(AbstractField):
    def components(self, lat):
        return []

    def strengths(self):
        return []

    def __eq__(self, o):
        return isinstance(o, NoField)


class LandauGauge(AbstractField):
    def __init__(self, B):
        self.B = float(B)

    def components(self, lat):
        return [(_lib.FIELD_LANDAU, self.B, 0.0, 0.0)]

    def strengths(self):
        return [self.B]

    def __repr__(self):
        return "LandauGauge(%r)" % self.B


class SymmetricGauge(AbstractField):
    def __init__(self, B):
        self.B = float(B)

    def components(self, lat):
        return [(_lib.FIELD_SYMMETRIC, self.B, 0.0, 0.0)]

    def strengths(self):
        return [self.B]


class PointFluxes(AbstractField):
    def __init__(self, fluxes, points, gauge="axial"):
        if gauge not in ("axial", "singular"):
            raise ValueError("Invalid gauge: %s; expected one of (:axial, :singular)" % gauge)
        fluxes = np.broadcast_to(np.asarray(fluxes, dtype=np.float64), (len(points),))
        if len(fluxes) != len(points):
            raise ValueError("Length of fluxes and points should be the same")
        self.fluxes = [float(f) for f in fluxes]
        self.points = [(float(p[0]), float(p[1])) for p in points]
        self.gauge = gauge

    def components(self, lat):
        gauge, fl, pts = self.gauge, self.fluxes, self.points
        if lat is not None and lat.boundaries:
            if gauge != "singular":
                warnings.warn("Setting flux gauge to singular for a lattice with periodic boundary conditions")
            gauge = "singular"
            new_fl, new_pts = [], []
            for r in lat.translation_vectors():
                new_fl.extend(fl)
                new_pts.extend((p[0] + r[0], p[1] + r[1]) for p in pts)
            fl, pts = new_fl, new_pts
        kind = _lib.FIELD_POINT_AXIAL if gauge == "axial" else _lib.FIELD_POINT_SINGULAR
        return [(kind, f, p[0], p[1]) for f, p in zip(fl, pts)]

    def strengths(self):
        return list(self.fluxes)


class PointFlux(PointFluxes):
    def __init__(self, flux, point=(0.0, 0.0), gauge="axial"):
        super().__init__([flux], [point], gauge)
        self.flux = float(flux)
        self.point = self.points[0]


class FieldSum(AbstractField):
    def __init__(self, *fields):
        flat = []
        for f in fields:
            flat.extend(f.fields if isinstance(f, FieldSum) else [f])
        self.fields = flat

    def components(self, lat):
        out = []
        for f in self.fields:
            out.extend(f.components(lat))
        return out

    def strengths(self):
        out = []
        for f in self.fields:
            out.extend(f.strengths())
        return out


def strength_index(field, lat):
    """For every lowered component, the index of the field strength it carries
    (fluxes of PointFluxes repeat over the periodic images; FieldSum
    concatenates its members)."""
    field = field or NoField()
    if isinstance(field, FieldSum):
        out, off = [], 0
        for f in field.fields:
            out.extend(off + i for i in strength_index(f, lat))
            off += len(f.strengths())
        return out
    ncomp = len(field.components(lat))
    nstr = len(field.strengths())
    return [c % nstr for c in range(ncomp)] if nstr else []


def strength_gradient(field, lat, grad_components):
    """dL/d(strengths) from dL/d(component strengths) (chain rule over images)."""
    idx = strength_index(field, lat)
    out = np.zeros(len((field or NoField()).strengths()))
    for c, i in enumerate(idx):
        out[i] += grad_components[c]
    return out


def lower(field, lat):
    """Field -> (kinds int32[m], params float64[m, 3]) for lmh_set_field."""
    comps = (field or NoField()).components(lat)
    kinds = np.array([c[0] for c in comps], dtype=np.int32)
    params = np.array([c[1:] for c in comps], dtype=np.float64).reshape(len(comps), 3)
    return kinds, params
