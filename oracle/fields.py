"""Oracle: gauge fields and their line integrals.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates:
  src/operators/magneticfield.jl:43-44     NoField (line integral 0)
  src/operators/magneticfield.jl:100-121   FieldSum (left-to-right sum), adapt_field
  src/zoo/magneticfields.jl:11-15          LandauGauge: (x1+x2)(y2-y1)B/2
  src/zoo/magneticfields.jl:27-31          SymmetricGauge: (x1 y2 - x2 y1)/2 B
  src/zoo/magneticfields.jl:66-85          PointFlux{:axial}
  src/zoo/magneticfields.jl:89-104         PointFlux{:singular}
  src/zoo/magneticfields.jl:119-184        PointFluxes
  src/zoo/magneticfields.jl:226-235        periodic_fluxes
  src/zoo/magneticfields.jl:237-266        adapt_field on periodic lattices

Pins (tests/test_oracle_fields.py): PointFluxes line integral 0.2
(test/test_field.jl:43-59); image counts 25 / 9 / 225 (test/test_field.jl:80-97);
Landau line integral equals the 100-step trapezoid rule of A=(0, xB)
(test/test_field.jl:23-41); plaquette flux of the Landau gauge = B and winding
of a point flux = its flux (closed forms).
"""
import math

from .geometry import cartesian_indices

TWO_PI = 2 * math.pi


def _sign(x):
    return (x > 0) - (x < 0)


class NoField:
    def line_integral(self, p1, p2):
        return 0.0

    def adapt(self, lat):
        return self


class LandauGauge:
    def __init__(self, B):
        self.B = float(B)

    def line_integral(self, p1, p2):
        return (p1[0] + p2[0]) * (p2[1] - p1[1]) * self.B / 2

    def adapt(self, lat):
        return self


class SymmetricGauge:
    def __init__(self, B):
        self.B = float(B)

    def line_integral(self, p1, p2):
        return (p1[0] * p2[1] - p2[0] * p1[1]) / 2 * self.B

    def adapt(self, lat):
        return self


def _axial(flux, point, p1, p2):
    x1, y1 = p1[0] - point[0], p1[1] - point[1]
    x2, y2 = p2[0] - point[0], p2[1] - point[1]
    n1 = math.sqrt(x1 * x1 + y1 * y1)
    n2 = math.sqrt(x2 * x2 + y2 * y2)
    if n1 < 1e-11 or n2 < 1e-11:
        return 0.0
    nnorm = n1 * n2
    det = x1 * y2 - x2 * y1
    anglecos = (x1 * x2 + y1 * y2) / nnorm / (1 + 1e-11)
    angle = math.acos(anglecos) * _sign(det)
    return angle * flux / TWO_PI


def _singular(flux, point, p1, p2):
    x1, y1 = p1[0] - point[0], p1[1] - point[1]
    x2, y2 = p2[0] - point[0], p2[1] - point[1]
    sg = x2 - x1
    if abs(sg) < 1e-11:
        return 0.0
    if x1 * x2 > 0 or max(x1, x2) == 0:
        return 0.0
    yint = (-y1 * x2 + y2 * x1) / (x1 - x2)
    return 0.0 if yint > 0 else flux * _sign(sg)


class PointFlux:
    def __init__(self, flux, point=(0.0, 0.0), gauge="axial"):
        if gauge not in ("axial", "singular"):
            raise ValueError("Invalid gauge: %s" % gauge)
        self.flux = float(flux)
        self.point = (float(point[0]), float(point[1]))
        self.gauge = gauge

    def line_integral(self, p1, p2):
        f = _axial if self.gauge == "axial" else _singular
        return f(self.flux, self.point, p1, p2)

    def adapt(self, lat):
        if not lat.boundaries:
            return self
        return PointFluxes([self.flux], [self.point], self.gauge).adapt(lat)


class PointFluxes:
    def __init__(self, fluxes, points, gauge="axial"):
        if len(fluxes) != len(points):
            raise ValueError("Length of fluxes and points should be the same")
        self.fluxes = [float(f) for f in fluxes]
        self.points = [(float(p[0]), float(p[1])) for p in points]
        self.gauge = gauge

    def line_integral(self, p1, p2):
        f = _axial if self.gauge == "axial" else _singular
        acc = None
        for flux, pt in zip(self.fluxes, self.points):
            v = f(flux, pt, p1, p2)
            acc = v if acc is None else acc + v
        return 0.0 if acc is None else acc

    def adapt(self, lat):
        """adapt_field(::PointFluxes, lat): singular gauge, one image of every
        point per boundary-shift tuple in cartesian order (magneticfields.jl:237-261)."""
        if not lat.boundaries:
            return self
        m = len(lat.boundaries)
        fl, pts = [], []
        p0 = lat.pointers[0]
        c0 = lat.coords[0]
        for tup in cartesian_indices(1, m):
            ptr = p0
            for k in range(m):
                ptr = lat.shift(ptr, lat.boundaries[k].translation, tup[k])
            c = lat.uc.site_coords(ptr[0], ptr[1])
            r = (c[0] - c0[0], c[1] - c0[1])
            fl.extend(self.fluxes)
            pts.extend([(p[0] + r[0], p[1] + r[1]) for p in self.points])
        return PointFluxes(fl, pts, "singular")


class FieldSum:
    def __init__(self, *fields):
        flat = []
        for f in fields:
            flat.extend(f.fields if isinstance(f, FieldSum) else [f])
        self.fields = flat

    def line_integral(self, p1, p2):
        acc = None
        for f in self.fields:
            v = f.line_integral(p1, p2)
            acc = v if acc is None else acc + v
        return acc

    def adapt(self, lat):
        return FieldSum(*[f.adapt(lat) for f in self.fields])


def periodic_fluxes(lat, fl):
    """periodic_fluxes(l, fl) (magneticfields.jl:226-235): one copy of the flux
    per lattice site, shifted to the unit cell of the site (a set)."""
    (a1x, a2x), (a1y, a2y) = lat.uc.t
    det = a1x * a2y - a2x * a1y
    px, py = fl.point
    l1 = round((a2y * px - a2x * py) / det)
    l2 = round((-a1y * px + a1x * py) / det)
    p0 = (px - (a1x * l1 + a2x * l2), py - (a1y * l1 + a2y * l2))
    pts = set()
    for (lc, _b) in lat.pointers:
        pts.add((p0[0] + a1x * lc[0] + a2x * lc[1], p0[1] + a1y * lc[0] + a2y * lc[1]))
    pts = sorted(pts)
    return PointFluxes([fl.flux] * len(pts), pts, fl.gauge)
