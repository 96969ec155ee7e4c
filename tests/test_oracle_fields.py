"""Pins of the oracle's gauge fields (CPU only): the reference's field tests
(test/test_field.jl) plus closed forms (plaquette flux, point-flux winding)."""
import math

import numpy as np

from oracle import fields as F
from oracle import geometry as G


def trapezoid(A, p1, p2, n):
    """line_integral(field, p1, p2, n_steps) (src/operators/magneticfield.jl:26-35)."""
    d = ((p2[0] - p1[0]) / n, (p2[1] - p1[1]) / n)
    p = (p1[0] + 0.5 * d[0], p1[1] + 0.5 * d[1])
    acc = 0.0
    for _ in range(n):
        a = A(p)
        acc += d[0] * a[0] + d[1] * a[1]
        p = (p[0] + d[0], p[1] + d[1])
    return acc


def test_landau_matches_trapezoid():
    # test/test_field.jl:24-29: A = (0, x B)
    la = F.LandauGauge(0.1)
    p1, p2 = (1.0, 2.0), (3.0, 4.0)
    assert math.isclose(la.line_integral(p1, p2), trapezoid(lambda p: (0.0, p[0] * 0.1), p1, p2, 100), rel_tol=1e-12)


def test_symmetric_matches_trapezoid():
    sym = F.SymmetricGauge(0.1)
    p1, p2 = (1.0, 2.0), (3.0, 4.0)
    A = lambda p: (-p[1] * 0.1 / 2, p[0] * 0.1 / 2)
    assert math.isclose(sym.line_integral(p1, p2), trapezoid(A, p1, p2, 1), rel_tol=1e-12)


def test_axial_pointflux_matches_trapezoid():
    # test/test_field.jl:37: PointFlux(0.1) at the origin vs 1000-step trapezoid, atol 1e-8
    flx = F.PointFlux(0.1)
    p1, p2 = (1.0, 2.0), (3.0, 4.0)
    A = lambda p: (-p[1] / (p[0] ** 2 + p[1] ** 2) / (2 * math.pi) * 0.1, p[0] / (p[0] ** 2 + p[1] ** 2) / (2 * math.pi) * 0.1)
    assert abs(flx.line_integral(p1, p2) - trapezoid(A, p1, p2, 1000)) < 1e-8


def test_point_fluxes_known_value():
    # test/test_field.jl:43-59
    p1, p2 = (0.0, 0.0), (4.0, 0.0)
    ps1 = F.PointFluxes([0.1, 0.1], [(1, 2), (3, 4)], gauge="singular")
    assert abs(ps1.line_integral(p1, p2) - 0.2) < 1e-8
    pf1, pf2 = F.PointFlux(0.1, (1, 2)), F.PointFlux(0.1, (3, 4))
    ps3 = F.PointFluxes([0.1, 0.1], [(1, 2), (3, 4)])
    assert abs(F.FieldSum(pf1, pf2).line_integral(p1, p2) - ps3.line_integral(p1, p2)) < 1e-8


def test_flux_images_on_periodic_lattices():
    # test/test_field.jl:80-97: 25 / 9 / 225 points
    lnb = G.SquareLattice(5, 5)
    lwb = G.SquareLattice(5, 5, periodic=(True, True))
    flx = F.PointFlux(0.1, (0.5, 0.5), gauge="singular")
    flxs = F.periodic_fluxes(lnb, flx)
    assert flxs.points == F.periodic_fluxes(lwb, flx).points
    assert isinstance(flx.adapt(lnb), F.PointFlux)
    assert len(flxs.adapt(lnb).points) == 25
    assert len(flx.adapt(lwb).points) == 9
    assert len(flxs.adapt(lwb).points) == 225


def test_periodic_fluxes_positions():
    # test/test_field.jl:68-77
    l = G.SquareLattice(10, 10)
    a = sorted(F.periodic_fluxes(l, F.PointFlux(0.1, (0.4, 0.4))).points)
    b = sorted(F.periodic_fluxes(l, F.PointFlux(0.1, (1.4, 0.4))).points)
    assert all(math.dist(p, q) < 1e-10 for p, q in zip(a, b))


def _loop(field, pts):
    return sum(field.line_integral(pts[k], pts[(k + 1) % len(pts)]) for k in range(len(pts)))


def test_plaquette_flux_closed_form():
    # counter-clockwise unit plaquette encloses B flux quanta for both uniform gauges
    for fld in (F.LandauGauge(0.037), F.SymmetricGauge(0.037)):
        for (x, y) in ((1, 1), (3, 7), (-2, 5)):
            sq = [(x, y), (x + 1, y), (x + 1, y + 1), (x, y + 1)]
            assert math.isclose(_loop(fld, sq), 0.037, rel_tol=1e-12)


def test_point_flux_winding():
    for gauge in ("axial", "singular"):
        f = F.PointFlux(0.23, (2.5, 3.5), gauge=gauge)
        around = [(2, 3), (3, 3), (3, 4), (2, 4)]
        away = [(5, 5), (6, 5), (6, 6), (5, 6)]
        assert abs(_loop(f, around) - 0.23) < 1e-12
        assert abs(_loop(f, away)) < 1e-12


def test_singular_and_axial_agree_modulo_flux_on_spectrum():
    # test/test_operators.jl: d1 and d2 of PointFlux axial/singular give the same spectrum
    from oracle import hamiltonian as H
    l = G.SquareLattice(5, 5)
    e1 = np.linalg.eigvalsh(H.tightbinding_hamiltonian(l, field=F.PointFlux(0.1, (3.5, 3.5))).dense())
    e2 = np.linalg.eigvalsh(H.tightbinding_hamiltonian(l, field=F.PointFlux(0.1, (3.5, 3.5), gauge="singular")).dense())
    # the axial formula's (1 + 1e-11) shrink breaks exact gauge equivalence at ~1e-11;
    # the reference compares with `≈` (rtol = sqrt(eps))
    assert np.max(np.abs(e1 - e2)) < 1.5e-8 * np.max(np.abs(e1))
