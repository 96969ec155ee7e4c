"""AdjacencyMatrix and SiteDistance bonds (src/core/bonds.jl:136-345) in the
oracle and the host mirror (CPU): the reference's own testset
(test/test_lattice.jl:234-262) and doctest (src/core/bonds.jl:176-199), the
Hamiltonian of an adjacency term against the translation-built one on an
open lattice (same bonds, same phases) and against the direct-coordinate
formula on a periodic one, and the mirror's bonds against the oracle's."""
import numpy as np

import lmhip
from oracle import fields as OF
from oracle import geometry as OG
from oracle import hamiltonian as OH


def _axis(ax):
    tr = [0, 0]
    tr[ax] = 1
    return OG.Translation((0, 0), tr)


def test_reference_adjacency_testset():
    l = OG.SquareLattice(2, 2)
    # sites (1,1), (1,2), (2,1), (2,2): indices 0..3 (last axis fastest)
    am = OG.AdjacencyMatrix(l, _axis(0), _axis(1))
    am1 = OG.AdjacencyMatrix(l, _axis(0)).union(OG.AdjacencyMatrix(l, _axis(1)))
    assert am.pairs == am1.pairs
    assert am.get(0, 2) and not am.get(0, 3)
    assert set(am.adjacentsites(0)) == {1, 2}
    am2 = OG.AdjacencyMatrix(l)
    for i, j in ((0, 1), (0, 2), (1, 3), (2, 3)):
        am2.set(i, j)
    assert am.pairs == am2.pairs
    am3 = OG.AdjacencyMatrix.from_predicate(l, lambda a, b: abs(a[0] - b[0]) + abs(a[1] - b[1]) <= 1)
    assert am3.pairs == am.pairs
    assert OG.AdjacencyMatrix(l, OG.SiteDistance(lambda d: d <= 1)).pairs == am.pairs
    assert OG.AdjacencyMatrix(l, OG.SiteDistance(lambda d: 0 <= d <= 1)).pairs == am.pairs


def test_reference_doctest_pattern():
    """a[site1, site2] = a[site2, site4] = a[site3, site4] = true: 6 stored
    entries in the symmetric matrix."""
    l = OG.SquareLattice(2, 2)
    a = OG.AdjacencyMatrix(l)
    a.set(0, 1)
    a.set(1, 3)
    a.set(2, 3)
    m = np.zeros((4, 4), dtype=int)
    for i, j in a.pairs:
        m[i, j] = m[j, i] = 1
    assert m.sum() == 6
    assert m.tolist() == [[0, 1, 0, 0], [1, 0, 0, 1], [0, 0, 0, 1], [0, 1, 1, 0]]


def test_adjacency_hamiltonian_equals_translation_one_on_open_lattice():
    l = OG.SquareLattice(6, 5)
    f = OF.PointFlux(0.3, (3.3, 2.7))
    tr = l.nn_translations(1)
    Ht = OH.OracleHamiltonian(l, [OH.Hopping(1.0, tr)], f)
    Ha = OH.OracleHamiltonian(l, [OH.Hopping(1.0, [OG.AdjacencyMatrix(l, *tr)])], f)
    assert Ht.entries.keys() == Ha.entries.keys()
    assert max(abs(Ht.entries[k] - Ha.entries[k]) for k in Ht.entries) <= 1e-15


def test_adjacency_on_periodic_lattice_uses_direct_coordinates():
    """A periodic lattice's wrap bonds become plain index pairs: the phase runs
    between the two sites' own coordinates (no unwrapped image, no twist)."""
    l = OG.SquareLattice(5, 4, periodic=(True, True))
    f = OF.LandauGauge(0.07)
    am = OG.AdjacencyMatrix(l, *l.nn_translations(1))
    H = OH.OracleHamiltonian(l, [OH.Hopping(0.8, [am])], f)
    ff = H.field
    for (i, j) in am.pairs:
        ri, rj = l.coords[i], l.coords[j]
        want = 0.8 * OH.field_factor(ff.line_integral(ri, rj))
        assert abs(H.entries[(i, j)] - want) <= 1e-15
        assert abs(H.entries[(j, i)] - want.conjugate()) <= 1e-15


def test_site_distance_uses_nearest_images():
    l = OG.SquareLattice(5, 5, periodic=(True, False))
    am = OG.AdjacencyMatrix(l, OG.SiteDistance(lambda d: d <= 1.0 + 1e-9))
    # periodic along axis 1 only: NN bonds (axis 1 wraps) = 5*5 + 5*4
    assert len(am.pairs) == 25 + 20
    assert am.get(0, 20)              # (1,1) - (5,1) across the axis-1 boundary


def test_mirror_matches_oracle():
    for periodic in ((False, False), (True, True)):
        lo = OG.SquareLattice(6, 5, periodic=periodic)
        bnd = tuple((ax, True) for ax, p in zip(("axis1", "axis2"), periodic) if p)
        lm = lmhip.SquareLattice(6, 5, boundaries=bnd)
        sd_o = OG.AdjacencyMatrix(lo, OG.SiteDistance(lambda d: d <= 1.5))
        sd_m = lmhip.AdjacencyMatrix(lm, lmhip.SiteDistance(lambda d: d <= 1.5))
        assert set(sd_m.pairs()) == sd_o.pairs
        tr_o = OG.AdjacencyMatrix(lo, *lo.nn_translations(1))
        tr_m = lmhip.AdjacencyMatrix(lm, lmhip.NearestNeighbor(1))
        assert set(tr_m.pairs()) == tr_o.pairs
        H = lmhip.construct_hamiltonian(lm, (1.0, sd_m), field=lmhip.LandauGauge(0.05))
        b = H.geometry()[0]
        ob = sd_o.bonds()
        assert [(int(i), int(j)) for i, j in zip(b.i, b.j)] == [(i, j) for i, j, _a, _b, _t in ob]
        assert np.allclose(b.rj, [r for _i, _j, _a, r, _t in ob], rtol=0, atol=0)
        assert np.all(b.twist == 1.0)


def test_mirror_edits_and_operator_pattern():
    lm = lmhip.SquareLattice(4, 4)
    am = lmhip.AdjacencyMatrix(lm, lmhip.NearestNeighbor(1))
    n0 = len(am)
    am[0, 1] = False
    am[0, 5] = True
    assert len(am) == n0 and not am[1, 0] and am[5, 0]
    H = lmhip.construct_hamiltonian(lm, (1.0, am))
    back = lmhip.AdjacencyMatrix.from_operator(H)
    assert back.pairs() == am.pairs()
