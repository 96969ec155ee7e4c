"""Pins of the oracle's geometry against the reference's doctests and closed
forms (CPU only)."""
import math

from oracle import geometry as G


def test_site_order_doctest():
    # src/core/latticevalue.jl:60-82: x of SquareLattice(2,2), y of SquareLattice(3,3)
    l = G.SquareLattice(2, 2)
    assert [c[0] for c in l.coords] == [1.0, 1.0, 2.0, 2.0]
    l3 = G.SquareLattice(3, 3)
    assert [c[1] for c in l3.coords] == [1.0, 2.0, 3.0] * 3


def test_square_nn_shells_doctest():
    # src/zoo/lattices.jl:52-61 (and src/core/lattice.jl:61-70)
    lens, lists = G.default_nnbonds(G.square_unitcell())
    assert [round(x, 5) for x in lens] == [1.0, 1.41421, 2.0]
    assert [[t.tr for t in s] for s in lists] == [[(1, 0), (0, 1)], [(1, -1), (1, 1)], [(2, 0), (0, 2)]]
    assert all(t.si == (0, 0) for s in lists for t in s)


def test_honeycomb_nn_doctest():
    # src/lattices/bravais/nearestneighbor.jl:146-153
    lat = G.HoneycombLattice(5, 5)
    assert len(lat) == 50
    nn = lat.nn_translations(1)
    assert [(t.si, t.tr) for t in nn] == [((1, 2), (0, -1)), ((1, 2), (-1, 0)), ((1, 2), (0, 0))]
    assert math.isclose(lat.nn_lens[0], 1 / math.sqrt(3), rel_tol=1e-12)


def test_honeycomb_coordinates():
    lat = G.HoneycombLattice(3, 2)
    # i = 2((j1-1) L2 + (j2-1)) + b ; r = b_b + j1 (1,0) + j2 (1/2, sqrt3/2)
    for idx, ((j1, j2), b) in enumerate(lat.pointers):
        assert idx == 2 * ((j1 - 1) * 2 + (j2 - 1)) + (b - 1)
        bx, by = (0.0, 0.0) if b == 1 else (0.5, math.sqrt(3) / 6)
        x, y = lat.coords[idx]
        assert math.isclose(x, bx + j1 + 0.5 * j2) and math.isclose(y, by + j2 * math.sqrt(3) / 2)


def test_bond_counts_open_and_periodic():
    for Lx, Ly in ((4, 4), (5, 3), (7, 2)):
        l = G.SquareLattice(Lx, Ly)
        nb = sum(len(l.translation_bonds(t)) for t in l.nn_translations(1))
        assert nb == Lx * (Ly - 1) + Ly * (Lx - 1)
        lp = G.SquareLattice(Lx, Ly, periodic=(True, True))
        nbp = sum(len(lp.translation_bonds(t)) for t in lp.nn_translations(1))
        assert nbp == 2 * Lx * Ly


def test_periodic_bonds_use_unwrapped_coordinates():
    l = G.SquareLattice(4, 3, periodic=(True, True))
    bonds = l.translation_bonds(G.Translation((0, 0), (1, 0)))
    # site (4, 2) -> unwrapped (5, 2), index of (1, 2)
    src = l.index[((4, 2), 1)]
    (b,) = [b for b in bonds if b[0] == src]
    assert b[1] == l.index[((1, 2), 1)]
    assert b[3] == (5.0, 2.0)
    assert b[4] == 1 + 0j


def test_twisted_boundary_factor():
    l = G.SquareLattice(3, 3, periodic=(True, False), twists=(0.7, 0.0))
    bonds = l.translation_bonds(G.Translation((0, 0), (1, 0)))
    wrapped = [b for b in bonds if b[3][0] == 4.0]
    assert wrapped and all(abs(b[4] - complex(math.cos(0.7), math.sin(0.7))) < 1e-15 for b in wrapped)
