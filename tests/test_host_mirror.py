"""Host-side logic of the lmhip mirror (CPU only): bond lists, unwrapped
coordinates, field lowering and term parsing agree with the oracle's
restatement of the reference."""
import numpy as np
import pytest

import lmhip
from helpers import Model, fields_pair
from oracle import geometry as OG

MODELS = [
    ("square_open", Model("square", (7, 5)), []),
    ("square_pbc", Model("square", (6, 4), (True, True)), []),
    ("square_pbc_y_nnn", Model("square", (5, 6), (False, True), t2=0.3), []),
    ("square_tiny_pbc", Model("square", (2, 3), (True, True)), []),
    ("honeycomb_open", Model("honeycomb", (4, 3), t2=0.1, m=0.2), []),
    ("honeycomb_pbc", Model("honeycomb", (3, 4), (True, True), t2=0.1, m=0.2), []),
]


@pytest.mark.parametrize("name,m,fs", MODELS, ids=[x[0] for x in MODELS])
def test_bonds_match_oracle_order_and_coordinates(name, m, fs):
    Ho = m.oracle_h(fs)
    Hd = m.device_h(fs)
    b = Hd.bonds
    assert [(int(i), int(j)) for i, j in zip(b.i, b.j)] == Ho.bonds
    # unwrapped destination coordinates and twists, translation by translation
    ref = []
    for term_translations in ([t for t in m.olat.nn_translations(1)],):
        for t in term_translations:
            ref.extend(m.olat.translation_bonds(t))
    k = len(ref)
    assert np.allclose(b.rj[:k], np.array([r[3] for r in ref]), atol=0, rtol=0)
    assert np.allclose(b.ri[:k], np.array([r[2] for r in ref]), atol=0, rtol=0)


def test_nn_shells_match_oracle_detection():
    for kind, uc in (("square", OG.square_unitcell()), ("honeycomb", OG.honeycomb_unitcell())):
        lat = lmhip.BravaisLattice(kind, (3, 3))
        _lens, lists = OG.default_nnbonds(uc)
        for order in (1, 2, 3):
            assert [(t.si, t.tr) for t in lat.nn_shell(order)] == [(t.si, t.tr) for t in lists[order - 1]]


def test_field_lowering_periodic_images():
    lat = lmhip.SquareLattice(5, 5, boundaries=(("axis1", True), ("axis2", True)))
    with pytest.warns(UserWarning):
        kinds, params = lmhip.field.lower(lmhip.PointFlux(0.1, (0.5, 0.5)), lat)
    assert len(kinds) == 9 and set(kinds) == {lmhip._lib.FIELD_POINT_SINGULAR}
    pts = sorted(map(tuple, params[:, 1:]))
    assert pts == sorted((0.5 + dx, 0.5 + dy) for dx in (-5, 0, 5) for dy in (-5, 0, 5))
    # oracle adapt_field gives the same image set
    from oracle import fields as OF
    o = OF.PointFlux(0.1, (0.5, 0.5)).adapt(OG.SquareLattice(5, 5, periodic=(True, True)))
    assert sorted(o.points) == pts
    k2, p2 = lmhip.field.lower(lmhip.LandauGauge(0.2) + lmhip.PointFlux(0.3, (1.5, 2.5)), lmhip.SquareLattice(4, 4))
    assert list(k2) == [lmhip._lib.FIELD_LANDAU, lmhip._lib.FIELD_POINT_AXIAL]
    assert p2[1].tolist() == [0.3, 1.5, 2.5]


def test_term_parsing_and_zero_terms():
    lat = lmhip.SquareLattice(3, 3)
    V = np.arange(9.0)
    h = lmhip.tightbinding_hamiltonian(lat, V, (2.0, V), t2=0.0)
    assert np.array_equal(h.onsite, 3 * V)
    assert len(h.hoppings) == 1 and h.hoppings[0][0] == 1.0
    h0 = lmhip.construct_hamiltonian(lat, (0.0, lmhip.NearestNeighbor(1)))
    assert h0.hoppings == [] and h0.onsite is None
    with pytest.raises(ValueError):
        lmhip.construct_hamiltonian(lat, np.zeros(5))
    with pytest.raises(ValueError):
        lmhip.haldane(lat, 1.0, 0.1)


def test_same_model_same_fields():
    of, lf = fields_pair([("landau", 0.1, 0, 0), ("axial", 0.2, 1.5, 1.5)])
    assert len(of.fields) == len(lf.fields) == 2


def test_internal_states_host_construction():
    """Spinful Hamiltonians on the host: matrix blocks expand to one bond per
    nonzero element at matrix indices 2 i + a (row-major per lattice bond),
    Hermitian on-site matrices give a real diagonal plus same-site bonds, and
    non-Hermitian on-site matrices or wrong sizes are rejected."""
    import lmhip
    lat = lmhip.SquareLattice(3, 2)
    h = lmhip.qwz(lat, 0.5)
    b, amp = h.geometry()
    sb = h.bonds
    assert h.n == 12 and len(sb) == 7 and len(b) == 4 * 7
    assert np.all(b.i // 2 == sb.i[h.bond_map]) and np.all(b.j // 2 == sb.j[h.bond_map])
    k0 = np.nonzero(h.bond_map == 0)[0]
    assert list(zip(b.i[k0] % 2, b.j[k0] % 2)) == [(0, 0), (0, 1), (1, 0), (1, 1)]
    assert np.allclose(h.onsite, np.tile([0.5, -0.5], 6))
    sx = np.array([[0, 1], [1, 0]])
    hc = lmhip.construct_hamiltonian(lat, (sx, np.arange(6.0)), (1.0, lmhip.NearestNeighbor(1)),
                                     internal=lmhip.SpinBasis(0.5))
    bc, ac = hc.geometry()
    same = bc.i // 2 == bc.j // 2
    assert np.all(hc.bond_map[same] == -1) and np.all(bc.i[same] % 2 == 0) and np.all(bc.j[same] % 2 == 1)
    assert np.allclose(ac[same], np.arange(1.0, 6.0))          # site 0 has value 0: no bond
    with pytest.raises(ValueError):
        lmhip.construct_hamiltonian(lat, (np.array([[0, 1], [0, 0]]), np.ones(6)), internal=2)
    with pytest.raises(ValueError):
        lmhip.construct_hamiltonian(lat, (np.eye(3), lmhip.NearestNeighbor(1)), internal=2)
