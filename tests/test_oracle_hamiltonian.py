"""Pins of the oracle's Hamiltonian assembly (CPU only): reference doctests and
closed forms."""
import math

import numpy as np

from oracle import fields as F
from oracle import geometry as G
from oracle import hamiltonian as H


def test_stored_entry_counts_doctests():
    # src/operators/system.jl:346-371 and src/operators/builder.jl:197-224
    assert H.tightbinding_hamiltonian(G.SquareLattice(4, 4)).csc.nnz == 48
    assert H.tightbinding_hamiltonian(G.SquareLattice(5, 5)).csc.nnz == 80
    assert H.tightbinding_hamiltonian(G.SquareLattice(5, 5), field=F.LandauGauge(0.1)).csc.nnz == 80


def test_4x4_spectrum_range_doctest():
    # src/spectrum.jl:88-106: "Eigenvalues in range -3.23607 .. 3.23607"
    e = np.linalg.eigvalsh(H.tightbinding_hamiltonian(G.SquareLattice(4, 4)).dense())
    assert round(e[0], 5) == -3.23607 and round(e[-1], 5) == 3.23607


def test_open_rectangle_spectrum_closed_form():
    for Lx, Ly in ((5, 7), (8, 3)):
        e = np.linalg.eigvalsh(H.tightbinding_hamiltonian(G.SquareLattice(Lx, Ly)).dense())
        exact = sorted(2 * math.cos(math.pi * k / (Lx + 1)) + 2 * math.cos(math.pi * l / (Ly + 1))
                       for k in range(1, Lx + 1) for l in range(1, Ly + 1))
        assert np.max(np.abs(e - np.array(exact))) < 1e-12


def test_periodic_ring_spectrum_closed_form():
    Lx, Ly = 6, 5
    e = np.linalg.eigvalsh(H.tightbinding_hamiltonian(G.SquareLattice(Lx, Ly, periodic=(True, True))).dense())
    exact = sorted(2 * math.cos(2 * math.pi * k / Lx) + 2 * math.cos(2 * math.pi * l / Ly)
                   for k in range(Lx) for l in range(Ly))
    assert np.max(np.abs(e - np.array(exact))) < 1e-12


def test_hermitian_and_onsite():
    l = G.SquareLattice(6, 5, periodic=(True, False))
    V = np.random.default_rng(0).uniform(-1, 1, 30)
    h = H.tightbinding_hamiltonian(l, onsite=V, t2=0.3, field=F.FieldSum(F.LandauGauge(0.07), F.PointFlux(0.2, (2.5, 2.5))))
    d = h.dense()
    assert np.array_equal(d, d.conj().T)
    assert np.array_equal(np.diag(d).real, V)


def test_landau_plaquette_phase_product():
    B = 0.083
    l = G.SquareLattice(6, 6)
    h = H.tightbinding_hamiltonian(l, field=F.LandauGauge(B))
    idx = lambda x, y: l.index[((x, y), 1)]
    for x, y in ((1, 1), (3, 2), (5, 5)):
        a, b, c, d = idx(x, y), idx(x + 1, y), idx(x + 1, y + 1), idx(x, y + 1)
        # H[i,j] carries exp(-2πi ∫_{r_i}^{r_j} A·dl); around the counter-clockwise
        # loop a -> b -> c -> d -> a the product is exp(-2πi B)
        prod = h[(a, b)] * h[(b, c)] * h[(c, d)] * h[(d, a)]
        assert abs(prod - complex(math.cos(-2 * math.pi * B), math.sin(-2 * math.pi * B))) < 1e-12


def test_haldane_structure():
    l = G.HoneycombLattice(4, 4)
    h = H.haldane(l, 1.0, 0.1, 0.2)
    d = h.dense()
    assert np.array_equal(d, d.conj().T)
    for i, (_lc, b) in enumerate(l.pointers):
        assert d[i, i] == (0.2 if b == 1 else -0.2)
    # NNN entries are purely imaginary ±0.1 i without field
    nnn = [d[i, j] for (i, j) in h.bonds if abs(abs(d[i, j]) - 0.1) < 1e-15]
    assert nnn and all(v.real == 0.0 and abs(v.imag) == 0.1 for v in nnn)


def test_haldane_bloch_spectrum_periodic():
    """Real-space PBC Haldane model vs the momentum-space blocks built from the
    same hopping table: checks the periodic routing and index bookkeeping."""
    L = 6
    l = G.HoneycombLattice(L, L, periodic=(True, True))
    h = H.haldane(l, 1.0, 0.1, 0.2)
    e = np.sort(np.linalg.eigvalsh(h.dense()))
    hops = [(t, 1.0) for t in l.nn_translations(1)] + [(t, 0.1j) for t in G.HONEYCOMB_2NN]
    ek = []
    for k1 in range(L):
        for k2 in range(L):
            q = (2 * math.pi * k1 / L, 2 * math.pi * k2 / L)
            hk = np.diag([0.2, -0.2]).astype(complex)
            for t, a in hops:
                ph = np.exp(1j * (q[0] * t.tr[0] + q[1] * t.tr[1]))
                pairs = [(b, b) for b in (1, 2)] if t.si == (0, 0) else [t.si]
                for (si, sj) in pairs:
                    # H[i_site(si), j_site(sj) at +tr] = a  -> Bloch: h[si,sj] += a e^{iq.tr}
                    hk[si - 1, sj - 1] += a * ph
                    hk[sj - 1, si - 1] += np.conj(a * ph)
            ek.extend(np.linalg.eigvalsh(hk))
    assert np.max(np.abs(e - np.sort(ek))) < 1e-12
