"""Parametric Hamiltonians of the host mirror.

A `Hamiltonian` here is a lightweight specification -- lattice, hopping terms,
on-site energies and gauge field -- not an assembled sparse matrix.  Handing
it to `HIPExp` uploads only the parameters; the device builds H(t) with its
Peierls phases (the "parametric" entry mode of the drop-in boundary: the
reference allows any AbstractMatrix as `Hamiltonian(sys, data)`,
src/operators/system.jl:380-389).

Term semantics follow src/operators/constructoperator.jl:4-56 and 146-194:
  * an array of N numbers (LatticeValue) or `(op, array)` is an on-site term;
  * `(amp, NearestNeighbor(k))`, `(amp, Bravais[...])` or `(amp, [translations])`
    is a hopping term whose Hermitian partner is added automatically;
  * terms whose operator is zero are skipped;
  * on-site terms of one site accumulate in term order.
`tightbinding_hamiltonian` and `haldane` follow constructoperator.jl:190-194
and src/zoo/models.jl:164-171.
"""
import numpy as np

from .field import NoField, lower
from .lattice import Bonds, Translation


class NearestNeighbor:
    def __init__(self, order=1):
        self.order = int(order)


HONEYCOMB_2NN = [
    # src/zoo/models.jl:139-145
    Translation((1, 1), (1, 0)),
    Translation((2, 2), (-1, 0)),
    Translation((1, 1), (0, -1)),
    Translation((2, 2), (0, 1)),
    Translation((1, 1), (-1, 1)),
    Translation((2, 2), (1, -1)),
]


class Hamiltonian:
    def __init__(self, lat, hoppings, onsite, field):
        self.lattice = lat
        self.hoppings = hoppings          # [(complex amp, [Translation])]
        self.onsite = onsite              # float64[N] or None
        self.field = field if field is not None else NoField()
        self._geom = None

    @property
    def n(self):
        return len(self.lattice)

    def geometry_key(self):
        return (id(self.lattice), tuple((a, tuple(t)) for a, t in self.hoppings))

    def geometry(self):
        """(Bonds, amp complex[nb]) in the reference's bond iteration order."""
        if self._geom is None:
            parts, amps = [], []
            for amp, trs in self.hoppings:
                b = self.lattice.bonds(trs)
                parts.append(b)
                amps.append(np.full(len(b), amp, dtype=np.complex128))
            bonds = Bonds.concat(parts)
            amp = np.concatenate(amps) if amps else np.zeros(0, dtype=np.complex128)
            self._geom = (bonds, amp)
        return self._geom

    @property
    def bonds(self):
        return self.geometry()[0]

    def field_params(self):
        return lower(self.field, self.lattice)

    def with_field(self, field):
        h = Hamiltonian(self.lattice, self.hoppings, self.onsite, field)
        h._geom = self._geom
        return h

    def __repr__(self):
        return "Hamiltonian(dim=%dx%d)" % (self.n, self.n)


def _as_translations(lat, bonds):
    if isinstance(bonds, NearestNeighbor):
        return lat.nn_shell(bonds.order)
    if isinstance(bonds, Translation):
        return [bonds]
    return list(bonds)


def _is_bonds(x):
    if isinstance(x, (NearestNeighbor, Translation)):
        return True
    return isinstance(x, (list, tuple)) and len(x) > 0 and all(isinstance(t, Translation) for t in x)


def construct_hamiltonian(lat, *terms, field=None):
    n = len(lat)
    hoppings = []
    onsite = None
    for term in terms:
        if isinstance(term, tuple) and len(term) == 2 and _is_bonds(term[1]):
            op, bonds = term
            if op == 0:
                continue
            hoppings.append((complex(op), _as_translations(lat, bonds)))
            continue
        if isinstance(term, tuple) and len(term) == 2 and np.ndim(term[0]) == 0 and np.ndim(term[1]) == 1:
            op, values = term
        else:
            op, values = 1.0, term
        if op == 0:
            continue
        v = np.asarray(values, dtype=np.float64)
        if v.shape != (n,):
            raise ValueError("on-site term of length %d on a %d-site lattice" % (v.size, n))
        v = op * v
        onsite = v.copy() if onsite is None else onsite + v
    return Hamiltonian(lat, hoppings, onsite, field)


def tightbinding_hamiltonian(lat, *terms, t1=1.0, t2=0.0, t3=0.0, field=None):
    return construct_hamiltonian(lat, *terms, (t1, NearestNeighbor(1)), (t2, NearestNeighbor(2)),
                                 (t3, NearestNeighbor(3)), field=field)


def haldane(lat, t1, t2, m=0.0, *terms, field=None):
    if lat.kind != "honeycomb":
        raise ValueError("Invalid lattice type; expected HoneycombLattice")
    mass = np.where(lat.b == 1, m, -m).astype(np.float64)
    return construct_hamiltonian(lat, mass, *terms, (t1, NearestNeighbor(1)), (1j * t2, HONEYCOMB_2NN), field=field)
