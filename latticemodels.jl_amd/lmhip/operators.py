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
from .lattice import AdjacencyMatrix, Bonds, SiteDistance, Translation


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


class SpinBasis:
    """SpinBasis(s): 2s + 1 internal states per site."""

    def __init__(self, spin):
        self.spin = spin
        self.n = int(round(2 * float(spin))) + 1

    def __len__(self):
        return self.n


class Hamiltonian:
    """Lattice, hopping terms, on-site energies and field.  With `internal` =
    N > 1 states per site the matrix index of state a on site i is i N + a
    (the basis internal ⊗ lattice, src/operators/system.jl:16); a hopping
    term's N x N block becomes the scalar bonds of its nonzero elements
    (row-major, the order of builder.jl:58-63), and the off-diagonal elements
    of a Hermitian on-site matrix become bonds within a site (no Peierls
    phase: their two ends coincide)."""

    def __init__(self, lat, hoppings, onsite, field, internal=1, onsite_offdiag=()):
        self.lattice = lat
        self.hoppings = hoppings          # [(complex amp or N x N matrix, [Translation])]
        self.onsite = onsite              # float64[n] (diagonal) or None
        self.field = field if field is not None else NoField()
        self.internal = int(internal)
        self.onsite_offdiag = list(onsite_offdiag)   # [(term index, a, b, complex[n_sites])], a < b
        self._geom = None
        self._site_bonds = None

    @property
    def n(self):
        return len(self.lattice) * self.internal

    @property
    def coords(self):
        c = self.lattice.coords
        return c if self.internal == 1 else np.repeat(c, self.internal, axis=0)

    def geometry_key(self):
        hop = tuple((np.asarray(a).tobytes() if np.ndim(a) else a,
                     t.key() if isinstance(t, AdjacencyMatrix) else tuple(t)) for a, t in self.hoppings)
        off = tuple((k, a, b, v.tobytes()) for k, a, b, v in self.onsite_offdiag)
        return (id(self.lattice), self.internal, hop, off)

    def _hopping_parts(self):
        """Per hopping term: (its position among the terms, lattice Bonds)."""
        return [(k, trs.bonds() if isinstance(trs, AdjacencyMatrix) else self.lattice.bonds(trs))
                for k, (_amp, trs) in enumerate(self.hoppings)]

    def geometry(self):
        """(Bonds, amp complex[nb]) of matrix indices, in the reference's write
        order; for N = 1 the lattice bonds themselves."""
        if self._geom is None:
            N = self.internal
            parts, amps, maps = [], [], []
            lb0 = 0
            hop = self._hopping_parts()
            pieces = [("hop", k, b) for k, b in hop] + [("off", k, q) for q, (k, _a, _b, _v) in
                                                       enumerate(self.onsite_offdiag)]
            pieces.sort(key=lambda x: x[1])
            for kind, k, obj in pieces:
                if kind == "hop":
                    b = obj
                    amp = self.hoppings[k][0]
                    if N == 1:
                        parts.append(b)
                        amps.append(np.full(len(b), amp, dtype=np.complex128))
                        maps.append(lb0 + np.arange(len(b)))
                    else:
                        B = np.asarray(amp, dtype=np.complex128)
                        if B.ndim == 0:
                            B = B * np.eye(N)
                        ab = [(x, y) for x in range(N) for y in range(N) if B[x, y] != 0]
                        if ab:
                            av = np.array([x for x, _ in ab])
                            bv = np.array([y for _, y in ab])
                            m = len(ab)
                            parts.append(Bonds((b.i[:, None] * N + av[None, :]).ravel(),
                                               (b.j[:, None] * N + bv[None, :]).ravel(),
                                               np.repeat(b.ri, m, axis=0), np.repeat(b.rj, m, axis=0),
                                               np.repeat(b.twist, m)))
                            amps.append(np.tile(np.array([B[x, y] for x, y in ab]), len(b)))
                            maps.append(np.repeat(lb0 + np.arange(len(b)), m))
                    lb0 += len(b)
                else:
                    _k, a, bb, vals = self.onsite_offdiag[obj]
                    sites = np.nonzero(vals != 0)[0]
                    c = self.lattice.coords[sites]
                    parts.append(Bonds(sites * N + a, sites * N + bb, c, c, np.ones(len(sites), dtype=np.complex128)))
                    amps.append(vals[sites].astype(np.complex128))
                    maps.append(np.full(len(sites), -1))
            bonds = Bonds.concat(parts)
            amp = np.concatenate(amps) if amps else np.zeros(0, dtype=np.complex128)
            self.bond_map = np.concatenate(maps) if maps else np.zeros(0, dtype=np.int64)
            self._geom = (bonds, amp)
        return self._geom

    @property
    def bonds(self):
        """The lattice bonds (site pairs) in the reference's iteration order."""
        if self.internal == 1:
            return self.geometry()[0]
        if self._site_bonds is None:
            self._site_bonds = Bonds.concat([b for _k, b in self._hopping_parts()])
        return self._site_bonds

    def field_params(self):
        return lower(self.field, self.lattice)

    def with_field(self, field):
        h = Hamiltonian(self.lattice, self.hoppings, self.onsite, field, self.internal, self.onsite_offdiag)
        h._geom = self._geom
        h._site_bonds = self._site_bonds
        if self._geom is not None:
            h.bond_map = self.bond_map
        return h

    def __repr__(self):
        return "Hamiltonian(dim=%dx%d)" % (self.n, self.n)


def _as_translations(lat, bonds):
    """A hopping term's bonds: a list of Bravais translations, or an
    AdjacencyMatrix (a SiteDistance becomes one, src/core/bonds.jl:316-319)."""
    if isinstance(bonds, NearestNeighbor):
        return lat.nn_shell(bonds.order)
    if isinstance(bonds, Translation):
        return [bonds]
    if isinstance(bonds, AdjacencyMatrix):
        if bonds.lat is not lat:
            raise ValueError("adjacency matrix of another lattice")
        return bonds
    if isinstance(bonds, SiteDistance):
        return AdjacencyMatrix(lat, bonds)
    return list(bonds)


def _is_bonds(x):
    if isinstance(x, (NearestNeighbor, Translation, AdjacencyMatrix, SiteDistance)):
        return True
    return isinstance(x, (list, tuple)) and len(x) > 0 and all(isinstance(t, Translation) for t in x)


class SparseHamiltonian:
    """The Hamiltonian as a sparse matrix in compressed-column form: the
    SparseMatrixCSC{ComplexF64} that the reference's builders produce
    (to_operator, src/operators/builder.jl:317-318; Hamiltonian.data,
    src/operators/system.jl:380-386) and that Evolution passes to
    update_solver! (src/evolution.jl:244).  colptr[n+1] and rowval[nnz] are
    0-based (rows strictly increasing within a column), nzval complex[nnz].

    On the device the pattern's entries above the diagonal are the bonds (in
    column order: `bond_ends()`), the values their amplitudes; H must be
    Hermitian exactly (stored H[j,i] = conj(H[i,j]), real diagonal).
    `==` is the reference's equality of two sparse matrices restricted to
    one stored form: same size, pattern and values."""

    internal = 1
    onsite = None

    def __init__(self, n, colptr, rowval, nzval):
        self.n = int(n)
        self.colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        self.rowval = np.ascontiguousarray(rowval, dtype=np.int64)
        self.nzval = np.ascontiguousarray(nzval, dtype=np.complex128)
        if self.colptr.shape != (self.n + 1,) or self.rowval.shape != self.nzval.shape or \
                len(self.rowval) != self.colptr[-1] - self.colptr[0]:
            raise ValueError("inconsistent CSC arrays")

    @classmethod
    def from_scipy(cls, m):
        """From a scipy.sparse matrix (converted to CSC, indices sorted)."""
        c = m.tocsc(copy=True)
        c.sort_indices()
        return cls(c.shape[0], c.indptr, c.indices, c.data)

    @property
    def nnz(self):
        return len(self.nzval)

    def bond_ends(self):
        """(i, j) of the device bonds: the stored entries H[i, j] with i < j,
        in column order."""
        cols = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.colptr))
        up = self.rowval < cols
        return self.rowval[up], cols[up]

    def __eq__(self, other):
        if not isinstance(other, SparseHamiltonian):
            return NotImplemented
        return (self.n == other.n and np.array_equal(self.colptr, other.colptr)
                and np.array_equal(self.rowval, other.rowval) and np.array_equal(self.nzval, other.nzval))

    __hash__ = None


def _is_zero(op):
    """Every element of the operator (a number or a matrix) is zero."""
    if isinstance(op, (int, float, complex)):
        return op == 0
    return bool(np.all(np.asarray(op) == 0))


def construct_hamiltonian(lat, *terms, field=None, internal=None):
    """construct_hamiltonian(l[, internal], terms...; field)
    (src/operators/constructoperator.jl:4-56,146-167).  `internal`: a
    SpinBasis or a number of states per site; then operators may be N x N
    matrices (a number means that number times the identity), and a bare
    matrix is an on-site term with factor 1 on every site."""
    n = len(lat)
    N = 1 if internal is None else (len(internal) if hasattr(internal, "__len__") else int(internal))
    hoppings = []
    onsite = None
    offdiag = []

    def add_onsite(op, v, k):
        nonlocal onsite
        if N == 1:
            d = float(np.real(op)) * v          # a new array: the caller's v is not kept
            onsite = d if onsite is None else onsite + d
            return
        B = np.asarray(op, dtype=np.complex128)
        if B.ndim == 0:
            B = B * np.eye(N)
        if B.shape != (N, N):
            raise ValueError("matrix size does not match on-site dims")
        if np.max(np.abs(B - B.conj().T)) != 0:
            raise ValueError("on-site operators must be Hermitian (the device keeps a real diagonal)")
        d = np.zeros(n * N)
        for a in range(N):
            if B[a, a] != 0:                  # zero elements are not written (builder.jl:61)
                d[a::N] = B[a, a].real * v
        onsite = d if onsite is None else onsite + d
        for a in range(N):
            for b in range(a + 1, N):
                if B[a, b] != 0:
                    offdiag.append((k, a, b, B[a, b] * v))

    for k, term in enumerate(terms):
        if isinstance(term, tuple) and len(term) == 2 and _is_bonds(term[1]):
            op, bonds = term
            if _is_zero(op):
                continue
            if N == 1:
                if np.ndim(op) != 0:
                    raise ValueError("matrix operator on a lattice without internal states")
                hoppings.append((complex(op), _as_translations(lat, bonds)))
            else:
                B = np.asarray(op, dtype=np.complex128)
                if B.ndim and B.shape != (N, N):
                    raise ValueError("matrix size does not match on-site dims")
                hoppings.append((B if B.ndim else complex(op), _as_translations(lat, bonds)))
            continue
        if isinstance(term, tuple) and len(term) == 2 and np.ndim(term[1]) <= 1:
            op, values = term
        elif np.ndim(term) == 2:
            op, values = term, 1.0
        else:
            op, values = 1.0, term
        if _is_zero(op):
            continue
        v = np.broadcast_to(np.asarray(values, dtype=np.float64), (n,)).copy() if np.ndim(values) == 0 else \
            np.asarray(values, dtype=np.float64)
        if v.shape != (n,):
            raise ValueError("on-site term of length %d on a %d-site lattice" % (v.size, n))
        add_onsite(op, v, k)
    return Hamiltonian(lat, hoppings, onsite, field, N, offdiag)


def tightbinding_hamiltonian(lat, *terms, t1=1.0, t2=0.0, t3=0.0, field=None):
    return construct_hamiltonian(lat, *terms, (t1, NearestNeighbor(1)), (t2, NearestNeighbor(2)),
                                 (t3, NearestNeighbor(3)), field=field)


def haldane(lat, t1, t2, m=0.0, *terms, field=None):
    if lat.kind != "honeycomb":
        raise ValueError("Invalid lattice type; expected HoneycombLattice")
    mass = np.where(lat.b == 1, m, -m).astype(np.float64)
    return construct_hamiltonian(lat, mass, *terms, (t1, NearestNeighbor(1)), (1j * t2, HONEYCOMB_2NN), field=field)


QWZ_SZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)
QWZ_X = np.array([[1, -1j], [-1j, -1]], dtype=np.complex128) / 2
QWZ_Y = np.array([[1, -1], [1, -1]], dtype=np.complex128) / 2


def qwz(lat, m=1.0, *terms, field=None):
    """qwz(l, m) (src/zoo/models.jl:130-136): sigma_z m on every site,
    [1 -i; -i -1]/2 along axis 1 and [1 -1; 1 -1]/2 along axis 2, one
    spin-1/2 per site (matrix index 2 i + a)."""
    if lat.kind != "square":
        raise ValueError("Invalid lattice type; expected SquareLattice")
    return construct_hamiltonian(lat, (QWZ_SZ, m), (QWZ_X, Translation((0, 0), (1, 0))),
                                 (QWZ_Y, Translation((0, 0), (0, 1))), *terms, field=field, internal=SpinBasis(0.5))
