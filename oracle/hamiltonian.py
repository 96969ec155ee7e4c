"""Oracle: Hamiltonian assembly with the reference's operator-builder semantics.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates:
  src/operators/builder.jl:60-66      elementwise setindex! (entries of the on-site
                                      matrix that are zero are skipped)
  src/operators/builder.jl:116-151    UniformMatrixBuilder: first write stores,
                                      later writes accumulate; CSC with sorted rows
  src/operators/builder.jl:282-285    expand_bond: t * twist_j * exp(-2πi ∫A·dl) * conj(twist_i)
                                      with the UNWRAPPED destination coordinates
  src/operators/builder.jl:296-309    auto-Hermitian partner H[j,i] += conj(...)
  src/operators/builder.jl:165-174    adapt_field for periodic lattices
  src/operators/constructoperator.jl:4-56  add_term! per term kind; zero terms skipped
  src/operators/constructoperator.jl:190-194  tightbinding_hamiltonian (t1,t2,t3 NN orders)
  src/zoo/models.jl:164-171           haldane: +m / -m on sublattices 1 / 2,
                                      t1 on NN(1), im*t2 on honeycomb_2nn
  src/zoo/models.jl:130-136           qwz: internal spin-1/2 blocks (site-major index
                                      i N + a, builder.jl:255-309 block writes)

Pins (tests/test_oracle_hamiltonian.py): stored-entry counts 48 (4x4) and 80
(5x5, with and without LandauGauge(0.1)) from the doctests at
src/operators/system.jl:346-371 and src/operators/builder.jl:197-224; the
4x4 eigenvalue range ±3.23607 (src/spectrum.jl:88-106); the open-rectangle
spectrum 2cos(πk/(Lx+1)) + 2cos(πl/(Ly+1)); Hermiticity; plaquette phase
products equal exp(-2πi·flux).
"""
import math

import numpy as np

from .fields import NoField
from .geometry import HONEYCOMB_2NN, AdjacencyMatrix

MINUS_TWO_PI = -(2 * math.pi)


def field_factor(x):
    """exp(-2π·im·x): the angle is fl(fl(-2π)·x) (builder.jl:283)."""
    th = MINUS_TWO_PI * x
    if th == 0.0:
        return complex(1.0, 0.0)
    return complex(math.cos(th), math.sin(th))


class Onsite:
    """`op => LatticeValue` on-site term (constructoperator.jl:21-30); op a
    number or, with internal degrees of freedom, an N x N matrix."""

    def __init__(self, values, op=1.0):
        self.values = [float(v) for v in values]
        self.op = op


class Hopping:
    """`amp => bonds` hopping term over a list of Bravais translations or
    adjacency matrices (constructoperator.jl:14-20,31-38); amp a number or an
    N x N matrix."""

    def __init__(self, amp, translations):
        self.amp = amp
        self.translations = list(translations)


def _block(op, N):
    """op_to_matrix (builder.jl:255-266): a number times the N x N identity, or
    the matrix itself (checked square of size N)."""
    if np.ndim(op) == 0:
        return [[complex(op) if a == b else 0j for b in range(N)] for a in range(N)]
    m = [[complex(x) for x in row] for row in np.asarray(op)]
    if len(m) != N or any(len(r) != N for r in m):
        raise ValueError("matrix size does not match on-site dims")
    return m


class CSCMatrix:
    def __init__(self, n, entries):
        self.n = n
        cols = [[] for _ in range(n)]
        for (i, j), v in entries.items():
            cols[j].append((i, v))
        colptr = [0]
        rowval, nzval = [], []
        for j in range(n):
            for i, v in sorted(cols[j]):
                rowval.append(i)
                nzval.append(v)
            colptr.append(len(rowval))
        self.colptr = np.asarray(colptr, dtype=np.int64)
        self.rowval = np.asarray(rowval, dtype=np.int64)
        self.nzval = np.asarray(nzval, dtype=np.complex128)

    @property
    def nnz(self):
        return len(self.rowval)

    def todense(self):
        d = np.zeros((self.n, self.n), dtype=np.complex128)
        for j in range(self.n):
            for k in range(self.colptr[j], self.colptr[j + 1]):
                d[self.rowval[k], j] = self.nzval[k]
        return d


class OracleHamiltonian:
    """Assembled H(t) plus the bond list (for currents), in reference order.

    `internal` = N internal states per site (a SpinBasis(1//2) has N = 2): the
    matrix index of state a on site i is i N + a (the basis internal ⊗ lattice,
    src/operators/system.jl:16, constructoperator.jl:24-28); every term writes
    an N x N block, element by element in row-major order, zero elements
    skipped (builder.jl:58-63), the Hermitian partner of a hopping block being
    conj(factor) * B' (builder.jl:296-309)."""

    def __init__(self, lat, terms, field=None, internal=1):
        field = NoField() if field is None else field
        self.lat = lat
        self.internal = N = int(internal)
        self.field = field.adapt(lat)        # auto_pbc_field=true (builder.jl:168-170)
        entries = {}

        def put(i, j, val):
            if (i, j) in entries:
                entries[(i, j)] = entries[(i, j)] + val
            else:
                entries[(i, j)] = val

        self.bonds = []                      # site pairs (i, j) in iteration order, first term first
        for term in terms:
            if isinstance(term, Onsite):
                if np.all(np.asarray(term.op) == 0):
                    continue
                if N == 1 and np.ndim(term.op) == 0:
                    for i, v in enumerate(term.values):
                        put(i, i, term.op * v)
                    continue
                B = _block(term.op, N)
                for i, v in enumerate(term.values):
                    for a in range(N):
                        for b in range(N):
                            if B[a][b] != 0:
                                put(i * N + a, i * N + b, B[a][b] * v)
            elif isinstance(term, Hopping):
                if np.all(np.asarray(term.amp) == 0):
                    continue
                scalar = N == 1 and np.ndim(term.amp) == 0
                t = complex(term.amp) if scalar else None
                B = None if scalar else _block(term.amp, N)
                for tr in term.translations:
                    bonds = tr.bonds() if isinstance(tr, AdjacencyMatrix) else lat.translation_bonds(tr)
                    for (i, j, ri, rj, twist) in bonds:
                        ff = field_factor(self.field.line_integral(ri, rj))
                        total = twist * ff               # * conj(s1.factor) == * 1
                        if scalar:
                            put(i, j, t * total)
                            if i != j:
                                put(j, i, t.conjugate() * total.conjugate())
                        else:
                            for a in range(N):
                                for b in range(N):
                                    if B[a][b] != 0:
                                        put(i * N + a, j * N + b, B[a][b] * total)
                            if i != j:
                                for b in range(N):
                                    for a in range(N):
                                        if B[a][b] != 0:
                                            put(j * N + b, i * N + a, B[a][b].conjugate() * total.conjugate())
                        self.bonds.append((i, j))
            else:
                raise TypeError("unknown term %r" % (term,))
        self.entries = entries
        self.csc = CSCMatrix(len(lat) * N, entries)

    @property
    def n(self):
        return len(self.lat) * self.internal

    def block(self, i, j):
        """_block(H, i, j) (src/zoo/currents.jl:24-27): the N x N block of sites i, j."""
        N = self.internal
        return np.array([[self.entries.get((i * N + a, j * N + b), 0j) for b in range(N)] for a in range(N)])

    def dense(self):
        return self.csc.todense()

    def __getitem__(self, ij):
        return self.entries.get(ij, 0j)


def tightbinding_terms(lat, extra=(), t1=1.0, t2=0.0, t3=0.0):
    """tightbinding_hamiltonian(args...; t1, t2, t3): user terms first, then
    t1 => NN(1), t2 => NN(2), t3 => NN(3) (constructoperator.jl:190-194)."""
    terms = list(extra)
    for order, t in ((1, t1), (2, t2), (3, t3)):
        if t != 0:
            terms.append(Hopping(t, lat.nn_translations(order)))
    return terms


def tightbinding_hamiltonian(lat, onsite=None, t1=1.0, t2=0.0, t3=0.0, field=None):
    extra = [] if onsite is None else [Onsite(onsite)]
    return OracleHamiltonian(lat, tightbinding_terms(lat, extra, t1, t2, t3), field)


def haldane_terms(lat, t1, t2, m=0.0, onsite=None):
    """haldane(lat, t1, t2, m) (zoo/models.jl:164-171), optional extra on-site V
    added after the sublattice mass."""
    mass = [m if b == 1 else -m for (_lc, b) in lat.pointers]
    terms = [Onsite(mass)]
    if onsite is not None:
        terms.append(Onsite(onsite))
    terms.append(Hopping(t1, lat.nn_translations(1)))
    terms.append(Hopping(complex(0.0, 1.0) * t2, HONEYCOMB_2NN))
    return terms


def haldane(lat, t1, t2, m=0.0, onsite=None, field=None):
    return OracleHamiltonian(lat, haldane_terms(lat, t1, t2, m, onsite), field)


SIGMA_Z = [[1, 0], [0, -1]]
QWZ_X = [[0.5, -0.5j], [-0.5j, -0.5]]        # [1 -im; -im -1] / 2
QWZ_Y = [[0.5, -0.5], [0.5, -0.5]]           # [1 -1; 1 -1] / 2


def qwz(lat, m=1.0, field=None):
    """qwz(l, m) (src/zoo/models.jl:130-136): sigma_z m on site, [1 -i; -i -1]/2
    along axis 1 and [1 -1; 1 -1]/2 along axis 2, one spin-1/2 per site."""
    from .geometry import Translation
    mv = [float(m)] * len(lat) if np.ndim(m) == 0 else [float(x) for x in m]
    terms = [Onsite(mv, SIGMA_Z), Hopping(QWZ_X, [Translation((0, 0), (1, 0))]),
             Hopping(QWZ_Y, [Translation((0, 0), (0, 1))])]
    return OracleHamiltonian(lat, terms, field, internal=2)
