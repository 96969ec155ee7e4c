"""Bravais lattices of the host mirror: site order, coordinates, bonds.

Mirrors LatticeModels.jl's lattice layer as far as the evolution path needs:
  SquareLattice / HoneycombLattice     src/zoo/lattices.jl:131,161
  site order (last lattice axis fastest, then the basis index)
                                       src/lattices/bravais/lattice.jl:101-111
  coordinates b + j1 a1 + j2 a2        src/lattices/bravais/unitcell.jl:119-122
  default NN shells                    src/lattices/bravais/nearestneighbor.jl:97-135,174-178
  `boundaries=(:axis1 => true, ...)`   src/core/boundaries.jl:46-65,138-148
  bond iteration with boundary routing src/core/bonds.jl:415-422,
                                       src/core/boundaries.jl:251-289
Everything is vectorised over sites with numpy (no per-site Python loops), so
building bond lists at N = 10^5 takes milliseconds.
"""
import math

import numpy as np

SQRT3 = math.sqrt(3.0)


class Translation:
    """Bravais translation `site_indices => translate_uc`
    (src/lattices/bravais/bonds.jl:10-25); site_indices (0, 0) keeps the basis
    index."""

    __slots__ = ("si", "tr")

    def __init__(self, si, tr):
        self.si = (int(si[0]), int(si[1]))
        self.tr = (int(tr[0]), int(tr[1]) if len(tr) > 1 else 0)

    def __eq__(self, o):
        return isinstance(o, Translation) and (self.si, self.tr) == (o.si, o.tr)

    def __hash__(self):
        return hash((self.si, self.tr))

    def __repr__(self):
        if self.si == (0, 0):
            return "Bravais[%d, %d]" % self.tr
        return "%d => %d, [%d, %d]" % (self.si + self.tr)


def Bravais(d1, d2=0):
    return Translation((0, 0), (d1, d2))


# Default nearest-neighbour shells produced by the reference's detect_nnhops
# for these unit cells (the square list is the doctest at
# src/zoo/lattices.jl:52-61, the honeycomb first shell the doctest at
# src/lattices/bravais/nearestneighbor.jl:146-153).
_SHELLS = {
    "square": [
        [Translation((0, 0), (1, 0)), Translation((0, 0), (0, 1))],
        [Translation((0, 0), (1, -1)), Translation((0, 0), (1, 1))],
        [Translation((0, 0), (2, 0)), Translation((0, 0), (0, 2))],
    ],
    "honeycomb": [
        [Translation((1, 2), (0, -1)), Translation((1, 2), (-1, 0)), Translation((1, 2), (0, 0))],
        [Translation((0, 0), (1, -1)), Translation((0, 0), (1, 0)), Translation((0, 0), (0, 1))],
        [Translation((1, 2), (-1, -1)), Translation((1, 2), (1, -1)), Translation((1, 2), (-1, 1))],
    ],
}

_UNITCELLS = {
    # translations (columns a1, a2) and basis sites
    "square": (np.array([[1.0, 0.0], [0.0, 1.0]]), np.array([[0.0, 0.0]])),
    "honeycomb": (np.array([[1.0, 0.5], [0.0, SQRT3 / 2]]), np.array([[0.0, 0.0], [0.5, SQRT3 / 6]])),
}


class Bonds:
    """Bond list in iteration order: i, j (0-based), r_i, unwrapped r_j, twist."""

    def __init__(self, i, j, ri, rj, twist):
        self.i, self.j, self.ri, self.rj, self.twist = i, j, ri, rj, twist

    def __len__(self):
        return len(self.i)

    @staticmethod
    def concat(parts):
        if not parts:
            z = np.zeros(0, dtype=np.int64)
            return Bonds(z, z, np.zeros((0, 2)), np.zeros((0, 2)), np.zeros(0, dtype=np.complex128))
        return Bonds(*(np.concatenate([getattr(p, a) for p in parts]) for a in ("i", "j", "ri", "rj", "twist")))


class BravaisLattice:
    def __init__(self, kind, dims, boundaries=()):
        if kind not in _UNITCELLS:
            raise ValueError("unsupported lattice kind %r" % kind)
        self.kind = kind
        self.dims = (int(dims[0]), int(dims[1]))
        if min(self.dims) < 1:
            raise ValueError("Invalid range: lattice dimensions must be positive")
        self.T, self.B = _UNITCELLS[kind]
        self.nb = len(self.B)
        L1, L2 = self.dims
        j1, j2, b = np.meshgrid(np.arange(1, L1 + 1), np.arange(1, L2 + 1), np.arange(1, self.nb + 1), indexing="ij")
        self.j1, self.j2, self.b = j1.ravel(), j2.ravel(), b.ravel()
        self.coords = self.site_coords(self.j1, self.j2, self.b)
        # boundaries: {axis: twist angle}; `True` is periodic (twist 0)
        self.boundaries = {}
        for spec in boundaries:
            ax, val = spec
            axis = {"axis1": 0, "axis2": 1, 0: 0, 1: 1}[ax]
            if val is False or val is None:
                continue
            self.boundaries[axis] = 0.0 if val is True else float(val)

    def __len__(self):
        return len(self.j1)

    def site_coords(self, j1, j2, b):
        f1 = np.asarray(j1, dtype=np.float64)
        f2 = np.asarray(j2, dtype=np.float64)
        tx = self.T[0, 0] * f1 + self.T[0, 1] * f2
        ty = self.T[1, 0] * f1 + self.T[1, 1] * f2
        bb = self.B[np.asarray(b) - 1]
        return np.stack([bb[..., 0] + tx, bb[..., 1] + ty], axis=-1)

    def index_of(self, j1, j2, b):
        return ((np.asarray(j1) - 1) * self.dims[1] + (np.asarray(j2) - 1)) * self.nb + (np.asarray(b) - 1)

    def nn_shell(self, order):
        return list(_SHELLS[self.kind][order - 1])

    def translation_bonds(self, t):
        """Bonds of one translation over all sites in index order."""
        src = np.arange(len(self))
        if t.si != (0, 0):
            src = src[self.b == t.si[0]]
            nb = np.full(len(src), t.si[1])
        else:
            nb = self.b[src]
        d1 = self.j1[src] + t.tr[0]
        d2 = self.j2[src] + t.tr[1]
        ok = np.ones(len(src), dtype=bool)
        w = [d1.copy(), d2.copy()]
        tups = [np.zeros(len(src), dtype=np.int64), np.zeros(len(src), dtype=np.int64)]
        for ax, (dest, L) in enumerate(zip((d1, d2), self.dims)):
            inside = (dest >= 1) & (dest <= L)
            if ax in self.boundaries:
                # first of the shifts site - tup*L, tup = -1, 0, 1, landing inside
                got = inside.copy()
                for tup in (-1, 1):
                    cand = dest - tup * L
                    take = ~got & (cand >= 1) & (cand <= L)
                    w[ax] = np.where(take, cand, w[ax])
                    tups[ax] = np.where(take, tup, tups[ax])
                    got |= take
                ok &= got
            else:
                ok &= inside
        src, nb, d1, d2 = src[ok], nb[ok], d1[ok], d2[ok]
        w1, w2 = w[0][ok], w[1][ok]
        t1, t2 = tups[0][ok], tups[1][ok]
        j = self.index_of(w1, w2, nb)
        ang = self.boundaries.get(0, 0.0) * t1 + self.boundaries.get(1, 0.0) * t2
        wrapped = (t1 != 0) | (t2 != 0)
        twist = np.where(wrapped & (ang != 0), np.exp(1j * ang), 1.0 + 0j)
        return Bonds(src.astype(np.int64), j.astype(np.int64), self.coords[src],
                     self.site_coords(d1, d2, nb), twist.astype(np.complex128))

    def bonds(self, translations):
        return Bonds.concat([self.translation_bonds(t) for t in translations])

    def translation_vectors(self):
        """Boundary translation vectors r(tup) = -(tup1 L1 a1 + tup2 L2 a2) in
        cartesian order over tup in {-1,0,1}^m (src/zoo/magneticfields.jl:247-259)."""
        axes = sorted(self.boundaries)
        m = len(axes)
        out = []
        for flat in range(3 ** m):
            tup = [(flat // 3 ** k) % 3 - 1 for k in range(m)]
            r = np.zeros(2)
            for ax, tv in zip(axes, tup):
                r -= tv * self.dims[ax] * self.T[:, ax]
            out.append(r)
        return out

    def __repr__(self):
        return "%d-site %sLattice" % (len(self), self.kind.capitalize())


def SquareLattice(Lx, Ly, boundaries=()):
    return BravaisLattice("square", (Lx, Ly), boundaries)


def HoneycombLattice(L1, L2, boundaries=()):
    return BravaisLattice("honeycomb", (L1, L2), boundaries)


class SiteDistance:
    """SiteDistance(f) (src/core/bonds.jl:136-162): the pairs of sites whose
    distance d satisfies f(d) -- on a periodic lattice the distance to the
    nearest periodic image (`sitedistance`, :120-131)."""

    def __init__(self, f):
        self.f = f


def _min_image_distance(lat, i, js):
    """|r_j - r_i| for the sites js, minimised over the boundary images."""
    d = lat.coords[js] - lat.coords[i]
    best = np.sqrt(np.sum(d * d, axis=1))
    if lat.boundaries:
        for r in lat.translation_vectors():
            e = d + r
            best = np.minimum(best, np.sqrt(np.sum(e * e, axis=1)))
    return best


class AdjacencyMatrix:
    """AdjacencyMatrix(lat[, bonds...]) (src/core/bonds.jl:164-345): a
    symmetric set of bonds between sites of `lat`, no site bonded to itself.
    Built from Bravais translations (their resolved site pairs, boundary wraps
    included), a SiteDistance, other adjacency matrices, a boolean matrix
    (`mat=`), a predicate on site coordinates (`from_predicate`) or an
    operator's bond pattern (`from_operator`); edited with am[i, j] = v (site
    indices).  As hopping bonds (`amp => am`) it is iterated as (i, j) with
    j > i ascending for every i (`_destination_inds`, :247-257) and resolved by
    index: the Peierls phase runs between the two sites' own coordinates,
    with no boundary unwrapping or twist (`resolvestyle = ResolveIndices()`)."""

    def __init__(self, lat, *bonds, mat=None):
        self.lat = lat
        self._pairs = set()
        self._key = None
        if mat is not None:
            m = np.asarray(mat.todense() if hasattr(mat, "todense") else mat, dtype=bool)
            if m.shape != (len(lat), len(lat)):
                raise ValueError("matrix of size %s for %d sites" % (m.shape, len(lat)))
            ii, jj = np.nonzero(m | m.T)
            self._pairs.update((int(a), int(b)) for a, b in zip(ii, jj) if a < b)
        for b in bonds:
            self._add_bonds(b)

    def _add_bonds(self, b):
        self._key = None
        if isinstance(b, AdjacencyMatrix):
            if b.lat is not self.lat:
                raise ValueError("adjacency matrices on different lattices")
            self._pairs |= b._pairs
        elif isinstance(b, SiteDistance):
            n = len(self.lat)
            for i in range(n - 1):
                js = np.arange(i + 1, n)
                keep = np.array([bool(b.f(float(x))) for x in _min_image_distance(self.lat, i, js)], dtype=bool)
                self._pairs.update((i, int(j)) for j in js[keep])
        else:
            if isinstance(b, Translation):
                trs = [b]
            elif hasattr(b, "order"):              # NearestNeighbor(order)
                trs = self.lat.nn_shell(b.order)
            else:
                trs = list(b)
            bd = self.lat.bonds(trs)
            self._pairs.update((min(int(a), int(c)), max(int(a), int(c))) for a, c in zip(bd.i, bd.j) if a != c)

    @classmethod
    def from_predicate(cls, lat, f):
        """AdjacencyMatrix(f, lat) (:327-345): f(r_i, r_j) for every pair i < j."""
        am = cls(lat)
        xy = lat.coords
        am._key = None
        for i in range(len(lat)):
            for j in range(i + 1, len(lat)):
                if f(xy[i], xy[j]):
                    am._pairs.add((i, j))
        return am

    @classmethod
    def from_operator(cls, H):
        """AdjacencyMatrix(op) (src/operators/latticeutils.jl:78-98): the site
        pairs an operator couples."""
        am = cls(H.lattice)
        b = H.bonds
        am._pairs.update((min(int(a), int(c)), max(int(a), int(c))) for a, c in zip(b.i, b.j) if a != c)
        return am

    def __setitem__(self, ij, v):
        i, j = int(ij[0]), int(ij[1])
        if i == j:
            return
        self._key = None
        p = (min(i, j), max(i, j))
        if v:
            self._pairs.add(p)
        else:
            self._pairs.discard(p)

    def __getitem__(self, ij):
        i, j = int(ij[0]), int(ij[1])
        return (min(i, j), max(i, j)) in self._pairs

    def union(self, *others):
        out = AdjacencyMatrix(self.lat, self)
        for o in others:
            out._add_bonds(o)
        return out

    def adjacentsites(self, i):
        return sorted({b if a == i else a for a, b in self._pairs if i in (a, b)})

    def pairs(self):
        """(i, j), i < j, in the reference's iteration order."""
        return sorted(self._pairs)

    def __len__(self):
        return len(self._pairs)

    def key(self):
        if self._key is None:
            p = np.array(self.pairs(), dtype=np.int64).reshape(-1, 2)
            self._key = ("adjacency", id(self.lat), p.tobytes())
        return self._key

    def bonds(self):
        p = np.array(self.pairs(), dtype=np.int64).reshape(-1, 2)
        i, j = p[:, 0], p[:, 1]
        return Bonds(i, j, self.lat.coords[i], self.lat.coords[j], np.ones(len(i), dtype=np.complex128))
