"""Oracle: Bravais lattice geometry, boundary routing and bond iteration.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates (paths relative to the LatticeModels.jl source tree):
  src/lattices/bravais/unitcell.jl:113-132   BravaisPointer, site_coords, ordering
  src/lattices/bravais/lattice.jl:101-111    add_bravaispointers! (last axis fastest)
  src/lattices/bravais/lattice.jl:161-172    span_unitcells, :axisN default translations
  src/lattices/bravais/nearestneighbor.jl:97-135,174-178  detect_nnhops / getnnbonds
  src/lattices/bravais/bonds.jl:10-116       BravaisTranslation, _destination_bp,
                                             BravaisSiteMapping
  src/core/bonds.jl:415-422                  iterate(::AbstractTranslation)
  src/core/boundaries.jl:4-25,155-164,251-289  nshifts, mappings, route,
                                             findfactor, resolve_site
  src/core/utils.jl:24-26                    cartesian_indices (first index fastest)
  src/zoo/lattices.jl:131,161                SquareLattice, HoneycombLattice
  src/zoo/models.jl:139-145                  honeycomb_2nn

Pins (tests/test_oracle_geometry.py): the reference doctests for site order
(src/core/latticevalue.jl:60-82), the square NN/NNN/NNNN lists
(src/core/lattice.jl:46-72, src/zoo/lattices.jl:37-62) and the honeycomb NN
list (src/lattices/bravais/nearestneighbor.jl:146-153); bond counts checked
against closed-form counts of open/periodic rectangles.
"""
import math
import itertools

SQRT3 = math.sqrt(3.0)


def isapprox(x, y, rtol=math.sqrt(2.220446049250313e-16)):
    """Julia's default `≈` for Float64 scalars (atol = 0)."""
    return x == y or abs(x - y) <= rtol * max(abs(x), abs(y))


def cartesian_indices(depth, m):
    """CartesianIndex(-d,...):CartesianIndex(d,...), first index fastest
    (src/core/utils.jl:24-26)."""
    rng = range(-depth, depth + 1)
    for tup in itertools.product(*([rng] * m)):
        yield tuple(reversed(tup))


class UnitCell:
    """2-D unit cell: `translations` columns a1, a2; `basis` list of 2-vectors
    (src/lattices/bravais/unitcell.jl:28-38)."""

    def __init__(self, translations, basis):
        self.t = translations          # ((a1x, a2x), (a1y, a2y)) row-major 2x2
        self.basis = basis             # [(bx, by), ...]

    def site_coords(self, latcoords, b):
        # basissites[:, b] + translations * latcoords  (unitcell.jl:119-122)
        j1, j2 = float(latcoords[0]), float(latcoords[1])
        tx = self.t[0][0] * j1 + self.t[0][1] * j2
        ty = self.t[1][0] * j1 + self.t[1][1] * j2
        bx, by = self.basis[b - 1]
        return (bx + tx, by + ty)

    def unitvec_times(self, c):
        return (self.t[0][0] * c[0] + self.t[0][1] * c[1],
                self.t[1][0] * c[0] + self.t[1][1] * c[1])


def square_unitcell():
    # @bravaisdef SquareLattice N -> UnitCell(SMatrix{N,N}(I))  (zoo/lattices.jl:131)
    return UnitCell(((1.0, 0.0), (0.0, 1.0)), [(0.0, 0.0)])


def honeycomb_unitcell():
    # UnitCell([1 0.5; 0 √3/2], [0 0.5; 0 √3/6])  (zoo/lattices.jl:161)
    return UnitCell(((1.0, 0.5), (0.0, SQRT3 / 2)), [(0.0, 0.0), (0.5, SQRT3 / 6)])


class Translation:
    """BravaisTranslation(site_indices, translate_uc)  (bravais/bonds.jl:10-25).
    site_indices == (0, 0) means 'any basis index, unchanged'."""

    def __init__(self, site_indices, tr):
        self.si = tuple(site_indices)
        self.tr = tuple(int(v) for v in tr)

    def key(self):
        return (self.si, self.tr)

    def __eq__(self, other):
        return self.key() == other.key()

    def __repr__(self):
        return "Translation(%s, %s)" % (self.si, self.tr)

    def inv(self):
        a, b = self.si
        return Translation((b, a), tuple(-v for v in self.tr))

    def destination(self, ptr):
        """_destination_bp (bravais/bonds.jl:93-102): ptr = (latcoords, b)."""
        lat, b = ptr
        if self.si == (0, 0):
            nb = b
        else:
            if self.si[0] != b:
                return None
            nb = self.si[1]
        tr = self.tr + (0,) * (len(lat) - len(self.tr))
        if len(self.tr) > len(lat) and any(self.tr[len(lat):]):
            return None
        return (tuple(l + t for l, t in zip(lat, tr)), nb)


def norm2(v):
    return math.sqrt(v[0] * v[0] + v[1] * v[1])


def detect_nnhops(uc, depth=2, limit=3):
    """nearestneighbor.jl:97-131 restated: distance shells and translation lists."""
    lens = []
    lists = []
    nb = len(uc.basis)
    for c in cartesian_indices(depth, 2):
        U = uc.unitvec_times(c)
        for i in range(1, nb + 1):
            for j in range(i, nb + 1):
                if i == j:
                    if i > 2:
                        continue
                    nzc = next((k for k, v in enumerate(c) if v != 0), None)
                    if nzc is None:
                        continue
                    if c[nzc] < 1:
                        continue
                bi, bj = uc.basis[i - 1], uc.basis[j - 1]
                R = (U[0] - bi[0] + bj[0], U[1] - bi[1] + bj[1])
                r = norm2(R)
                k = next((l for l in range(len(lens)) if lens[l] > r or isapprox(lens[l], r)), None)
                if k is None or not isapprox(lens[k], r):
                    if k is None:
                        k = len(lens)
                    if k + 1 > limit:
                        continue
                    lens.insert(k, r)
                    lists.insert(k, [])
                t = Translation((0, 0), c) if i == j else Translation((i, j), c)
                if t not in lists[k]:
                    lists[k].append(t)
    return lens, lists


def default_nnbonds(uc):
    """getnnbonds(::BravaisLattice)  (nearestneighbor.jl:174-178)."""
    lens, lists = detect_nnhops(uc)
    m = min(3, len(lens))
    return lens[:m], lists[:m]


# honeycomb_2nn (zoo/models.jl:139-145); `axis=a, dist=d` -> one-hot * d
HONEYCOMB_2NN = [
    Translation((1, 1), (1, 0)),
    Translation((2, 2), (-1, 0)),
    Translation((1, 1), (0, -1)),
    Translation((2, 2), (0, 1)),
    Translation((1, 1), (-1, 1)),
    Translation((2, 2), (1, -1)),
]


class Boundary:
    """TwistedBoundary(translation, theta)  (core/boundaries.jl:46-59)."""

    def __init__(self, translation, theta=0.0):
        self.translation = translation
        self.theta = float(theta)


class BravaisLattice:
    """A rectangular span of unit cells with optional twisted/periodic
    boundaries along the :axis1/:axis2 default translations."""

    def __init__(self, uc, dims, periodic=(False, False), twists=(0.0, 0.0), kind=""):
        self.uc = uc
        self.kind = kind
        self.dims = tuple(dims)
        nb = len(uc.basis)
        # add_bravaispointers! : CartesianIndices(reverse(axes)) -> last axis fastest,
        # then basis index (bravais/lattice.jl:101-111)
        self.pointers = []
        for j1 in range(1, dims[0] + 1):
            for j2 in range(1, dims[1] + 1):
                for b in range(1, nb + 1):
                    self.pointers.append(((j1, j2), b))
        self.index = {p: i for i, p in enumerate(self.pointers)}
        self.coords = [uc.site_coords(p[0], p[1]) for p in self.pointers]
        # default translations :axis1 => Bravais[L1,0], :axis2 => Bravais[0,L2]
        # (bravais/lattice.jl:167-170); boundaries keep only `true`/twist ones
        # (core/boundaries.jl:62-65,128-129)
        self.boundaries = []
        for ax in range(2):
            if periodic[ax]:
                tr = [0, 0]
                tr[ax] = dims[ax]
                self.boundaries.append(Boundary(Translation((0, 0), tr), twists[ax]))
        self.nn_lens, self.nn_lists = default_nnbonds(uc)

    def __len__(self):
        return len(self.pointers)

    def site_index(self, ptr):
        return self.index.get(ptr)

    def shift(self, ptr, tr, n):
        """nshifts(site, tr, n): site -= tr, |n| times; tr negated if n < 0
        (core/boundaries.jl:4-11)."""
        if n == 0:
            return ptr
        if n < 0:
            tr = tr.inv()
        for _ in range(abs(n)):
            ptr = tr.inv().destination(ptr)
            if ptr is None:
                return None
        return ptr

    def resolve(self, ptr):
        """resolve_site(::LatticeWithMetadata, site) (core/boundaries.jl:280-289):
        returns (index, unwrapped_ptr, twist_factor) or None."""
        if ptr is None:
            return None
        i = self.site_index(ptr)
        if i is not None:
            return i, ptr, complex(1.0, 0.0)
        m = len(self.boundaries)
        if m == 0:
            return None
        for tup in cartesian_indices(1, m):          # mappings(bcs, site)
            new = ptr
            for k in range(m):
                new = self.shift(new, self.boundaries[k].translation, tup[k])
                if new is None:
                    break
            if new is None:
                continue
            i = self.site_index(new)
            if i is not None:
                if all(t == 0 for t in tup):
                    return i, ptr, complex(1.0, 0.0)
                # findfactor for TwistedBoundary: exp(im * sum(theta_k * tup_k))
                ang = 0.0
                first = True
                for k in range(m):
                    v = self.boundaries[k].theta * tup[k]
                    ang = v if first else ang + v
                    first = False
                fac = complex(math.cos(ang), math.sin(ang)) if ang != 0.0 else complex(1.0, 0.0)
                return i, ptr, fac
        return None

    def translation_bonds(self, t):
        """iterate(::AbstractTranslation) (core/bonds.jl:415-422): for each site
        in index order, yield (i, j, r_i, r_j_unwrapped, twist_j)."""
        out = []
        for i, p in enumerate(self.pointers):
            dest = t.destination(p)
            r = self.resolve(dest)
            if r is None:
                continue
            j, uptr, fac = r
            out.append((i, j, self.coords[i], self.uc.site_coords(uptr[0], uptr[1]), fac))
        return out

    def nn_translations(self, order):
        """adapt_bonds(NearestNeighbor(order), lat) -> its BravaisSiteMapping
        translations (nearestneighbor.jl:87-93,133-135)."""
        if order <= len(self.nn_lists):
            return list(self.nn_lists[order - 1])
        lens, lists = detect_nnhops(self.uc, 1 + math.ceil(math.sqrt(order)), order)
        return list(lists[order - 1])


def SquareLattice(Lx, Ly, periodic=(False, False), twists=(0.0, 0.0)):
    return BravaisLattice(square_unitcell(), (Lx, Ly), periodic, twists, kind="square")


def HoneycombLattice(L1, L2, periodic=(False, False), twists=(0.0, 0.0)):
    return BravaisLattice(honeycomb_unitcell(), (L1, L2), periodic, twists, kind="honeycomb")


class SiteDistance:
    """SiteDistance(f) (src/core/bonds.jl:136-162): pairs at a distance d with
    f(d); on a lattice with boundaries the distance to the nearest image
    (`sitedistance`, :120-131 -> `translate_to_nearest`)."""

    def __init__(self, f):
        self.f = f


def site_distance(lat, i, j):
    """sitedistance(lat, site_i, site_j) with the boundary images."""
    ri, rj = lat.coords[i], lat.coords[j]
    best = math.hypot(rj[0] - ri[0], rj[1] - ri[1])
    if lat.boundaries:
        m = len(lat.boundaries)
        for tup in cartesian_indices(1, m):
            sx, sy = rj[0], rj[1]
            for k in range(m):
                v = lat.uc.unitvec_times(lat.boundaries[k].translation.tr)
                sx += tup[k] * v[0]
                sy += tup[k] * v[1]
            best = min(best, math.hypot(sx - ri[0], sy - ri[1]))
    return best


class AdjacencyMatrix:
    """AdjacencyMatrix (src/core/bonds.jl:164-345): a symmetric boolean
    relation on the sites, diagonal dropped (:203-212).  Constructors: from
    bonds (Bravais translations, each resolved pair, boundary wraps included;
    SiteDistance; other adjacency matrices) (:300-319), from a predicate
    f(r_i, r_j) over i < j (:327-345), from pairs set one by one
    (setindex!, :258-268).  As hopping bonds it is iterated as (i, j), j > i
    ascending for each i (`_destination_inds`, :247-257), and resolved by index
    (`resolvestyle = ResolveIndices()`, :246): r_j is the site's own
    coordinate and there is no twist."""

    def __init__(self, lat, *bonds):
        self.lat = lat
        self.pairs = set()
        for b in bonds:
            if isinstance(b, AdjacencyMatrix):
                self.pairs |= b.pairs
            elif isinstance(b, SiteDistance):
                for i in range(len(lat)):
                    for j in range(i + 1, len(lat)):
                        if b.f(site_distance(lat, i, j)):
                            self.pairs.add((i, j))
            else:
                for (i, j, _ri, _rj, _tw) in lat.translation_bonds(b):
                    if i != j:
                        self.pairs.add((min(i, j), max(i, j)))

    @classmethod
    def from_predicate(cls, lat, f):
        am = cls(lat)
        for i in range(len(lat)):
            for j in range(i + 1, len(lat)):
                if f(lat.coords[i], lat.coords[j]):
                    am.pairs.add((i, j))
        return am

    def set(self, i, j, v=True):
        if i == j:
            return
        p = (min(i, j), max(i, j))
        if v:
            self.pairs.add(p)
        else:
            self.pairs.discard(p)

    def get(self, i, j):
        return (min(i, j), max(i, j)) in self.pairs

    def union(self, *others):
        return AdjacencyMatrix(self.lat, self, *others)

    def adjacentsites(self, i):
        return sorted({b if a == i else a for a, b in self.pairs if i in (a, b)})

    def bonds(self):
        """(i, j, r_i, r_j, twist) in iteration order."""
        return [(i, j, self.lat.coords[i], self.lat.coords[j], complex(1.0, 0.0)) for i, j in sorted(self.pairs)]
