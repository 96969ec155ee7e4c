"""Shared builders that make the same physical model on the oracle side and on
the lmhip (device) side from one parameter set."""
import numpy as np

import lmhip
from oracle import fields as OF
from oracle import geometry as OG
from oracle import hamiltonian as OH


def lattices(kind, dims, periodic=(False, False)):
    bnd = tuple((ax, True) for ax, p in zip(("axis1", "axis2"), periodic) if p)
    if kind == "square":
        return OG.SquareLattice(*dims, periodic=periodic), lmhip.SquareLattice(*dims, boundaries=bnd)
    return OG.HoneycombLattice(*dims, periodic=periodic), lmhip.HoneycombLattice(*dims, boundaries=bnd)


def fields_pair(spec):
    """spec: list of (kind, a, px, py) with kind in landau/symmetric/axial/singular."""
    of, lf = [], []
    for kind, a, px, py in spec:
        if kind == "landau":
            of.append(OF.LandauGauge(a))
            lf.append(lmhip.LandauGauge(a))
        elif kind == "symmetric":
            of.append(OF.SymmetricGauge(a))
            lf.append(lmhip.SymmetricGauge(a))
        else:
            of.append(OF.PointFlux(a, (px, py), gauge=kind))
            lf.append(lmhip.PointFlux(a, (px, py), gauge=kind))
    if not of:
        return OF.NoField(), lmhip.NoField()
    if len(of) == 1:
        return of[0], lf[0]
    return OF.FieldSum(*of), lmhip.FieldSum(*lf)


class Model:
    """Tight-binding (square) or Haldane (honeycomb) model on both sides."""

    def __init__(self, kind, dims, periodic=(False, False), onsite=None, t1=1.0, t2=0.0, m=0.0):
        self.kind = kind
        self.olat, self.llat = lattices(kind, dims, periodic)
        self.onsite = None if onsite is None else np.asarray(onsite, dtype=np.float64)
        self.t1, self.t2, self.m = t1, t2, m

    @property
    def n(self):
        return len(self.llat)

    def oracle_h(self, field_spec):
        of, _ = fields_pair(field_spec)
        if self.kind == "square":
            return OH.tightbinding_hamiltonian(self.olat, onsite=self.onsite, t1=self.t1, t2=self.t2, field=of)
        return OH.haldane(self.olat, self.t1, self.t2, self.m, onsite=self.onsite, field=of)

    def device_h(self, field_spec):
        _, lf = fields_pair(field_spec)
        extra = () if self.onsite is None else (self.onsite,)
        if self.kind == "square":
            return lmhip.tightbinding_hamiltonian(self.llat, *extra, t1=self.t1, t2=self.t2, field=lf)
        return lmhip.haldane(self.llat, self.t1, self.t2, self.m, *extra, field=lf)


def bond_dict(bonds_i, bonds_j, values):
    return {(int(i), int(j)): v for i, j, v in zip(bonds_i, bonds_j, values)}
