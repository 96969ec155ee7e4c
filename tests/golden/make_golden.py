"""Writes tests/golden/cfg1.npz with the oracle (the reference's algorithm
restated in oracle/): the repository's own small CPU example,
SquareLattice(10, 10), open boundaries, t1 = 1, PointFlux through (5.5, 5.5)
ramped as B(t) = 0.2 min(t, 10)/10, P0 = densitymatrix(h(0), mu = 0) at T = 0,
times 0:0.1:20 (200 steps).

Stored: P0 (the committed initial state: the spectrum has ten E = 0 modes whose
filling depends on LAPACK rounding, so both sides must start from this array),
the time grid, bond list, K per step, localdensity and DensityCurrents after
every step, and P after the last step.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fields, geometry, hamiltonian, numerics  # noqa: E402


def main():
    lat = geometry.SquareLattice(10, 10)
    tau = 10.0

    def h(t):
        return hamiltonian.tightbinding_hamiltonian(lat, field=fields.PointFlux(0.2 * min(t, tau) / tau, (5.5, 5.5)))

    P0, _E = numerics.densitymatrix(h(0.0).dense(), mu=0.0)
    ts = numerics.julia_range(0.0, 0.1, 20.0)
    ev = numerics.Evolution(h, P0)
    rhos, Js, Ks = [], [], []
    bonds = None
    for k, (H, P, t) in enumerate(ev(ts)):
        if k == 0:
            continue
        rhos.append(numerics.localdensity(P))
        Js.append(numerics.density_currents(H, P))
        Ks.append(ev.K)
        bonds = H.bonds
    np.savez_compressed(os.path.join(HERE, "cfg1.npz"), P0=P0, ts=np.array(ts), K=np.array(Ks, dtype=np.int32),
                        rho=np.array(rhos), J=np.array(Js), P_final=np.array(ev.P),
                        bonds=np.array(bonds, dtype=np.int64))


if __name__ == "__main__":
    main()
