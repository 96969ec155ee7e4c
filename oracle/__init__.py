"""CPU oracle for density-matrix time evolution of tight-binding lattice models.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library and the
``lmhip`` host package) may import, link or execute anything in this folder.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker (or the timed CPU baseline),
never as the thing measured or shipped.

The oracle is a restatement of LatticeModels.jl 1.0.7 (pure Julia; the
reference cannot be built or run here because no Julia toolchain exists in
this image).  Each function cites the reference file:line it follows.

Modules
-------
geometry     Bravais lattices (square, honeycomb), site order, coordinates,
             nearest-neighbour detection, boundary routing, bond iteration.
fields       Gauge fields: Landau, PointFlux (axial / singular), PointFluxes,
             field sums, periodic-boundary adaptation.
hamiltonian  Operator-builder semantics (accumulation, auto-Hermitian), the
             on-site / hopping terms, tightbinding_hamiltonian, haldane.
numerics     ctypes bindings to ``lmoracle.c``: the reference's sparse
             Taylor exponential (myexp!), its U*P*U' step, plus observables,
             the T=0 Fermi projector and the Evolution iterator semantics.

Parity pins for each module are listed in the module docstrings; functions
without a pin say "parity unpinned".
"""
