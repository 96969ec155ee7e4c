"""lmhip: host mirror of LatticeModels.jl's density-matrix evolution API over
liblmhip.so (MI355X / gfx950 HIP kernels).

    from lmhip import *
    l = SquareLattice(10, 10)
    h = lambda B: tightbinding_hamiltonian(l, field=PointFlux(B, (5.5, 5.5)))
    P0 = densitymatrix(h(0), mu=0)
    for (P, H, t) in Evolution(lambda t: h(0.2 * min(t, 10) / 10), P0)(times):
        rho = localdensity(P); J = DensityCurrents(H, P).values()

The package requires the built library (python latticemodels.jl_amd/build.py)
and a HIP device; there is no CPU fallback.
"""
from ._lib import LMHError, lib
from .lattice import (AdjacencyMatrix, Bravais, BravaisLattice, HoneycombLattice, SiteDistance, SquareLattice,
                      Translation)
from .field import (FieldSum, LandauGauge, NoField, PointFlux, PointFluxes, SymmetricGauge, strength_gradient,
                    strength_index)
from .operators import (HONEYCOMB_2NN, Hamiltonian, NearestNeighbor, SparseHamiltonian, SpinBasis,
                        construct_hamiltonian, haldane, qwz, tightbinding_hamiltonian)
from .timesequence import DeviceSeries, TimeSequence, currentsfrom, currentsfromto
from .adjoint import (EvolutionTape, Gradients, SparseCotangent, currents_cotangent, hamiltonian_derivative,
                      localdensity_cotangent, rrule_evolution)
from .evolution import (BatchedEvolution, BoseEinstein, Context, DensityCurrents, DensityMatrix, Eigensystem,
                        Evolution, EvolutionSolver, EvolutionTimestamp, FermiDirac, HIPExp, HIPKrylovExp, Kets,
                        default_context, densitymatrix, diagonalize, localdensity, observables)
