"""MI355X-native (gfx950) single-device Chorin projection step and rollout,
a drop-in for the single-core path of
helmhlotz/ns_solver_jax_differential_distributed (``NS_solver.py``).

The solver API lives in ``ns_solver_jax_differential_distributed_amd.NS_solver``
and is re-exported here.
"""
from .NS_solver import (  # noqa: F401
    CONFIG, BatchedNSSolver, AdvectionScheme, CentralAdvection, ChorinProjector, DifferentialOps,
    GaussianForcingField, Grid, PeriodicDifferentialOps, Phys, SimConfig,
    SingleCoreNSSolver, SpectralPoissonSolver, Time, UpwindAdvection, load_config,
    make_advection_scheme, make_grid, setup_logger,
)

__all__ = [
    "CONFIG", "BatchedNSSolver", "AdvectionScheme", "CentralAdvection", "ChorinProjector", "DifferentialOps",
    "GaussianForcingField", "Grid", "PeriodicDifferentialOps", "Phys", "SimConfig",
    "SingleCoreNSSolver", "SpectralPoissonSolver", "Time", "UpwindAdvection", "load_config",
    "make_advection_scheme", "make_grid", "setup_logger",
]
