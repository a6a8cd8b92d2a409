"""gfx950 (MI355X) implementation of tackandr/CFD-demo's degree-4 nodal DG
Euler hot path: RHS, troubled-cell and positivity limiting, CFL step and
SSP-RK3 stages as hand-written HIP kernels behind the C ABI of libcfdhip.so.

Exports the reference's public names (``src/cfd_demo/__init__.py:10-32``).
"""

from . import batched  # noqa: F401  (registers the batched C-ABI entry points)
from .adjoint import DiscreteAdjoint, evolve
from .boundary import FixedState, Periodic, SlipWall, Transmissive
from .cases import SodShockTube, four_quadrant, isentropic_vortex, sod_shock_tube, vortex_bumps
from .mesh import TriangleMesh, rectangle_mesh
from .physics import INTERFACE_FLUXES, LIMITER_FUNCTIONS, get_interface_flux, get_limiter_function
from .solver import EulerSolver, RunResult, SolverConfig, SolverError, StepReport, apply_limiters

__all__ = [
    "DiscreteAdjoint",
    "INTERFACE_FLUXES",
    "LIMITER_FUNCTIONS",
    "EulerSolver",
    "FixedState",
    "Periodic",
    "RunResult",
    "SlipWall",
    "SodShockTube",
    "SolverConfig",
    "SolverError",
    "StepReport",
    "Transmissive",
    "TriangleMesh",
    "apply_limiters",
    "evolve",
    "four_quadrant",
    "get_interface_flux",
    "get_limiter_function",
    "isentropic_vortex",
    "rectangle_mesh",
    "sod_shock_tube",
    "vortex_bumps",
]
