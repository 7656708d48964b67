"""hydroflux_amd: hydroflux solver-2d's second-order time step on MI355X (gfx950).

The compute path is hand-written HIP in hydroflux_amd/csrc behind the C ABI
declared in include/hydroflux_amd.h; this package is the host-side mirror of
the reference's API over that ABI.
"""
from .solver2d import (  # noqa: F401
    GRAVITY,
    H_DRY,
    H_VEL,
    Boundaries2D,
    Boundary,
    DegenerateDt,
    DeviceError,
    Grid2D,
    Integrator,
    InvalidConfig,
    Mesh2D,
    PointSource,
    ShapeMismatch,
    SimError,
    Simulation,
    SimulationConfig,
    StepBudgetExhausted,
    max_depth_exceeds_relief,
)

__all__ = [
    "GRAVITY", "H_DRY", "H_VEL", "Boundaries2D", "Boundary", "DegenerateDt", "DeviceError", "Grid2D",
    "Integrator", "InvalidConfig", "Mesh2D", "PointSource", "ShapeMismatch", "SimError", "Simulation",
    "SimulationConfig", "StepBudgetExhausted", "max_depth_exceeds_relief",
]
