"""pyshield_amd -- MI355X-native GFDL cloud microphysics column physics.

Drop-in for pySHiELD's microphysics path (``Microphysics(state, timestep)``)
and the column-local state-update stencils around it, on hand-written HIP
kernels for gfx950 behind the C ABI in include/shield_mp.h.
"""

from ._config import PHYSICS_PACKAGES, PhysicsConfig
from .microphysics import (GridData, Microphysics, MicrophysicsFunction, MicrophysicsState, QuantityFactory,
                           StencilFactory, adjoint, forward, forward_update, microphysics_ad)
from .physics import Physics, physics_prologue

__all__ = [
    "PHYSICS_PACKAGES", "PhysicsConfig", "Microphysics", "MicrophysicsState",
    "StencilFactory", "QuantityFactory", "GridData", "forward", "forward_update", "Physics",
    "physics_prologue", "adjoint", "MicrophysicsFunction", "microphysics_ad",
]
__version__ = "0.2.0"
