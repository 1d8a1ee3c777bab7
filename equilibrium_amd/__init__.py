"""equilibrium_amd -- MI355X-native Stam stable-fluids frame step.

The compute path is libeqfluid.so (HIP kernels for gfx950 behind the C ABI of
include/eqfluid.h); this package is its Python front end.
"""
from . import _abi  # noqa: F401
from .fluid import (  # noqa: F401
    ContainerWall, CurrentSimulation, Fluid, FluidConfigs, GsMode, ObstaclesType, Rectangle,
    SimulationConfigs)
