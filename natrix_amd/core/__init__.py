from .fluid_simulator import DeviceBuffer, FluidSimulator  # noqa: F401
from .particles_area import SmoothParticlesArea  # noqa: F401
