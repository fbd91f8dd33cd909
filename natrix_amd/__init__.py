"""natrix_amd: the Natrix fluid step as hand-written gfx950 HIP kernels.

``natrix_amd.core.FluidSimulator`` keeps the reference's
``natrix.core.fluid_simulator.FluidSimulator`` API; the compute runs in
``libnatrix_hip.so`` (built by ``python -m natrix_amd.build``).
"""
from .core import DeviceBuffer, FluidSimulator, SmoothParticlesArea  # noqa: F401
