"""``SmoothParticlesArea``: the demo's dye field over the HIP simulator.

Mirrors demo/smooth_particles_area.py:15-211: a scalar field on its own
(width x height) grid, advected by a FluidSimulator's velocity
(shader.AdvectParticle.comp) and fed by cone splats clamped to [0, 255]
(shader.AddParticle.comp).  The field lives inside the simulator object;
``update`` runs the advection kernel on the simulator's stream, after the
fluid step, as the demo does (demo/simulation_demo.py:222-223).
"""
from __future__ import annotations

from .. import _lib
from .fluid_simulator import FluidSimulator


class SmoothParticlesArea:
    PARTICLES_IN = 0
    PARTICLES_OUT = 1

    _speed = 500.0
    _dissipation = 1.0

    simulate = True

    def __init__(self, width: int, height: int,
                 fluid_simulation: FluidSimulator, vertex_layout=None):
        self.fluid_simulation = fluid_simulation
        self.vertex_layout = vertex_layout
        self._width = int(width)
        self._height = int(height)
        fluid_simulation.enable_dye(self._width, self._height)
        # the demo advects the dye itself, after FluidSimulator.update
        fluid_simulation._dye_in_step = False

    @property
    def speed(self):
        return self._speed

    @speed.setter
    def speed(self, value):
        if value > 0:
            self._speed = value
        else:
            raise ValueError("'Speed' should be greater than zero")

    @property
    def dissipation(self):
        return self._dissipation

    @dissipation.setter
    def dissipation(self, value):
        if value > 0:
            self._dissipation = value
        else:
            raise ValueError("'Dissipation' should be grater than zero")

    def add_particles(self, position: tuple, radius: float, strength: float):
        """smooth_particles_area.py:70-83."""
        if self.simulate:
            self.fluid_simulation.add_dye(position, radius, strength)

    def update(self, time_delta: float):
        """smooth_particles_area.py:85-104."""
        if self.simulate:
            sim = self.fluid_simulation
            sim._dye_speed = self._speed
            sim._dye_dissipation = self._dissipation
            sim._push_params()
            _lib.call("nx_stage_advect_dye", sim._alive(), time_delta)

    def download(self):
        return self.fluid_simulation.download("dye")

    def destroy(self):
        pass
