"""Dam-break setups of the benchmark workloads, assembled like
``examples/fluid/dam_break_2d.jl`` / ``dam_break_3d.jl``:
WendlandC2 kernel with h = 1.2 dx, StateEquationCole(rho0 = 1000, c = 20 sqrt(g H), exponent
7), ArtificialViscosityMonaghan(alpha = 0.02), DensityDiffusionMolteniColagrossi(delta = 0.1),
ContinuityDensity, AdamiPressureExtrapolation walls with clip_negative_pressure = true,
gravity -9.81 e_y, a RectangularTank with 3 dummy-particle layers and an open top, and a
lattice fluid block with hydrostatic density plus seeded uniform jitter of +-0.01 dx."""
from __future__ import annotations

import math
from dataclasses import dataclass

from .schemes import (AdamiPressureExtrapolation, ArtificialViscosityMonaghan,
                      BoundaryModelDummyParticles, ContinuityDensity,
                      DensityDiffusionMolteniColagrossi, StateEquationAdaptiveCole,
                      StateEquationCole, WallBoundarySystem, WeaklyCompressibleSPHSystem,
                      WendlandC2Kernel)
from .setups import RectangularTank

__all__ = ["DamBreak", "dam_break", "marrone_dam_break_2d"]


@dataclass
class DamBreak:
    tank: RectangularTank
    fluid_system: WeaklyCompressibleSPHSystem
    boundary_system: WallBoundarySystem
    state_equation: StateEquationCole
    particle_spacing: float

    @property
    def n_fluid(self):
        return self.fluid_system.n_particles

    @property
    def n_wall(self):
        return self.boundary_system.n_particles


def dam_break(ndims=2, particle_spacing=0.01, fluid_size=None, tank_size=None, *,
              alpha=0.02, delta=0.1, gravity=9.81, n_layers=3, open_top=True,
              jitter=0.01, seed=0, smoothing_factor=1.2, exponent=7.0, water_height=0.6,
              clip_wall_pressure=True, adaptive_sound_speed=False, wall_viscosity=None,
              density_diffusion=None, density_calculator=None):
    """Dam break with the workload defaults: 2D 1.2 x 0.6 m water in a 3.2 x 1.0 m tank, 3D
    1.0 x 0.5 x 0.5 m water in a 3.2 x 1.0 x 1.0 m tank.  ``adaptive_sound_speed`` selects
    the StateEquationAdaptiveCole of examples/fluid/dam_break_3d.jl (its defaults: Mach 0.1,
    10 to 100 m/s) in place of StateEquationCole with c = 20 sqrt(g H).  ``wall_viscosity``
    is the wall model's viscosity (no-slip walls; ``None`` = free slip, as in the examples).
    ``density_diffusion`` replaces DensityDiffusionMolteniColagrossi(delta), e.g. by
    DensityDiffusionAntuono(delta=0.1), the alternative examples/fluid/dam_break_2d.jl:62-66
    names.  ``density_calculator`` replaces ContinuityDensity(), e.g. by SummationDensity()
    (whose state has no density rows and whose kick has no density diffusion term)."""
    if fluid_size is None:
        fluid_size = (1.2, 0.6) if ndims == 2 else (1.0, 0.5, 0.5)
    if tank_size is None:
        tank_size = (3.2, 1.0) if ndims == 2 else (3.2, 1.0, 1.0)
    sound_speed = 20 * math.sqrt(gravity * water_height)
    if adaptive_sound_speed:
        state_equation = StateEquationAdaptiveCole(reference_density=1000.0, exponent=exponent)
    else:
        state_equation = StateEquationCole(sound_speed=sound_speed, reference_density=1000.0,
                                           exponent=exponent, clip_negative_pressure=False)
    acceleration = (0.0, -gravity) if ndims == 2 else (0.0, -gravity, 0.0)
    faces = (True, True, True, not open_top) if ndims == 2 else \
        (True, True, True, not open_top, True, True)
    tank = RectangularTank(particle_spacing, fluid_size, tank_size, 1000.0, n_layers=n_layers,
                           acceleration=acceleration, state_equation=state_equation,
                           faces=faces, coordinates_perturbation=jitter, seed=seed)
    h = smoothing_factor * particle_spacing
    kernel = WendlandC2Kernel(ndims)
    fluid_system = WeaklyCompressibleSPHSystem(
        tank.fluid, ContinuityDensity() if density_calculator is None else density_calculator,
        state_equation, kernel, h,
        acceleration=acceleration, viscosity=ArtificialViscosityMonaghan(alpha=alpha, beta=0.0),
        density_diffusion=DensityDiffusionMolteniColagrossi(delta=delta)
        if density_diffusion is None else density_diffusion)
    boundary_model = BoundaryModelDummyParticles(
        tank.boundary.density, tank.boundary.mass, AdamiPressureExtrapolation(), kernel, h,
        state_equation=state_equation, clip_negative_pressure=clip_wall_pressure,
        viscosity=wall_viscosity)
    boundary_system = WallBoundarySystem(tank.boundary, boundary_model)
    return DamBreak(tank, fluid_system, boundary_system, state_equation, particle_spacing)


def marrone_dam_break_2d(particles_per_height=40, *, alpha=0.01):
    """The 2D dam-break validation case (validation/dam_break_2d/setup_marrone_2011.jl with
    examples/fluid/dam_break_2d.jl): water H = 0.6 m high and 2H wide in a closed tank
    floor(5.366 H / dx) dx wide and 4 m high with 4 wall layers, dx = H /
    particles_per_height, WendlandC2 with h = 2 dx, StateEquationCole(c = 20 sqrt(g H),
    exponent 1, no clipping), Monaghan viscosity, Molteni-Colagrossi delta = 0.1, Adami walls
    with clipping, no jitter.  ``alpha`` defaults to the 0.01 recorded in the metadata of the
    reference's validation series."""
    H = 0.6
    dx = H / particles_per_height
    tank_size = (math.floor(5.366 * H / dx) * dx, 4.0)
    return dam_break(2, dx, fluid_size=(2 * H, H), tank_size=tank_size, alpha=alpha, delta=0.1,
                     n_layers=4, open_top=False, jitter=0.0, smoothing_factor=2.0,
                     exponent=1.0, water_height=H)

