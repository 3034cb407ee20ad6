"""System and model types of the WCSPH right-hand side, named and parameterised like the
reference's constructors.  They hold parameters and initial data only; all per-RHS compute
runs in the HIP library (``csrc/``).

Reference constructors mirrored here:

* ``WendlandC2Kernel{NDIMS}`` / ``SchoenbergCubicSplineKernel{NDIMS}`` --
  ``src/general/smoothing_kernels.jl:397-458, 161-227``;
* ``StateEquationCole`` -- ``src/schemes/fluid/weakly_compressible_sph/state_equations.jl:100-139``;
* ``ArtificialViscosityMonaghan`` -- ``src/schemes/fluid/viscosity.jl:68-132``;
* ``DensityDiffusionMolteniColagrossi`` -- ``weakly_compressible_sph/density_diffusion.jl:41-55``;
* ``DensityDiffusionAntuono`` -- ``weakly_compressible_sph/density_diffusion.jl:89-211``;
* ``ContinuityDensity`` / ``SummationDensity`` -- ``src/general/density_calculators.jl:1-48``;
* ``WeaklyCompressibleSPHSystem`` -- ``weakly_compressible_sph/system.jl:65-174``;
* ``AdamiPressureExtrapolation`` / ``BoundaryModelDummyParticles`` --
  ``src/schemes/boundary/wall_boundary/dummy_particles.jl:52-150``;
* ``WallBoundarySystem`` -- ``src/schemes/boundary/wall_boundary/system.jl:22-100``.

Options outside the hot path (other kernels, viscosity models, density calculators,
corrections, surface tension, shifting, moving walls) are rejected with ``ValueError``
/ ``NotImplementedError`` at construction, the way the reference raises ``ArgumentError``
for unsupported combinations.
"""
from __future__ import annotations

import math

import numpy as np

from .setups import InitialCondition

__all__ = [
    "WendlandC2Kernel", "SchoenbergCubicSplineKernel", "StateEquationCole",
    "ArtificialViscosityMonaghan", "DensityDiffusionMolteniColagrossi",
    "DensityDiffusionAntuono", "ContinuityDensity", "SummationDensity",
    "WeaklyCompressibleSPHSystem", "AdamiPressureExtrapolation",
    "BoundaryModelDummyParticles", "WallBoundarySystem",
]


class _Kernel:
    kernel_id = -1

    def __init__(self, ndims: int):
        if ndims not in (2, 3):
            raise ValueError("only 2D and 3D kernels are supported")
        self.ndims = ndims

    def compact_support(self, h):
        return 2 * h

    def __repr__(self):
        return f"{type(self).__name__}{{{self.ndims}}}()"


class WendlandC2Kernel(_Kernel):
    """``WendlandC2Kernel{NDIMS}()`` (smoothing_kernels.jl:397-458)."""

    kernel_id = 0

    def normalization_factor(self, h_inv):
        if self.ndims == 2:
            return (7 / (4 * math.pi)) * h_inv * h_inv
        return (21 / (16 * math.pi)) * h_inv * h_inv * h_inv

    def kernel(self, r, h):
        """``kernel(kernel, r, h)`` (smoothing_kernels.jl:30-34, 436-441), elementwise:
        ``sigma (1 - q/2)^4 (2q + 1)`` for ``r < 2h``, else 0."""
        r = np.asarray(r, dtype=np.float64)
        h_inv = 1 / h
        q = r * h_inv
        w = self.normalization_factor(h_inv) * (1 - q / 2) ** 4 * (2 * q + 1)
        return np.where(r < self.compact_support(h), w, 0.0)


class SchoenbergCubicSplineKernel(_Kernel):
    """``SchoenbergCubicSplineKernel{NDIMS}()`` (smoothing_kernels.jl:161-227)."""

    kernel_id = 1

    def normalization_factor(self, h_inv):
        if self.ndims == 2:
            return 10 / (7 * math.pi) * h_inv * h_inv
        return 1 / math.pi * h_inv * h_inv * h_inv

    def kernel(self, r, h):
        """``kernel(kernel, r, h)`` (smoothing_kernels.jl:30-34, 192-199), elementwise:
        ``sigma ((2 - q)^3 / 4 - (q < 1)(1 - q)^3)`` for ``r < 2h``, else 0."""
        r = np.asarray(r, dtype=np.float64)
        h_inv = 1 / h
        q = r * h_inv
        w = self.normalization_factor(h_inv) * ((2 - q) ** 3 / 4 - (q < 1) * (1 - q) ** 3)
        return np.where(r < self.compact_support(h), w, 0.0)


class StateEquationCole:
    """``StateEquationCole(; sound_speed, reference_density, exponent,
    background_pressure=0.0, clip_negative_pressure=false)`` (state_equations.jl:100-139)."""

    def __init__(self, *, sound_speed, reference_density, exponent, background_pressure=0.0,
                 clip_negative_pressure=False):
        self.sound_speed = float(sound_speed)
        self.reference_density = float(reference_density)
        self.exponent = float(exponent)
        self.background_pressure = float(background_pressure)
        self.clip_negative_pressure = bool(clip_negative_pressure)

    def __call__(self, density):
        B = self.reference_density * self.sound_speed ** 2 / self.exponent
        p = B * ((density / self.reference_density) ** self.exponent - 1) + \
            self.background_pressure
        return max(0.0, p) if self.clip_negative_pressure else p

    def inverse(self, pressure):
        """``inverse_state_equation`` (state_equations.jl:131-139)."""
        B = self.reference_density * self.sound_speed ** 2 / self.exponent
        tmp = (pressure - self.background_pressure) / B + 1
        return self.reference_density * tmp ** (1 / self.exponent)


class StateEquationAdaptiveCole(StateEquationCole):
    """``StateEquationAdaptiveCole(; mach_number_target=0.1, min_sound_speed=10.0,
    reference_density, max_sound_speed=100.0, exponent, background_pressure=0.0,
    clip_negative_pressure=false)`` (state_equations.jl:1-76).  The sound speed starts at
    ``min_sound_speed``; every kick sets it to ``clamp(max |v| / mach_number_target,
    min_sound_speed, max_sound_speed)`` over the fluid's current velocities
    (``update_speed_of_sound!``, wcsph/system.jl:307-320).  That update runs on the device;
    ``sound_speed`` here is the host's copy (the start value until
    ``Semidiscretization.sound_speed()`` reads the current one)."""

    def __init__(self, *, reference_density, exponent, mach_number_target=0.1,
                 min_sound_speed=10.0, max_sound_speed=100.0, background_pressure=0.0,
                 clip_negative_pressure=False):
        super().__init__(sound_speed=min_sound_speed, reference_density=reference_density,
                         exponent=exponent, background_pressure=background_pressure,
                         clip_negative_pressure=clip_negative_pressure)
        self.mach_number_target = float(mach_number_target)
        self.min_sound_speed = float(min_sound_speed)
        self.max_sound_speed = float(max_sound_speed)

    def update_sound_speed(self, max_velocity):
        """The sound speed ``update_speed_of_sound!`` sets for a maximum fluid speed."""
        self.sound_speed = min(self.max_sound_speed,
                               max(self.min_sound_speed, max_velocity / self.mach_number_target))
        return self.sound_speed


class ArtificialViscosityMonaghan:
    """``ArtificialViscosityMonaghan(; alpha, beta=0.0, epsilon=0.01)`` (viscosity.jl:68-83)."""

    def __init__(self, *, alpha, beta=0.0, epsilon=0.01):
        self.alpha = float(alpha)
        self.beta = float(beta)
        self.epsilon = float(epsilon)


class ViscosityAdami:
    """``ViscosityAdami(; nu, epsilon=0.01)`` (viscosity.jl:202-217).  The HIP right-hand side
    uses it as the viscosity of a wall model (no-slip walls)."""

    def __init__(self, *, nu, epsilon=0.01):
        self.nu = float(nu)
        self.epsilon = float(epsilon)


class DensityDiffusionMolteniColagrossi:
    """``DensityDiffusionMolteniColagrossi(; delta)`` (density_diffusion.jl:41-47)."""

    kind = 1  # enum tpsph_density_diffusion

    def __init__(self, *, delta):
        self.delta = float(delta)


class DensityDiffusionAntuono:
    """``DensityDiffusionAntuono(; delta)`` (density_diffusion.jl:89-125): the
    Molteni-Colagrossi term with the renormalised density gradients of both particles
    subtracted, which every kick computes in a pass of its own (``update!``,
    density_diffusion.jl:176-211).  It runs with the Verlet lists or the cell walk."""

    kind = 2  # enum tpsph_density_diffusion

    def __init__(self, *, delta):
        self.delta = float(delta)

    def __repr__(self):
        # Base.show, density_diffusion.jl:143-149
        return f"DensityDiffusionAntuono({_julia_float(self.delta)})"


def _julia_float(x: float) -> str:
    """Julia's shortest round-trip text of a Float64 (``repr(0.1) == "0.1"``)."""
    r = repr(float(x))
    return r if ("e" in r or "." in r or "n" in r) else r + ".0"


class ContinuityDensity:
    """``ContinuityDensity()`` marker (density_calculators.jl:14-24): the density is
    integrated (row ndims+1 of ``v_ode``)."""

    def __repr__(self):
        return "ContinuityDensity()"


class SummationDensity:
    """``SummationDensity()`` marker (density_calculators.jl:1-11): the density is summed over
    the fluid and wall neighbours every kick (``summation_density!``,
    density_calculators.jl:26-48); ``v_ode`` holds the velocity alone
    (wcsph/system.jl:236-238) and the pressure acceleration is
    ``pressure_acceleration_summation_density`` (pressure_acceleration.jl:8-15)."""

    def __repr__(self):
        return "SummationDensity()"


class AdamiPressureExtrapolation:
    """``AdamiPressureExtrapolation(; pressure_offset=0, allow_loop_flipping=true)``
    (dummy_particles.jl:142-149).  The HIP pass always runs the wall-parallel loop order
    (the reference's GPU choice, ``parallelize = is_gpu || ...`` at :523-528)."""

    def __init__(self, *, pressure_offset=0.0, allow_loop_flipping=True):
        self.pressure_offset = float(pressure_offset)
        self.allow_loop_flipping = bool(allow_loop_flipping)


class BoundaryModelDummyParticles:
    """``BoundaryModelDummyParticles(initial_density, hydrodynamic_mass, density_calculator,
    smoothing_kernel, smoothing_length; viscosity=nothing, state_equation=nothing,
    correction=nothing, clip_negative_pressure=false)`` (dummy_particles.jl:52-110)."""

    def __init__(self, initial_density, hydrodynamic_mass, density_calculator,
                 smoothing_kernel, smoothing_length, *, viscosity=None, state_equation=None,
                 correction=None, clip_negative_pressure=False, reference_particle_spacing=0.0):
        initial_density = np.asarray(initial_density, dtype=np.float64)
        hydrodynamic_mass = np.asarray(hydrodynamic_mass, dtype=np.float64)
        if initial_density.shape != hydrodynamic_mass.shape:
            raise ValueError("length(initial_density) != length(hydrodynamic_mass)")
        if not isinstance(density_calculator, AdamiPressureExtrapolation):
            raise NotImplementedError("the HIP wall pass implements AdamiPressureExtrapolation")
        if viscosity is not None and \
                not isinstance(viscosity, (ArtificialViscosityMonaghan, ViscosityAdami)):
            raise NotImplementedError("the HIP wall pass implements ArtificialViscosityMonaghan "
                                      "and ViscosityAdami as wall viscosity")
        if correction is not None:
            raise NotImplementedError("kernel corrections are not implemented")
        if state_equation is None:
            raise ValueError("AdamiPressureExtrapolation needs a `state_equation`")
        self.initial_density = initial_density
        self.hydrodynamic_mass = hydrodynamic_mass
        self.density_calculator = density_calculator
        self.smoothing_kernel = smoothing_kernel
        self.smoothing_length = float(smoothing_length)
        self.viscosity = viscosity
        self.state_equation = state_equation
        self.correction = correction
        self.clip_negative_pressure = bool(clip_negative_pressure)
        # ``pressure`` and ``cache.density`` of the reference; filled by the library on demand
        self.pressure = np.zeros_like(initial_density)
        self.density = initial_density.copy()

    def __repr__(self):
        return (f"BoundaryModelDummyParticles({type(self.density_calculator).__name__}, "
                f"{type(self.viscosity).__name__ if self.viscosity is not None else 'Nothing'})")


class WeaklyCompressibleSPHSystem:
    """``WeaklyCompressibleSPHSystem(initial_condition; density_calculator, state_equation,
    smoothing_kernel, smoothing_length, acceleration, viscosity=nothing,
    density_diffusion=nothing, correction=nothing, ...)`` (wcsph/system.jl:90-174)."""

    def __init__(self, initial_condition: InitialCondition, density_calculator=None,
                 state_equation=None, smoothing_kernel=None, smoothing_length=None, *,
                 acceleration=None, viscosity=None, density_diffusion=None,
                 correction=None, pressure_acceleration=None, surface_tension=None,
                 shifting_technique=None, reference_particle_spacing=0, buffer_size=None,
                 source_terms=None):
        nd = initial_condition.ndims
        if smoothing_kernel is None or smoothing_length is None or state_equation is None:
            raise ValueError("smoothing_kernel, smoothing_length and state_equation are required")
        if smoothing_kernel.ndims != nd:
            raise ValueError(f"smoothing kernel dimensionality must be {nd} for a {nd}D problem")
        if acceleration is None:
            acceleration = (0.0,) * nd
        if len(acceleration) != nd:
            raise ValueError(f"`acceleration` must be of length {nd} for a {nd}D problem")
        if density_calculator is None:
            density_calculator = ContinuityDensity()
        if not isinstance(density_calculator, (ContinuityDensity, SummationDensity)):
            raise NotImplementedError("the HIP right-hand side implements ContinuityDensity "
                                      "and SummationDensity")
        if not isinstance(state_equation, StateEquationCole):
            raise NotImplementedError("the HIP right-hand side implements StateEquationCole "
                                      "and StateEquationAdaptiveCole")
        if viscosity is not None and not isinstance(viscosity, ArtificialViscosityMonaghan):
            raise NotImplementedError("the HIP right-hand side implements "
                                      "ArtificialViscosityMonaghan")
        if density_diffusion is not None and not isinstance(
                density_diffusion, (DensityDiffusionMolteniColagrossi, DensityDiffusionAntuono)):
            raise NotImplementedError("the HIP right-hand side implements "
                                      "DensityDiffusionMolteniColagrossi and "
                                      "DensityDiffusionAntuono")
        for name, opt in (("correction", correction),
                          ("pressure_acceleration", pressure_acceleration),
                          ("surface_tension", surface_tension),
                          ("shifting_technique", shifting_technique),
                          ("buffer_size", buffer_size), ("source_terms", source_terms)):
            if opt is not None:
                raise NotImplementedError(f"`{name}` is outside the implemented hot path")
        self.initial_condition = initial_condition
        self.density_calculator = density_calculator
        self.state_equation = state_equation
        self.smoothing_kernel = smoothing_kernel
        self.smoothing_length = float(smoothing_length)
        self.acceleration = tuple(float(a) for a in acceleration)
        self.viscosity = viscosity
        self.density_diffusion = density_diffusion
        self.mass = np.asarray(initial_condition.mass, dtype=np.float64).copy()
        # ``system.pressure`` and, with SummationDensity, ``system.cache.density``: filled by
        # the library on demand (reference order)
        self.pressure = np.zeros(initial_condition.n_particles)
        self.density = np.asarray(initial_condition.density, dtype=np.float64).copy()

    @property
    def ndims(self):
        return self.initial_condition.ndims

    @property
    def n_particles(self):
        return self.initial_condition.n_particles

    @property
    def summation(self):
        return isinstance(self.density_calculator, SummationDensity)

    def v_nvariables(self):
        """``v_nvariables`` (wcsph/system.jl:232-242): velocity + density rows with
        ContinuityDensity, velocity rows alone with SummationDensity."""
        return self.ndims + (0 if self.summation else 1)

    def compact_support(self):
        return self.smoothing_kernel.compact_support(self.smoothing_length)


class WallBoundarySystem:
    """``WallBoundarySystem(initial_condition, boundary_model)`` (wall_boundary/system.jl:22-100).
    Static walls only (``prescribed_motion=nothing``)."""

    def __init__(self, initial_condition: InitialCondition, boundary_model, *,
                 prescribed_motion=None, adhesion_coefficient=0.0):
        if prescribed_motion is not None:
            raise NotImplementedError("moving walls are outside the implemented hot path")
        if initial_condition.n_particles != boundary_model.hydrodynamic_mass.size:
            raise ValueError("boundary model size does not match the initial condition")
        self.initial_condition = initial_condition
        self.coordinates = np.asarray(initial_condition.coordinates, dtype=np.float64)
        self.boundary_model = boundary_model
        self.adhesion_coefficient = float(adhesion_coefficient)

    @property
    def ndims(self):
        return self.initial_condition.ndims

    @property
    def n_particles(self):
        return self.initial_condition.n_particles

    def v_nvariables(self):
        return 0
