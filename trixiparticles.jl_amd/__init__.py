"""MI355X-native WCSPH right-hand side with TrixiParticles.jl's interface.

Python host side of the HIP library ``libtpsph.so`` (C ABI in ``include/tpsph.h``): the
reference's system/model constructors, setups and its ``Semidiscretization`` /
``semidiscretize`` / ``kick!`` / ``drift!`` ODE surface.  The directory name contains a dot,
so it is imported through :func:`load_package` (registered as ``trixiparticles_jl_amd``)."""
from .examples import DamBreak, dam_break, marrone_dam_break_2d
from .schemes import (AdamiPressureExtrapolation, ArtificialViscosityMonaghan,
                      BoundaryModelDummyParticles, ContinuityDensity,
                      DensityDiffusionAntuono, DensityDiffusionMolteniColagrossi,
                      SchoenbergCubicSplineKernel,
                      StateEquationAdaptiveCole, StateEquationCole, SummationDensity,
                      ViscosityAdami,
                      WallBoundarySystem, WeaklyCompressibleSPHSystem, WendlandC2Kernel)
from .semidiscretization import (CK2N54_A, CK2N54_B, CK2N54_C, CarpenterKennedy2N54,
                                 DynamicalODEProblem, Semidiscretization, calculate_dt, drift,
                                 kick, kick_pairs, semidiscretize, solve, solve_tstops)
from .callbacks import SortingCallback
from .postprocess import compute_pressure, interpolate_line, interpolate_points, max_x_coord
from .setups import InitialCondition, RectangularShape, RectangularTank, splitmix64_uniform
from . import native

__all__ = [
    "AdamiPressureExtrapolation", "ArtificialViscosityMonaghan", "BoundaryModelDummyParticles",
    "ContinuityDensity", "DensityDiffusionAntuono", "DensityDiffusionMolteniColagrossi",
    "SchoenbergCubicSplineKernel",
    "StateEquationAdaptiveCole", "StateEquationCole", "SummationDensity", "ViscosityAdami", "WallBoundarySystem", "WeaklyCompressibleSPHSystem", "WendlandC2Kernel",
    "CK2N54_A", "CK2N54_B", "CK2N54_C", "CarpenterKennedy2N54", "DynamicalODEProblem",
    "Semidiscretization", "calculate_dt", "drift", "kick", "kick_pairs", "semidiscretize", "solve",
    "solve_tstops", "InitialCondition",
    "RectangularShape", "RectangularTank", "splitmix64_uniform", "DamBreak", "dam_break",
    "marrone_dam_break_2d", "compute_pressure", "interpolate_line", "interpolate_points",
    "max_x_coord", "SortingCallback",
    "native",
]
