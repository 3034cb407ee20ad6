"""Particle setups: ``RectangularShape`` and ``RectangularTank``.

Host-side (numpy, FP64) restatement of the reference's setup generators, kept in the
reference's particle order so that arrays are interchangeable with the reference's own
initial conditions:

* ``rectangular_shape_coords`` -- ``src/setups/rectangular_shape.jl:198-224``
  (``x = min + dx * (i - 1/2)``, loop-order permutations ``:158-196``);
* hydrostatic initial pressure/density -- ``rectangular_shape.jl:114-140, 226-267``;
* ``RectangularTank`` -- ``src/setups/rectangular_tank.jl:99-200`` with the wall layout of
  ``initialize_boundaries`` (2D ``:530-689``, 3D ``:692+``);
* ``InitialCondition`` mass ``m = dx^D * rho`` -- ``src/general/initial_condition.jl:199-205``.

The reference perturbs lattices with Julia's ``Random.seed!``/``rand``
(``rectangular_shape.jl:107-112``), a stream that is not reproducible outside Julia.  The
jitter here is drawn from SplitMix64 (Steele, Lea & Flood 2014) in counter mode so the
same seed gives the same bytes everywhere.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "InitialCondition",
    "RectangularShape",
    "RectangularTank",
    "splitmix64_uniform",
    "rectangular_shape_coords",
]

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """``n`` doubles in [0, 1): SplitMix64 outputs 1..n of state ``seed``, top 53 bits."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed % (1 << 64)) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class InitialCondition:
    """Particle data of one system (``src/general/initial_condition.jl``).

    ``coordinates`` and ``velocity`` are ``(ndims, n)`` arrays (the reference's column-major
    ``ndims x n`` layout, one column per particle)."""

    coordinates: np.ndarray
    velocity: np.ndarray
    mass: np.ndarray
    density: np.ndarray
    pressure: np.ndarray
    particle_spacing: float
    normals: np.ndarray | None = field(default=None, repr=False)

    @property
    def ndims(self) -> int:
        return int(self.coordinates.shape[0])

    @property
    def n_particles(self) -> int:
        return int(self.coordinates.shape[1])


def _loop_permutation(loop_order, ndims):
    """``loop_permutation`` (rectangular_shape.jl:158-196), 0-based axes."""
    if ndims == 1:
        if loop_order in (None, "x_first"):
            return (0,)
    elif ndims == 2:
        if loop_order in (None, "y_first"):
            return (0, 1)
        if loop_order == "x_first":
            return (1, 0)
    elif ndims == 3:
        if loop_order in (None, "z_first"):
            return (0, 1, 2)
        if loop_order == "y_first":
            return (1, 0, 2)
        if loop_order == "x_first":
            return (2, 1, 0)
    raise ValueError(f"{loop_order} is not a valid loop order")


def _permuted_indices(n_per_dim, loop_order):
    """1-based cartesian indices in the reference's iteration order: ``(ndims, n)``."""
    nd = len(n_per_dim)
    perm = _loop_permutation(loop_order, nd)
    grids = np.meshgrid(*[np.arange(1, n + 1, dtype=np.int64) for n in n_per_dim], indexing="ij")
    return np.stack([np.transpose(g, perm).ravel(order="F") for g in grids])


def rectangular_shape_coords(particle_spacing, n_per_dim, min_coordinates, loop_order=None):
    """``rectangular_shape_coords`` (rectangular_shape.jl:198-224)."""
    idx = _permuted_indices(n_per_dim, loop_order)
    mins = np.asarray(min_coordinates, dtype=np.float64).reshape(-1, 1)
    return mins + particle_spacing * (idx.astype(np.float64) - 0.5)


def _hydrostatic_pressure(particle_spacing, acceleration, density_fun, n_per_dim):
    """``initialize_pressure!`` (rectangular_shape.jl:226-267), default loop order."""
    acc = [float(a) for a in acceleration]
    if sum(abs(a) > np.finfo(float).eps for a in acc) > 1:
        raise ValueError("hydrostatic pressure calculation is not supported with "
                         "diagonal acceleration")
    accel_dim = next(i for i, a in enumerate(acc) if abs(a) > np.finfo(float).eps)
    factor = particle_spacing * abs(acc[accel_dim])
    n = n_per_dim[accel_dim]
    p1 = np.zeros(n)
    p1[0] = 0.5 * factor * density_fun(0.0)
    for i in range(n - 1):
        p1[i + 1] = p1[i] + factor * density_fun(p1[i])
    if acc[accel_dim] < 0:
        p1 = p1[::-1].copy()
    idx = _permuted_indices(n_per_dim, None)
    return p1[idx[accel_dim] - 1]


def RectangularShape(particle_spacing, n_particles_per_dimension, min_coordinates, *,
                     velocity=None, mass=None, density=None, pressure=0.0,
                     acceleration=None, state_equation=None, loop_order=None,
                     coordinates_perturbation=None, seed=0):
    """Rectangular particle block (rectangular_shape.jl:79-154).

    ``coordinates_perturbation``: uniform jitter of amplitude ``perturbation * spacing`` per
    coordinate, drawn with :func:`splitmix64_uniform` (``seed``)."""
    if particle_spacing < np.finfo(float).eps:
        raise ValueError(f"`particle_spacing` needs to be positive and larger than "
                         f"{np.finfo(float).eps}")
    nd = len(n_particles_per_dimension)
    if len(min_coordinates) != nd:
        raise ValueError(f"`min_coordinates` must be of length {nd} for a {nd}D problem")
    n = int(np.prod(n_particles_per_dimension))
    coords = rectangular_shape_coords(particle_spacing, n_particles_per_dimension,
                                      min_coordinates, loop_order)
    if coordinates_perturbation is not None:
        amp = coordinates_perturbation * particle_spacing
        u = splitmix64_uniform(seed, nd * n).reshape(n, nd).T
        coords = coords + (2.0 * u - 1.0) * amp

    if acceleration is not None:
        if pressure != 0.0:
            raise ValueError("`pressure` cannot be used together with `acceleration` and "
                             "`state_equation` (hydrostatic pressure gradient)")
        if state_equation is None:
            density_fun = lambda p: density  # noqa: E731
        else:
            if density is not None:
                raise ValueError("`density` cannot be used together with `acceleration` and "
                                 "`state_equation` (hydrostatic pressure gradient)")
            density_fun = state_equation.inverse
        pressure_arr = _hydrostatic_pressure(particle_spacing, acceleration, density_fun,
                                             n_particles_per_dimension)
        if state_equation is not None:
            density_arr = np.array([state_equation.inverse(p) for p in pressure_arr])
        else:
            density_arr = np.full(n, float(density))
    else:
        if state_equation is not None:
            raise ValueError("`state_equation` must be used together with `acceleration`")
        if density is None:
            raise ValueError("`density` must be specified when not using `acceleration` and "
                             "`state_equation` (hydrostatic pressure gradient)")
        density_arr = np.full(n, float(density))
        pressure_arr = np.full(n, float(pressure))

    if np.any(density_arr < np.finfo(float).eps):
        raise ValueError("density must be positive and larger than `eps()`")
    if velocity is None:
        vel = np.zeros((nd, n))
    else:
        vel = np.repeat(np.asarray(velocity, dtype=np.float64).reshape(nd, 1), n, axis=1)
    if mass is None:
        mass_arr = (particle_spacing ** nd) * density_arr
    else:
        mass_arr = np.full(n, float(mass))
    return InitialCondition(coordinates=coords, velocity=vel, mass=mass_arr,
                            density=density_arr, pressure=pressure_arr,
                            particle_spacing=float(particle_spacing))


def _round_n(size, spacing):
    """``round_n_particles`` (rectangular_tank.jl:406-416): round-half-even like Julia."""
    n = int(np.rint(size / spacing))
    return n, n * spacing


def _isapprox(a, b):
    return abs(a - b) <= math.sqrt(np.finfo(float).eps) * max(abs(a), abs(b))


def _boundary_coords(spacing, tank_size, n_b, n_layers, faces):
    """``initialize_boundaries`` (rectangular_tank.jl:530-689 for 2D, 692+ for 3D)."""
    nd = len(tank_size)
    off = -n_layers * spacing
    parts = []
    rs = lambda n, m, lo=None: rectangular_shape_coords(spacing, n, m, lo)  # noqa: E731
    if nd == 2:
        nx, ny = n_b
        L, Rt, B, T = faces
        if L: parts.append(rs((n_layers, ny), (off, 0.0), "x_first"))
        if Rt: parts.append(rs((n_layers, ny), (tank_size[0], 0.0), "x_first"))
        if B: parts.append(rs((nx, n_layers), (0.0, off), "y_first"))
        if T: parts.append(rs((nx, n_layers), (0.0, tank_size[1]), "y_first"))
        if L and B: parts.append(rs((n_layers, n_layers), (off, off)))
        if L and T: parts.append(rs((n_layers, n_layers), (off, tank_size[1])))
        if Rt and B: parts.append(rs((n_layers, n_layers), (tank_size[0], off)))
        if Rt and T: parts.append(rs((n_layers, n_layers), (tank_size[0], tank_size[1])))
    else:
        nx, ny, nz = n_b
        L, Rt, B, T, F, K = faces
        X, Y, Z = tank_size
        nl = n_layers
        if L: parts.append(rs((nl, ny, nz), (off, 0.0, 0.0), "x_first"))
        if Rt: parts.append(rs((nl, ny, nz), (X, 0.0, 0.0), "x_first"))
        if B: parts.append(rs((nx, nl, nz), (0.0, off, 0.0), "y_first"))
        if T: parts.append(rs((nx, nl, nz), (0.0, Y, 0.0), "y_first"))
        if F: parts.append(rs((nx, ny, nl), (0.0, 0.0, off), "z_first"))
        if K: parts.append(rs((nx, ny, nl), (0.0, 0.0, Z), "z_first"))
        # edges
        if L and B: parts.append(rs((nl, nl, nz), (off, off, 0.0)))
        if L and T: parts.append(rs((nl, nl, nz), (off, Y, 0.0)))
        if Rt and B: parts.append(rs((nl, nl, nz), (X, off, 0.0)))
        if Rt and T: parts.append(rs((nl, nl, nz), (X, Y, 0.0)))
        if F and B: parts.append(rs((nx, nl, nl), (0.0, off, off)))
        if F and T: parts.append(rs((nx, nl, nl), (0.0, Y, off)))
        if K and B: parts.append(rs((nx, nl, nl), (0.0, off, Z)))
        if K and T: parts.append(rs((nx, nl, nl), (0.0, Y, Z)))
        if L and F: parts.append(rs((nl, ny, nl), (off, 0.0, off)))
        if L and K: parts.append(rs((nl, ny, nl), (off, 0.0, Z)))
        if Rt and F: parts.append(rs((nl, ny, nl), (X, 0.0, off)))
        if Rt and K: parts.append(rs((nl, ny, nl), (X, 0.0, Z)))
        # corners
        if L and B and F: parts.append(rs((nl, nl, nl), (off, off, off)))
        if L and T and F: parts.append(rs((nl, nl, nl), (off, Y, off)))
        if L and B and K: parts.append(rs((nl, nl, nl), (off, off, Z)))
        if L and T and K: parts.append(rs((nl, nl, nl), (off, Y, Z)))
        if Rt and B and F: parts.append(rs((nl, nl, nl), (X, off, off)))
        if Rt and T and F: parts.append(rs((nl, nl, nl), (X, Y, off)))
        if Rt and B and K: parts.append(rs((nl, nl, nl), (X, off, Z)))
        if Rt and T and K: parts.append(rs((nl, nl, nl), (X, Y, Z)))
    if not parts:
        return np.zeros((nd, 0))
    return np.concatenate(parts, axis=1)


class RectangularTank:
    """Rectangular tank with a fluid block resting in its corner (rectangular_tank.jl:99-200).

    Attributes ``fluid`` and ``boundary`` are :class:`InitialCondition` objects."""

    def __init__(self, particle_spacing, fluid_size, tank_size, fluid_density, *,
                 velocity=None, fluid_mass=None, pressure=0.0, acceleration=None,
                 state_equation=None, boundary_density=None, n_layers=1, spacing_ratio=1,
                 min_coordinates=None, faces=None, coordinates_perturbation=None, seed=0):
        nd = len(fluid_size)
        if particle_spacing < np.finfo(float).eps:
            raise ValueError(f"`particle_spacing` needs to be positive and larger than "
                             f"{np.finfo(float).eps}.")
        if fluid_density < np.finfo(float).eps:
            raise ValueError(f"`fluid_density` needs to be positive and larger than "
                             f"{np.finfo(float).eps}.")
        if len(tank_size) != nd:
            raise ValueError(f"`tank_size` must be of length {nd} for a {nd}D problem")
        if faces is None:
            faces = (True,) * (2 * nd)
        if min_coordinates is None:
            min_coordinates = (0.0,) * nd
        if boundary_density is None:
            boundary_density = fluid_density

        n_fluid, fluid_size_ = zip(*[_round_n(float(s), particle_spacing) for s in fluid_size])
        n_fluid, fluid_size_ = list(n_fluid), list(fluid_size_)
        tank_size_ = [float(t) for t in tank_size]
        for d in range(nd):
            if _isapprox(float(fluid_size[d]), float(tank_size[d])):
                tank_size_[d] = fluid_size_[d]
        boundary_spacing = particle_spacing / spacing_ratio
        n_b, tank_size_ = zip(*[_round_n(t, boundary_spacing) for t in tank_size_])
        n_b, tank_size_ = list(n_b), list(tank_size_)

        bcoords = _boundary_coords(boundary_spacing, tank_size_, n_b, n_layers, faces)
        nbnd = bcoords.shape[1]
        bmass = boundary_density * boundary_spacing ** nd * np.ones(nbnd)
        bdens = boundary_density * np.ones(nbnd)

        # check_tank_overlap (rectangular_tank.jl:476-528)
        for d in range(nd):
            if tank_size_[d] < fluid_size_[d] - 1e-5 * particle_spacing:
                n_fluid[d] -= 1
                fluid_size_[d] = n_fluid[d] * particle_spacing

        mins = np.asarray(min_coordinates, dtype=np.float64).reshape(-1, 1)
        self.boundary = InitialCondition(coordinates=bcoords + mins,
                                         velocity=np.zeros_like(bcoords), mass=bmass,
                                         density=bdens, pressure=np.zeros(nbnd),
                                         particle_spacing=boundary_spacing)
        if state_equation is not None:
            fluid = RectangularShape(particle_spacing, tuple(n_fluid), (0.0,) * nd,
                                     velocity=velocity, pressure=pressure,
                                     acceleration=acceleration, state_equation=state_equation,
                                     mass=fluid_mass,
                                     coordinates_perturbation=coordinates_perturbation,
                                     seed=seed)
        else:
            fluid = RectangularShape(particle_spacing, tuple(n_fluid), (0.0,) * nd,
                                     density=fluid_density, velocity=velocity,
                                     pressure=pressure, acceleration=acceleration,
                                     mass=fluid_mass,
                                     coordinates_perturbation=coordinates_perturbation,
                                     seed=seed)
        fluid.coordinates = fluid.coordinates + mins
        self.fluid = fluid
        self.fluid_size = tuple(fluid_size_)
        self.tank_size = tuple(tank_size_)
        self.faces = tuple(faces)
        self.particle_spacing = particle_spacing
        self.spacing_ratio = spacing_ratio
        self.n_layers = n_layers
        self.n_particles_per_dimension = tuple(n_fluid)
