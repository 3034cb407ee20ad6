"""ODE surface: ``Semidiscretization``, ``semidiscretize``, ``kick!``/``drift!`` and a
``CarpenterKennedy2N54`` driver, mirroring ``src/general/semidiscretization.jl``.

The ODE state keeps the reference's layout (``semidiscretization.jl:128-135``):
``v_ode`` is ``(ndims+1) x n_fluid`` (velocity + density, ``wcsph/system.jl:240-242``;
``ndims x n_fluid`` with SummationDensity, ``:236-238``) and ``u_ode`` is
``ndims x n_fluid``; the wall block of both is empty.  Arrays are stored
column-major like Julia's, i.e. as C-contiguous arrays of shape ``(n_fluid, ndims+1)`` /
``(n_fluid, ndims)`` whose rows are the reference's columns.

Two backends:

* ``"hip"`` (default): state in device memory (torch tensors on ``cuda:0`` as plain device
  buffers); ``kick``/``drift`` launch the HIP kernels through the C ABI, asynchronously.
* ``"hip_host"``: state in host numpy arrays; the library stages through device buffers.

There is no CPU compute path here: both backends run the HIP library.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import native
from .schemes import (StateEquationAdaptiveCole, ViscosityAdami, WallBoundarySystem,
                      WeaklyCompressibleSPHSystem)

__all__ = ["Semidiscretization", "semidiscretize", "DynamicalODEProblem", "kick", "drift",
           "kick_pairs",
           "calculate_dt", "solve_tstops",
           "CarpenterKennedy2N54", "solve", "CK2N54_A", "CK2N54_B", "CK2N54_C"]

# Carpenter & Kennedy (1994), NASA TM-109112, five-stage fourth-order 2N-storage scheme.
CK2N54_A = (0.0, -567301805773 / 1357537059087, -2404267990393 / 2016746695238,
            -3550918686646 / 2091501179385, -1275806237668 / 842570457699)
CK2N54_B = (1432997174477 / 9575080441755, 5161836677717 / 13612068292357,
            1720146321549 / 2090206949498, 3134564353537 / 4481467310338,
            2277821191437 / 14882151754819)
CK2N54_C = (0.0, 1432997174477 / 9575080441755, 2526269341429 / 6820363962896,
            2006345519317 / 3224310063776, 2802321613138 / 2924317926251)


def _ptr(a):
    """Raw address of a numpy array or torch tensor (None for None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous or a.dtype != np.float64:
            raise ValueError("arrays must be C-contiguous float64")
        return a.ctypes.data
    if not a.is_contiguous():
        raise ValueError("tensors must be contiguous")
    return a.data_ptr()


class Semidiscretization:
    """``Semidiscretization(systems...; neighborhood_search, parallelization_backend)``
    (semidiscretization.jl:106-155) for one ``WeaklyCompressibleSPHSystem`` and at most one
    static ``WallBoundarySystem``.

    ``domain``: optional ``(min_corner, max_corner)`` of the cell grid (the reference's
    ``FullGridCellList`` box).  By default it is the bounding box of both systems, grown
    by ``domain_margin`` times its extent on every side.

    ``precision``: "fp64" (the reference's Float64 throughout) or "mixed" (FP32 pair
    arithmetic on brick-relative coordinates with compensated density-rate sums; FP64
    state, pressures, wall sums, stage update and integrals).

    ``cell_divisions`` / ``x_cell_divisions`` (cells per compact support along y, z / x),
    ``skin`` (fraction of the compact support) and ``pair_search`` ("auto", "lists" or
    "cells") tune the device neighbour search; they change speed, never the
    neighbour relation."""

    def __init__(self, *systems, neighborhood_search=None, parallelization_backend="hip",
                 domain=None, domain_margin=0.05, device=0, cell_divisions=0,
                 x_cell_divisions=0, skin=-1.0, pair_search="auto", precision="fp64"):
        fluids = [s for s in systems if isinstance(s, WeaklyCompressibleSPHSystem)]
        walls = [s for s in systems if isinstance(s, WallBoundarySystem)]
        if len(fluids) != 1 or len(walls) > 1 or len(fluids) + len(walls) != len(systems):
            raise NotImplementedError("the HIP right-hand side handles one "
                                      "WeaklyCompressibleSPHSystem and at most one "
                                      "WallBoundarySystem")
        if parallelization_backend not in ("hip", "hip_host"):
            raise ValueError("parallelization_backend must be 'hip' or 'hip_host'")
        self.systems = tuple(systems)
        self.fluid = fluids[0]
        self.wall = walls[0] if walls else None
        self.backend = parallelization_backend
        self.device = device
        nd = self.fluid.ndims
        if self.wall is not None:
            if self.wall.ndims != nd:
                raise ValueError("all systems must have the same number of dimensions")
            bm = self.wall.boundary_model
            if type(bm.smoothing_kernel) is not type(self.fluid.smoothing_kernel) or \
                    not math.isclose(bm.smoothing_length, self.fluid.smoothing_length,
                                     rel_tol=0, abs_tol=0):
                raise NotImplementedError("the wall model must use the fluid's smoothing "
                                          "kernel and smoothing length")
        if domain is None:
            pts = [self.fluid.initial_condition.coordinates]
            if self.wall is not None:
                pts.append(self.wall.coordinates)
            allp = np.concatenate(pts, axis=1)
            lo, hi = allp.min(axis=1), allp.max(axis=1)
            ext = hi - lo
            domain = (lo - domain_margin * ext, hi + domain_margin * ext)
        self.domain = (np.asarray(domain[0], dtype=np.float64),
                       np.asarray(domain[1], dtype=np.float64))
        self.neighborhood_search = neighborhood_search
        self.cell_divisions = int(cell_divisions)
        self.x_cell_divisions = int(x_cell_divisions)
        self.skin = float(skin)
        if pair_search not in native.PAIR_SEARCH:
            raise ValueError(f"pair_search must be one of {sorted(native.PAIR_SEARCH)}")
        self.pair_search = pair_search
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self._handle = None

    # -- parameters of the C ABI
    def params(self) -> native.TpsphParams:
        f = self.fluid
        se = f.state_equation
        p = native.TpsphParams()
        p.ndims = f.ndims
        p.kernel = 0
        p.memory = native.TPSPH_MEM_DEVICE if self.backend == "hip" else native.TPSPH_MEM_HOST
        p.cell_divisions = self.cell_divisions
        p.x_cell_divisions = self.x_cell_divisions
        p.skin = self.skin
        p.pair_search = native.PAIR_SEARCH[self.pair_search]
        p.precision = native.PRECISION_MIXED if self.precision == "mixed" else \
            native.PRECISION_FP64
        p.density_calculator = native.DENSITY_SUMMATION if f.summation else \
            native.DENSITY_CONTINUITY
        p.clip_fluid_eos = int(se.clip_negative_pressure)
        p.smoothing_length = f.smoothing_length
        p.reference_density = se.reference_density
        p.sound_speed = se.sound_speed
        p.exponent = se.exponent
        p.background_pressure = se.background_pressure
        if isinstance(se, StateEquationAdaptiveCole):
            p.adaptive_sound_speed = 1
            p.mach_number_target = se.mach_number_target
            p.min_sound_speed = se.min_sound_speed
            p.max_sound_speed = se.max_sound_speed
        if f.viscosity is not None:
            p.use_viscosity = 1
            p.alpha, p.beta, p.epsilon = f.viscosity.alpha, f.viscosity.beta, f.viscosity.epsilon
        if f.density_diffusion is not None:
            p.use_density_diffusion = f.density_diffusion.kind
            p.delta = f.density_diffusion.delta
        for d in range(f.ndims):
            p.acceleration[d] = f.acceleration[d]
            p.domain_min[d] = self.domain[0][d]
            p.domain_max[d] = self.domain[1][d]
        if self.wall is not None:
            bm = self.wall.boundary_model
            p.clip_wall_pressure = int(bm.clip_negative_pressure)
            p.pressure_offset = bm.density_calculator.pressure_offset
            if bm.viscosity is not None:
                # the fluid side of the wall terms uses the fluid's kinematic viscosity
                # (wall_boundary/system.jl:286-289), defined here for Monaghan's model
                if f.viscosity is None:
                    raise NotImplementedError("a wall viscosity needs the fluid's "
                                              "ArtificialViscosityMonaghan")
                p.wall_epsilon = bm.viscosity.epsilon
                if isinstance(bm.viscosity, ViscosityAdami):
                    p.wall_viscosity, p.wall_nu = 2, bm.viscosity.nu
                else:
                    p.wall_viscosity = 1
                    p.wall_alpha, p.wall_beta = bm.viscosity.alpha, bm.viscosity.beta
            wse = bm.state_equation
            if isinstance(se, StateEquationAdaptiveCole):
                # the reference shares one adaptive object between the two systems
                if wse is not se:
                    raise NotImplementedError("the wall model must use the fluid's adaptive "
                                              "state equation object")
            elif (wse.sound_speed, wse.reference_density, wse.exponent,
                    wse.background_pressure) != (se.sound_speed, se.reference_density,
                                                 se.exponent, se.background_pressure):
                raise NotImplementedError("the wall model must use the fluid's state equation")
        return p

    def sound_speed(self):
        """``system_sound_speed`` of the fluid (wcsph/system.jl:285): for an adaptive state
        equation the value the last kick set (read from the library; the start value
        before the first kick), copied into the host's state equation object."""
        se = self.fluid.state_equation
        if isinstance(se, StateEquationAdaptiveCole) and self._handle is not None:
            c = C.c_double()
            native.check(native.load().tpsph_get_sound_speed(self._handle, C.byref(c)),
                         self._handle)
            se.sound_speed = c.value
        return se.sound_speed

    # -- array helpers for the configured backend
    def _zeros(self, *shape):
        if self.backend == "hip":
            import torch
            return torch.zeros(shape, dtype=torch.float64, device=f"cuda:{self.device}")
        return np.zeros(shape, dtype=np.float64)

    def _array(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if self.backend == "hip":
            import torch
            return torch.from_numpy(a).to(f"cuda:{self.device}")
        return a

    @property
    def handle(self):
        if self._handle is None:
            lib = native.load()
            p = self.params()
            nf = self.fluid.n_particles
            nw = self.wall.n_particles if self.wall is not None else 0
            fm = self._array(self.fluid.mass)
            if self.wall is not None:
                wx = self._array(self.wall.coordinates.T)
                wm = self._array(self.wall.boundary_model.hydrodynamic_mass)
            else:
                wx = wm = None
            out = C.c_void_p()
            rc = lib.tpsph_create(C.byref(p), nf, nw, _ptr(fm), _ptr(wx), _ptr(wm), None,
                                  C.byref(out))
            native.check(rc, None)
            self._handle = out
            self._keep = (fm, wx, wm)
        return self._handle

    def close(self):
        if self._handle is not None:
            native.load().tpsph_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state
    def initial_state(self):
        """``write_v0!``/``write_u0!`` (fluid.jl:79-97, wcsph/system.jl:440-445)."""
        ic = self.fluid.initial_condition
        if self.fluid.summation:  # the velocity alone (wcsph/system.jl:236-238)
            v0 = ic.velocity.T
        else:
            v0 = np.concatenate([ic.velocity, ic.density.reshape(1, -1)], axis=0).T
        u0 = ic.coordinates.T
        return self._array(v0), self._array(u0)

    def integrals(self, v_ode, u_ode):
        """FP64 integrals of a fluid state (``tpsph_integrals``): total mass, momentum (3),
        kinetic energy, potential energy, fluid volume sum m/rho and sum rho, as a dict."""
        out = np.zeros(8)
        if self.backend == "hip":
            import torch
            out_d = torch.zeros(8, dtype=torch.float64, device=f"cuda:{self.device}")
            native.check(native.load().tpsph_integrals(self.handle, _ptr(v_ode), _ptr(u_ode),
                                                       out_d.data_ptr()), self.handle)
            out = out_d.cpu().numpy()
        else:
            native.check(native.load().tpsph_integrals(self.handle, _ptr(v_ode), _ptr(u_ode),
                                                       out.ctypes.data), self.handle)
        return {"mass": out[0], "momentum": out[1:4].copy(), "kinetic_energy": out[4],
                "potential_energy": out[5], "volume": out[6], "density_sum": out[7]}

    def synchronize(self):
        native.check(native.load().tpsph_synchronize(self.handle), self.handle)

    # -- quantities read by callbacks
    def update_pressures(self):
        """Copy ``system.pressure``, ``boundary_model.pressure`` and
        ``boundary_model.cache.density`` of the last kick to the host objects."""
        f, w = self.fluid, self.wall
        pw = w.boundary_model.pressure if w is not None else None
        rw = w.boundary_model.density if w is not None else None
        if self.backend == "hip_host":
            native.check(native.load().tpsph_get_pressures(
                self.handle, f.pressure.ctypes.data,
                pw.ctypes.data if pw is not None else None,
                rw.ctypes.data if rw is not None else None), self.handle)
        else:
            self._get_pressures_device()
        return f.pressure, pw, rw

    def density(self):
        """``current_density(v, system)`` of the last kick in the caller's order
        (``tpsph_get_density``): the summed densities with SummationDensity (also copied to
        ``fluid.density``, the reference's ``system.cache.density``)."""
        f = self.fluid
        if self.backend == "hip":
            import torch
            d = torch.empty(f.n_particles, dtype=torch.float64, device=f"cuda:{self.device}")
            native.check(native.load().tpsph_get_density(self.handle, d.data_ptr()), self.handle)
            f.density[:] = d.cpu().numpy()
        else:
            native.check(native.load().tpsph_get_density(self.handle, f.density.ctypes.data),
                         self.handle)
        return f.density.copy()

    def _get_pressures_device(self):
        import torch
        f, w = self.fluid, self.wall
        dev = f"cuda:{self.device}"
        pf = torch.empty(f.n_particles, dtype=torch.float64, device=dev)
        nw = w.n_particles if w is not None else 0
        pw = torch.empty(max(nw, 1), dtype=torch.float64, device=dev)
        rw = torch.empty(max(nw, 1), dtype=torch.float64, device=dev)
        native.check(native.load().tpsph_get_pressures(
            self.handle, pf.data_ptr(), pw.data_ptr() if nw else None,
            rw.data_ptr() if nw else None), self.handle)
        f.pressure[:] = pf.cpu().numpy()
        if nw:
            w.boundary_model.pressure[:] = pw.cpu().numpy()
            w.boundary_model.density[:] = rw.cpu().numpy()

    # -- SortingCallback
    def sort_particles(self, v_ode, u_ode):
        """``sort_particles!`` + ``sort_system!`` for the fluid system
        (callbacks/sorting.jl:118-158) through ``tpsph_sort_particles``: ``v_ode`` and
        ``u_ode`` are put into the cell-sorted order of the current positions in place, and
        the host copies the reference permutes (``pressure``, ``density``) are permuted with
        them, as is the fluid ``mass`` (which the reference leaves in the old order,
        sorting.jl:131).  ``particle_ids`` follows the particles: entry j is the index in the
        initial condition of the particle now at row j.  Returns the permutation (numpy
        int64): row j of the sorted state was row ``perm[j]``."""
        f = self.fluid
        nf = f.n_particles
        lib = native.load()
        if self.backend == "hip":
            import torch
            perm_d = torch.empty(nf, dtype=torch.int64, device=f"cuda:{self.device}")
            native.check(lib.tpsph_sort_particles(self.handle, _ptr(v_ode), _ptr(u_ode),
                                                  perm_d.data_ptr()), self.handle)
            perm = perm_d.cpu().numpy()
        else:
            perm = np.empty(nf, dtype=np.int64)
            native.check(lib.tpsph_sort_particles(self.handle, _ptr(v_ode), _ptr(u_ode),
                                                  perm.ctypes.data), self.handle)
        f.mass[:] = f.mass[perm]
        f.pressure[:] = f.pressure[perm]
        f.density[:] = f.density[perm]
        self.particle_ids = self.particle_ids[perm]
        return perm

    @property
    def particle_ids(self):
        if getattr(self, "_particle_ids", None) is None:
            self._particle_ids = np.arange(self.fluid.n_particles, dtype=np.int64)
        return self._particle_ids

    @particle_ids.setter
    def particle_ids(self, ids):
        self._particle_ids = ids

    def neighbors(self, u_ode, pair=native.PAIR_FLUID_FLUID):
        """Neighbour CSR ``(offsets, indices)`` of the current positions (host arrays)."""
        lib = native.load()
        nf = self.fluid.n_particles
        nw = self.wall.n_particles if self.wall is not None else 0
        nrows = nw if pair == native.PAIR_WALL_FLUID else nf
        offsets = np.zeros(nrows + 1, dtype=np.int64)
        if self.backend == "hip":
            import torch
            off_d = torch.zeros(nrows + 1, dtype=torch.int64, device=f"cuda:{self.device}")
            native.check(lib.tpsph_neighbors_csr(self.handle, _ptr(u_ode), pair,
                                                 off_d.data_ptr(), None, 0), self.handle)
            offsets = off_d.cpu().numpy()
            idx_d = torch.zeros(max(int(offsets[-1]), 1), dtype=torch.int32,
                                device=f"cuda:{self.device}")
            native.check(lib.tpsph_neighbors_csr(self.handle, _ptr(u_ode), pair,
                                                 off_d.data_ptr(), idx_d.data_ptr(),
                                                 idx_d.numel()), self.handle)
            return offsets, idx_d.cpu().numpy()[: int(offsets[-1])]
        native.check(lib.tpsph_neighbors_csr(self.handle, _ptr(u_ode), pair,
                                             offsets.ctypes.data, None, 0), self.handle)
        idx = np.zeros(max(int(offsets[-1]), 1), dtype=np.int32)
        native.check(lib.tpsph_neighbors_csr(self.handle, _ptr(u_ode), pair,
                                             offsets.ctypes.data, idx.ctypes.data, idx.size),
                     self.handle)
        return offsets, idx[: int(offsets[-1])]


def _csr_from_library(semi, call, nrows):
    """Run a CSR export of the library twice (offsets, then indices) into host arrays."""
    if semi.backend == "hip":
        import torch
        dev = f"cuda:{semi.device}"
        off_d = torch.zeros(nrows + 1, dtype=torch.int64, device=dev)
        native.check(call(off_d.data_ptr(), None, 0), semi.handle)
        offsets = off_d.cpu().numpy()
        idx_d = torch.zeros(max(int(offsets[-1]), 1), dtype=torch.int32, device=dev)
        native.check(call(off_d.data_ptr(), idx_d.data_ptr(), idx_d.numel()), semi.handle)
        return offsets, idx_d.cpu().numpy()[: int(offsets[-1])]
    offsets = np.zeros(nrows + 1, dtype=np.int64)
    native.check(call(offsets.ctypes.data, None, 0), semi.handle)
    idx = np.zeros(max(int(offsets[-1]), 1), dtype=np.int32)
    native.check(call(offsets.ctypes.data, idx.ctypes.data, idx.size), semi.handle)
    return offsets, idx[: int(offsets[-1])]


def kick_pairs(semi, pair=native.PAIR_FLUID_FLUID):
    """The pairs the kernels of the last ``kick`` summed, as host CSR ``(offsets, indices)``
    (``tpsph_kick_pairs_csr``): for the parity tests of the pair search."""
    lib = native.load()
    nf = semi.fluid.n_particles
    nw = semi.wall.n_particles if semi.wall is not None else 0
    nrows = nw if pair == native.PAIR_WALL_FLUID else nf
    return _csr_from_library(
        semi, lambda off, idx, cap: lib.tpsph_kick_pairs_csr(semi.handle, pair, off, idx, cap),
        nrows)


def kick(dv_ode, v_ode, u_ode, p, t):
    """``kick!(dv_ode, v_ode, u_ode, p, t)`` (semidiscretization.jl:589-612)."""
    semi = p["semi"]
    native.check(native.load().tpsph_kick(semi.handle, _ptr(dv_ode), _ptr(v_ode), _ptr(u_ode),
                                          float(t)), semi.handle)
    return dv_ode


def drift(du_ode, v_ode, u_ode, p, t):
    """``drift!(du_ode, v_ode, u_ode, p, t)`` (semidiscretization.jl:522-536)."""
    semi = p["semi"]
    native.check(native.load().tpsph_drift(semi.handle, _ptr(du_ode), _ptr(v_ode),
                                           _ptr(u_ode), float(t)), semi.handle)
    return du_ode


@dataclass
class DynamicalODEProblem:
    """``DynamicalODEProblem(kick!, drift!, v0, u0, tspan, p)`` (semidiscretization.jl:394-395)."""

    f1: object
    f2: object
    v0: object
    u0: object
    tspan: tuple
    p: dict


def semidiscretize(semi: Semidiscretization, tspan) -> DynamicalODEProblem:
    """``semidiscretize(semi, tspan)`` (semidiscretization.jl:293-396): initial state in the
    backend's memory, static data copied to the device once (handle creation)."""
    v0, u0 = semi.initial_state()
    _ = semi.handle
    return DynamicalODEProblem(kick, drift, v0, u0, tuple(tspan),
                               {"semi": semi, "split_integration_data": None})


@dataclass
class CarpenterKennedy2N54:
    """``CarpenterKennedy2N54(williamson_condition=false)`` (OrdinaryDiffEqLowStorageRK)."""

    williamson_condition: bool = False


def _split_callbacks(callback):
    """Plain step functions ``f(v, u, t, step)`` and discrete callbacks (objects with
    ``initialize`` / ``condition`` / ``affect``, such as ``SortingCallback``), in order."""
    if callback is None:
        return [], []
    items = list(callback) if isinstance(callback, (list, tuple)) else [callback]
    discrete = [c for c in items if hasattr(c, "condition") and hasattr(c, "affect")]
    fns = [c for c in items if not (hasattr(c, "condition") and hasattr(c, "affect"))]
    return fns, discrete


def solve(ode: DynamicalODEProblem, alg=None, *, dt, nsteps=None, callback=None,
          stage_update="fused"):
    """Fixed-step ``solve(ode, CarpenterKennedy2N54(); dt, adaptive=false)``.

    ``callback``: ``None``, a step function ``f(v, u, t, step)``, a discrete callback such
    as ``SortingCallback``, or a list of these (the reference's ``CallbackSet``), run after
    every step in list order.

    Without step functions the steps run on the device (``tpsph_solve_ck2n54``), in one
    library call per stretch of steps between two steps at which a discrete callback fires.
    With step functions every step is driven from here through ``kick``/``drift`` and the
    stage update, the sequence OrdinaryDiffEq runs: ``stage_update="fused"`` updates each
    register and state in one ``tpsph_rk_stage``; ``"broadcasts"`` issues the single
    broadcasts of LowStorageRK2N's ``perform_step!`` (``tmp = dt k``, ``u = u + B tmp``,
    ``tmp = A tmp``, ``tmp = tmp + dt k``) as ``tpsph_axpby`` calls, as a device-array type
    does under OrdinaryDiffEq.  Returns the final ``(v, u)`` (copies of the initial state,
    advanced)."""
    from .callbacks import Integrator
    if stage_update not in ("fused", "broadcasts"):
        raise ValueError("stage_update must be 'fused' or 'broadcasts'")
    semi = ode.p["semi"]
    t0, t1 = ode.tspan
    if nsteps is None:
        nsteps = int(round((t1 - t0) / dt))
    nsteps = int(nsteps)
    v = ode.v0.clone() if hasattr(ode.v0, "clone") else ode.v0.copy()
    u = ode.u0.clone() if hasattr(ode.u0, "clone") else ode.u0.copy()
    lib = native.load()
    fns, discrete = _split_callbacks(callback)
    integ = Integrator(ode.p, (v, u), float(t0))
    for cb in discrete:
        cb.initialize(integ)

    def after_step(n):
        """Integrator state after accepted step n; the discrete callbacks that fire."""
        integ.t = t0 + n * dt
        integ.stats.naccept = n
        integ.isfinished = n == nsteps
        for cb in discrete:
            if cb.condition(None, integ.t, integ):
                cb.affect(integ)

    if not fns:
        step = 0
        while step < nsteps:
            # the next step at which a callback fires (the conditions read t and naccept only)
            n = step + 1
            while n < nsteps:
                probe = Integrator(ode.p, (v, u), t0 + n * dt)
                probe.stats.naccept = n
                if any(cb.condition(None, probe.t, probe) for cb in discrete):
                    break
                n += 1
            native.check(lib.tpsph_solve_ck2n54(semi.handle, _ptr(v), _ptr(u),
                                                float(t0 + step * dt), float(dt), n - step),
                         semi.handle)
            step = n
            after_step(step)
        return v, u
    dv, du = semi._zeros(*v.shape), semi._zeros(*u.shape)
    tv, tu = semi._zeros(*v.shape), semi._zeros(*u.shape)
    t = t0

    def axpby(y, a, x, b):
        native.check(lib.tpsph_axpby(semi.handle, int(np.prod(y.shape)), _ptr(y), float(a),
                                     _ptr(x), float(b)), semi.handle)

    for step in range(nsteps):
        for i in range(5):
            ts = t + CK2N54_C[i] * dt
            if stage_update == "broadcasts" and i > 0:
                axpby(tv, CK2N54_A[i], tv, 0.0)  # tmp = A tmp
                axpby(tu, CK2N54_A[i], tu, 0.0)
            ode.f1(dv, v, u, ode.p, ts)
            ode.f2(du, v, u, ode.p, ts)
            for y, tmp, k in ((v, tv, dv), (u, tu, du)):
                if stage_update == "broadcasts":
                    axpby(tmp, 0.0 if i == 0 else 1.0, k, dt)  # tmp = dt k / tmp + dt k
                    axpby(y, 1.0, tmp, CK2N54_B[i])             # u = u + B tmp
                else:
                    native.check(lib.tpsph_rk_stage(semi.handle, _ptr(y), _ptr(tmp), _ptr(k),
                                                    CK2N54_A[i], CK2N54_B[i], float(dt),
                                                    int(np.prod(y.shape)), int(i == 0)),
                                 semi.handle)
        t = t0 + (step + 1) * dt
        for f in fns:
            f(v, u, t, step + 1)
        after_step(step + 1)
    return v, u


def calculate_dt(v_ode, u_ode, cfl_number, semi: Semidiscretization):
    """``calculate_dt(v_ode, u_ode, cfl_number, semi)`` (semidiscretization.jl:510-520) as the
    ``StepsizeCallback(cfl)`` uses it: the minimum over the systems, a wall system giving
    ``Inf`` (wall_boundary/system.jl:291-293).  For the fluid (fluid.jl:199-231):
    ``min(0.125 h^2 / nu, 0.25 sqrt(h / |g|), cfl h / c)`` with the Monaghan kinematic
    viscosity ``nu = alpha h c / (2 ndims + 4)`` (viscosity.jl:82-87).  ``c`` is
    ``system_sound_speed``: constant for StateEquationCole, the value of the last kick for
    StateEquationAdaptiveCole."""
    f = semi.fluid
    h = f.smoothing_length
    c = semi.sound_speed()
    dt_viscosity = math.inf
    if f.viscosity is not None:
        nu = f.viscosity.alpha * h * c / (2 * f.ndims + 4)
        dt_viscosity = 0.125 * h ** 2 / nu if nu > 0 else math.inf
    g = math.sqrt(sum(a * a for a in f.acceleration))
    dt_acceleration = 0.25 * math.sqrt(h / g) if g > 0 else math.inf
    dt_sound_speed = cfl_number * h / c
    return min(dt_viscosity, dt_acceleration, dt_sound_speed)


def solve_tstops(ode: DynamicalODEProblem, *, dt, tstops, callback=None):
    """Fixed-step CarpenterKennedy2N54 through a sorted list of stop times, the way
    OrdinaryDiffEq steps a non-adaptive integrator with ``tstops`` (the stops the
    reference's periodic callbacks add): every step is ``min(dt, tstop - t)``, a step that
    reaches a stop lands on it exactly, and the next one is ``dt`` again.  The steps run on
    the device (``tpsph_solve_ck2n54``, one step per call).  ``callback(v, u, t)`` is called
    at every stop.  Returns the final ``(v, u)``."""
    semi = ode.p["semi"]
    lib = native.load()
    v = ode.v0.clone() if hasattr(ode.v0, "clone") else ode.v0.copy()
    u = ode.u0.clone() if hasattr(ode.u0, "clone") else ode.u0.copy()
    t = float(ode.tspan[0])
    for stop in sorted(float(x) for x in tstops):
        if stop <= t:
            continue
        while t < stop:
            step = min(dt, stop - t)
            native.check(lib.tpsph_solve_ck2n54(semi.handle, _ptr(v), _ptr(u), t, float(step), 1),
                         semi.handle)
            t = stop if step == stop - t else t + step
        if callback is not None:
            callback(v, u, t)
    return v, u
