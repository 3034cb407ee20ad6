"""Ensembles of independent dam breaks advanced together (``include/tpsph_batch.h``).

The reference runs one set-up per ``semidiscretize`` / ``solve`` (its 2D template is
``examples/fluid/dam_break_2d.jl:18-127``); an ensemble of such problems is a loop over
them.  :class:`Ensemble` packs many of them -- of uneven size, each with its own geometry,
sound speed, alpha and delta -- into one set of kernels per stage on the device, and
captures each CarpenterKennedy2N54 step in a HIP graph.  Each case's result equals, bit for
bit, the same case run alone.

:func:`dam_break_ensemble` draws the workload of uneven 2D dam breaks from one seed;
:func:`autotune` picks the launch shapes of the ensemble's kernels by timing them and keeps
the choice in ``tuning/ensemble_<arch>.json``."""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time
from dataclasses import dataclass

import numpy as np

from . import native
from .examples import dam_break
from .semidiscretization import Semidiscretization
from .setups import splitmix64_uniform

__all__ = ["TpsphCase", "EnsembleCase", "dam_break_ensemble", "Ensemble", "autotune",
           "load_tuning", "TUNING_DIR"]

TUNING_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")

BATCH_FLUID_BLOCK = 0
BATCH_WALL_BLOCK = 1


class TpsphCase(C.Structure):
    """Mirror of ``tpsph_case`` (include/tpsph_batch.h)."""

    _fields_ = [
        ("n_fluid", C.c_int64),
        ("n_wall", C.c_int64),
        ("alpha", C.c_double),
        ("delta", C.c_double),
        ("sound_speed", C.c_double),
        ("domain_min", C.c_double * 3),
        ("domain_max", C.c_double * 3),
    ]


def _bind(lib):
    if getattr(lib, "_tpsph_batch_bound", False):
        return lib
    P = C.c_void_p
    lib.tpsph_batch_create.argtypes = [C.POINTER(native.TpsphParams), C.c_int32,
                                       C.POINTER(TpsphCase), P, P, P, P, C.POINTER(C.c_void_p)]
    lib.tpsph_batch_kick.argtypes = [P, P, P, P, C.c_double]
    lib.tpsph_batch_solve_ck2n54.argtypes = [P, P, P, C.c_double, C.c_double, C.c_int64,
                                             C.c_int32]
    lib.tpsph_batch_set_launch.argtypes = [P, C.c_int32, C.c_int32]
    lib.tpsph_batch_synchronize.argtypes = [P]
    lib.tpsph_batch_last_error.argtypes = [P]
    lib.tpsph_batch_last_error.restype = C.c_char_p
    lib.tpsph_batch_destroy.argtypes = [P]
    lib.tpsph_batch_destroy.restype = None
    for name in ("tpsph_batch_create", "tpsph_batch_kick", "tpsph_batch_solve_ck2n54",
                 "tpsph_batch_set_launch", "tpsph_batch_synchronize"):
        getattr(lib, name).restype = C.c_int
    lib._tpsph_batch_bound = True
    return lib


def _check(lib, code, handle):
    if code != native.TPSPH_OK:
        raise native.TpsphError(code, lib.tpsph_batch_last_error(handle).decode(errors="replace"))


@dataclass
class EnsembleCase:
    """One member: its dam break, its Semidiscretization (grid box, parameters) and the
    drawn parameters."""

    dam_break: object
    semi: Semidiscretization
    width: float
    aspect: float
    alpha: float
    delta: float

    @property
    def n_fluid(self):
        return self.dam_break.n_fluid

    @property
    def n_wall(self):
        return self.dam_break.n_wall


def dam_break_ensemble(n_cases=256, *, seed=1, particle_spacing=0.01, width=(0.3, 1.2),
                       aspect=(0.5, 2.0), alpha=(0.01, 0.05), delta=(0.05, 0.15),
                       tank_width_factor=2.5, tank_height_factor=5.0 / 3.0,
                       cell_divisions=None, x_cell_divisions=None):
    """Uneven 2D dam breaks drawn from one seed (SplitMix64, four draws per case): water
    column width w ~ U(width), height aspect * w with aspect ~ U(aspect), alpha ~ U(alpha),
    delta ~ U(delta); the open tank is tank_width_factor * w wide and tank_height_factor *
    height high (the 3.2 x 1.0 tank of the 1.2 x 0.6 template keeps 5/3 of the height); the
    sound speed is 20 sqrt(g height) as in the template.  Sizes are rounded to the particle
    spacing by RectangularTank.  Case k jitters its lattice with seed + 1 + k.
    ``cell_divisions`` / ``x_cell_divisions``: cells per compact support of every case's grid
    (the cell tiles of the packed cell walk); ``None`` takes the autotuner's committed choice
    (library defaults without one).  The tiling changes the order of each case's sums, so a
    case compares bit for bit with itself alone on the same tiling."""
    tuned = load_tuning() or {}
    if cell_divisions is None:
        cell_divisions = tuned.get("cell_divisions", 0)
    if x_cell_divisions is None:
        x_cell_divisions = tuned.get("x_cell_divisions", 0)
    u = splitmix64_uniform(seed, 4 * n_cases).reshape(n_cases, 4)
    cases = []
    for k in range(n_cases):
        w = width[0] + (width[1] - width[0]) * u[k, 0]
        a = aspect[0] + (aspect[1] - aspect[0]) * u[k, 1]
        al = alpha[0] + (alpha[1] - alpha[0]) * u[k, 2]
        de = delta[0] + (delta[1] - delta[0]) * u[k, 3]
        H = a * w
        db = dam_break(2, particle_spacing, fluid_size=(w, H),
                       tank_size=(tank_width_factor * w, tank_height_factor * H),
                       alpha=al, delta=de, water_height=H, seed=seed + 1 + k)
        semi = Semidiscretization(db.fluid_system, db.boundary_system, pair_search="cells",
                                  skin=0.0, cell_divisions=cell_divisions,
                                  x_cell_divisions=x_cell_divisions)
        cases.append(EnsembleCase(db, semi, w, a, al, de))
    return cases


def load_tuning(arch="gfx950"):
    """The autotuner's choice for `arch` (tuning/ensemble_<arch>.json), or None."""
    path = os.path.join(TUNING_DIR, f"ensemble_{arch}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


class Ensemble:
    """All cases on the device in packed arrays (``tpsph_batch_create``).

    ``v``, ``u``: the cases' initial states concatenated (torch tensors on cuda:0);
    ``slices[k]``: the fluid rows of case k.  ``tuning``: launch shapes (a dict as written
    by :func:`autotune`; ``"default"`` loads the committed choice, ``None`` keeps the
    library's defaults)."""

    def __init__(self, cases, tuning="default"):
        import torch
        self.cases = list(cases)
        self.lib = _bind(native.load())
        semi0 = self.cases[0].semi
        self.params = semi0.params()
        self.ndims = semi0.fluid.ndims
        arr = (TpsphCase * len(self.cases))()
        vs, us, ms, wxs, wms, self.slices = [], [], [], [], [], []
        row = 0
        for k, c in enumerate(self.cases):
            p = c.semi.params()
            a = arr[k]
            a.n_fluid, a.n_wall = c.n_fluid, c.n_wall
            a.alpha, a.delta, a.sound_speed = p.alpha, p.delta, p.sound_speed
            for d in range(self.ndims):
                a.domain_min[d] = p.domain_min[d]
                a.domain_max[d] = p.domain_max[d]
            ic = c.semi.fluid.initial_condition
            vs.append(np.concatenate([ic.velocity, ic.density.reshape(1, -1)], axis=0).T)
            us.append(ic.coordinates.T)
            ms.append(c.semi.fluid.mass)
            wxs.append(c.semi.wall.coordinates.T)
            wms.append(c.semi.wall.boundary_model.hydrodynamic_mass)
            self.slices.append(slice(row, row + c.n_fluid))
            row += c.n_fluid
        dev = "cuda:0"
        cat = lambda xs: torch.from_numpy(np.ascontiguousarray(np.concatenate(xs))).to(dev)
        self.v0, self.u0 = cat(vs), cat(us)
        self._mass, self._wx, self._wm = cat(ms), cat(wxs), cat(wms)
        self.n_fluid, self.n_wall = int(self.v0.shape[0]), int(self._wx.shape[0])
        out = C.c_void_p()
        torch.cuda.synchronize()
        rc = self.lib.tpsph_batch_create(C.byref(self.params), len(self.cases), arr,
                                         self._mass.data_ptr(), self._wx.data_ptr(),
                                         self._wm.data_ptr(), None, C.byref(out))
        _check(self.lib, rc, None)
        self.handle = out
        if tuning == "default":
            tuning = load_tuning()
        if tuning:
            self.set_launch(fluid_block=tuning.get("fluid_block"),
                            wall_block=tuning.get("wall_block"))

    def set_launch(self, fluid_block=None, wall_block=None):
        if fluid_block:
            _check(self.lib, self.lib.tpsph_batch_set_launch(self.handle, BATCH_FLUID_BLOCK,
                                                             int(fluid_block)), self.handle)
        if wall_block:
            _check(self.lib, self.lib.tpsph_batch_set_launch(self.handle, BATCH_WALL_BLOCK,
                                                             int(wall_block)), self.handle)

    def kick(self, dv, v, u, t=0.0):
        _check(self.lib, self.lib.tpsph_batch_kick(self.handle, dv.data_ptr(), v.data_ptr(),
                                                   u.data_ptr(), float(t)), self.handle)
        return dv

    def solve(self, dt, nsteps, *, graph=True, v=None, u=None):
        """``nsteps`` fixed CarpenterKennedy2N54 steps of every case from (v, u) (default:
        copies of the initial state); returns the advanced (v, u)."""
        v = self.v0.clone() if v is None else v
        u = self.u0.clone() if u is None else u
        _check(self.lib, self.lib.tpsph_batch_solve_ck2n54(self.handle, v.data_ptr(),
                                                           u.data_ptr(), 0.0, float(dt),
                                                           int(nsteps), int(bool(graph))),
               self.handle)
        return v, u

    def close(self):
        if getattr(self, "handle", None) is not None:
            self.lib.tpsph_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(ens: Ensemble, *, dt=1e-5, steps=3, repeats=3, arch="gfx950", write=True):
    """Time every launch shape of the ensemble's fluid and Adami passes on `steps` graph
    steps (best of `repeats`), keep the fastest and, with ``write``, store it in
    tuning/ensemble_<arch>.json.  The shapes do not change any result (each case's sums
    follow its own cell order), so the choice is free."""
    import torch
    results = []
    for fb in (64, 128, 256):
        for wb in (64, 128, 256):
            ens.set_launch(fluid_block=fb, wall_block=wb)
            v, u = ens.v0.clone(), ens.u0.clone()
            ens.solve(dt, 1, v=v, u=u)  # capture
            best = math.inf
            for _ in range(repeats):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                ens.solve(dt, steps, v=v, u=u)
                best = min(best, (time.perf_counter() - t0) / steps)
            results.append({"fluid_block": fb, "wall_block": wb, "ms_per_step": best * 1e3})
    choice = min(results, key=lambda r: r["ms_per_step"])
    out = {"arch": arch, "fluid_block": choice["fluid_block"],
           "wall_block": choice["wall_block"], "ms_per_step": choice["ms_per_step"],
           "workload": {"cases": len(ens.cases), "n_fluid": ens.n_fluid, "n_wall": ens.n_wall},
           "candidates": results}
    ens.set_launch(fluid_block=choice["fluid_block"], wall_block=choice["wall_block"])
    if write:
        os.makedirs(TUNING_DIR, exist_ok=True)
        with open(os.path.join(TUNING_DIR, f"ensemble_{arch}.json"), "w") as f:
            json.dump(out, f, indent=1)
    return out
