"""Reverse mode of the HIP right-hand side: the pullback of ``kick!`` and of a
CarpenterKennedy2N54 trajectory with step checkpoints, wrapped as ``torch.autograd.Function``s.

They are the Python counterparts of ChainRulesCore ``rrule``s: ``Kick`` pairs
``tpsph_kick`` with ``tpsph_kick_vjp``; ``Trajectory`` pairs ``tpsph_adjoint_forward``
(nsteps steps, the state kept every ``checkpoint_every`` steps) with ``tpsph_adjoint_reverse``
(each segment recomputed from its checkpoint, then swept backwards).  A loss written with
torch operations on the final state is differentiated by autograd down to the initial state
and to the parameters ``[alpha, delta, g]`` (Monaghan alpha, Molteni-Colagrossi delta and
the gravity magnitude; the reference has no automatic differentiation through its right-hand
side, src/schemes/fluid/weakly_compressible_sph/rhs.jl:5-127).

The parameters live in the library handle, so the ``params`` tensor passed to the functions
must hold the handle's values (it is checked); it is an input only so that autograd carries
its gradient.  ``with_parameters`` builds the Semidiscretization of other values."""
from __future__ import annotations

import copy
import ctypes as C
import math
import time

import numpy as np

from . import native
from .schemes import ArtificialViscosityMonaghan, DensityDiffusionMolteniColagrossi
from .semidiscretization import Semidiscretization, _ptr

__all__ = ["parameters", "with_parameters", "kick_vjp", "adjoint_forward", "adjoint_reverse",
           "Kick", "Trajectory", "mse_loss", "gradient", "differentiation_case", "benchmark"]


def _torch():
    import torch
    return torch


def parameters(semi):
    """``[alpha, delta, g]`` of a Semidiscretization (g = |acceleration|)."""
    f = semi.fluid
    alpha = f.viscosity.alpha if f.viscosity is not None else 0.0
    delta = f.density_diffusion.delta if f.density_diffusion is not None else 0.0
    g = math.sqrt(sum(a * a for a in f.acceleration))
    return np.array([alpha, delta, g])


def with_parameters(semi, alpha=None, delta=None, g=None):
    """A Semidiscretization of the same systems, grid box and search with other alpha,
    delta or gravity magnitude g (along the same direction); the sound speed is unchanged."""
    f = copy.copy(semi.fluid)
    if alpha is not None:
        v = f.viscosity
        f.viscosity = ArtificialViscosityMonaghan(alpha=alpha, beta=v.beta, epsilon=v.epsilon)
    if delta is not None:
        f.density_diffusion = DensityDiffusionMolteniColagrossi(delta=delta)
    if g is not None:
        a = np.asarray(f.acceleration, dtype=np.float64)
        f.acceleration = tuple(float(x) for x in a / np.linalg.norm(a) * g)
    f.pressure = np.zeros(f.n_particles)
    systems = (f,) if semi.wall is None else (f, semi.wall)
    return Semidiscretization(*systems, domain=semi.domain, device=semi.device,
                              cell_divisions=semi.cell_divisions,
                              x_cell_divisions=semi.x_cell_divisions, skin=semi.skin,
                              pair_search=semi.pair_search, precision=semi.precision)


def _check_params(semi, params):
    if params is None:
        return
    got = params.detach().cpu().numpy() if hasattr(params, "detach") else np.asarray(params)
    if not np.array_equal(got, parameters(semi)):
        raise ValueError(f"params {got} differ from the Semidiscretization's "
                         f"{parameters(semi)}: use with_parameters for other values")


def kick_vjp(semi, dv_bar, v, u, t=0.0):
    """The pullback of ``kick!`` at ``(v, u)`` (device tensors): returns ``(v_bar, u_bar,
    param_bar)`` with ``param_bar = [dL/dalpha, dL/ddelta, dL/dg]`` (numpy)."""
    torch = _torch()
    v_bar, u_bar = torch.empty_like(v), torch.empty_like(u)
    pb = np.zeros(3)
    native.check(native.load().tpsph_kick_vjp(semi.handle, _ptr(dv_bar.contiguous()), _ptr(v),
                                              _ptr(u), float(t), _ptr(v_bar), _ptr(u_bar),
                                              pb.ctypes.data), semi.handle)
    return v_bar, u_bar, pb


def adjoint_forward(semi, v0, u0, dt, nsteps, checkpoint_every=10, t0=0.0):
    """``nsteps`` CK2N54 steps from ``(v0, u0)`` keeping checkpoints; returns ``(v, u)``."""
    torch = _torch()
    v, u = torch.empty_like(v0), torch.empty_like(u0)
    native.check(native.load().tpsph_adjoint_forward(
        semi.handle, _ptr(v0.contiguous()), _ptr(u0.contiguous()), float(t0), float(dt),
        int(nsteps), int(checkpoint_every), _ptr(v), _ptr(u)), semi.handle)
    return v, u


def adjoint_reverse(semi, v_bar, u_bar):
    """The pullback of the last ``adjoint_forward``: ``(v0_bar, u0_bar, param_bar)``."""
    torch = _torch()
    v0_bar, u0_bar = torch.empty_like(v_bar), torch.empty_like(u_bar)
    pb = np.zeros(3)
    native.check(native.load().tpsph_adjoint_reverse(
        semi.handle, _ptr(v_bar.contiguous()), _ptr(u_bar.contiguous()), _ptr(v0_bar),
        _ptr(u0_bar), pb.ctypes.data), semi.handle)
    return v0_bar, u0_bar, pb


def _autograd_functions():
    torch = _torch()

    class Kick(torch.autograd.Function):
        """``dv = kick!(v, u)`` with its pullback (``tpsph_kick_vjp``)."""

        @staticmethod
        def forward(ctx, v, u, params, semi):
            _check_params(semi, params)
            dv = torch.empty_like(v)
            native.check(native.load().tpsph_kick(semi.handle, _ptr(dv), _ptr(v.contiguous()),
                                                  _ptr(u.contiguous()), 0.0), semi.handle)
            ctx.semi = semi
            ctx.save_for_backward(v, u)
            return dv

        @staticmethod
        def backward(ctx, dv_bar):
            v, u = ctx.saved_tensors
            v_bar, u_bar, pb = kick_vjp(ctx.semi, dv_bar, v, u)
            return v_bar, u_bar, torch.from_numpy(pb).to(v.device), None

    class Trajectory(torch.autograd.Function):
        """``(v, u)`` after ``nsteps`` CK2N54 steps from ``(v0, u0)``, with the checkpointed
        reverse sweep as its pullback."""

        @staticmethod
        def forward(ctx, v0, u0, params, semi, dt, nsteps, checkpoint_every=10):
            _check_params(semi, params)
            ctx.semi = semi
            v, u = adjoint_forward(semi, v0, u0, dt, nsteps, checkpoint_every)
            return v, u

        @staticmethod
        def backward(ctx, v_bar, u_bar):
            v0_bar, u0_bar, pb = adjoint_reverse(ctx.semi, v_bar, u_bar)
            return (v0_bar, u0_bar, torch.from_numpy(pb).to(v0_bar.device),
                    None, None, None, None)

    return Kick, Trajectory


_FUNCTIONS = None


def __getattr__(name):
    """``Kick`` and ``Trajectory`` are built on first use (importing torch then)."""
    global _FUNCTIONS
    if name in ("Kick", "Trajectory"):
        if _FUNCTIONS is None:
            _FUNCTIONS = _autograd_functions()
        return _FUNCTIONS[0] if name == "Kick" else _FUNCTIONS[1]
    raise AttributeError(name)


def mse_loss(u, target):
    """The trajectory loss of the differentiation workload: the mean over fluid particles of
    the squared distance of the final positions to the targets."""
    return ((u - target) ** 2).sum() / u.shape[0]


def gradient(semi, v0, u0, target, dt, nsteps, checkpoint_every=10):
    """Loss and its gradient: ``(L, v0_bar, u0_bar, param_bar)`` for ``mse_loss`` after
    ``nsteps`` steps, through the checkpointed reverse sweep."""
    torch = _torch()
    v, u = adjoint_forward(semi, v0, u0, dt, nsteps, checkpoint_every)
    n = u.shape[0]
    loss = float(((u - target) ** 2).sum() / n)
    u_bar = 2.0 * (u - target) / n
    v_bar = torch.zeros_like(v)
    v0_bar, u0_bar, pb = adjoint_reverse(semi, v_bar, u_bar)
    return loss, v0_bar, u0_bar, pb


def differentiation_case(ndims=3, particle_spacing=0.01, seed=2, vscale=0.05):
    """The differentiation workload: the dam break at ``particle_spacing`` (3D 0.01 m:
    100 x 50 x 50 = 250,000 fluid particles) with seeded initial velocities N(0, vscale^2)
    per component (numpy's generator, seed 2), FP64.  Returns ``(db, semi, v0, u0)``."""
    torch = _torch()
    from .examples import dam_break
    db = dam_break(ndims, particle_spacing)
    semi = Semidiscretization(db.fluid_system, db.boundary_system)
    v0, u0 = semi.initial_state()
    rng = np.random.default_rng(seed)
    vel = rng.normal(0.0, vscale, size=(db.n_fluid, ndims))
    v0[:, :ndims] = torch.from_numpy(vel).to(v0.device)
    return db, semi, v0, u0


def benchmark(ndims=3, particle_spacing=0.01, dt=2e-5, nsteps=100, checkpoint_every=10,
              target_params=(0.03, 0.12, 9.6), repeats=3):
    """Forward run against the gradient of the differentiation workload: 100 CK2N54 steps of
    the 250k-particle 3D dam break at dt = 2e-5; loss = mean squared distance of the final
    positions to targets from a run with alpha, delta, g = 0.03, 0.12, 9.6; checkpoints every
    10 steps.  Times the plain forward run (tpsph_solve_ck2n54) and the gradient (forward
    with checkpoints + reverse sweep), best of ``repeats``; reports their ratio."""
    torch = _torch()
    db, semi, v0, u0 = differentiation_case(ndims, particle_spacing)
    tsemi = with_parameters(semi, *target_params)
    vt, ut = v0.clone(), u0.clone()
    native.check(native.load().tpsph_solve_ck2n54(tsemi.handle, _ptr(vt), _ptr(ut), 0.0, dt,
                                                  nsteps), tsemi.handle)
    torch.cuda.synchronize()
    target = ut.clone()
    tsemi.close()
    lib = native.load()
    t_fwd, t_grad, t_rev = [], [], []
    for _ in range(repeats):
        v, u = v0.clone(), u0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        native.check(lib.tpsph_solve_ck2n54(semi.handle, _ptr(v), _ptr(u), 0.0, dt, nsteps),
                     semi.handle)
        torch.cuda.synchronize()
        t_fwd.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        vT, uT = adjoint_forward(semi, v0, u0, dt, nsteps, checkpoint_every)
        n = uT.shape[0]
        u_bar = 2.0 * (uT - target) / n
        v_bar = torch.zeros_like(vT)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        v0_bar, u0_bar, pb = adjoint_reverse(semi, v_bar, u_bar)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_grad.append(t2 - t0)
        t_rev.append(t2 - t1)
    loss = float(((uT - target) ** 2).sum() / uT.shape[0])
    finite = bool(torch.isfinite(v0_bar).all() and torch.isfinite(u0_bar).all() and
                  np.isfinite(pb).all())
    semi.close()
    tf, tg, tr = min(t_fwd), min(t_grad), min(t_rev)
    return {
        "workload": f"{ndims}D dam break at dx = {particle_spacing}: {db.n_fluid} fluid + "
                    f"{db.n_wall} wall particles, v0 ~ N(0, 0.05^2) (seed 2), {nsteps} "
                    f"CK2N54 steps at dt = {dt}, checkpoints every {checkpoint_every} steps, "
                    f"targets from alpha, delta, g = {target_params}",
        "forward_s": tf,
        "gradient_s": tg,
        "reverse_sweep_s": tr,
        "adjoint_over_forward": tg / tf,
        "reverse_over_forward": tr / tf,
        "loss": loss,
        "dL_dparams": {"alpha": pb[0], "delta": pb[1], "g": pb[2]},
        "gradient_finite": finite,
    }
