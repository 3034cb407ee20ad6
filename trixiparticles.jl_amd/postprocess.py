"""Host-side post-processing of the device state, as the reference's callbacks do it after
``transfer2cpu`` (callbacks/post_process.jl:144-200): SPH interpolation at points and along
lines (general/interpolation.jl:381-760) and the ``max_x_coord`` quantity of the dam-break
validation (validation/dam_break_2d/sensors.jl).  The inputs are the ODE arrays in the
reference layout (numpy arrays or device tensors); the work is a few thousand kernel
evaluations per sensor, so it runs on the host like the reference's."""
from __future__ import annotations

import numpy as np

__all__ = ["compute_pressure", "interpolate_points", "interpolate_line", "max_x_coord"]


def _host(a):
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def max_x_coord(u_ode, ndims):
    """Largest x coordinate of the fluid (sensors.jl ``max_x_coord``)."""
    return float(_host(u_ode).reshape(-1, ndims)[:, 0].max())


def compute_pressure(semi, v_ode):
    """``compute_pressure!`` (wcsph/system.jl:426-438) on the host: ``system.pressure`` of the
    fluid from the density rows of ``v_ode`` through its StateEquationCole
    (state_equations.jl:115-129), the update the reference's post-processing callback runs
    before its quantities (post_process.jl:157-162).  Returns the pressure array."""
    fluid = semi.fluid
    eos = fluid.state_equation
    rho = _host(v_ode).reshape(-1, fluid.ndims + 1)[:, fluid.ndims]
    B = eos.reference_density * eos.sound_speed ** 2 / eos.exponent
    p = B * ((rho / eos.reference_density) ** eos.exponent - 1) + eos.background_pressure
    fluid.pressure = np.maximum(0.0, p) if eos.clip_negative_pressure else p
    return fluid.pressure


def interpolate_points(point_coords, semi, ref_system, v_ode, u_ode, *,
                       smoothing_length=None, cut_off_bnd=True, clip_negative_pressure=False):
    """``interpolate_points(point_coords, semi, ref_system, v_ode, u_ode; smoothing_length,
    cut_off_bnd, clip_negative_pressure)`` (interpolation.jl:480-490, 544-658, 702-720) for
    the fluid system: with V_b = m_b / rho_b and W_ab the reference system's kernel at
    ``smoothing_length`` (default its own h), over fluid neighbours b within the compact
    support of every point a

        density_a  = (sum_b rho_b V_b W_ab) / (sum_b V_b W_ab)
        pressure_a = (sum_b p_b V_b W_ab) / (sum_b V_b W_ab)   (p_b clipped at 0 on request)
        velocity_a = (sum_b v_b V_b W_ab) / (sum_b V_b W_ab)

    with p_b = ``ref_system.pressure`` (``current_pressure``; :func:`compute_pressure`
    refreshes it from the state, as the callback does first).  A point is cut off (NaN values, no neighbours) when
    ``sum_b m_b W_ab`` over the fluid is below eps, or, with ``cut_off_bnd``, below the same
    sum over the wall particles.  ``point_coords`` is ndims x n_points.  Returns a dict with
    ``computed_density``, ``point_coords``, ``neighbor_count``, ``velocity``, ``pressure``
    and ``density``."""
    fluid = semi.fluid
    if ref_system is not fluid:
        raise NotImplementedError("interpolation is implemented for the fluid system")
    nd = fluid.ndims
    h = fluid.smoothing_length if smoothing_length is None else float(smoothing_length)
    kern = fluid.smoothing_kernel
    radius = kern.compact_support(h)
    pts = np.asarray(point_coords, dtype=np.float64).reshape(nd, -1).T
    v = _host(v_ode).reshape(-1, nd + 1)
    x = _host(u_ode).reshape(-1, nd)
    rho = v[:, nd]
    m = fluid.mass
    vol = m / rho
    p = np.asarray(fluid.pressure, dtype=np.float64)
    if clip_negative_pressure:
        p = np.maximum(0.0, p)
    n = pts.shape[0]
    out = {"computed_density": np.zeros(n), "point_coords": pts.T.copy(),
           "neighbor_count": np.zeros(n, dtype=np.int64), "velocity": np.zeros((nd, n)),
           "pressure": np.zeros(n), "density": np.zeros(n)}
    wall = semi.wall if cut_off_bnd else None
    for a in range(n):
        d = x - pts[a]
        dist = np.sqrt((d * d).sum(axis=1))
        nb = np.nonzero(dist <= radius)[0]
        w = kern.kernel(dist[nb], h)
        vw = vol[nb] * w
        shepard = vw.sum()
        computed = (m[nb] * w).sum()
        count = nb.size
        other = 0.0
        if wall is not None and wall.n_particles:
            xw = wall.coordinates.T
            dw = xw - pts[a]
            distw = np.sqrt((dw * dw).sum(axis=1))
            nbw = np.nonzero(distw <= radius)[0]
            other = (wall.boundary_model.hydrodynamic_mass[nbw] * kern.kernel(distw[nbw], h)).sum()
            count += nbw.size
        if computed < np.finfo(np.float64).eps or (cut_off_bnd and other > computed):
            out["computed_density"][a] = np.nan
            out["velocity"][:, a] = np.nan
            out["pressure"][a] = np.nan
            out["density"][a] = np.nan
            continue
        out["computed_density"][a] = computed
        out["neighbor_count"][a] = count
        out["velocity"][:, a] = (v[nb, :nd] * vw[:, None]).sum(axis=0) / shepard
        out["pressure"][a] = (p[nb] * vw).sum() / shepard
        out["density"][a] = (rho[nb] * vw).sum() / shepard
    return out


def _lin_range(a, b, n):
    """Points of Julia's ``LinRange(a, b, n)`` (what ``range`` gives for static vectors):
    point j is ``(1 - j/d) a + (j/d) b`` with d = n - 1, evaluated in that order."""
    if n == 1:
        return a[:, None].copy()
    d = n - 1
    t = np.arange(n) / d
    return (1 - t)[None, :] * a[:, None] + t[None, :] * b[:, None]


def interpolate_line(start, end, n_points, semi, ref_system, v_ode, u_ode, *, endpoint=True,
                     smoothing_length=None, cut_off_bnd=True, clip_negative_pressure=False):
    """``interpolate_line(start, end_, n_points, semi, ref_system, v_ode, u_ode; endpoint,
    smoothing_length, cut_off_bnd, clip_negative_pressure)`` (interpolation.jl:392-410):
    ``n_points`` points of ``range(start, end, length=n_points)`` (the two ends dropped when
    ``endpoint`` is false), interpolated as in :func:`interpolate_points`."""
    start = np.asarray(start, dtype=np.float64)
    end = np.asarray(end, dtype=np.float64)
    pts = _lin_range(start, end, int(n_points))
    if not endpoint:  # a LinRange slice is the LinRange between its new end points
        pts = _lin_range(pts[:, 1], pts[:, -2], int(n_points) - 2)
    return interpolate_points(pts, semi, ref_system, v_ode, u_ode,
                              smoothing_length=smoothing_length, cut_off_bnd=cut_off_bnd,
                              clip_negative_pressure=clip_negative_pressure)
