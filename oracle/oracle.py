"""ctypes wrapper of the CPU oracle ``liboracle_sph.so`` (built from ``sph_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke()`` and by the
bench's ``cpu_baseline`` leg, as the checker and the CPU baseline; never by the product.

Array conventions match the C oracle: coordinates ``(n, ndims)`` C-contiguous (= the
reference's column-major ``ndims x n``), fluid state ``(n, ndims+1)`` (``(n, ndims)`` with
SummationDensity)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle_sph.so")


class OracleParams(C.Structure):
    _fields_ = [
        ("ndims", C.c_int32),
        ("kernel", C.c_int32),
        ("clip_fluid_eos", C.c_int32),
        ("clip_wall_pressure", C.c_int32),
        ("use_viscosity", C.c_int32),
        ("use_density_diffusion", C.c_int32),
        ("smoothing_length", C.c_double),
        ("reference_density", C.c_double),
        ("sound_speed", C.c_double),
        ("exponent", C.c_double),
        ("background_pressure", C.c_double),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("epsilon", C.c_double),
        ("delta", C.c_double),
        ("gravity", C.c_double * 3),
        ("pressure_offset", C.c_double),
        ("adaptive_sound_speed", C.c_int32),
        ("density_calculator", C.c_int32),
        ("mach_number_target", C.c_double),
        ("min_sound_speed", C.c_double),
        ("max_sound_speed", C.c_double),
        ("wall_viscosity", C.c_int32),
        ("reserved1", C.c_int32),
        ("wall_alpha", C.c_double),
        ("wall_beta", C.c_double),
        ("wall_epsilon", C.c_double),
        ("wall_nu", C.c_double),
    ]


def build():
    """Compile the oracle (always recompiles)."""
    subprocess.run(["make", "-B", "-C", ORACLE_DIR, "liboracle_sph.so"], check=True,
                   capture_output=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        d = C.c_double
        i64 = C.c_int64
        L.oracle_kernel.argtypes = [C.c_int, C.c_int, d, d]
        L.oracle_kernel.restype = d
        L.oracle_kernel_deriv.argtypes = [C.c_int, C.c_int, d, d]
        L.oracle_kernel_deriv.restype = d
        L.oracle_kernel_deriv_div_r.argtypes = [C.c_int, C.c_int, d, d]
        L.oracle_kernel_deriv_div_r.restype = d
        L.oracle_state_equation.argtypes = [d, d, d, d, C.c_int, d]
        L.oracle_state_equation.restype = d
        L.oracle_inverse_state_equation.argtypes = [d, d, d, d, d]
        L.oracle_inverse_state_equation.restype = d
        L.oracle_av_monaghan.argtypes = [P, P, d, d, d, d, P, P, P, P]
        L.oracle_av_monaghan.restype = None
        L.oracle_neighbors_bruteforce.argtypes = [C.c_int, d, i64, P, i64, P, P, P, P]
        L.oracle_neighbors_bruteforce.restype = None
        L.oracle_neighbors_celllist.argtypes = [C.c_int, d, i64, P, i64, P, P, P, P]
        L.oracle_compute_pressure.argtypes = [P, i64, P, P]
        L.oracle_compute_pressure.restype = None
        L.oracle_adami.argtypes = [P, i64, P, i64, P, P, P, P, P, P]
        L.oracle_interact_fluid_fluid.argtypes = [P, i64, P, P, P, P, P, P]
        L.oracle_antuono_gradient.argtypes = [P, i64, P, P, P, P, P]
        L.oracle_interact_fluid_wall.argtypes = [P, i64, P, P, P, i64, P, P, P, P, P, P, P]
        L.oracle_wall_velocity.argtypes = [P, i64, P, i64, P, P, P]
        L.oracle_av_adami.argtypes = [C.c_int, d, d, d, d, P, d, P, d, d, d, d, P, P, P]
        L.oracle_av_adami.restype = None
        L.oracle_kick.argtypes = [P, i64, P, P, P, i64, P, P, P, P, P, P]
        L.oracle_kick_rho.argtypes = [P, i64, P, P, P, i64, P, P, P, P, P, P, P]
        L.oracle_summation_density.argtypes = [P, i64, P, P, i64, P, P, P]
        L.oracle_adaptive_sound_speed.argtypes = [P, i64, P]
        L.oracle_adaptive_sound_speed.restype = d
        L.oracle_drift.argtypes = [C.c_int, i64, P, P]
        L.oracle_drift.restype = None
        L.oracle_drift_rows.argtypes = [C.c_int, C.c_int, i64, P, P]
        L.oracle_drift_rows.restype = None
        L.oracle_solve_ck2n54.argtypes = [P, i64, P, P, P, i64, P, P, d, i64]
        L.oracle_dual_kick.argtypes = [P, i64, P, P, P, i64, P, P, P, P, P, P, P]
        L.oracle_dual_solve_ck2n54.argtypes = [P, i64, P, P, P, i64, P, P, d, i64, P, P, P, P,
                                               P, P, P]
        L.oracle_ck2n54_coefficients.argtypes = [P, P, P]
        L.oracle_ck2n54_coefficients.restype = None
        L.oracle_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags.c_contiguous and a.dtype in (np.float64, np.int64), a.dtype
    return a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def params_from_semi(semi) -> OracleParams:
    """Oracle parameters equal to those a ``Semidiscretization`` hands to the library."""
    f = semi.fluid
    se = f.state_equation
    P = OracleParams()
    P.ndims = f.ndims
    P.kernel = f.smoothing_kernel.kernel_id
    P.clip_fluid_eos = int(se.clip_negative_pressure)
    P.smoothing_length = f.smoothing_length
    P.reference_density = se.reference_density
    P.sound_speed = se.sound_speed
    P.exponent = se.exponent
    P.background_pressure = se.background_pressure
    if semi.wall is not None and semi.wall.boundary_model.viscosity is not None:
        wv = semi.wall.boundary_model.viscosity
        P.wall_epsilon = wv.epsilon
        if hasattr(wv, "nu"):  # ViscosityAdami
            P.wall_viscosity, P.wall_nu = 2, wv.nu
        else:
            P.wall_viscosity, P.wall_alpha, P.wall_beta = 1, wv.alpha, wv.beta
    if type(f.density_calculator).__name__ == "SummationDensity":
        P.density_calculator = 1
    if hasattr(se, "mach_number_target"):  # StateEquationAdaptiveCole
        P.adaptive_sound_speed = 1
        P.mach_number_target = se.mach_number_target
        P.min_sound_speed = se.min_sound_speed
        P.max_sound_speed = se.max_sound_speed
    if f.viscosity is not None:
        P.use_viscosity = 1
        P.alpha, P.beta, P.epsilon = f.viscosity.alpha, f.viscosity.beta, f.viscosity.epsilon
    if f.density_diffusion is not None:
        # DensityDiffusionAntuono carries the correction-matrix pass (kind 2)
        P.use_density_diffusion = getattr(f.density_diffusion, "kind", 1)
        P.delta = f.density_diffusion.delta
    for d in range(f.ndims):
        P.gravity[d] = f.acceleration[d]
    if semi.wall is not None:
        bm = semi.wall.boundary_model
        P.clip_wall_pressure = int(bm.clip_negative_pressure)
        P.pressure_offset = bm.density_calculator.pressure_offset
    return P


def kernel(which, nd, r, h):
    return lib().oracle_kernel(which, nd, r, h)


def kernel_deriv(which, nd, r, h):
    return lib().oracle_kernel_deriv(which, nd, r, h)


def state_equation(c, rho0, exponent, p_bg, clip, rho):
    return lib().oracle_state_equation(c, rho0, exponent, p_bg, int(clip), rho)


def inverse_state_equation(c, rho0, exponent, p_bg, p):
    return lib().oracle_inverse_state_equation(c, rho0, exponent, p_bg, p)


def neighbors(nd, R, xa, xb, brute=False):
    """CSR (offsets, idx) of all b with |xa - xb|^2 <= R^2; idx ascending per row."""
    xa, xb = _f64(xa), _f64(xb)
    na, nb = xa.shape[0], xb.shape[0]
    counts = np.zeros(na, dtype=np.int64)
    fn = lib().oracle_neighbors_bruteforce if brute else lib().oracle_neighbors_celllist
    fn(nd, R, na, _p(xa), nb, _p(xb), _p(counts), None, None)
    offsets = np.zeros(na + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    idx = np.zeros(max(int(offsets[-1]), 1), dtype=np.int64)
    fn(nd, R, na, _p(xa), nb, _p(xb), _p(counts), _p(offsets), _p(idx))
    idx = idx[: int(offsets[-1])]
    for a in range(na):
        idx[offsets[a]:offsets[a + 1]].sort()
    return offsets, idx


def compute_pressure(P, v):
    v = _f64(v)
    p = np.zeros(v.shape[0])
    lib().oracle_compute_pressure(C.byref(P), v.shape[0], _p(v), _p(p))
    return p


def adami(P, xw, x, v, pf):
    xw, x, v, pf = _f64(xw), _f64(x), _f64(v), _f64(pf)
    nw = xw.shape[0]
    pw, rw, vol = np.zeros(nw), np.zeros(nw), np.zeros(nw)
    rc = lib().oracle_adami(C.byref(P), nw, _p(xw), x.shape[0], _p(x), _p(v), _p(pf), _p(pw),
                            _p(rw), _p(vol))
    assert rc == 0
    return pw, rw, vol


def interact_fluid_fluid(P, x, v, m, pf, dv=None, ddgrad=None):
    """Fluid-fluid terms; ``ddgrad`` (nf x nd) are the DensityDiffusionAntuono gradients,
    needed when ``P.use_density_diffusion == 2``."""
    x, v, m, pf = _f64(x), _f64(v), _f64(m), _f64(pf)
    dv = np.zeros_like(v) if dv is None else dv
    if P.use_density_diffusion == 2:
        assert ddgrad is not None, "DensityDiffusionAntuono needs its density gradients"
        ddgrad = _f64(ddgrad)
    rc = lib().oracle_interact_fluid_fluid(C.byref(P), x.shape[0], _p(x), _p(v), _p(m),
                                           _p(pf), _p(dv), _p(ddgrad))
    assert rc == 0
    return dv


def antuono_gradient(P, x, v, m):
    """DensityDiffusionAntuono update!: (normalised density gradients nf x nd, inverted
    correction matrices nf x nd x nd with [a, i, j] = L_a[i, j])."""
    x, v, m = _f64(x), _f64(v), _f64(m)
    nf, nd = x.shape[0], P.ndims
    grad = np.zeros((nf, nd))
    corr = np.zeros((nf, nd, nd))  # column-major per particle: [a, j, i] in C order
    rc = lib().oracle_antuono_gradient(C.byref(P), nf, _p(x), _p(v), _p(m), _p(grad),
                                       _p(corr))
    assert rc == 0
    return grad, np.ascontiguousarray(corr.transpose(0, 2, 1))


def interact_fluid_wall(P, x, v, pf, xw, mw, pw, rhow, dv=None, m=None, wvel=None):
    """Fluid-wall interact!; ``m`` (fluid masses) and ``wvel`` (no-slip wall velocities,
    n_wall x ndims) only with a wall viscosity."""
    x, v, pf, xw, mw, pw, rhow = map(_f64, (x, v, pf, xw, mw, pw, rhow))
    m = _f64(m) if m is not None else None
    wvel = _f64(wvel) if wvel is not None else None
    dv = np.zeros_like(v) if dv is None else dv
    rc = lib().oracle_interact_fluid_wall(C.byref(P), x.shape[0], _p(x), _p(v), _p(pf),
                                          xw.shape[0], _p(xw), _p(mw), _p(pw), _p(rhow),
                                          _p(dv), _p(m), _p(wvel))
    assert rc == 0
    return dv


def wall_velocity(P, xw, x, v):
    """No-slip wall velocities (n_wall x ndims) for the fluid state (x, v)."""
    xw, x, v = map(_f64, (xw, x, v))
    out = np.zeros((xw.shape[0], P.ndims))
    rc = lib().oracle_wall_velocity(C.byref(P), xw.shape[0], _p(xw), x.shape[0], _p(x), _p(v),
                                    _p(out))
    assert rc == 0
    return out


def av_adami(nd, nu_a, nu_b, epsilon, h, pos_diff, grad_kernel, m_a, m_b, rho_a, rho_b, v_a,
             v_b):
    """ViscosityAdami's pair acceleration (adami_viscosity_force!)."""
    pos_diff, grad_kernel, v_a, v_b = map(_f64, (pos_diff, grad_kernel, v_a, v_b))
    dv = np.zeros(nd)
    lib().oracle_av_adami(nd, nu_a, nu_b, epsilon, h, _p(pos_diff),
                          float(np.sqrt(pos_diff @ pos_diff)), _p(grad_kernel), m_a, m_b, rho_a,
                          rho_b, _p(v_a), _p(v_b), _p(dv))
    return dv


def kick(P, v, u, m, xw=None, mw=None, return_rho=False):
    """kick!: returns (dv, fluid pressure, wall pressure, wall density[, fluid density]).
    v and dv are (nf, nd+1) with ContinuityDensity and (nf, nd) with SummationDensity
    (``P.density_calculator == 1``), whose summed density ``return_rho`` adds."""
    v, u, m = _f64(v), _f64(u), _f64(m)
    nf = v.shape[0]
    assert v.shape[1] == P.ndims + (0 if P.density_calculator == 1 else 1), v.shape
    nw = 0 if xw is None else xw.shape[0]
    xw = _f64(xw) if nw else None
    mw = _f64(mw) if nw else None
    dv = np.zeros_like(v)
    pf = np.zeros(nf)
    rho = np.zeros(nf)
    pw, rw = np.zeros(max(nw, 1)), np.zeros(max(nw, 1))
    rc = lib().oracle_kick_rho(C.byref(P), nf, _p(v), _p(u), _p(m), nw, _p(xw), _p(mw), _p(dv),
                               _p(pf), _p(pw), _p(rw), _p(rho))
    assert rc == 0
    if return_rho:
        return dv, pf, pw[:nw], rw[:nw], rho
    return dv, pf, pw[:nw], rw[:nw]


def summation_density(P, u, m, xw=None, mw=None):
    """summation_density! of the fluid at positions u (fluid, then wall neighbours)."""
    u, m = _f64(u), _f64(m)
    nw = 0 if xw is None else xw.shape[0]
    xw = _f64(xw) if nw else None
    mw = _f64(mw) if nw else None
    rho = np.zeros(u.shape[0])
    rc = lib().oracle_summation_density(C.byref(P), u.shape[0], _p(u), _p(m), nw, _p(xw),
                                        _p(mw), _p(rho))
    assert rc == 0
    return rho


def adaptive_sound_speed(P, v):
    """The sound speed update_speed_of_sound! sets for the fluid velocities in v."""
    v = _f64(v)
    return lib().oracle_adaptive_sound_speed(C.byref(P), v.shape[0], _p(v))


def drift(nd, v):
    """drift!: the velocity rows of v (nd + 1 rows with ContinuityDensity, nd with
    SummationDensity)."""
    v = _f64(v)
    du = np.zeros((v.shape[0], nd))
    lib().oracle_drift_rows(nd, v.shape[1], v.shape[0], _p(v), _p(du))
    return du


def solve_ck2n54(P, v, u, m, xw, mw, dt, nsteps):
    v, u = _f64(v).copy(), _f64(u).copy()
    m = _f64(m)
    nw = 0 if xw is None else xw.shape[0]
    xw = _f64(xw) if nw else None
    mw = _f64(mw) if nw else None
    rc = lib().oracle_solve_ck2n54(C.byref(P), v.shape[0], _p(v), _p(u), _p(m), nw, _p(xw),
                                   _p(mw), dt, nsteps)
    assert rc == 0
    return v, u


DUAL_TANGENTS = 4


def _dual_seeds(nf, nd, seed_params, seed_v, seed_u):
    T = DUAL_TANGENTS
    sp = None if seed_params is None else np.ascontiguousarray(seed_params, dtype=np.float64).reshape(3, T)
    sv = None if seed_v is None else np.ascontiguousarray(seed_v, dtype=np.float64).reshape(T, nf, nd + 1)
    su = None if seed_u is None else np.ascontiguousarray(seed_u, dtype=np.float64).reshape(T, nf, nd)
    return sp, sv, su


def dual_kick(P, v, u, m, xw, mw, seed_params=None, seed_v=None, seed_u=None):
    """kick! on dual numbers (sph_dual.cpp): returns (dv, tangents of dv), the tangents of
    shape (4, nf, nd+1) along the seeds: seed_params (3 x 4: alpha, delta, g), seed_v
    (4 x nf x (nd+1)), seed_u (4 x nf x nd)."""
    v, u, m = _f64(v), _f64(u), _f64(m)
    nf, nd = v.shape[0], v.shape[1] - 1
    nw = 0 if xw is None else xw.shape[0]
    xw = _f64(xw) if nw else None
    mw = _f64(mw) if nw else None
    sp, sv, su = _dual_seeds(nf, nd, seed_params, seed_v, seed_u)
    dv = np.zeros_like(v)
    tan = np.zeros((DUAL_TANGENTS,) + v.shape)
    rc = lib().oracle_dual_kick(C.byref(P), nf, _p(v), _p(u), _p(m), nw, _p(xw), _p(mw),
                                _p(sp), _p(sv), _p(su), _p(dv), _p(tan))
    assert rc == 0
    return dv, tan


def dual_solve_ck2n54(P, v, u, m, xw, mw, dt, nsteps, seed_params=None, seed_v=None,
                      seed_u=None, u_target=None):
    """CarpenterKennedy2N54 on dual numbers: returns (v, u, v tangents, u tangents, loss),
    loss = (L, dL along the 4 seeds) for L = sum |u - u_target|^2 / nf, or None."""
    v, u = _f64(v).copy(), _f64(u).copy()
    m = _f64(m)
    nf, nd = v.shape[0], v.shape[1] - 1
    nw = 0 if xw is None else xw.shape[0]
    xw = _f64(xw) if nw else None
    mw = _f64(mw) if nw else None
    sp, sv, su = _dual_seeds(nf, nd, seed_params, seed_v, seed_u)
    tv = np.zeros((DUAL_TANGENTS,) + v.shape)
    tu = np.zeros((DUAL_TANGENTS,) + u.shape)
    ut = None if u_target is None else _f64(u_target)
    loss = np.zeros(1 + DUAL_TANGENTS)
    rc = lib().oracle_dual_solve_ck2n54(C.byref(P), nf, _p(v), _p(u), _p(m), nw, _p(xw),
                                        _p(mw), dt, nsteps, _p(sp), _p(sv), _p(su), _p(tv),
                                        _p(tu), _p(ut), _p(loss) if ut is not None else None)
    assert rc == 0
    return v, u, tv, tu, (loss if ut is not None else None)


def ck2n54_coefficients():
    A, B, Cc = np.zeros(5), np.zeros(5), np.zeros(5)
    lib().oracle_ck2n54_coefficients(_p(A), _p(B), _p(Cc))
    return A, B, Cc


def num_threads():
    return lib().oracle_num_threads()
