"""StateEquationAdaptiveCole (state_equations.jl:1-76): the sound speed of every kick is
clamp(max |v| / mach_number_target, min_sound_speed, max_sound_speed) over the fluid's
current velocities (update_speed_of_sound!, wcsph/system.jl:307-320), used by the state
equation of fluid and walls, the viscosity and the density diffusion.  The reference's tests
hold no known answers for it; the oracle is pinned by the closed-form cases of the clamp and
by agreeing bit for bit with its own fixed-speed kick at the speed the clamp gives.  The GPU
tests compare the library with that oracle (1e-12 norm-wise relative, as every kick)."""
import math

import numpy as np
import pytest

from conftest import gpu_available


def _case(tp, nd, dx, **kw):
    db = tp.dam_break(nd, dx, adaptive_sound_speed=True)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend="hip_host", **kw)
    return db, semi


def _host(a):
    return a.cpu().numpy().copy() if hasattr(a, "cpu") else np.array(a, copy=True)


def _moving_state(semi, seed, scale=0.5):
    v, u = (_host(a) for a in semi.initial_state())
    nd = semi.fluid.ndims
    rng = np.random.default_rng(seed)
    v[:, :nd] = rng.normal(0.0, scale, size=(v.shape[0], nd))
    return v, u


def test_host_state_equation_starts_at_the_minimum(tp):
    se = tp.StateEquationAdaptiveCole(reference_density=1000.0, exponent=7)
    assert (se.sound_speed, se.mach_number_target, se.min_sound_speed, se.max_sound_speed) == \
        (10.0, 0.1, 10.0, 100.0)
    assert se.update_sound_speed(0.0) == 10.0
    assert se.update_sound_speed(1e6) == 100.0
    assert se.update_sound_speed(4.0) == pytest.approx(40.0, rel=1e-15)


def test_oracle_sound_speed_closed_forms(tp, orc):
    db, semi = _case(tp, 2, 0.04)
    P = orc.params_from_semi(semi)
    assert P.adaptive_sound_speed == 1
    n = 5
    v = np.zeros((n, 3))
    assert orc.adaptive_sound_speed(P, v) == 10.0           # fluid at rest: c_min
    v[3, :2] = (3.0, 4.0)                                     # |v| = 5 exactly
    assert orc.adaptive_sound_speed(P, v) == 5.0 / 0.1
    v[1, :2] = (-3.0, -4.0)                                   # a tie in any order
    assert orc.adaptive_sound_speed(P, v[::-1].copy()) == 5.0 / 0.1
    v[0, :2] = (0.0, 1e3)                                     # beyond c_max
    assert orc.adaptive_sound_speed(P, v) == 100.0
    v[:, 2] = 1e9                                             # the density row is not a velocity
    v[0, :2] = 0.0
    assert orc.adaptive_sound_speed(P, v) == 5.0 / 0.1


@pytest.mark.parametrize("nd,dx", [(2, 0.02), (3, 0.05)])
def test_oracle_adaptive_kick_is_the_fixed_kick_at_that_speed(tp, orc, nd, dx):
    db, semi = _case(tp, nd, dx)
    v, u = _moving_state(semi, 3)
    P = orc.params_from_semi(semi)
    c = orc.adaptive_sound_speed(P, v)
    assert 10.0 < c < 100.0
    args = (v, u, db.fluid_system.mass, db.boundary_system.coordinates.T,
            db.boundary_system.boundary_model.hydrodynamic_mass)
    got = orc.kick(P, *args)
    Q = orc.params_from_semi(semi)
    Q.adaptive_sound_speed = 0
    Q.sound_speed = c
    want = orc.kick(Q, *args)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    # and a different speed changes the result (the speed is used at all)
    Q.sound_speed = c * 1.5
    assert not np.array_equal(orc.kick(Q, *args)[0], got[0])


gpu = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]


def rel_err(got, ref):
    scale = np.abs(ref).max()
    return np.abs(got - ref).max() / (scale if scale > 0 else 1.0)


@pytest.mark.parametrize("search", ["lists", "cells"])
@pytest.mark.parametrize("nd,dx", [(2, 0.01), (3, 0.025)])
@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_adaptive_kick_matches_oracle(tp, orc, nd, dx, search):
    import torch
    db = tp.dam_break(nd, dx, adaptive_sound_speed=True)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system, pair_search=search)
    v, u = _moving_state(semi, 11)
    vd, ud = torch.from_numpy(v).cuda(), torch.from_numpy(u).cuda()
    dv = torch.full_like(vd, float("nan"))
    tp.kick(dv, vd, ud, {"semi": semi}, 0.0)
    P = orc.params_from_semi(semi)
    c_ref = orc.adaptive_sound_speed(P, v)
    assert semi.sound_speed() == pytest.approx(c_ref, rel=1e-15)
    ref, pf, pw, rw = orc.kick(P, v, u, db.fluid_system.mass, db.boundary_system.coordinates.T,
                               db.boundary_system.boundary_model.hydrodynamic_mass)
    got = dv.cpu().numpy()
    for col in range(nd + 1):
        assert rel_err(got[:, col], ref[:, col]) <= 1e-12, col
    semi.update_pressures()
    assert rel_err(db.fluid_system.pressure, pf) <= 1e-12


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_adaptive_trajectory_matches_oracle(tp, orc):
    """20 CK2N54 steps from a seeded velocity field: the sound speed follows the fluid
    stage by stage."""
    import torch
    db = tp.dam_break(2, 0.01, adaptive_sound_speed=True)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system)
    ode = tp.semidiscretize(semi, (0.0, 20e-5))
    v0, _ = _moving_state(semi, 5)
    ode.v0.copy_(torch.from_numpy(v0))
    v, u = tp.solve(ode, dt=1e-5, nsteps=20)
    P = orc.params_from_semi(semi)
    vr, ur = orc.solve_ck2n54(P, ode.v0.cpu().numpy(), ode.u0.cpu().numpy(),
                              db.fluid_system.mass, db.boundary_system.coordinates.T,
                              db.boundary_system.boundary_model.hydrodynamic_mass, 1e-5, 20)
    vg = v.cpu().numpy()
    for col in range(3):
        assert rel_err(vg[:, col], vr[:, col]) <= 1e-11, col
    assert rel_err(u.cpu().numpy(), ur) <= 1e-12
    # the last kick saw the state of the last stage, a little before the final state
    c = semi.sound_speed()
    assert 10.0 < c < 100.0
    assert c == pytest.approx(orc.adaptive_sound_speed(P, vg), rel=1e-2)
