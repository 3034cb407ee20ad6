"""No-slip walls: the wall model's viscosity (BoundaryModelDummyParticles(; viscosity)) acting
on fluid-wall pairs with the wall velocity the Adami pass extrapolates
(dummy_particles.jl:716-756): ArtificialViscosityMonaghan (viscosity.jl:89-132) or
ViscosityAdami (viscosity.jl:202-280).

Pins of the oracle: the reference's known answer for ViscosityAdami
(test/schemes/fluid/viscosity.jl:74-102), the closed forms of the extrapolated wall velocity
(a uniform fluid velocity u gives -u, no fluid gives 0), and a wall at rest under fluid at
rest, where every viscous term vanishes.  The GPU tests compare kicks with the oracle (1e-12
norm-wise relative)."""
import math

import numpy as np
import pytest

from conftest import gpu_available


def test_oracle_adami_viscosity_known_answer(orc):
    """viscosity.jl:74-102: SchoenbergCubicSpline 2D, h = 0.24, nu = 7e-3 on both sides."""
    h = 1.2 * 0.2
    pos_diff = np.array([-0.25 * h, 0.375 * h])
    r = math.sqrt(pos_diff @ pos_diff)
    grad = orc.kernel_deriv(1, 2, r, h) / r * pos_diff  # smoothing_kernel_grad
    dv = orc.av_adami(2, 7e-3, 7e-3, 0.01, h, pos_diff, grad, 0.01, 0.01, 1000.0, 1000.0,
                      np.array([0.3, -1.0]), np.zeros(2))
    assert dv[0] == pytest.approx(-1.089560204803541e-5, abs=6e-15)
    assert dv[1] == pytest.approx(3.6318673493451364e-5, abs=6e-15)


def _case(tp, nd, dx, wall_viscosity, backend="hip_host", **kw):
    db = tp.dam_break(nd, dx, wall_viscosity=wall_viscosity)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend=backend, **kw)
    return db, semi


def _host(a):
    return a.cpu().numpy().copy() if hasattr(a, "cpu") else np.array(a, copy=True)


def test_wall_velocity_closed_forms(tp, orc):
    db, semi = _case(tp, 2, 0.02, tp.ViscosityAdami(nu=1e-3))
    v, u = (_host(a) for a in semi.initial_state())
    P = orc.params_from_semi(semi)
    assert P.wall_viscosity == 2
    xw = db.boundary_system.coordinates.T.copy()
    v[:, 0], v[:, 1] = 0.7, -0.2
    wv = orc.wall_velocity(P, xw, u, v)
    near = np.abs(wv).sum(axis=1) > 0
    assert near.sum() > 100 and (~near).sum() > 100   # walls by the fluid, walls far from it
    np.testing.assert_allclose(wv[near], np.tile([-0.7, 0.2], (near.sum(), 1)), rtol=1e-15)


@pytest.mark.parametrize("model", ["monaghan", "adami"])
def test_oracle_wall_viscosity_vanishes_at_rest(tp, orc, model):
    visc = tp.ArtificialViscosityMonaghan(alpha=0.05) if model == "monaghan" else \
        tp.ViscosityAdami(nu=1e-3)
    db, semi = _case(tp, 2, 0.02, visc)
    v, u = (_host(a) for a in semi.initial_state())
    args = (v, u, db.fluid_system.mass, db.boundary_system.coordinates.T,
            db.boundary_system.boundary_model.hydrodynamic_mass)
    P = orc.params_from_semi(semi)
    Q = orc.params_from_semi(semi)
    Q.wall_viscosity = 0
    for a, b in zip(orc.kick(P, *args), orc.kick(Q, *args)):
        np.testing.assert_array_equal(a, b)
    v[:, :2] = np.random.default_rng(2).normal(0.0, 0.5, size=(v.shape[0], 2))
    assert not np.array_equal(orc.kick(P, *args)[0], orc.kick(Q, *args)[0])


def test_wall_viscosity_needs_the_fluid_viscosity(tp):
    db = tp.dam_break(2, 0.05, alpha=0.02, wall_viscosity=tp.ViscosityAdami(nu=1e-3))
    db.fluid_system.viscosity = None
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend="hip_host")
    with pytest.raises(NotImplementedError):
        semi.params()


def rel_err(got, ref):
    scale = np.abs(ref).max()
    return np.abs(got - ref).max() / (scale if scale > 0 else 1.0)


@pytest.mark.parametrize("search", ["lists", "cells"])
@pytest.mark.parametrize("nd,dx", [(2, 0.01), (3, 0.025)])
@pytest.mark.parametrize("model", ["monaghan", "adami"])
@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_no_slip_kick_matches_oracle(tp, orc, model, nd, dx, search):
    import torch
    visc = tp.ArtificialViscosityMonaghan(alpha=0.05, beta=0.5) if model == "monaghan" else \
        tp.ViscosityAdami(nu=1e-3)
    db, semi = _case(tp, nd, dx, visc, backend="hip", pair_search=search)
    v, u = (_host(a) for a in semi.initial_state())
    v[:, :nd] = np.random.default_rng(13).normal(0.0, 0.5, size=(v.shape[0], nd))
    vd, ud = torch.from_numpy(v).cuda(), torch.from_numpy(u).cuda()
    dv = torch.full_like(vd, float("nan"))
    tp.kick(dv, vd, ud, {"semi": semi}, 0.0)
    P = orc.params_from_semi(semi)
    ref = orc.kick(P, v, u, db.fluid_system.mass, db.boundary_system.coordinates.T,
                   db.boundary_system.boundary_model.hydrodynamic_mass)[0]
    got = dv.cpu().numpy()
    for col in range(nd + 1):
        assert rel_err(got[:, col], ref[:, col]) <= 1e-12, col
    # the wall terms matter: free slip differs
    Q = orc.params_from_semi(semi)
    Q.wall_viscosity = 0
    free = orc.kick(Q, v, u, db.fluid_system.mass, db.boundary_system.coordinates.T,
                    db.boundary_system.boundary_model.hydrodynamic_mass)[0]
    assert rel_err(free[:, :nd], ref[:, :nd]) > 1e-9
