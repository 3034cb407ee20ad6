"""DensityDiffusionAntuono (density_diffusion.jl:89-211): the Molteni-Colagrossi term with
the renormalised density gradients of both particles subtracted; the gradients come from a
gradient-correction pass over the fluid (corrections.jl:329-354, 425-457) on every kick.

Pins of the oracle (the reference's own test only checks ``show``, which is pinned too):
- a linear density field rho = rho0 + g.x: the renormalised gradient is g exactly for every
  particle, whatever the particle layout, since L_a inverts the very sum that the gradient
  sum reduces to (-sum_b V_b grad W_ab (x) r_ab) g;
- on that field psi_ab = 2 g.r_ab - (g + g).r_ab = 0, so the Antuono term vanishes while
  the Molteni-Colagrossi term does not;
- two isolated particles (a collinear, singular correction matrix): the identity replaces
  the inverse (|det| < 1f-9), so the gradient is the uncorrected sum.
The GPU tests compare kicks with the oracle (1e-12 norm-wise relative)."""
import math

import numpy as np
import pytest

from conftest import gpu_available


def _patch(nd, n=7, dx=0.1, jitter=0.2, seed=3):
    rng = np.random.default_rng(seed)
    axes = [np.arange(n) * dx] * nd
    x = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, nd)
    return x + rng.uniform(-jitter * dx, jitter * dx, size=x.shape)


def _params(orc, nd, h, kind=2, delta=0.1, c=10.0):
    P = orc.OracleParams()
    P.ndims = nd
    P.kernel = 0
    P.smoothing_length = h
    P.reference_density = 1000.0
    P.sound_speed = c
    P.exponent = 7.0
    P.use_density_diffusion = kind
    P.delta = delta
    return P


def _state(x, rho):
    nf, nd = x.shape
    v = np.zeros((nf, nd + 1))
    v[:, nd] = rho
    return v


def test_show_matches_reference(tp):
    # test/schemes/fluid/weakly_compressible_sph/density_diffusion.jl:4-8
    assert repr(tp.DensityDiffusionAntuono(delta=0.1)) == "DensityDiffusionAntuono(0.1)"


@pytest.mark.parametrize("nd", [2, 3])
def test_gradient_of_a_linear_density_field_is_exact(orc, nd):
    dx = 0.1
    x = _patch(nd, n=7 if nd == 2 else 5, dx=dx)
    g = np.array([3.0, -7.0, 2.0])[:nd]
    rho = 1000.0 + x @ g
    m = np.full(x.shape[0], 1000.0 * dx**nd) * (1.0 + 0.1 * np.sin(np.arange(x.shape[0])))
    P = _params(orc, nd, 1.5 * dx)
    grad, corr = orc.antuono_gradient(P, x, _state(x, rho), m)
    # no matrix is singular here (every particle has neighbours off every line / plane)
    assert not np.any(np.all(corr == np.eye(nd), axis=(1, 2)))
    # a wrong sign or a dropped volume in the correction matrix breaks the identity
    np.testing.assert_allclose(grad, np.tile(g, (x.shape[0], 1)), rtol=0, atol=1e-11 * 7.0)


@pytest.mark.parametrize("nd", [2, 3])
def test_antuono_term_vanishes_on_a_linear_density_field(orc, nd):
    dx = 0.1
    x = _patch(nd, n=7 if nd == 2 else 5, dx=dx)
    g = np.array([3.0, -7.0, 2.0])[:nd]
    rho = 1000.0 + x @ g
    m = np.full(x.shape[0], 1000.0 * dx**nd)
    v = _state(x, rho)
    pf = np.zeros(x.shape[0])
    h = 1.5 * dx
    P, P0, PM = _params(orc, nd, h), _params(orc, nd, h, kind=0), _params(orc, nd, h, kind=1)
    grad, _ = orc.antuono_gradient(P, x, v, m)
    d_ant = orc.interact_fluid_fluid(P, x, v, m, pf, ddgrad=grad)[:, nd]
    d_none = orc.interact_fluid_fluid(P0, x, v, m, pf)[:, nd]
    d_mol = orc.interact_fluid_fluid(PM, x, v, m, pf)[:, nd]
    mol = np.abs(d_mol - d_none).max()
    assert mol > 1.0
    assert np.abs(d_ant - d_none).max() <= 1e-10 * mol


def test_singular_correction_matrix_falls_back_to_identity(orc):
    h = 0.15
    x = np.array([[0.0, 0.0], [0.1, 0.05]])
    rho = np.array([1000.0, 1003.0])
    m = np.array([10.0, 10.5])
    P = _params(orc, 2, h)
    grad, corr = orc.antuono_gradient(P, x, _state(x, rho), m)
    np.testing.assert_array_equal(corr, np.tile(np.eye(2), (2, 1, 1)))
    r = x[0] - x[1]
    d = math.sqrt(r @ r)
    kdr = orc.kernel_deriv(0, 2, d, h) / d
    # grad rho_a = -(rho_a - rho_b) V_b grad W_ab, with grad W_ba = -grad W_ab
    np.testing.assert_allclose(grad[0], -(rho[0] - rho[1]) * (m[1] / rho[1]) * kdr * r,
                               rtol=1e-14)
    np.testing.assert_allclose(grad[1], -(rho[1] - rho[0]) * (m[0] / rho[0]) * kdr * -r,
                               rtol=1e-14)


def test_antuono_options_are_validated(tp):
    db = tp.dam_break(2, 0.05, density_diffusion=tp.DensityDiffusionAntuono(delta=0.1))
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend="hip_host", pair_search="lists")
    p = semi.params()
    assert p.use_density_diffusion == 2 and p.delta == 0.1
    with pytest.raises(ValueError, match="pair_search"):
        tp.Semidiscretization(db.fluid_system, db.boundary_system,
                              parallelization_backend="hip_host", pair_search="tiles")


def rel_err(got, ref):
    scale = np.abs(ref).max()
    return np.abs(got - ref).max() / (scale if scale > 0 else 1.0)


def _host(a):
    return a.cpu().numpy().copy() if hasattr(a, "cpu") else np.array(a, copy=True)


@pytest.mark.parametrize("search", ["lists", "cells"])
@pytest.mark.parametrize("nd,dx", [(2, 0.01), (3, 0.025)])
@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_antuono_kick_matches_oracle(tp, orc, nd, dx, adaptive, search):
    import torch
    db = tp.dam_break(nd, dx, density_diffusion=tp.DensityDiffusionAntuono(delta=0.1),
                      adaptive_sound_speed=adaptive)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend="hip", pair_search=search)
    v, u = (_host(a) for a in semi.initial_state())
    rng = np.random.default_rng(17)
    v[:, :nd] = rng.normal(0.0, 0.5, size=(v.shape[0], nd))
    v[:, nd] *= 1.0 + rng.uniform(-2e-3, 2e-3, size=v.shape[0])
    vd, ud = torch.from_numpy(v).cuda(), torch.from_numpy(u).cuda()
    dv = torch.full_like(vd, float("nan"))
    tp.kick(dv, vd, ud, {"semi": semi}, 0.0)
    P = orc.params_from_semi(semi)
    assert P.use_density_diffusion == 2
    args = (v, u, db.fluid_system.mass, db.boundary_system.coordinates.T,
            db.boundary_system.boundary_model.hydrodynamic_mass)
    ref = orc.kick(P, *args)[0]
    got = dv.cpu().numpy()
    for col in range(nd + 1):
        assert rel_err(got[:, col], ref[:, col]) <= 1e-12, col
    # the correction matters: Molteni-Colagrossi differs
    Q = orc.params_from_semi(semi)
    Q.use_density_diffusion = 1
    mol = orc.kick(Q, *args)[0]
    assert rel_err(mol[:, nd], ref[:, nd]) > 1e-6


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_antuono_with_overflowing_lists_matches_oracle(tp, orc):
    """Forty-fold masses size the Verlet lists for a fortieth of the candidates: every list
    overflows and the gradient pass and the fluid pass walk the cells instead."""
    import torch
    db = tp.dam_break(3, 0.025, density_diffusion=tp.DensityDiffusionAntuono(delta=0.1))
    db.fluid_system.mass = db.fluid_system.mass * 40.0
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend="hip", pair_search="lists")
    v, u = (_host(a) for a in semi.initial_state())
    v[:, 3] *= 1.0 + np.random.default_rng(5).uniform(-2e-3, 2e-3, size=v.shape[0])
    vd, ud = torch.from_numpy(v).cuda(), torch.from_numpy(u).cuda()
    dv = torch.full_like(vd, float("nan"))
    tp.kick(dv, vd, ud, {"semi": semi}, 0.0)
    ref = orc.kick(orc.params_from_semi(semi), v, u, db.fluid_system.mass,
                   db.boundary_system.coordinates.T,
                   db.boundary_system.boundary_model.hydrodynamic_mass)[0]
    got = dv.cpu().numpy()
    for col in range(4):
        assert rel_err(got[:, col], ref[:, col]) <= 1e-12, col


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_antuono_steps_match_oracle(tp, orc):
    """Twenty CK2N54 steps of the 2D dam break with Antuono on the device solver against
    the oracle's integrator."""
    import torch
    db = tp.dam_break(2, 0.02, density_diffusion=tp.DensityDiffusionAntuono(delta=0.1))
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend="hip")
    ode = tp.semidiscretize(semi, (0.0, 1.0))
    v0, u0 = _host(ode.v0), _host(ode.u0)
    dt, n = 2e-5, 20
    v1, u1 = (_host(a) for a in tp.solve(ode, tp.CarpenterKennedy2N54(), dt=dt, nsteps=n))
    P = orc.params_from_semi(semi)
    vr, ur = orc.solve_ck2n54(P, v0, u0, db.fluid_system.mass,
                              db.boundary_system.coordinates.T,
                              db.boundary_system.boundary_model.hydrodynamic_mass, dt, n)
    for col in range(3):
        assert rel_err(v1[:, col], vr[:, col]) <= 1e-11, col
    assert rel_err(u1, ur) <= 1e-12
