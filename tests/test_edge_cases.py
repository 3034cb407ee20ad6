"""Edge cases of the HIP right-hand side against the oracle, through the C ABI: a fluid
without walls (the brick kernel writes every rate, k_fluid_walls is not launched), blocks
with fewer bricks than XCDs, a single fluid particle with and without walls, and the device
solver on a fluid without walls.  Tolerance as in test_gpu_parity.py: 1e-12 norm-wise."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

RTOL = 1e-12


def rel_err(got, ref):
    scale = np.abs(ref).max()
    return np.abs(got - ref).max() / (scale if scale > 0 else 1.0)


def fluid_block(tp, nd, n_per_dim, dx=0.02, seed=3):
    """A jittered block of water at rest (no tank) with the dam-break physics."""
    import math
    g = (0.0, -9.81) if nd == 2 else (0.0, -9.81, 0.0)
    se = tp.StateEquationCole(sound_speed=20 * math.sqrt(9.81 * 0.6), reference_density=1000.0,
                              exponent=7.0, clip_negative_pressure=False)
    ic = tp.RectangularShape(dx, n_per_dim, (0.0,) * nd, density=1000.0,
                             coordinates_perturbation=0.01, seed=seed)
    return tp.WeaklyCompressibleSPHSystem(
        ic, tp.ContinuityDensity(), se, tp.WendlandC2Kernel(nd), 1.2 * dx, acceleration=g,
        viscosity=tp.ArtificialViscosityMonaghan(alpha=0.02, beta=0.0),
        density_diffusion=tp.DensityDiffusionMolteniColagrossi(delta=0.1))


def kick_both(tp, orc, semi, fluid, wall, v, u):
    import torch
    P = orc.params_from_semi(semi)
    xw = wall.coordinates.T if wall is not None else None
    mw = wall.boundary_model.hydrodynamic_mass if wall is not None else None
    dv_ref, *_ = orc.kick(P, v, u, fluid.mass, xw, mw)
    vd = torch.from_numpy(np.ascontiguousarray(v)).cuda()
    ud = torch.from_numpy(np.ascontiguousarray(u)).cuda()
    dv = torch.full_like(vd, float("nan"))
    tp.kick(dv, vd, ud, {"semi": semi}, 0.0)
    semi.synchronize()
    return dv.cpu().numpy(), dv_ref


def random_velocities(v, nd, seed=21):
    rng = np.random.default_rng(seed)
    v = v.copy()
    v[:, :nd] = rng.normal(0.0, 0.5, size=(v.shape[0], nd))
    return v


@pytest.mark.parametrize("search", ["lists", "cells"])
@pytest.mark.parametrize("nd,n", [(2, (40, 30)), (3, (14, 12, 10)), (3, (3, 3, 3))])
def test_fluid_without_walls(tp, orc, nd, n, search):
    """No WallBoundarySystem: every rate comes from the fluid kernel, gravity included.
    The 3 x 3 x 3 block has fewer bricks than the device has XCDs."""
    fluid = fluid_block(tp, nd, n)
    semi = tp.Semidiscretization(fluid, parallelization_backend="hip", pair_search=search)
    v0, u0 = semi.initial_state()
    v = random_velocities(v0.cpu().numpy(), nd)
    got, ref = kick_both(tp, orc, semi, fluid, None, v, u0.cpu().numpy())
    assert np.isfinite(got).all()
    for col in range(ref.shape[1]):
        assert rel_err(got[:, col], ref[:, col]) <= RTOL, f"dv column {col}"


@pytest.mark.parametrize("with_walls", [False, True])
@pytest.mark.parametrize("nd", [2, 3])
def test_single_fluid_particle(tp, orc, nd, with_walls):
    """One fluid particle has no fluid neighbour: its velocity rate is gravity plus the wall
    terms, its density rate the wall continuity term only."""
    if with_walls:
        db = tp.dam_break(nd, 0.05)
        wall = db.boundary_system
        fluid = db.fluid_system
        # keep the particle nearest the tank floor's centre, in its hydrostatic state
        x = fluid.initial_condition.coordinates
        k = int(np.argmin(x[1]))
        fluid = fluid_block(tp, nd, (1,) * nd, dx=0.05)
        fluid.initial_condition.coordinates[:, 0] = x[:, k]
    else:
        wall = None
        fluid = fluid_block(tp, nd, (1,) * nd)
    systems = (fluid, wall) if wall is not None else (fluid,)
    semi = tp.Semidiscretization(*systems, parallelization_backend="hip")
    v0, u0 = semi.initial_state()
    v = random_velocities(v0.cpu().numpy(), nd)
    got, ref = kick_both(tp, orc, semi, fluid, wall, v, u0.cpu().numpy())
    g = np.array(fluid.acceleration)
    if not with_walls:
        assert np.array_equal(got[0, :nd], g) and got[0, nd] == 0.0
    # one particle: a velocity component can be a near cancellation of the wall terms (the
    # particle sits over the middle of the floor), so the velocity rates are compared on the
    # scale of the largest one, the density rate on its own
    scale = np.abs(ref[:, :nd]).max()
    assert np.abs(got[:, :nd] - ref[:, :nd]).max() <= RTOL * scale
    assert rel_err(got[:, nd], ref[:, nd]) <= RTOL


def test_solver_without_walls_matches_oracle(tp, orc):
    """20 CK2N54 steps of a 2D block without walls on the device solver vs the oracle."""
    fluid = fluid_block(tp, 2, (40, 30))
    semi = tp.Semidiscretization(fluid, parallelization_backend="hip")
    ode = tp.semidiscretize(semi, (0.0, 20e-5))
    v, u = tp.solve(ode, tp.CarpenterKennedy2N54(), dt=1e-5, nsteps=20)
    P = orc.params_from_semi(semi)
    vr, ur = orc.solve_ck2n54(P, ode.v0.cpu().numpy(), ode.u0.cpu().numpy(), fluid.mass,
                              None, None, 1e-5, 20)
    got_v = v.cpu().numpy()
    for col in range(vr.shape[1]):
        assert rel_err(got_v[:, col], vr[:, col]) <= 1e-10, f"v column {col}"
    du = u.cpu().numpy() - ode.u0.cpu().numpy()
    dur = ur - ode.u0.cpu().numpy()
    for col in range(dur.shape[1]):
        assert rel_err(du[:, col], dur[:, col]) <= 1e-10, f"u column {col}"
