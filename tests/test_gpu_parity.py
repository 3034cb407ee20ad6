"""Parity of the HIP right-hand side against the CPU oracle, through the C ABI.

Tolerances (from the workload definition): neighbour sets bit-exact as sorted index sets;
FP64 dv/dt and drho/dt within 1e-12 norm-wise relative error (max |gpu - oracle| over the
column divided by max |oracle| over the column)."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

RTOL = 1e-12


def rel_err(got, ref):
    scale = np.abs(ref).max()
    return np.abs(got - ref).max() / (scale if scale > 0 else 1.0)


def setup_case(tp, nd, dx, backend="hip", grid=None, **kw):
    db = tp.dam_break(nd, dx, **kw)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend=backend, **(grid or {}))
    return db, semi


def oracle_kick(orc, semi, db, v, u):
    P = orc.params_from_semi(semi)
    return orc.kick(P, v, u, db.fluid_system.mass, db.boundary_system.coordinates.T,
                    db.boundary_system.boundary_model.hydrodynamic_mass)


def gpu_kick(tp, semi, v, u):
    import torch
    vd = torch.from_numpy(np.ascontiguousarray(v)).cuda()
    ud = torch.from_numpy(np.ascontiguousarray(u)).cuda()
    dv = torch.full_like(vd, float("nan"))
    tp.kick(dv, vd, ud, {"semi": semi}, 0.0)
    semi.synchronize()
    return dv.cpu().numpy()


def check_columns(got, ref, rtol=RTOL, what="dv"):
    for col in range(ref.shape[1]):
        err = rel_err(got[:, col], ref[:, col])
        assert err <= rtol, f"{what} column {col}: norm-wise relative error {err:.3e}"


GRIDS = [
    {},
    {"cell_divisions": 1, "x_cell_divisions": 1, "skin": 0.0},
    {"cell_divisions": 3, "x_cell_divisions": 5, "skin": 0.3},
    {"cell_divisions": 2, "x_cell_divisions": 8, "skin": 0.1, "domain_margin": 0.0},
]


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("nd,dx", [(2, 0.01), (3, 0.02)])
def test_neighbor_sets_match_bruteforce(tp, orc, nd, dx, grid):
    db, semi = setup_case(tp, nd, dx, grid=grid)
    v0, u0 = semi.initial_state()
    u = u0.cpu().numpy()
    xw = db.boundary_system.coordinates.T.copy()
    R = 2 * db.fluid_system.smoothing_length
    for pair, xa, xb in ((tp.native.PAIR_FLUID_FLUID, u, u),
                         (tp.native.PAIR_FLUID_WALL, u, xw),
                         (tp.native.PAIR_WALL_FLUID, xw, u)):
        off, idx = semi.neighbors(u0, pair)
        roff, ridx = orc.neighbors(nd, R, xa, xb, brute=(nd == 2))
        assert np.array_equal(off, roff), f"pair {pair}: neighbour counts differ"
        assert np.array_equal(idx.astype(np.int64), ridx), f"pair {pair}: neighbour sets differ"


@pytest.mark.parametrize("nd,dx,grid", [(2, 0.01, {}), (2, 0.02, GRIDS[1]), (3, 0.025, {}),
                                         (3, 0.0125, {}), (3, 0.025, GRIDS[2])])
def test_kick_matches_oracle(tp, orc, nd, dx, grid):
    db, semi = setup_case(tp, nd, dx, grid=grid)
    v0, u0 = semi.initial_state()
    v, u = v0.cpu().numpy(), u0.cpu().numpy()
    # a non-trivial velocity field so viscosity, continuity and diffusion all contribute
    rng = np.random.default_rng(7)
    v = v.copy()
    v[:, :nd] = rng.normal(0.0, 0.5, size=(v.shape[0], nd))
    v[:, nd] *= 1.0 + rng.uniform(-2e-3, 2e-3, size=v.shape[0])
    dv_ref, pf_ref, pw_ref, rw_ref = oracle_kick(orc, semi, db, v, u)
    got = gpu_kick(tp, semi, v, u)
    check_columns(got, dv_ref)
    pf, pw, rw = semi.update_pressures()
    assert rel_err(pf, pf_ref) <= RTOL
    assert rel_err(pw, pw_ref) <= RTOL
    assert rel_err(rw, rw_ref) <= RTOL


@pytest.mark.parametrize("search", ["lists", "cells"])
@pytest.mark.parametrize("nd,dx", [(2, 0.01), (3, 0.02)])
def test_kick_matches_oracle_for_every_pair_search(tp, orc, nd, dx, search):
    """The Verlet lists (brick kernel) and the per-kick cell walk give the oracle's
    right-hand side."""
    db, semi = setup_case(tp, nd, dx, grid={"pair_search": search})
    v0, u0 = semi.initial_state()
    v, u = v0.cpu().numpy(), u0.cpu().numpy().copy()
    rng = np.random.default_rng(11)
    v = v.copy()
    v[:, :nd] = rng.normal(0.0, 0.5, size=(v.shape[0], nd))
    dv_ref, *_ = oracle_kick(orc, semi, db, v, u)
    check_columns(gpu_kick(tp, semi, v, u), dv_ref)


@pytest.mark.parametrize("viscosity", ["none", "alpha_beta"])
@pytest.mark.parametrize("nd,dx", [(2, 0.01), (3, 0.025)])
def test_brick_kernel_viscosity_forms(tp, orc, nd, dx, viscosity):
    """The brick kernel's pair loop is compiled once per viscosity form of the fluid
    (none, ArtificialViscosityMonaghan with beta = 0, with beta != 0); the forms the other
    tests do not reach match the oracle too, with approaching and receding pairs."""
    db = tp.dam_break(nd, dx)
    db.fluid_system.viscosity = (None if viscosity == "none" else
                                 tp.ArtificialViscosityMonaghan(alpha=0.5, beta=0.3))
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend="hip", pair_search="lists")
    v0, u0 = semi.initial_state()
    v, u = v0.cpu().numpy().copy(), u0.cpu().numpy()
    rng = np.random.default_rng(13)
    v[:, :nd] = rng.normal(0.0, 0.5, size=(v.shape[0], nd))
    dv_ref, *_ = oracle_kick(orc, semi, db, v, u)
    check_columns(gpu_kick(tp, semi, v, u), dv_ref)


def test_list_overflow_falls_back_to_the_cell_walk(tp, orc):
    """List capacities come from the mean particle volume m / rho0; with 40x the masses
    they are far too small, every list overflows and each particle walks the cells."""
    db = tp.dam_break(3, 0.025)
    db.fluid_system.mass = db.fluid_system.mass * 40.0
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend="hip", pair_search="lists")
    v0, u0 = semi.initial_state()
    v, u = v0.cpu().numpy(), u0.cpu().numpy()
    dv_ref, *_ = oracle_kick(orc, semi, db, v, u)
    check_columns(gpu_kick(tp, semi, v, u), dv_ref)


def test_kick_host_memory_mode(tp, orc):
    db, semi = setup_case(tp, 2, 0.02, backend="hip_host")
    v, u = semi.initial_state()
    dv = np.full_like(v, np.nan)
    tp.kick(dv, v, u, {"semi": semi}, 0.0)
    dv_ref, *_ = oracle_kick(orc, semi, db, v, u)
    check_columns(dv, dv_ref)
    du = np.zeros_like(u)
    tp.drift(du, v, u, {"semi": semi}, 0.0)
    assert np.array_equal(du, orc.drift(2, v))


def test_drift_is_exact_copy(tp, orc):
    import torch
    db, semi = setup_case(tp, 3, 0.025)
    v0, u0 = semi.initial_state()
    v = v0.clone()
    v[:, :3] = torch.randn_like(v[:, :3])
    du = torch.full_like(u0, float("nan"))
    tp.drift(du, v, u0, {"semi": semi}, 0.0)
    semi.synchronize()
    assert np.array_equal(du.cpu().numpy(), orc.drift(3, v.cpu().numpy()))


def test_kick_is_deterministic(tp):
    db, semi = setup_case(tp, 3, 0.02)
    v0, u0 = semi.initial_state()
    v, u = v0.cpu().numpy(), u0.cpu().numpy()
    a = gpu_kick(tp, semi, v, u)
    b = gpu_kick(tp, semi, v, u)
    assert np.array_equal(a, b)


def test_trajectory_matches_oracle(tp, orc):
    """20 CK2N54 steps of the 2D dam break (dt = 1e-5) on the device vs the oracle."""
    db, semi = setup_case(tp, 2, 0.01)
    ode = tp.semidiscretize(semi, (0.0, 20e-5))
    v, u = tp.solve(ode, tp.CarpenterKennedy2N54(), dt=1e-5, nsteps=20)
    P = orc.params_from_semi(semi)
    vr, ur = orc.solve_ck2n54(P, ode.v0.cpu().numpy(), ode.u0.cpu().numpy(),
                              db.fluid_system.mass, db.boundary_system.coordinates.T,
                              db.boundary_system.boundary_model.hydrodynamic_mass, 1e-5, 20)
    check_columns(v.cpu().numpy(), vr, rtol=1e-10, what="v")
    # positions: compare the displacement from the initial state
    du_got = u.cpu().numpy() - ode.u0.cpu().numpy()
    du_ref = ur - ode.u0.cpu().numpy()
    check_columns(du_got, du_ref, rtol=1e-10, what="u displacement")


@pytest.mark.parametrize("stage_update", ["fused", "broadcasts"])
def test_callback_driver_matches_native_solver(tp, stage_update):
    """kick!/drift! driven from the host with the stage update of tpsph_rk_stage, or with
    LowStorageRK2N's single broadcasts through tpsph_axpby, equal the device solver bit for
    bit."""
    db, semi = setup_case(tp, 2, 0.02)
    ode = tp.semidiscretize(semi, (0.0, 5e-5))
    v1, u1 = tp.solve(ode, dt=1e-5, nsteps=5)
    seen = []
    v2, u2 = tp.solve(ode, dt=1e-5, nsteps=5, callback=lambda v, u, t, s: seen.append(s),
                      stage_update=stage_update)
    assert seen == [1, 2, 3, 4, 5]
    assert np.array_equal(v1.cpu().numpy(), v2.cpu().numpy())
    assert np.array_equal(u1.cpu().numpy(), u2.cpu().numpy())


@pytest.mark.parametrize("nd,dx,wall_viscosity", [(2, 0.02, None), (3, 0.04, None),
                                                  (3, 0.04, "adami")])
def test_native_solver_in_motion_matches_callback_driver(tp, nd, dx, wall_viscosity):
    """The device solver's stage kernel writes the records of the new state and raises the
    next kick's rebuild flag.  From a state in motion that rebuilds the lists every few
    steps it equals, bit for bit, the callback driver's kick!/drift! with the stage update of
    tpsph_rk_stage, and a kick from the caller after the solve sees the current records."""
    import ctypes as C
    import torch
    kw = {} if wall_viscosity is None else {"wall_viscosity": tp.ViscosityAdami(nu=1e-3)}
    db, semi = setup_case(tp, nd, dx, **kw)
    nsteps, dt = 12, 1e-4
    ode = tp.semidiscretize(semi, (0.0, nsteps * dt))
    rng = np.random.default_rng(5)
    v0 = ode.v0.clone()
    v0[:, :nd] = torch.from_numpy(rng.normal(0.0, 1.0, size=(v0.shape[0], nd))).to(v0)
    ode.v0 = v0
    lib = tp.native.load()
    count = C.c_int64()
    lib.tpsph_rebuild_count(semi.handle, C.byref(count))
    r0 = count.value
    v1, u1 = tp.solve(ode, dt=dt, nsteps=nsteps)
    semi.synchronize()
    lib.tpsph_rebuild_count(semi.handle, C.byref(count))
    assert count.value - r0 >= 3, "the state should leave its skin several times"
    v2, u2 = tp.solve(ode, dt=dt, nsteps=nsteps, callback=lambda v, u, t, s: None)
    assert np.array_equal(v1.cpu().numpy(), v2.cpu().numpy())
    assert np.array_equal(u1.cpu().numpy(), u2.cpu().numpy())
    # a kick from the caller after the solve gives the rates a fresh handle gives for the
    # solver's final state (in its own cell order: within 1e-12)
    dv = torch.full_like(v1, float("nan"))
    tp.kick(dv, v1, u1, {"semi": semi}, 0.0)
    _, semi2 = setup_case(tp, nd, dx, **kw)
    dv2 = torch.full_like(v1, float("nan"))
    tp.kick(dv2, v1, u1, {"semi": semi2}, 0.0)
    semi.synchronize()
    semi2.synchronize()
    check_columns(dv.cpu().numpy(), dv2.cpu().numpy())


@pytest.mark.parametrize("nd", [2, 3])
def test_particles_outside_domain_are_clamped(tp, orc, nd):
    """Particles leaving the grid box are clamped into border cells; the search stays exact,
    also for pairs that are both outside and close to each other."""
    db, semi = setup_case(tp, nd, 0.02 if nd == 2 else 0.04)
    v0, u0 = semi.initial_state()
    v, u = v0.cpu().numpy(), u0.cpu().numpy().copy()
    h = db.fluid_system.smoothing_length
    u[:5, 1] += 50.0  # far above the tank and the domain box
    u[5:8, 0] -= 30.0
    u[8, :] = u[5, :] + 0.7 * h  # a close pair outside the box
    dv_ref, *_ = oracle_kick(orc, semi, db, v, u)
    got = gpu_kick(tp, semi, v, u)
    check_columns(got, dv_ref)
    R = 2 * h
    import torch
    off, idx = semi.neighbors(torch.from_numpy(u).cuda())
    roff, ridx = orc.neighbors(nd, R, u, u, brute=True)
    assert np.array_equal(off, roff) and np.array_equal(idx.astype(np.int64), ridx)


@pytest.mark.parametrize("search", ["lists", "cells"])
def test_skin_rebuild_keeps_results_exact(tp, orc, search):
    """Kicks at positions that moved less and more than half the skin after the last
    rebuild both match the oracle (the cell order and lists are reused, then rebuilt)."""
    db, semi = setup_case(tp, 3, 0.025, grid={"skin": 0.2, "pair_search": search})
    v0, u0 = semi.initial_state()
    v, u = v0.cpu().numpy(), u0.cpu().numpy()
    R = 2 * db.fluid_system.smoothing_length
    rng = np.random.default_rng(3)
    gpu_kick(tp, semi, v, u)  # builds the cell list
    for amp in (0.04 * R, 0.09 * R, 0.4 * R, 0.01 * R):
        u = u + rng.uniform(-amp, amp, size=u.shape)
        dv_ref, *_ = oracle_kick(orc, semi, db, v, u)
        got = gpu_kick(tp, semi, v, u)
        check_columns(got, dv_ref)


def test_full_size_3d_kick_matches_oracle(tp, orc):
    """The 1.02M-particle 3D dam break (dx = 6.25e-3) at full size, with a random velocity
    field (so the viscosity and the continuity term contribute) and densities perturbed
    around the hydrostatic ones (so the diffusion term does)."""
    db, semi = setup_case(tp, 3, 6.25e-3)
    v0, u0 = semi.initial_state()
    v, u = v0.cpu().numpy().copy(), u0.cpu().numpy()
    rng = np.random.default_rng(17)
    v[:, :3] = rng.normal(0.0, 0.5, size=(v.shape[0], 3))
    v[:, 3] *= 1.0 + rng.uniform(-2e-3, 2e-3, size=v.shape[0])
    dv_ref, pf_ref, pw_ref, rw_ref = oracle_kick(orc, semi, db, v, u)
    got = gpu_kick(tp, semi, v, u)
    check_columns(got, dv_ref)
    pf, pw, rw = semi.update_pressures()
    assert rel_err(pw, pw_ref) <= RTOL and rel_err(rw, rw_ref) <= RTOL


def test_overflowed_lists_between_rebuilds(tp, orc):
    """With 40x the masses every list overflows and the kicks walk the cells over the SoA
    coordinates, which the gather keeps current between rebuilds only while some list has
    overflowed.  Kicks at u, at u moved by less than half the skin (no rebuild: the walk
    must read the moved positions) and then 5 native CK2N54 steps all match the oracle."""
    db = tp.dam_break(3, 0.025)
    db.fluid_system.mass = db.fluid_system.mass * 40.0
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend="hip", pair_search="lists")
    v0, u0 = semi.initial_state()
    v, u = v0.cpu().numpy().copy(), u0.cpu().numpy().copy()
    rng = np.random.default_rng(19)
    v[:, :3] = rng.normal(0.0, 0.2, size=(v.shape[0], 3))
    R = 2 * db.fluid_system.smoothing_length
    dv_ref, *_ = oracle_kick(orc, semi, db, v, u)
    check_columns(gpu_kick(tp, semi, v, u), dv_ref)
    u2 = u + rng.uniform(-0.005 * R, 0.005 * R, size=u.shape)
    dv_ref, *_ = oracle_kick(orc, semi, db, v, u2)
    check_columns(gpu_kick(tp, semi, v, u2), dv_ref)
    import torch
    ode = tp.DynamicalODEProblem(tp.kick, tp.drift, torch.from_numpy(v).cuda(),
                                 torch.from_numpy(u2).cuda(), (0.0, 5e-5), {"semi": semi})
    vs, us = tp.solve(ode, dt=1e-5, nsteps=5)
    P = orc.params_from_semi(semi)
    vr, ur = orc.solve_ck2n54(P, v, u2, db.fluid_system.mass, db.boundary_system.coordinates.T,
                              db.boundary_system.boundary_model.hydrodynamic_mass, 1e-5, 5)
    check_columns(vs.cpu().numpy(), vr, rtol=1e-11, what="v")
    check_columns(us.cpu().numpy() - u2, ur - u2, rtol=1e-11, what="u displacement")
