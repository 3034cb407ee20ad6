"""SummationDensity (density_calculators.jl:11, 26-48; fluid.jl:133-139; wcsph/system.jl:
236-273): the density of every fluid particle is summed over its fluid and wall neighbours
each kick, v_ode holds the velocity alone, and the pressure acceleration is
pressure_acceleration_summation_density (pressure_acceleration.jl:8-15).

Oracle pins (CPU):
  - an isolated particle has rho = m W(0) = m sigma h^-d (the closed-form normalisation of
    WendlandC2, 7/(4 pi) in 2D and 21/(16 pi) in 3D), a wall neighbour adds m_w W(r);
  - the summed densities equal a brute-force O(N^2) enumeration with the oracle's kernel
    (itself pinned by its integral and derivative) over fluid and wall particles, self
    included, on a jittered block next to a wall;
  - on a uniform lattice an interior particle gets rho0 up to the lattice quadrature error
    of the kernel (the sum of W over the lattice times dx^d tends to the integral, 1);
  - the reference's conservation tests with SummationDensity (test/schemes/fluid/rhs.jl:
    110-264): linear and angular momentum and the total energy with the summation form of
    the density rate, which the continuity-density pressure form does not conserve;
  - the pair force is antisymmetric (rhs.jl:4-107).
GPU parity (tests marked gpu): the kick of the HIP path (dv and the summed density) against
the oracle at 1e-12 in 2D and 3D with random velocities, with and without walls, in device
and host memory, and a 20-step CK2N54 trajectory."""
import numpy as np
import pytest

from conftest import gpu_available
from test_oracle_pins import (WC2, _angular_terms, _params, _pack, assert_cancels,
                              rectangular_patch)


def sum_params(orc, nd, h, **kw):
    P = _params(orc, nd, WC2, h, **kw)
    P.density_calculator = 1
    return P


@pytest.mark.parametrize("nd,sigma", [(2, 7.0 / (4.0 * np.pi)), (3, 21.0 / (16.0 * np.pi))])
def test_isolated_particle_density_is_mass_times_kernel_at_zero(orc, nd, sigma):
    h = 0.012
    P = sum_params(orc, nd, h)
    x = np.zeros((1, nd))
    m = np.array([0.37])
    rho = orc.summation_density(P, x, m)
    assert rho[0] == pytest.approx(0.37 * sigma / h ** nd, rel=1e-15)
    # a wall neighbour at r = 0.9 h adds m_w W(r); one beyond 2h adds nothing
    xw = np.zeros((2, nd))
    xw[0, 0], xw[1, 0] = 0.9 * h, 2.0 * h
    mw = np.array([0.5, 7.0])
    rho_w = orc.summation_density(P, x, m, xw, mw)
    q = 0.9
    w = sigma / h ** nd * (1 - q / 2) ** 4 * (2 * q + 1)
    assert rho_w[0] == pytest.approx(0.37 * sigma / h ** nd + 0.5 * w, rel=1e-14)


def _brute(orc, nd, h, x, m, xw, mw):
    R2 = (2 * h) ** 2
    rho = np.zeros(len(x))
    for src, mass in ((x, m), (xw, mw)):
        for a in range(len(x)):
            d = x[a] - src
            d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + (d[:, 2] * d[:, 2] if nd == 3 else 0.0)
            for b in np.nonzero(d2 <= R2)[0]:
                rho[a] += mass[b] * orc.kernel(WC2, nd, float(np.sqrt(d2[b])), h)
    return rho


@pytest.mark.parametrize("nd", [2, 3])
def test_summed_density_matches_brute_force(orc, nd):
    dx = 0.02
    h = 1.2 * dx
    rng = np.random.default_rng(8)
    n = 9 if nd == 2 else 5
    g = np.stack(np.meshgrid(*([np.arange(n)] * nd), indexing="ij"), -1).reshape(-1, nd)
    x = dx * (g + 0.5) + rng.uniform(-0.2 * dx, 0.2 * dx, size=g.shape)
    m = 1000.0 * dx ** nd * (1.0 + rng.uniform(-0.05, 0.05, size=len(x)))
    # three wall layers below the block
    gw = np.stack(np.meshgrid(*([np.arange(-1, n + 1)] * (nd - 1) + [np.arange(3)]),
                              indexing="ij"), -1).reshape(-1, nd)
    xw = np.zeros((len(gw), nd))
    xw[:, :nd - 1] = dx * (gw[:, :nd - 1] + 0.5)
    xw[:, nd - 1] = -dx * (gw[:, nd - 1] + 0.5)
    mw = np.full(len(xw), 1000.0 * dx ** nd)
    P = sum_params(orc, nd, h)
    got = orc.summation_density(P, x, m, xw, mw)
    ref = _brute(orc, nd, h, x, m, xw, mw)
    assert np.abs(got - ref).max() <= 1e-13 * ref.max()
    # the wall terms matter next to the wall and the self terms everywhere
    fluid_only = orc.summation_density(P, x, m)
    assert (got - fluid_only).max() > 0.1 * ref.max()
    assert fluid_only.min() > m.min() * orc.kernel(WC2, nd, 0.0, h)


@pytest.mark.parametrize("nd", [2, 3])
def test_uniform_lattice_interior_density_converges_to_reference_density(orc, nd):
    """sum_b m W(x_b) = rho0 sum_b W(x_b) dx^d is a quadrature of the kernel's integral (1):
    an interior particle of a uniform lattice gets rho0 up to the lattice error, which
    shrinks as h / dx grows (measured 1.8 %, 0.3-0.5 %, 0.05-0.12 % at h / dx = 1.2, 1.5, 2);
    every interior particle gets the same value (the lattice sum is translation invariant)
    and a corner particle less."""
    dx, rho0 = 0.01, 1000.0
    n = 21 if nd == 2 else 13
    g = np.stack(np.meshgrid(*([np.arange(n)] * nd), indexing="ij"), -1).reshape(-1, nd)
    x = dx * (g + 0.5)
    m = np.full(len(x), rho0 * dx ** nd)
    interior = np.all((g >= 4) & (g <= n - 5), axis=1)
    errs = []
    for factor, tol in ((1.2, 2.5e-2), (1.5, 6e-3), (2.0, 1.5e-3)):
        rho = orc.summation_density(sum_params(orc, nd, factor * dx), x, m)
        ri = rho[interior]
        errs.append(np.abs(ri / rho0 - 1.0).max())
        assert errs[-1] <= tol, (factor, errs[-1])
        assert np.ptp(ri) <= 1e-12 * rho0
        assert rho[0] < ri.min()
    assert errs[0] > errs[1] > errs[2]


def _drho_summation(orc, nd, h, x, m, v):
    """The rate of the summed density, sum_b m_b v_ab . grad W_ab (rhs.jl:226-246)."""
    drho = np.zeros(len(x))
    for a in range(len(x)):
        for b in range(len(x)):
            pd = x[a] - x[b]
            r = float(np.sqrt(np.sum(pd * pd)))
            if r < np.sqrt(np.finfo(float).eps) or r >= 2 * h:
                continue
            grad = orc.lib().oracle_kernel_deriv_div_r(WC2, nd, r, h) * pd
            drho[a] += m[b] * np.dot(v[a] - v[b], grad)
    return drho


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_summation_density_conserves_momentum_and_energy(orc, nd, seed):
    dx = 0.1
    h = 1.2 * dx
    x, m, rho, p, v = rectangular_patch(nd, dx, seed=seed)
    P = sum_params(orc, nd, h)
    dv = orc.interact_fluid_fluid(P, x, _pack(v, rho), m, p)
    acc = dv[:, :nd]
    assert np.all(dv[:, nd] == 0.0)  # no continuity equation
    for d in range(nd):
        assert_cancels(m * acc[:, d], f"linear momentum {d}")
    ang = _angular_terms(m, x, acc)
    for d in ([2] if nd == 2 else [0, 1, 2]):
        assert_cancels(ang[:, d], f"angular momentum {d}")
    drho = _drho_summation(orc, nd, h, x, m, v)
    energy = np.concatenate([(m[:, None] * v * acc).ravel(), m * p / rho ** 2 * drho])
    assert_cancels(energy, "energy")
    # the continuity-density pressure form does not conserve this energy
    Pc = _params(orc, nd, WC2, h)
    accc = orc.interact_fluid_fluid(Pc, x, _pack(v, rho), m, p)[:, :nd]
    wrong = np.concatenate([(m[:, None] * v * accc).ravel(), m * p / rho ** 2 * drho])
    assert abs(wrong.sum()) > 1e-6 * np.abs(wrong).sum()


@pytest.mark.parametrize("masses", [(0.01, 0.01), (0.73, 0.31)])
@pytest.mark.parametrize("dens", [(1000.0, 1000.0), (900.0, 1201.0), (1003.0, 353.4)])
@pytest.mark.parametrize("press", [(0.0, 0.0), (10.0, 10_000.0), (1000.0, -1000.0)])
def test_summation_pressure_acceleration_is_antisymmetric(orc, masses, dens, press):
    P = sum_params(orc, 2, 0.12)
    x = np.array([[0.0, 0.0], [0.07, -0.05]])
    m = np.array(masses)
    dv = orc.interact_fluid_fluid(P, x, _pack(np.zeros((2, 2)), np.array(dens)), m,
                                  np.array(press))
    f_ab, f_ba = m[0] * dv[0, :2], m[1] * dv[1, :2]
    assert np.allclose(f_ab, -f_ba, rtol=2 * np.finfo(float).eps, atol=0.0)


def test_oracle_kick_layout_and_solve(tp, orc):
    """The oracle kick with SummationDensity takes v with nd rows, returns dv with nd rows
    and the summed density; a CK2N54 run keeps the layout."""
    db = tp.dam_break(2, 0.04, density_calculator=tp.SummationDensity())
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system)
    P = orc.params_from_semi(semi)
    assert P.density_calculator == 1
    ic = db.fluid_system.initial_condition
    v = np.ascontiguousarray(ic.velocity.T)
    u = np.ascontiguousarray(ic.coordinates.T)
    m, xw = db.fluid_system.mass, db.boundary_system.coordinates.T
    mw = db.boundary_system.boundary_model.hydrodynamic_mass
    dv, pf, pw, rw, rho = orc.kick(P, v, u, m, xw, mw, return_rho=True)
    assert dv.shape == v.shape == (db.n_fluid, 2)
    assert np.array_equal(rho, orc.summation_density(P, u, m, xw, mw))
    rho0 = db.state_equation.reference_density
    # the hydrostatic masses and the walls give densities near rho0 inside the water
    assert np.median(rho) == pytest.approx(rho0, rel=0.02)
    v2, u2 = orc.solve_ck2n54(P, v, u, m, xw, mw, 1e-4, 3)
    assert v2.shape == v.shape and np.isfinite(v2).all()
    semi.close()


# ------------------------------------------------------------------------------------------
# GPU parity
# ------------------------------------------------------------------------------------------
needs_gpu = pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")


def rel_err(got, ref):
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))


def _state(db, seed):
    ic = db.fluid_system.initial_condition
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, 0.3, size=ic.velocity.T.shape)
    return np.ascontiguousarray(v), np.ascontiguousarray(ic.coordinates.T)


@pytest.mark.gpu
@needs_gpu
@pytest.mark.parametrize("search", ["lists", "cells"])
@pytest.mark.parametrize("nd,dx,walls,backend", [(2, 0.01, True, "hip"), (2, 0.02, False, "hip"),
                                                  (3, 0.02, True, "hip"),
                                                  (2, 0.02, True, "hip_host")])
def test_summation_kick_matches_oracle(tp, orc, nd, dx, walls, backend, search):
    """The summation pass and the fluid pass over the Verlet lists (the brick kernel) or the
    cell walk, against the oracle."""
    import torch
    db = tp.dam_break(nd, dx, density_calculator=tp.SummationDensity())
    systems = (db.fluid_system, db.boundary_system) if walls else (db.fluid_system,)
    semi = tp.Semidiscretization(*systems, parallelization_backend=backend, pair_search=search)
    v, u = _state(db, 3)
    P = orc.params_from_semi(semi)
    xw = db.boundary_system.coordinates.T if walls else None
    mw = db.boundary_system.boundary_model.hydrodynamic_mass if walls else None
    dv_ref, pf_ref, pw_ref, rw_ref, rho_ref = orc.kick(P, v, u, db.fluid_system.mass, xw, mw,
                                                       return_rho=True)
    if backend == "hip":
        vd, ud = torch.from_numpy(v).cuda(), torch.from_numpy(u).cuda()
        dv = torch.full_like(vd, float("nan"))
        tp.kick(dv, vd, ud, {"semi": semi}, 0.0)
        got = dv.cpu().numpy()
    else:
        got = np.full_like(v, np.nan)
        tp.kick(got, v, u, {"semi": semi}, 0.0)
    assert got.shape == (db.n_fluid, nd)
    for d in range(nd):
        assert rel_err(got[:, d], dv_ref[:, d]) <= 1e-12, d
    rho = semi.density()
    assert rel_err(rho, rho_ref) <= 1e-13
    pf, pw, rw = semi.update_pressures()
    assert rel_err(pf, pf_ref) <= 1e-12
    if walls:
        assert rel_err(pw, pw_ref) <= 1e-12 and rel_err(rw, rw_ref) <= 1e-13
    du = semi._zeros(*u.shape)
    tp.drift(du, vd if backend == "hip" else v, ud if backend == "hip" else u,
             {"semi": semi}, 0.0)
    du = du.cpu().numpy() if backend == "hip" else du
    assert np.array_equal(du, v)
    semi.close()


@pytest.mark.gpu
@needs_gpu
@pytest.mark.parametrize("search", ["lists", "cells"])
def test_summation_trajectory_matches_oracle(tp, orc, search):
    db = tp.dam_break(2, 0.02, density_calculator=tp.SummationDensity())
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system, pair_search=search)
    ode = tp.semidiscretize(semi, (0.0, 20e-5))
    v, u = tp.solve(ode, dt=1e-5, nsteps=20)
    P = orc.params_from_semi(semi)
    vr, ur = orc.solve_ck2n54(P, ode.v0.cpu().numpy(), ode.u0.cpu().numpy(),
                              db.fluid_system.mass, db.boundary_system.coordinates.T,
                              db.boundary_system.boundary_model.hydrodynamic_mass, 1e-5, 20)
    assert rel_err(v.cpu().numpy(), vr) <= 1e-11
    assert np.abs(u.cpu().numpy() - ur).max() <= 1e-13
    semi.close()
