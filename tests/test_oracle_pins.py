"""Pins of the CPU oracle (oracle/sph_oracle.c) against what the reference's own tests fix
and against what the mathematics fixes.

Each test names the reference test it follows.  Where the reference draws random input with
Julia's RNG (test/rectangular_patch.jl), the same construction is drawn from numpy's seeded
generator: the asserted properties (conservation, antisymmetry, round trips) hold for any
draw, so the reference's tolerances carry over."""
import math

import numpy as np
import pytest

WC2, CUBIC = 0, 1


def julia_eps(x):
    return np.nextafter(x, np.inf) - x


# ---------------------------------------------------------------------------------------
# Smoothing kernels (test/general/smoothing_kernels.jl:3-133)
# ---------------------------------------------------------------------------------------
def _radial_integral(orc, which, nd, pieces):
    """Integral of W over R^nd with h = 1, by Gauss-Legendre on each polynomial piece (exact
    for these piecewise polynomials up to rounding)."""
    xg, wg = np.polynomial.legendre.leggauss(12)
    total = 0.0
    for a, b in pieces:
        r = 0.5 * (b - a) * xg + 0.5 * (a + b)
        f = np.array([orc.kernel(which, nd, ri, 1.0) for ri in r])
        if nd == 1:
            g = 2.0 * f
        elif nd == 2:
            g = 2.0 * math.pi * r * f
        else:
            g = 4.0 * math.pi * r * r * f
        total += 0.5 * (b - a) * np.dot(wg, g)
    return total


@pytest.mark.parametrize("which,nd,pieces", [
    (WC2, 2, [(0.0, 2.0), (2.0, 2.2)]),
    (WC2, 3, [(0.0, 2.0), (2.0, 2.2)]),
    (CUBIC, 1, [(0.0, 1.0), (1.0, 2.0), (2.0, 2.2)]),
    (CUBIC, 2, [(0.0, 1.0), (1.0, 2.0), (2.0, 2.2)]),
    (CUBIC, 3, [(0.0, 1.0), (1.0, 2.0), (2.0, 2.2)]),
])
def test_kernel_integrates_to_one(orc, which, nd, pieces):
    """The normalisation factors (7/(4 pi) h^-2, 21/(16 pi) h^-3, 10/(7 pi) h^-2 ...) make
    the integral 1; the piece beyond the support checks the compact-support guard."""
    assert abs(_radial_integral(orc, which, nd, pieces) - 1.0) <= 1e-15


def _fd_derivative(f, r, d):
    """Eighth-order central difference: exact (up to rounding) for polynomials of degree
    <= 8, and each kernel is a polynomial of degree <= 5 on the piece around r."""
    c = [4 / 5, -1 / 5, 4 / 105, -1 / 280]
    return sum(ck * (f(r + (k + 1) * d) - f(r - (k + 1) * d)) for k, ck in enumerate(c)) / d


@pytest.mark.parametrize("which", [WC2, CUBIC])
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_kernel_derivative_is_derivative_of_kernel(orc, which, nd):
    if which == WC2 and nd == 1:
        pytest.skip("WendlandC2Kernel has no 1D normalisation")
    for h in (0.25, 0.5, 0.75, 1.0):
        cs = 2.0 * h
        d = 1e-3 * h
        scale = max(abs(orc.kernel_deriv(which, nd, r, h)) for r in np.linspace(0, cs, 41))
        for q in np.arange(0.05, 1.0, 0.1):
            r = q * cs
            if which == CUBIC and abs(r - h) < 5 * d:
                continue  # the cubic spline changes polynomial piece at r = h
            fd = _fd_derivative(lambda x: orc.kernel(which, nd, x, h), r, d)
            assert abs(orc.kernel_deriv(which, nd, r, h) - fd) <= 1e-10 * scale
            # kernel_deriv_div_r, used in the pair terms, is kernel_deriv / r
            kdr = orc.lib().oracle_kernel_deriv_div_r(which, nd, r, h)
            assert kdr == pytest.approx(orc.kernel_deriv(which, nd, r, h) / r, rel=1e-14)
        # zero at and beyond the compact support
        assert orc.kernel(which, nd, cs, h) == 0.0
        assert orc.kernel_deriv(which, nd, 1.01 * cs, h) == 0.0


# ---------------------------------------------------------------------------------------
# StateEquationCole (test/schemes/fluid/weakly_compressible_sph/state_equation.jl)
# ---------------------------------------------------------------------------------------
ATM = 101_325.0


@pytest.mark.parametrize("exponent,values", [
    # state_equation.jl:24-37
    (7.15, [(1, 998.34), (100, 1002.8323123356663), (500, 1019.8235062499685),
            (1000, 1038.8747989986027), (3000, 1099.0607413267035),
            (9000, 1210.4689472510186)]),
    # state_equation.jl:85-98
    (1.0, [(1, 998.34), (100, 1002.8949541016121), (500, 1021.298809057621),
           (1000, 1044.3036277526319), (3000, 1136.3229025326757),
           (9000, 1412.380726872807)]),
])
def test_inverse_state_equation_known_values(orc, exponent, values):
    """Bit-exact, as the reference asserts (==): the C library's pow rounds these
    Float64^Float64 powers like Julia's."""
    for p_atm, rho in values:
        got = orc.inverse_state_equation(1484.0, 998.34, exponent, 1.0 * ATM, p_atm * ATM)
        assert got == rho, (p_atm, got, rho)


def test_state_equation_background_pressure_and_clipping(orc):
    """state_equation.jl:41-61."""
    for p_bg in (0.0, 10_000.0, 100_000.0, 200_000.0):
        se = lambda rho: orc.state_equation(10.0, 1000.0, 7.0, p_bg, False, rho)  # noqa: E731
        assert se(1000.0) == p_bg
        assert se(1001.0) > p_bg + 10
        assert se(999.0) < p_bg - 10
    for rho in (999.0, 900.0):
        assert orc.state_equation(10.0, 1000.0, 7.0, 0.0, True, rho) == 0.0


@pytest.mark.parametrize("c,rho0,exponent,p_bg", [
    (1484.0, 998.34, 7.15, 101_325.0), (10.0, 1000.0, 7.0, 10_000.0),
    (10.0, 1000.0, 7.0, 0.0), (10.0, 1000.0, 7.0, -100_000.0),
    (1484.0, 998.34, 1.0, 101_325.0), (10.0, 1000.0, 1.0, 100_000.0),
    (10.0, 1000.0, 1.0, 90_000.0), (10.0, 1000.0, 1.0, 0.0), (10.0, 1000.0, 1.0, -100_000.0),
    (20 * math.sqrt(9.81 * 0.6), 1000.0, 7.0, 0.0),  # the dam-break equation of state
])
def test_state_equation_round_trip(orc, c, rho0, exponent, p_bg):
    """state_equation.jl:103-141: inverse(eos(rho)) ~ rho (Julia's default rtol sqrt(eps))
    and eos(inverse(p)) ~ p within atol 2e-7, rtol 1e-10."""
    for rho in (100.0, 500.0, 900.0, 990.0, 1000.0, 1005.0, 1100.0, 1600.0):
        p = orc.state_equation(c, rho0, exponent, p_bg, False, rho)
        assert orc.inverse_state_equation(c, rho0, exponent, p_bg, p) == \
            pytest.approx(rho, rel=math.sqrt(np.finfo(float).eps))
    for p in (-100.0, 0.0, 100.0, 10_000.0, 100_000.0, 100_000_000.0):
        rho = orc.inverse_state_equation(c, rho0, exponent, p_bg, p)
        back = orc.state_equation(c, rho0, exponent, p_bg, False, rho)
        assert abs(back - p) <= 2e-7 + 1e-10 * abs(p)


def test_cole_closed_form_for_integer_exponent(orc):
    """exponent 7 is an integer: p = B ((rho/rho0)^7 - 1) with B = rho0 c^2 / 7, evaluated
    here with exact rational arithmetic."""
    from fractions import Fraction as F
    c, rho0 = 20 * math.sqrt(9.81 * 0.6), 1000.0
    for rho in (990.0, 999.9, 1000.0, 1000.1, 1013.0):
        exact = F(rho0) * F(c) ** 2 / 7 * ((F(rho) / F(rho0)) ** 7 - 1)
        got = orc.state_equation(c, rho0, 7.0, 0.0, False, rho)
        assert abs(got - float(exact)) <= 4 * julia_eps(abs(float(exact))) + 1e-9


# ---------------------------------------------------------------------------------------
# ArtificialViscosityMonaghan (test/schemes/fluid/viscosity.jl:1-42)
# ---------------------------------------------------------------------------------------
def _params(orc, nd, which, h, c=0.0, alpha=0.0, beta=0.0, epsilon=0.01, rho0=1000.0,
            exponent=7.0, visc=False, delta=0.0, gravity=(0.0, 0.0, 0.0), clip_wall=False):
    P = orc.OracleParams()
    P.ndims, P.kernel, P.smoothing_length = nd, which, h
    P.sound_speed, P.reference_density, P.exponent = c, rho0, exponent
    P.alpha, P.beta, P.epsilon = alpha, beta, epsilon
    P.use_viscosity = int(visc)
    P.use_density_diffusion = int(delta > 0)
    P.delta = delta
    for d in range(3):
        P.gravity[d] = gravity[d]
    P.clip_wall_pressure = int(clip_wall)
    return P


def test_monaghan_viscosity_known_answer(orc):
    import ctypes as C
    h = 1.2 * 0.2
    c = 10 * math.sqrt(9.81 * 0.9)
    P = _params(orc, 2, CUBIC, h, c=c, alpha=0.02, beta=0.0, visc=True)
    pos_diff = np.array([-0.25 * h, 0.375 * h])
    distance = float(np.linalg.norm(pos_diff))
    grad = orc.lib().oracle_kernel_deriv_div_r(CUBIC, 2, distance, h) * pos_diff
    v_a, v_b = np.array([0.3, -1.0]), np.zeros(2)
    dv = np.zeros(2)
    orc.lib().oracle_av_monaghan(C.byref(P), pos_diff.ctypes.data, distance, 1.0, 1000.0,
                                 1000.0, v_a.ctypes.data, v_b.ctypes.data, grad.ctypes.data,
                                 dv.ctypes.data)
    assert abs(dv[0] - -0.02049217623299368) <= 6e-15
    assert abs(dv[1] - 0.03073826434949052) <= 6e-15


# ---------------------------------------------------------------------------------------
# Random patches (test/rectangular_patch.jl) and the conservation tests
# (test/schemes/fluid/rhs.jl:110-264, test/schemes/boundary/dummy_particles/rhs.jl)
# ---------------------------------------------------------------------------------------
def rectangular_patch(nd, dx=0.1, n=3, seed=1, density=1000.0):
    """n^nd lattice centred at the origin; positions, mass, density, pressure and velocity
    perturbed uniformly like the reference's patch (amplitudes 0.5 dx, 0.1 m, 0.1 rho,
    2000, 0.5 dx); the centre particle sits exactly at the origin."""
    rng = np.random.default_rng(seed)
    idx = np.stack(np.meshgrid(*([np.arange(n)] * nd), indexing="ij"), -1).reshape(-1, nd)
    idx = idx[:, ::-1]  # x fastest, like the reference's column-major lattice
    x = -dx / 2 * n + dx * (idx + 0.5)
    N = x.shape[0]
    pert = lambda a, shape: 2 * a * rng.random(shape) - a  # noqa: E731
    x = x + pert(0.5 * dx, x.shape)
    x[(N + 1) // 2 - 1] = 0.0
    m = dx ** nd * density * np.ones(N)
    m = m + pert(0.1 * m[0], N)
    rho = density + pert(0.1 * density, N)
    p = pert(2000.0, N)
    v = pert(0.5 * dx, (N, nd))
    return x, m, rho, p, v


def _pack(v, rho):
    return np.ascontiguousarray(np.column_stack([v, rho]))


def _energy_terms(m, v, dv, p, rho, drho):
    return np.concatenate([(m[:, None] * v * dv).ravel(), m * p / rho ** 2 * drho])


def assert_cancels(terms, what):
    """sum(terms) == 0 up to rounding: |sum| <= 16 eps sum|terms|.  The reference states
    absolute tolerances (5e-14, 4e-15, 6e-15) tuned to its own random draws; relative to
    the size of the summed terms they are of the same order.  A wrong term, sign or index
    leaves a residual of the order of the terms themselves."""
    terms = np.asarray(terms, dtype=np.float64).ravel()
    total, scale = abs(float(np.sum(terms))), float(np.sum(np.abs(terms)))
    assert scale > 1e-3, f"{what}: the terms vanish, nothing is tested"
    assert total <= 16 * np.finfo(float).eps * scale, f"{what}: {total:.3e} vs {scale:.3e}"


def _angular_terms(m, x, acc):
    nd = x.shape[1]
    r3 = np.zeros((len(m), 3))
    a3 = np.zeros((len(m), 3))
    r3[:, :nd], a3[:, :nd] = x, acc
    return m[:, None] * np.cross(r3, a3)


CONSERVATION = [(2, CUBIC), (2, WC2), (3, CUBIC), (3, WC2)]


@pytest.mark.parametrize("nd,which", CONSERVATION)
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fluid_fluid_conservation(orc, nd, which, seed):
    """Linear momentum, angular momentum and total energy of the fluid-fluid pass with
    pressure and continuity terms are conserved."""
    dx = 0.1
    P = _params(orc, nd, which, 1.2 * dx)
    x, m, rho, p, v = rectangular_patch(nd, dx, seed=seed)
    dv = orc.interact_fluid_fluid(P, x, _pack(v, rho), m, p)
    acc, drho = dv[:, :nd], dv[:, nd]
    for d in range(nd):
        assert_cancels(m * acc[:, d], f"linear momentum {d}")
    ang = _angular_terms(m, x, acc)
    for d in ([2] if nd == 2 else [0, 1, 2]):
        assert_cancels(ang[:, d], f"angular momentum {d}")
    assert_cancels(_energy_terms(m, v, acc, p, rho, drho), "energy")


@pytest.mark.parametrize("nd,which", CONSERVATION)
def test_viscosity_and_diffusion_conserve_momentum(orc, nd, which):
    """Monaghan viscosity acts along r_ab and density diffusion only on drho, so linear and
    angular momentum stay conserved with both switched on; diffusion conserves
    sum_a V_a^0 ... only approximately, so mass is not asserted."""
    dx = 0.1
    P = _params(orc, nd, which, 1.2 * dx, c=30.0, alpha=0.5, beta=0.3, visc=True, delta=0.1)
    x, m, rho, p, v = rectangular_patch(nd, dx, seed=4)
    v = v * 100.0  # make viscosity active on many pairs
    dv = orc.interact_fluid_fluid(P, x, _pack(v, rho), m, p)
    acc = dv[:, :nd]
    P0 = _params(orc, nd, which, 1.2 * dx)
    dv0 = orc.interact_fluid_fluid(P0, x, _pack(v, rho), m, p)
    assert np.abs(dv[:, :nd] - dv0[:, :nd]).max() > 1e-3  # viscosity contributed
    assert np.abs(dv[:, nd] - dv0[:, nd]).max() > 1e-3  # diffusion contributed
    for d in range(nd):
        assert_cancels(m * acc[:, d], f"linear momentum {d}")
    ang = _angular_terms(m, x, acc)
    for d in ([2] if nd == 2 else [0, 1, 2]):
        assert_cancels(ang[:, d], f"angular momentum {d}")


@pytest.mark.parametrize("masses", [(0.01, 0.01), (0.73, 0.31)])
@pytest.mark.parametrize("dens", [(1000.0, 1000.0), (900.0, 1201.0), (1003.0, 353.4)])
@pytest.mark.parametrize("press", [(0.0, 0.0), (10_000.0, 10_000.0), (10.0, 10_000.0),
                                   (1000.0, -1000.0)])
def test_pressure_acceleration_is_antisymmetric(orc, masses, dens, press):
    """test/schemes/fluid/rhs.jl:4-107: m_a dv_ab = -m_b dv_ba within rtol 2 eps."""
    h = 0.12
    P = _params(orc, 2, WC2, h)
    x = np.array([[0.0, 0.0], [0.07, -0.05]])
    v = np.zeros((2, 2))
    m = np.array(masses)
    dv = orc.interact_fluid_fluid(P, x, _pack(v, np.array(dens)), m, np.array(press))
    f_ab, f_ba = m[0] * dv[0, :2], m[1] * dv[1, :2]
    assert np.allclose(f_ab, -f_ba, rtol=2 * np.finfo(float).eps, atol=0.0)
    if press[0] + press[1] != 0.0:  # the pair force is proportional to p_a + p_b
        assert np.abs(f_ab).max() > 0
    else:
        assert np.all(f_ab == 0.0) and np.all(f_ba == 0.0)


@pytest.mark.parametrize("nd,which", CONSERVATION)
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fluid_wall_energy_conservation(orc, nd, which, seed):
    """test/schemes/boundary/dummy_particles/rhs.jl: the patch split at its centre particle
    into fluid and dummy wall (wall pressure and density given).  With the wall density
    following the continuity equation drho_w = sum_a (rho_w / rho_a) m_a (0 - v_a).grad W_wa
    the total energy rate vanishes."""
    dx = 0.1
    h = 1.2 * dx
    P = _params(orc, nd, which, h)
    x, m, rho, p, v = rectangular_patch(nd, dx, seed=seed)
    c = (len(m) + 1) // 2
    xf, mf, rf, pf, vf = x[:c], m[:c], rho[:c], p[:c], v[:c]
    xw, mw, rw, pw = x[c:], m[c:], rho[c:], p[c:]
    vv = _pack(vf, rf)
    dv = orc.interact_fluid_fluid(P, xf, vv, mf, pf)
    dv = orc.interact_fluid_wall(P, xf, vv, pf, xw, mw, pw, rw, dv=dv)
    terms = [_energy_terms(mf, vf, dv[:, :nd], pf, rf, dv[:, nd])]
    # wall continuity, written out here as the check's own term
    drho_w = np.zeros(len(mw))
    for w in range(len(mw)):
        for a in range(c):
            pd = xw[w] - xf[a]
            r = float(np.sqrt(np.sum(pd * pd)))
            if r < 2 * h and r * r > julia_eps(h * h):
                grad = orc.lib().oracle_kernel_deriv_div_r(which, nd, r, h) * pd
                drho_w[w] += rw[w] / rf[a] * mf[a] * np.dot(-vf[a], grad)
    terms.append(mw * pw / rw ** 2 * drho_w)
    assert np.abs(drho_w).max() > 0
    assert_cancels(np.concatenate(terms), "energy")


# ---------------------------------------------------------------------------------------
# Adami pressure extrapolation (test/schemes/boundary/dummy_particles/dummy_particles.jl)
# ---------------------------------------------------------------------------------------
def _adami_tank(tp, density, acceleration=None, state_equation=None, n_layers=2):
    return tp.RectangularTank(0.1, (1.0, 1.0), (1.0, 1.0), density, n_layers=n_layers,
                              faces=(True, True, True, False), acceleration=acceleration,
                              state_equation=state_equation)


def _adami(orc, P, tank, rho_fluid, pf):
    x = np.ascontiguousarray(tank.fluid.coordinates.T)
    xw = np.ascontiguousarray(tank.boundary.coordinates.T)
    v = _pack(np.zeros_like(x), rho_fluid)
    return orc.adami(P, xw, x, v, pf)


def _adami_params(orc, gravity=(0.0, 0.0, 0.0), clip=False):
    return _params(orc, 2, CUBIC, 1.5 * 0.1, c=10.0, rho0=257.0, exponent=7.0,
                   gravity=gravity, clip_wall=clip)


def test_adami_constant_zero_pressure(tp, orc):
    """dummy_particles.jl:338-368: rho = rho0 everywhere gives p_w == 0 exactly."""
    P = _adami_params(orc)
    tank = _adami_tank(tp, 257.0)
    rho = tank.fluid.density
    pf = orc.compute_pressure(P, _pack(np.zeros((len(rho), 2)), rho))
    assert np.all(pf == 0.0)
    pw, rw, vol = _adami(orc, P, tank, rho, pf)
    assert np.all(pw == 0.0)
    assert np.all(rw == 257.0)
    assert vol.max() > 0


def test_adami_constant_nonzero_pressure(tp, orc):
    """dummy_particles.jl:373-410: rho = 260 != rho0 gives p_w = p_f within 1e-12."""
    P = _adami_params(orc)
    tank = _adami_tank(tp, 260.0)
    rho = tank.fluid.density
    pf = orc.compute_pressure(P, _pack(np.zeros((len(rho), 2)), rho))
    assert np.allclose(pf, pf[0]) and pf[0] > 0
    pw, rw, _ = _adami(orc, P, tank, rho, pf)
    assert np.abs(pw - pf[0]).max() <= 1e-12
    assert np.abs(rw - 260.0).max() <= 1e-10


def test_adami_hydrostatic_pressure_gradient(tp, orc):
    """dummy_particles.jl:611-697: the extrapolated wall pressure continues the fluid's
    hydrostatic gradient; compared with a wall-less tank as large as fluid plus walls,
    within atol 4.0 (pressures reach ~3000 Pa)."""
    g = (0.0, -9.81)
    se = tp.StateEquationCole(sound_speed=10.0, reference_density=257.0, exponent=7.0)
    P = _adami_params(orc, gravity=(0.0, -9.81, 0.0))
    tank = _adami_tank(tp, 257.0, acceleration=g, state_equation=se)
    rho = tank.fluid.density
    pf = orc.compute_pressure(P, _pack(np.zeros((len(rho), 2)), rho))
    pw, _, _ = _adami(orc, P, tank, rho, pf)
    nl, n, dx = 2, 10, 0.1
    ref = tp.RectangularTank(dx, (dx * (n + 2 * nl), dx * (n + nl)),
                             (dx * (n + 2 * nl), dx * (n + nl)), 257.0, acceleration=g,
                             state_equation=se, n_layers=0, faces=(True, True, True, False))
    grid = np.full((n + 2 * nl, n + nl), np.nan)
    grid_ref = np.full_like(grid, np.nan)

    def put(target, coords, offset, values):
        idx = np.rint(coords / dx + offset).astype(int) - 1
        target[idx[0], idx[1]] = values

    put(grid, tank.boundary.coordinates, nl + 0.5, pw)
    put(grid, tank.fluid.coordinates, nl + 0.5, pf)
    pref = np.array([se(r) for r in ref.fluid.density])
    put(grid_ref, ref.fluid.coordinates, 0.5, pref)
    assert not np.isnan(grid).any() and not np.isnan(grid_ref).any()
    assert np.abs(grid - grid_ref).max() <= 4.0
    assert pref.max() > 2000.0


def test_adami_clips_negative_pressure(tp, orc):
    """dummy_particles.jl:13-68: with clip_negative_pressure a negative extrapolated
    pressure becomes 0 (and the wall density rho0); without it stays negative."""
    tank = _adami_tank(tp, 250.0)
    rho = tank.fluid.density
    for clip in (True, False):
        P = _adami_params(orc, clip=clip)
        pf = orc.compute_pressure(P, _pack(np.zeros((len(rho), 2)), rho))
        assert pf.max() < 0
        pw, rw, _ = _adami(orc, P, tank, rho, pf)
        near = pw != 0.0
        if clip:
            assert np.all(pw == 0.0) and np.all(rw == 257.0)
        else:
            assert near.any() and np.all(pw[near] < 0)


def test_adami_isolated_wall_particle_has_zero_pressure(orc):
    """A wall particle without fluid neighbours keeps volume 0 and pressure 0
    (dummy_particles.jl:547-569: p_w = P_w / V_w only if V_w > eps)."""
    P = _adami_params(orc)
    x = np.array([[0.0, 0.0]])
    xw = np.array([[5.0, 5.0], [0.1, 0.0]])
    v = _pack(np.zeros((1, 2)), np.array([260.0]))
    pf = orc.compute_pressure(P, v)
    pw, rw, vol = orc.adami(P, xw, x, v, pf)
    assert pw[0] == 0.0 and vol[0] == 0.0 and rw[0] == 257.0
    assert pw[1] == pytest.approx(pf[0], rel=1e-15)


# ---------------------------------------------------------------------------------------
# CarpenterKennedy2N54 (Carpenter & Kennedy 1994, solution 3, five stages, fourth order)
# ---------------------------------------------------------------------------------------
def test_ck2n54_linear_stability_polynomial(orc):
    """Applied to y' = z y the 2N-storage stages du = A du + z u; u += B du give
    u1 = R(z) u0 with R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24 + c5 z^5 (fourth order)."""
    A, B, _ = orc.ck2n54_coefficients()
    P = np.polynomial.Polynomial
    u, du = P([1.0]), P([0.0])
    z = P([0.0, 1.0])
    for i in range(5):
        du = A[i] * du + z * u
        u = u + B[i] * du
    coef = u.coef
    expected = [1.0, 1.0, 1 / 2, 1 / 6, 1 / 24]
    assert np.abs(coef[:5] - expected).max() <= 1e-15
    assert 0 < coef[5] < 1 / 120  # the fifth-order term is damped, not exact
    assert A[0] == 0.0


def test_ck2n54_stage_times_integrate_polynomials(orc):
    """With f(t) = t^k and u0 = 0 one step of size 1 from t = 0 gives 1/(k+1) for
    k <= 3, which fixes the stage times C."""
    A, B, Cc = orc.ck2n54_coefficients()
    for k in range(4):
        u, du = 0.0, 0.0
        for i in range(5):
            du = A[i] * du + Cc[i] ** k
            u = u + B[i] * du
        assert abs(u - 1.0 / (k + 1)) <= 4e-16, k


def test_ck2n54_free_fall_is_exact(orc):
    """Particles farther apart than the support only feel gravity; a fourth-order method
    reproduces v = v0 + g t and x = x0 + v0 t + g t^2 / 2 up to rounding."""
    nd, h = 3, 0.01
    P = _params(orc, nd, WC2, h, c=10.0, gravity=(0.0, -9.81, 0.5))
    x0 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    v0 = _pack(np.array([[0.1, 0.2, -0.3], [0.0, 1.0, 0.0], [-2.0, 0.0, 0.5]]),
               np.full(3, 1000.0))
    m = np.ones(3)
    dt, n = 1e-3, 50
    v, u = orc.solve_ck2n54(P, v0, x0, m, None, None, dt, n)
    T = dt * n
    g = np.array([0.0, -9.81, 0.5])
    assert np.allclose(v[:, :nd], v0[:, :nd] + g * T, rtol=0, atol=1e-13)
    assert np.allclose(u, x0 + v0[:, :nd] * T + 0.5 * g * T * T, rtol=0, atol=1e-13)
    assert np.all(v[:, nd] == 1000.0)


# ---------------------------------------------------------------------------------------
# Neighbour search (radius semantics |x_a - x_b|^2 <= R^2)
# ---------------------------------------------------------------------------------------
def _numpy_pairs(nd, R, xa, xb):
    d = xa[:, None, :] - xb[None, :, :]
    d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
    if nd == 3:
        d2 = d2 + d[..., 2] * d[..., 2]
    return d2 <= R * R


@pytest.mark.parametrize("nd", [2, 3])
def test_cell_list_matches_brute_force(orc, nd):
    rng = np.random.default_rng(11)
    R = 0.05
    xa = rng.uniform(-0.3, 0.3, size=(700, nd))
    # points at exactly R along an axis, duplicates and far outliers
    extra = xa[:20].copy()
    extra[:, 0] += R
    xb = np.concatenate([xa, extra, xa[:5], xa[:3] + 100.0])
    mask = _numpy_pairs(nd, R, xa, xb)
    for brute in (True, False):
        off, idx = orc.neighbors(nd, R, xa, xb, brute=brute)
        assert off[-1] == mask.sum()
        for a in range(len(xa)):
            assert np.array_equal(idx[off[a]:off[a + 1]], np.nonzero(mask[a])[0])
