"""SPH interpolation at points and along lines (host post-processing) against the known
answers of the reference's own tests, ``test/general/interpolation.jl:43-258`` (2D, no
boundary): a 10 x 10 patch of spacing 0.2 (``test/rectangular_patch.jl``, no perturbation,
particle 50 moved to the patch centre) with mass 1, density 666, pressure 2y and velocity
(5 + 3x, 0.1 y^2 + 0.1), SchoenbergCubicSpline with h = 0.12.  The reference compares with
``atol = 5e-4`` and exact neighbour counts and point coordinates; the same is used here."""
import math

import numpy as np
import pytest

ATOL = 5e-4  # compare_interpolation_result (interpolation.jl:2)


@pytest.fixture(scope="module")
def patch(tp):
    spacing, nx, ny = 0.2, 10, 10
    ic = tp.RectangularShape(spacing, (nx, ny), (-spacing / 2 * nx, -spacing / 2 * ny),
                             density=1000.0)
    center = math.ceil(nx * ny / 2) - 1  # rectangular_patch.jl:13-14 (1-based ceil(n/2))
    ic.coordinates[:, center] = 0.0
    ic.coordinates[1, :] += ny * 0.5 * spacing
    x, y = ic.coordinates
    ic.mass[:] = 1.0
    ic.density[:] = 666.0
    ic.pressure[:] = 2 * y
    ic.velocity[0] = 5 + 3 * x
    ic.velocity[1] = 0.1 * y ** 2 + 0.1
    eos = tp.StateEquationCole(sound_speed=10 * math.sqrt(9.81 * 0.9), reference_density=1000.0,
                               exponent=7, clip_negative_pressure=False)
    fluid = tp.WeaklyCompressibleSPHSystem(
        ic, tp.ContinuityDensity(), eos, tp.SchoenbergCubicSplineKernel(2), 1.2 * 0.5 * spacing,
        viscosity=tp.ArtificialViscosityMonaghan(alpha=0.02, beta=0.0), acceleration=(0.0, -9.81))
    fluid.pressure = ic.pressure.copy()  # the test overwrites system.pressure (:115-116)
    semi = tp.Semidiscretization(fluid, parallelization_backend="hip_host")
    v = np.vstack([ic.velocity, ic.density[None, :]]).T.copy()
    u = ic.coordinates.T.copy()
    return semi, fluid, v, u, spacing, ny


def check(actual, expected):
    n = len(expected["density"])
    assert actual["density"].shape == (n,)
    assert list(actual["neighbor_count"]) == list(expected["neighbor_count"])
    np.testing.assert_array_equal(actual["point_coords"], np.asarray(expected["point_coords"]))
    for key in ("density", "pressure"):
        got, exp = actual[key], np.asarray(expected[key], dtype=float)
        both_nan = np.isnan(got) & np.isnan(exp)
        assert np.all(both_nan | (np.abs(got - exp) <= ATOL)), (key, got, exp)
    got, exp = actual["velocity"], np.asarray(expected["velocity"], dtype=float)
    both_nan = np.isnan(got).all(axis=0) & np.isnan(exp).all(axis=0)
    assert np.all(both_nan | (np.abs(got - exp) <= ATOL).all(axis=0)), (got, exp)


@pytest.mark.parametrize("cut_off_bnd", [True, False])
def test_points_match_reference_known_answers(tp, patch, cut_off_bnd):
    semi, fluid, v, u, spacing, ny = patch
    at = lambda y: tp.interpolate_points(np.array([[0.0], [y]]), semi, fluid, v, u,  # noqa: E731
                                         cut_off_bnd=cut_off_bnd)
    check(at(0.0), dict(density=[665.9999999999999], neighbor_count=[2],
                        point_coords=[[0.0], [0.0]], velocity=[[5.0], [0.101]], pressure=[0.2]))
    pts = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 1.0]])
    check(tp.interpolate_points(pts, semi, fluid, v, u, cut_off_bnd=cut_off_bnd),
          dict(density=[666.0, 666.0000000000001, 666.0], neighbor_count=[2, 6, 5],
               point_coords=pts, velocity=[[5.0, 5.0, 5.0], [0.101, 0.125, 0.20035665520692278]],
               pressure=[0.19999999999999996, 1.0000000000000002, 2.0]))
    # above the last particle row the values of that row come back (:64-75)
    wall_height = ny * spacing - 0.5 * spacing
    y = ny * spacing
    hy = min(y, wall_height)
    check(at(y), dict(density=[666.0], neighbor_count=[2], point_coords=[[0.0], [y]],
                      velocity=[[5.0], [0.1 * hy ** 2 + 0.1]], pressure=[2 * hy]))


@pytest.mark.parametrize("cut_off_bnd", [True, False])
def test_fluid_edge_found_by_bisection_matches(tp, patch, cut_off_bnd):
    """binary_search_outside (:19-40): where the neighbour count drops to zero above and
    below the patch, to 1e-5, with the reference's bisection steps."""
    semi, fluid, v, u, spacing, ny = patch

    def count(y):
        return tp.interpolate_points(np.array([[0.0], [y]]), semi, fluid, v, u,
                                     cut_off_bnd=cut_off_bnd)["neighbor_count"][0]

    def bisect(inside, outside):
        original = inside
        while abs(outside - inside) > 1e-5:
            mid = (inside + outside) / 2
            if count(mid) == 0:
                outside = mid
            else:
                inside = mid
        return abs(outside - original)

    assert bisect(ny * spacing, (ny + 2) * spacing) == pytest.approx(0.11817626953124982, abs=1e-14)
    assert bisect(0.0, -2 * spacing) == pytest.approx(0.11817626953125, abs=1e-14)


@pytest.mark.parametrize("cut_off_bnd", [True, False])
def test_line_matches_reference_known_answers(tp, patch, cut_off_bnd):
    semi, fluid, v, u, _, _ = patch
    res = tp.interpolate_line([1.0, -0.05], [1.0, 1.0], 5, semi, fluid, v, u, endpoint=False,
                              cut_off_bnd=cut_off_bnd)
    check(res, dict(density=[666.0, 666.0, 666.0], neighbor_count=[2, 2, 1],
                    point_coords=[[1.0, 1.0, 1.0], [0.2125, 0.47500000000000003, 0.7375]],
                    velocity=[[7.699999999999999, 7.7, 7.7],
                              [0.10605429538320173, 0.12465095587703466, 0.14900000000000002]],
                    pressure=[0.4527147691600855, 0.9912738969258665, 1.4000000000000001]))
    res = tp.interpolate_line([1.0, -0.05], [1.0, 1.0], 5, semi, fluid, v, u, endpoint=True,
                              cut_off_bnd=cut_off_bnd)
    check(res, dict(density=[666.0] * 5, neighbor_count=[1, 2, 2, 1, 1],
                    point_coords=[[1.0] * 5, [-0.05, 0.2125, 0.475, 0.7375, 1.0]],
                    velocity=[[7.7, 7.699999999999999, 7.699999999999999, 7.7, 7.7],
                              [0.10100000000000002, 0.10605429538320173, 0.12465095587703466,
                               0.14900000000000002, 0.22100000000000006]],
                    pressure=[0.19999999999999998, 0.4527147691600855, 0.9912738969258663,
                              1.4000000000000001, 2.2]))


def test_compute_pressure_is_the_state_equation(tp, patch):
    semi, fluid, v, u, _, _ = patch
    saved = fluid.pressure.copy()
    try:
        p = tp.compute_pressure(semi, v)
        eos = fluid.state_equation
        np.testing.assert_allclose(p, [eos(r) for r in v[:, 2]], rtol=1e-15)
    finally:
        fluid.pressure = saved
