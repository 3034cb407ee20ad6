"""Host side of the reverse mode (no GPU): the parameter vector [alpha, delta, g] of a
Semidiscretization and the copies with other values that the gradient checks and the
differentiation workload's targets use."""
import math

import numpy as np


def test_parameters_and_with_parameters(tp):
    from trixiparticles_jl_amd import adjoint as ad
    db = tp.dam_break(2, 0.05)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system, skin=0.1)
    assert np.array_equal(ad.parameters(semi), [0.02, 0.1, 9.81])
    s2 = ad.with_parameters(semi, alpha=0.03, delta=0.12, g=9.6)
    assert np.array_equal(ad.parameters(s2), [0.03, 0.12, 9.6])
    # same direction of gravity, same sound speed, grid box and search settings
    assert s2.fluid.acceleration == (0.0, -9.6)
    assert s2.fluid.state_equation is semi.fluid.state_equation
    assert s2.skin == semi.skin and s2.pair_search == semi.pair_search
    assert all(np.array_equal(a, b) for a, b in zip(s2.domain, semi.domain))
    # the original is untouched
    assert np.array_equal(ad.parameters(semi), [0.02, 0.1, 9.81])
    p1, p2 = semi.params(), s2.params()
    assert (p2.alpha, p2.delta) == (0.03, 0.12) and (p1.alpha, p1.delta) == (0.02, 0.1)
    assert math.isclose(p2.acceleration[1], -9.6, rel_tol=0, abs_tol=0)
    # only the named values change
    s3 = ad.with_parameters(semi, delta=0.05)
    assert np.array_equal(ad.parameters(s3), [0.02, 0.05, 9.81])


def test_differentiation_workload_sizes(tp):
    """The differentiation workload: 100 x 50 x 50 = 250,000 fluid particles at dx = 0.01 in
    3D (359,268 wall particles, open top), 120 x 60 = 7,200 in its 2D copy."""
    db3 = tp.dam_break(3, 0.01)
    assert (db3.n_fluid, db3.n_wall) == (250000, 359268)
    db2 = tp.dam_break(2, 0.01)
    assert db2.n_fluid == 7200
