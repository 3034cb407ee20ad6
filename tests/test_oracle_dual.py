"""Pins of the oracle's dual-number restatement (oracle/sph_dual.cpp), the checker of the
adjoint's gradients.

- values: the dual kick and trajectory equal the plain oracle bit for bit (same expressions,
  same neighbour order);
- tangents: against central finite differences of the plain oracle for every seeded input
  (alpha, delta, g, a direction in the initial velocities and positions);
- closed forms: dv depends on g through gravity alone away from the walls (d dv / d g =
  the unit gravity direction, d drho / d g = 0), and drho is affine in delta
  (d drho / d delta = (drho(delta) - drho(0)) / delta)."""
import numpy as np
import pytest


def small_case(tp, orc, dx=0.05, seed=4):
    db = tp.dam_break(2, dx)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system)
    P = orc.params_from_semi(semi)
    ic = db.fluid_system.initial_condition
    v = np.concatenate([ic.velocity, ic.density.reshape(1, -1)], axis=0).T.copy()
    u = ic.coordinates.T.copy()
    rng = np.random.default_rng(seed)
    v[:, :2] = rng.normal(0.0, 0.3, size=(v.shape[0], 2))
    v[:, 2] *= 1.0 + rng.uniform(-2e-3, 2e-3, size=v.shape[0])
    xw = db.boundary_system.coordinates.T.copy()
    mw = db.boundary_system.boundary_model.hydrodynamic_mass
    return db, P, v, u, db.fluid_system.mass, xw, mw


def test_dual_values_equal_the_oracle(tp, orc):
    db, P, v, u, m, xw, mw = small_case(tp, orc)
    dv_ref, *_ = orc.kick(P, v, u, m, xw, mw)
    dv, _ = orc.dual_kick(P, v, u, m, xw, mw)
    assert np.array_equal(dv, dv_ref)
    vr, ur = orc.solve_ck2n54(P, v, u, m, xw, mw, 1e-4, 3)
    vd, ud, *_ = orc.dual_solve_ck2n54(P, v, u, m, xw, mw, 1e-4, 3)
    assert np.array_equal(vd, vr) and np.array_equal(ud, ur)


def _seeds(nf, rng):
    sp = np.zeros((3, 4))
    sp[0, 0] = sp[1, 1] = sp[2, 2] = 1.0  # tangents 0..2: alpha, delta, g
    sv = np.zeros((4, nf, 3))
    su = np.zeros((4, nf, 2))
    sv[3, :, :2] = rng.normal(size=(nf, 2))  # tangent 3: a direction in v0 and u0
    su[3] = 1e-3 * rng.normal(size=(nf, 2))
    return sp, sv, su


def test_dual_trajectory_tangents_match_central_differences(tp, orc):
    """Tangents of the loss after 10 steps against central differences of the plain oracle.
    The target is offset coherently so that every input moves the loss well above the
    rounding noise of the positions; the steps are 1e-4 of each parameter and 1e-5 along the
    state direction, where truncation and rounding errors both stay below 1e-7."""
    db, P, v, u, m, xw, mw = small_case(tp, orc)
    nf = v.shape[0]
    rng = np.random.default_rng(9)
    sp, sv, su = _seeds(nf, rng)
    dt, steps = 2e-4, 10
    target = u + np.array([1e-3, -2e-3]) + 1e-4 * rng.normal(size=u.shape)
    *_, loss = orc.dual_solve_ck2n54(P, v, u, m, xw, mw, dt, steps, sp, sv, su, target)

    def L(alpha=P.alpha, delta=P.delta, g=None, t=0.0):
        Q = type(P).from_buffer_copy(P)
        Q.alpha, Q.delta = alpha, delta
        if g is not None:
            Q.gravity[1] = -g
        _, uu = orc.solve_ck2n54(Q, v + t * sv[3], u + t * su[3], m, xw, mw, dt, steps)
        return np.sum((uu - target) ** 2) / nf

    g0 = -P.gravity[1]
    cases = [
        (lambda e: L(alpha=P.alpha + e), 1e-4 * P.alpha, 1),
        (lambda e: L(delta=P.delta + e), 1e-4 * P.delta, 2),
        (lambda e: L(g=g0 + e), 1e-4 * g0, 3),
        (lambda e: L(t=e), 1e-5, 4),
    ]
    for f, eps, k in cases:
        fd = (f(eps) - f(-eps)) / (2 * eps)
        assert abs(loss[k] - fd) <= 1e-6 * abs(fd), (k, loss[k], fd)


def test_dual_kick_closed_forms(tp, orc):
    db, P, v, u, m, xw, mw = small_case(tp, orc)
    nf = v.shape[0]
    sp = np.zeros((3, 4))
    sp[1, 0] = 1.0  # tangent 0: delta
    sp[2, 1] = 1.0  # tangent 1: g
    dv, tan = orc.dual_kick(P, v, u, m, xw, mw, seed_params=sp)
    # d drho / d delta: drho is affine in delta
    Q = type(P).from_buffer_copy(P)
    Q.delta = 0.0
    dv0, *_ = orc.kick(Q, v, u, m, xw, mw)
    ref = (dv[:, 2] - dv0[:, 2]) / P.delta
    assert np.abs(tan[0, :, 2] - ref).max() <= 1e-12 * np.abs(ref).max()
    # d dv / d g away from the walls: the gravity direction (0, -1); drho does not move
    R = 2 * db.fluid_system.smoothing_length
    d2 = ((u[:, None, :] - xw[None, :, :]) ** 2).sum(axis=2)
    interior = (d2 > (R * 1.01) ** 2).all(axis=1)
    assert interior.sum() > 10
    assert np.array_equal(tan[1, interior, 0], np.zeros(interior.sum()))
    assert np.array_equal(tan[1, interior, 1], -np.ones(interior.sum()))
    assert np.array_equal(tan[1, interior, 2], np.zeros(interior.sum()))
    # near the walls g also moves the Adami pressures
    assert np.abs(tan[1, ~interior, 1] + 1.0).max() > 0.0
