"""The ensemble runner (include/tpsph_batch.h): many uneven dam breaks in one set of kernels
per stage, each CK2N54 step captured in a HIP graph.

The bar, from the workload definition: every case of the ensemble equals, bit for bit, the
same case run alone.  "Alone" is checked two ways: the case as an ensemble of one, and the
case on its own tpsph_handle (cell search, skin 0, the same box), whose kick the parity
tests check against the oracle; one kick of the ensemble is also checked against the oracle
directly, per case, at 1e-12."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

DT = 1e-5


@pytest.fixture(scope="module")
def en(tp):
    from trixiparticles_jl_amd import ensemble
    return ensemble


def small_cases(en, n=10, seed=3):
    return en.dam_break_ensemble(n, seed=seed, width=(0.3, 0.5), aspect=(0.5, 1.0))


def alone_on_handle(tp, case, steps):
    """The case on its own handle: cell search, skin 0, the ensemble's box for the case."""
    ode = tp.semidiscretize(case.semi, (0.0, steps * DT))
    v, u = tp.solve(ode, dt=DT, nsteps=steps)
    case.semi.synchronize()
    return v.cpu().numpy(), u.cpu().numpy()


def test_ensemble_kick_matches_oracle_per_case(tp, orc, en):
    import torch
    cases = small_cases(en, 4)
    ens = en.Ensemble(cases, tuning=None)
    dv = torch.full_like(ens.v0, float("nan"))
    ens.kick(dv, ens.v0, ens.u0)
    torch.cuda.synchronize()
    got = dv.cpu().numpy()
    for k, c in enumerate(cases):
        P = orc.params_from_semi(c.semi)
        v0, u0 = c.semi.initial_state()
        dv_ref, *_ = orc.kick(P, v0.cpu().numpy(), u0.cpu().numpy(), c.semi.fluid.mass,
                              c.semi.wall.coordinates.T,
                              c.semi.wall.boundary_model.hydrodynamic_mass)
        g = got[ens.slices[k]]
        for col in range(3):
            scale = np.abs(dv_ref[:, col]).max()
            err = np.abs(g[:, col] - dv_ref[:, col]).max() / scale
            assert err <= 1e-12, f"case {k} column {col}: {err:.3e}"
    ens.close()


def test_ensemble_equals_each_case_alone(tp, en):
    cases = small_cases(en, 10)
    steps = 10
    ens = en.Ensemble(cases, tuning=None)
    v, u = ens.solve(DT, steps)
    v, u = v.cpu().numpy(), u.cpu().numpy()
    for k in (0, 3, 9):
        one = en.Ensemble([cases[k]], tuning=None)
        v1, u1 = one.solve(DT, steps)
        assert np.array_equal(v[ens.slices[k]], v1.cpu().numpy()), f"case {k}: v differs"
        assert np.array_equal(u[ens.slices[k]], u1.cpu().numpy()), f"case {k}: u differs"
        one.close()
        vh, uh = alone_on_handle(tp, cases[k], steps)
        assert np.array_equal(v[ens.slices[k]], vh), f"case {k}: v differs from its handle"
        assert np.array_equal(u[ens.slices[k]], uh), f"case {k}: u differs from its handle"
    ens.close()


def test_launch_shapes_do_not_change_results(en):
    cases = small_cases(en, 6)
    ens = en.Ensemble(cases, tuning=None)
    ref = [a.cpu().numpy() for a in ens.solve(DT, 3)]
    for fb, wb in ((64, 128), (128, 64), (256, 256)):
        ens.set_launch(fluid_block=fb, wall_block=wb)
        got = [a.cpu().numpy() for a in ens.solve(DT, 3)]
        assert all(np.array_equal(a, b) for a, b in zip(got, ref)), (fb, wb)
    ens.close()


def test_graph_replays_equal_eager_steps(en):
    """Each replay of the captured step, with eager work and synchronisations between the
    replays, equals the same step run eagerly."""
    import torch
    cases = small_cases(en, 5)
    ens = en.Ensemble(cases, tuning=None)
    ve, ue = ens.v0.clone(), ens.u0.clone()
    vg, ug = ens.v0.clone(), ens.u0.clone()
    for step in range(4):
        ens.solve(DT, 1, graph=False, v=ve, u=ue)
        ens.solve(DT, 1, graph=True, v=vg, u=ug)
        scratch = torch.randn(1 << 16, device="cuda", dtype=torch.float64).sum()  # eager work
        torch.cuda.synchronize()
        del scratch
        assert torch.equal(ve, vg) and torch.equal(ue, ug), f"replay {step} differs"
    ens.close()


def test_full_ensemble_subset_equals_cases_alone(tp, en):
    """The seeded 256-case workload (about 2M fluid particles): 5 steps of the whole
    ensemble, four of its cases checked against themselves alone."""
    cases = en.dam_break_ensemble(256)
    ens = en.Ensemble(cases)
    steps = 5
    v, u = ens.solve(DT, steps)
    v, u = v.cpu().numpy(), u.cpu().numpy()
    assert np.isfinite(v).all()
    sizes = np.array([c.n_fluid for c in cases])
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 200):
        one = en.Ensemble([cases[k]])
        v1, u1 = one.solve(DT, steps)
        assert np.array_equal(v[ens.slices[k]], v1.cpu().numpy()), f"case {k}: v differs"
        assert np.array_equal(u[ens.slices[k]], u1.cpu().numpy()), f"case {k}: u differs"
        one.close()
    ens.close()
