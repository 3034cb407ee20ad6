"""SortingCallback (src/callbacks/sorting.jl) and tpsph_sort_particles.

CPU tests: the callback's interface and decisions, mirroring test/callbacks/sorting.jl
("show", "initial_sort", "condition", "Illegal Input"), and the chunking of the device
solver between the steps at which it fires (with a recording stand-in for the library).

GPU tests: the sort through the C ABI on a shuffled caller order --
- the sorted state is the old state permuted, row by row, bit for bit, masses included;
- the order is the library's cell order: cell keys of the fluid grid (x fastest) never
  decrease, and sorting again is the identity;
- every particle's rates equal, bit for bit, those of the unsorted state (the sums run in
  the same order), for ContinuityDensity and SummationDensity, device and host memory;
- a run with a SortingCallback, mapped back to the initial order through particle_ids,
  follows the oracle's trajectory of the unsorted state within 1e-10."""
from types import SimpleNamespace

import numpy as np
import pytest

from conftest import gpu_available


# -- CPU: the callback ---------------------------------------------------------------------

def test_show_and_illegal_input(tp):
    with pytest.raises(ValueError, match="either `interval` or `dt` must be set to"):
        tp.SortingCallback()
    assert repr(tp.SortingCallback(interval=11)) == \
        "SortingCallback(interval=11, initial_sort=true)"
    assert repr(tp.SortingCallback(dt=1.2)) == "SortingCallback(dt=1.2, initial_sort=true)"
    assert repr(tp.SortingCallback(interval=3, initial_sort=False)) == \
        "SortingCallback(interval=3, initial_sort=false)"
    with pytest.raises(ValueError, match="setting both `interval` and `dt` is not supported"):
        tp.SortingCallback(dt=0.1, interval=1)


class _MockSemi:
    def __init__(self):
        self.sorted = []

    def sort_particles(self, v, u):
        self.sorted.append((v, u))


def test_initial_sort(tp):
    from trixiparticles_jl_amd.callbacks import Integrator
    semi = _MockSemi()
    integ = Integrator({"semi": semi}, ("v", "u"), 0.0)
    cb = tp.SortingCallback(interval=1, initial_sort=False)
    cb.initialize(integ)
    assert semi.sorted == []
    cb = tp.SortingCallback(interval=1, initial_sort=True)
    cb.initialize(integ)
    assert semi.sorted == [("v", "u")]


def test_condition(tp):
    cb = tp.SortingCallback(interval=3)
    integ = SimpleNamespace(t=0.1, stats=SimpleNamespace(naccept=2), isfinished=False)
    assert cb.condition(None, 0.1, integ) is False
    integ.stats.naccept = 3
    assert cb.condition(None, 0.1, integ) is True
    integ.isfinished = True
    assert cb.condition(None, 0.1, integ) is False

    cb = tp.SortingCallback(dt=0.02)
    integ = SimpleNamespace(isfinished=False)
    assert cb.condition(None, 0.019, integ) is False
    assert cb.condition(None, 0.021, integ) is True
    cb.last_t = 0.021  # don't trigger again until the time has advanced by another 0.02
    assert cb.condition(None, 0.022, integ) is False
    integ.isfinished = True
    assert cb.condition(None, 0.05, integ) is False


class _RecordingLib:
    """Stands in for the library in the driver test: records the solver's step counts."""

    def __init__(self):
        self.calls = []

    def tpsph_solve_ck2n54(self, handle, v, u, t0, dt, nsteps):
        self.calls.append(("solve", round(t0 / dt), nsteps))
        return 0


@pytest.mark.parametrize("cb_kw,nsteps,expect", [
    # sorts after steps 3 and 6 (not after the last step: the run is finished)
    ({"interval": 3}, 9, [("sort",), ("solve", 0, 3), ("sort",), ("solve", 3, 3), ("sort",),
                          ("solve", 6, 3)]),
    ({"interval": 4, "initial_sort": False}, 6, [("solve", 0, 4), ("sort",), ("solve", 4, 2)]),
    # dt: the first step after 2.5 dt has passed since the last sort (steps 3 and 6)
    ({"dt": 2.5e-5, "initial_sort": False}, 7, [("solve", 0, 3), ("sort",), ("solve", 3, 3),
                                                ("sort",), ("solve", 6, 1)]),
])
def test_device_solver_runs_in_stretches_between_sorts(tp, monkeypatch, cb_kw, nsteps, expect):
    import trixiparticles_jl_amd.semidiscretization as sd
    lib = _RecordingLib()
    monkeypatch.setattr(sd.native, "load", lambda *a: lib)
    semi = SimpleNamespace(handle=None,
                           sort_particles=lambda v, u: lib.calls.append(("sort",)))
    ode = sd.DynamicalODEProblem(None, None, np.zeros((2, 3)), np.zeros((2, 2)),
                                 (0.0, nsteps * 1e-5), {"semi": semi})
    cb = tp.SortingCallback(**cb_kw)
    tp.solve(ode, dt=1e-5, nsteps=nsteps, callback=cb)
    assert lib.calls == expect


# -- GPU: the sort -------------------------------------------------------------------------

needs_gpu = pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")


def shuffled_case(tp, nd, dx, seed, backend="hip", **kw):
    """A dam break whose fluid particles are in a seeded random order (initial condition,
    masses and host caches permuted together)."""
    db = tp.dam_break(nd, dx, **kw)
    fs = db.fluid_system
    sh = np.random.default_rng(seed).permutation(fs.n_particles)
    ic = fs.initial_condition
    ic.coordinates = np.ascontiguousarray(ic.coordinates[:, sh])
    ic.velocity = np.ascontiguousarray(ic.velocity[:, sh])
    ic.mass = np.ascontiguousarray(ic.mass[sh])
    ic.density = np.ascontiguousarray(ic.density[sh])
    ic.pressure = np.ascontiguousarray(ic.pressure[sh])
    fs.mass = np.ascontiguousarray(fs.mass[sh])
    fs.density = np.ascontiguousarray(fs.density[sh])
    fs.pressure = np.ascontiguousarray(fs.pressure[sh])
    semi = tp.Semidiscretization(fs, db.boundary_system, parallelization_backend=backend)
    return db, semi


def to_np(a):
    return a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)


def kick_np(tp, semi, v, u):
    dv = semi._zeros(*v.shape)
    tp.kick(dv, v, u, {"semi": semi}, 0.0)
    semi.synchronize()
    return to_np(dv).copy()


def library_cell_keys(semi, x):
    """Linear cell index of each point in the library's fluid grid, restated from
    tpsph_create (cells of span / divisions per axis around the domain box, x fastest)."""
    p = semi.params()
    nd = p.ndims
    R = 2.0 * p.smoothing_length
    div = p.cell_divisions if p.cell_divisions > 0 else 2
    xdiv = p.x_cell_divisions if p.x_cell_divisions > 0 else 2 * div
    skin = p.skin if p.skin >= 0.0 else 0.04
    lo = [p.domain_min[d] for d in range(nd)]
    hi = [p.domain_max[d] for d in range(nd)]
    extent = R
    for d in range(nd):
        extent = max(extent, max(abs(lo[d]), abs(hi[d])) + 4.0 * R)
    margin = 1e-9 * R + 64.0 * np.finfo(float).eps * extent
    hs_fluid = 0.5 * skin * R + margin
    span = (R * 1.00001 + 2.0 * hs_fluid) * (1.0 + 1e-9)
    key = np.zeros(x.shape[0], dtype=np.int64)
    stride = 1
    for d in range(nd):
        cs = span / (xdiv if d == 0 else div)
        inv = 1.0 / cs
        lo_d = lo[d] - cs
        n = int(np.ceil((hi[d] + cs - lo_d) * inv)) + 1
        c = np.clip(np.floor((x[:, d] - lo_d) * inv), 0, n - 1).astype(np.int64)
        key += c * stride
        stride *= n
    return key


@pytest.mark.gpu
@needs_gpu
@pytest.mark.parametrize("backend", ["hip", "hip_host"])
@pytest.mark.parametrize("nd,dx,summation", [(2, 0.01, False), (3, 0.025, False),
                                             (2, 0.02, True)])
def test_sort_permutes_state_and_keeps_rates(tp, nd, dx, summation, backend):
    kw = {"density_calculator": tp.SummationDensity()} if summation else {}
    _, semi_a = shuffled_case(tp, nd, dx, seed=7, backend=backend, **kw)
    _, semi_b = shuffled_case(tp, nd, dx, seed=7, backend=backend, **kw)
    va, ua = semi_a.initial_state()
    dva = kick_np(tp, semi_a, va, ua)
    pa = semi_a.update_pressures()
    fa = pa[0].copy()
    wa = (pa[1].copy(), pa[2].copy()) if pa[1] is not None else None

    vb, ub = semi_b.initial_state()
    v_in, u_in = to_np(vb).copy(), to_np(ub).copy()
    m_in = semi_b.fluid.mass.copy()
    perm = semi_b.sort_particles(vb, ub)
    n = v_in.shape[0]
    assert perm.dtype == np.int64 and np.array_equal(np.sort(perm), np.arange(n))
    assert not np.array_equal(perm, np.arange(n))  # the shuffled order was not sorted
    assert np.array_equal(to_np(vb), v_in[perm])
    assert np.array_equal(to_np(ub), u_in[perm])
    assert np.array_equal(semi_b.fluid.mass, m_in[perm])
    assert np.array_equal(semi_b.particle_ids, perm)

    # the pressures of the sorted state are available at once, the old ones permuted
    pb = semi_b.update_pressures()
    assert np.array_equal(pb[0], fa[perm])
    if wa is not None:
        assert np.array_equal(pb[1], wa[0]) and np.array_equal(pb[2], wa[1])

    # the library's cell order: keys never decrease, and sorting again changes nothing
    keys = library_cell_keys(semi_b, to_np(ub))
    assert np.all(np.diff(keys) >= 0)
    perm2 = semi_b.sort_particles(vb, ub)
    assert np.array_equal(perm2, np.arange(n))

    dvb = kick_np(tp, semi_b, vb, ub)
    assert np.array_equal(dvb, dva[perm]), "rates of the sorted state differ"
    if semi_b.fluid.summation:
        da = semi_a.density()
        assert np.array_equal(semi_b.density(), da[perm])
    semi_a.close()
    semi_b.close()


@pytest.mark.gpu
@needs_gpu
def test_sorted_run_follows_the_oracle(tp, orc):
    """20 CK2N54 steps from a shuffled order with SortingCallback(interval=5) (and the
    initial sort), mapped back through particle_ids, against the oracle's 20 steps of the
    unsorted state."""
    db, semi = shuffled_case(tp, 2, 0.01, seed=3)
    ode = tp.semidiscretize(semi, (0.0, 20e-5))
    v0, u0 = to_np(ode.v0).copy(), to_np(ode.u0).copy()
    m0 = semi.fluid.mass.copy()
    P = orc.params_from_semi(semi)
    vr, ur = orc.solve_ck2n54(P, v0, u0, m0, db.boundary_system.coordinates.T,
                              db.boundary_system.boundary_model.hydrodynamic_mass, 1e-5, 20)
    cb = tp.SortingCallback(interval=5)
    v, u = tp.solve(ode, dt=1e-5, nsteps=20, callback=cb)
    assert cb.sorts == 4  # initial sort + after steps 5, 10, 15
    ids = semi.particle_ids
    back_v = np.empty_like(v0)
    back_u = np.empty_like(u0)
    back_v[ids] = to_np(v)
    back_u[ids] = to_np(u)
    assert np.array_equal(m0[ids], semi.fluid.mass)

    def rel(a, b):
        return np.abs(a - b).max() / np.abs(b).max()

    for col in range(vr.shape[1]):
        assert rel(back_v[:, col], vr[:, col]) <= 1e-10, f"v column {col}"
    for col in range(ur.shape[1]):
        assert rel(back_u[:, col] - u0[:, col], ur[:, col] - u0[:, col]) <= 1e-10, \
            f"u displacement column {col}"
    semi.close()


@pytest.mark.gpu
@needs_gpu
def test_sort_keeps_mixed_precision_rates(tp):
    """The mixed-precision pass sums in the cell order too: its rates move with the rows."""
    _, semi_a = shuffled_case(tp, 3, 0.025, seed=5)
    _, semi_b = shuffled_case(tp, 3, 0.025, seed=5)
    for semi in (semi_a, semi_b):
        semi.precision = "mixed"
    va, ua = semi_a.initial_state()
    dva = kick_np(tp, semi_a, va, ua)
    vb, ub = semi_b.initial_state()
    perm = semi_b.sort_particles(vb, ub)
    dvb = kick_np(tp, semi_b, vb, ub)
    assert np.array_equal(dvb, dva[perm])
    semi_a.close()
    semi_b.close()


@pytest.mark.gpu
@needs_gpu
def test_sorting_with_step_functions_follows_the_oracle(tp, orc):
    """SortingCallback beside a step function: the host-driven steps (kick!, drift! and the
    stage update per stage) with sorts every 3 steps, mapped back through particle_ids,
    against the oracle's unsorted trajectory."""
    db, semi = shuffled_case(tp, 2, 0.02, seed=11)
    ode = tp.semidiscretize(semi, (0.0, 10e-5))
    v0, u0 = to_np(ode.v0).copy(), to_np(ode.u0).copy()
    m0 = semi.fluid.mass.copy()
    P = orc.params_from_semi(semi)
    vr, ur = orc.solve_ck2n54(P, v0, u0, m0, db.boundary_system.coordinates.T,
                              db.boundary_system.boundary_model.hydrodynamic_mass, 1e-5, 10)
    seen = []
    cb = tp.SortingCallback(interval=3)
    v, u = tp.solve(ode, dt=1e-5, nsteps=10,
                    callback=[lambda v, u, t, s: seen.append(s), cb])
    assert seen == list(range(1, 11)) and cb.sorts == 4  # initial + after steps 3, 6, 9
    ids = semi.particle_ids
    back_v = np.empty_like(v0)
    back_u = np.empty_like(u0)
    back_v[ids] = to_np(v)
    back_u[ids] = to_np(u)
    for col in range(vr.shape[1]):
        err = np.abs(back_v[:, col] - vr[:, col]).max() / np.abs(vr[:, col]).max()
        assert err <= 1e-10, (col, err)
    du = back_u - u0
    dur = ur - u0
    for col in range(ur.shape[1]):
        err = np.abs(du[:, col] - dur[:, col]).max() / np.abs(dur[:, col]).max()
        assert err <= 1e-10, (col, err)
    semi.close()
