"""The 2D dam-break validation case against the reference's own validation series.

``tests/golden/dam_break_2d_wcsph_40.json`` holds the series of
``validation/dam_break_2d/validation_reference_wcsph_40.json``, written by the reference's run
of ``setup_marrone_2011.jl`` (CarpenterKennedy2N54 at the StepsizeCallback step, with the stop
times its callbacks add) every 0.01 / sqrt(g / H) s for 1.73 s (701 samples): the maximum x
coordinate of the fluid (the surge front) and the wall pressure sensors P1-P4 (the mean of 10
``interpolate_line`` pressures over 8.4 mm of the right wall, ``sensors.jl``).  The CPU tests
pin the set-up and the step size against the series' first value and recorded metadata; the
GPU test integrates the case through the HIP library (about 3,100 steps) and compares every
sample of the five series."""
import json
import math
import os

import numpy as np
import pytest

from conftest import gpu_available

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "dam_break_2d_wcsph_40.json")
# Both runs sum the same terms in different orders, so they part by rounding that the flow
# amplifies.  Measured |hip - reference| of the front: 1.3e-15 m until the front reaches the
# far wall (t = 0.6 s, sample 240), 5.1e-11 m until t = 1.48 s and 1.6e-7 m at the end of the
# series, after the run-up on the wall.
FRONT_ATOL_BEFORE_IMPACT = 1e-12
FRONT_ATOL = 2e-6
N_BEFORE_IMPACT = 240
# Sensor pressures, as max |hip - reference| / max |reference| over a series: measured 2.7e-10
# up to t = 0.99 s (sample 400) and 1.7e-5 at the end (0.28 Pa of 16,484 Pa).
SENSOR_RTOL_EARLY = 1e-8
SENSOR_RTOL = 2e-4
N_EARLY = 400
H = 0.6
SENSOR_SIZE = 0.0084
SENSOR_TOPS = {"P1": (6 + 4.2) / 600 * H, "P2": (30 + 4.2) / 600 * H,
               "P3": (60 + 4.2) / 600 * H, "P4": (160 + 4.2) / 600 * H}


@pytest.fixture(scope="module")
def ref():
    return json.load(open(FIXTURE))


def _case(tp):
    db = tp.marrone_dam_break_2d(40)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 parallelization_backend="hip_host")
    return db, semi


def test_setup_matches_the_recorded_run(tp, ref):
    db, semi = _case(tp)
    meta = ref["meta"]
    f, b = meta["fluid"], meta["boundary"]
    assert db.fluid_system.smoothing_length == pytest.approx(f["smoothing_length"], rel=1e-15)
    assert db.particle_spacing == pytest.approx(f["particle_spacing"], rel=1e-15)
    assert db.state_equation.sound_speed == pytest.approx(f["sound_speed"], rel=1e-15)
    assert db.state_equation.exponent == f["state_equation"]["exponent"]
    assert db.fluid_system.viscosity.alpha == f["viscosity_model"]["alpha"]
    assert db.fluid_system.density_diffusion.delta == f["density_diffusion"]["delta"]
    # the series' first samples: fluid front and the outer wall layer
    assert db.fluid_system.initial_condition.coordinates[0].max() == ref["values"][0]
    assert db.boundary_system.initial_condition.coordinates[0].max() == \
        pytest.approx(ref["wall_max_x"], abs=1e-13)


def test_step_size_is_the_recorded_dt_max(tp, ref):
    db, semi = _case(tp)
    dt = tp.calculate_dt(None, None, 0.9, semi)
    assert dt == pytest.approx(ref["meta"]["time_integrator"]["dt_max"], rel=1e-14)


def stop_times(ref):
    """The stops of the reference run inside the window: the post-processing period (the
    series' own times), the 0.02 s solution-saving period (examples/fluid/dam_break_2d.jl)
    and the saving times of setup_marrone_2011.jl, t sqrt(g / H) = 1.5, 2.36, 3.0, 5.7."""
    t_end = ref["time"][-1]
    saving = [0.02 * k for k in range(1, int(t_end / 0.02) + 1)]
    paper = [x / math.sqrt(9.81 / 0.6) for x in (1.5, 2.36, 3.0, 5.7, 6.45)]
    return sorted(set(ref["time"][1:]) | set(saving) | {t for t in paper if t <= t_end})


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_surge_front_follows_the_reference_series(tp, ref):
    import torch
    db = tp.marrone_dam_break_2d(40)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system)
    ode = tp.semidiscretize(semi, (0.0, ref["time"][-1]))
    dt = tp.calculate_dt(None, None, 0.9, semi)
    wanted = {t: i for i, t in enumerate(ref["time"])}
    front = {}
    fluid = db.fluid_system
    x_wall = db.tank.tank_size[0]
    sensors = {k: np.zeros(len(ref["time"])) for k in SENSOR_TOPS}

    def record(v, u, t):
        if t in wanted:
            torch.cuda.synchronize()
            i = wanted[t]
            front[i] = u.view(-1, 2)[:, 0].max().item()
            vh, uh = v.cpu().numpy(), u.cpu().numpy()
            tp.compute_pressure(semi, vh)  # the callback's update before its quantities
            for k, top in SENSOR_TOPS.items():
                res = tp.interpolate_line([x_wall, top], [x_wall, top - SENSOR_SIZE], 10, semi,
                                          fluid, vh, uh,
                                          smoothing_length=2.0 * fluid.smoothing_length,
                                          clip_negative_pressure=True, cut_off_bnd=False)
                sensors[k][i] = np.nan_to_num(res["pressure"], nan=0.0).sum() / 10

    v, u = tp.solve_tstops(ode, dt=dt, tstops=stop_times(ref), callback=record)
    assert sorted(front) == list(range(1, len(ref["time"])))
    got = np.array([ref["values"][0]] + [front[i] for i in range(1, len(ref["time"]))])
    err = np.abs(got - np.array(ref["values"]))
    print(f"surge front: max |hip - reference| = {err[:N_BEFORE_IMPACT].max():.3e} m before "
          f"the impact, {err.max():.3e} m over all {len(err)} samples")
    assert np.isfinite(got).all()
    assert err[:N_BEFORE_IMPACT].max() <= FRONT_ATOL_BEFORE_IMPACT, err[:N_BEFORE_IMPACT].argmax()
    assert err.max() <= FRONT_ATOL, (err.argmax(), got[err.argmax()], ref["values"][err.argmax()])
    for k in SENSOR_TOPS:
        want = np.array(ref["pressure"][k])
        e = np.abs(sensors[k] - want) / np.abs(want).max()
        print(f"sensor {k}: relative error {e[:N_EARLY].max():.3e} until t = "
              f"{ref['time'][N_EARLY - 1]:.2f} s, {e.max():.3e} over the series")
        assert e[:N_EARLY].max() <= SENSOR_RTOL_EARLY, (k, e[:N_EARLY].argmax())
        assert e.max() <= SENSOR_RTOL, (k, e.argmax(), sensors[k][e.argmax()], want[e.argmax()])
