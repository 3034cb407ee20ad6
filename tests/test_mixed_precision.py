"""Mixed precision (TPSPH_PRECISION_MIXED): FP32 pair arithmetic on brick-relative
coordinates with Kahan-compensated density rates, FP64 state, pressures, wall sums, stage
update and integrals.

The bar, from the workload definition: density within 1e-5 (relative) of the FP64 result.
The FP64 path is the one the parity tests check against the oracle at 1e-12; here the mixed
path is compared with the oracle directly on small cases (one kick, a short trajectory) and
with the FP64 path at the 3D workload sizes (1.02M and 8.2M fluid particles), where the
oracle is too slow to run the trajectory.  Every comparison prints its errors:

  density   max |rho_mixed - rho_fp64| / rho0
  pressure  max |p_mixed - p_fp64| / max |p_fp64|
  position  max |x_mixed - x_fp64| / dx
  mass      drift of the fluid volume sum m/rho relative to the FP64 run (continuity
            density conserves sum m exactly; the volume integral shows the density drift)
"""
import json
import os

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

DENSITY_RTOL = 1e-5


def rel_err(got, ref):
    scale = np.abs(ref).max()
    return float(np.abs(got - ref).max() / (scale if scale > 0 else 1.0))


def report(name, values):
    """Print the errors and, with TPSPH_REPORT_DIR set, keep them as JSON."""
    print(f"{name}: " + ", ".join(f"{k} = {v:.3e}" if isinstance(v, float) else f"{k} = {v}"
                                  for k, v in values.items()), flush=True)
    out = os.environ.get("TPSPH_REPORT_DIR")
    if out:
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, f"{name}.json"), "w") as f:
            json.dump(values, f, indent=1)


def perturbed_state(semi, nd, seed, vscale=0.3):
    v0, u0 = semi.initial_state()
    v, u = v0.cpu().numpy().copy(), u0.cpu().numpy().copy()
    rng = np.random.default_rng(seed)
    v[:, :nd] = rng.normal(0.0, vscale, size=(v.shape[0], nd))
    v[:, nd] *= 1.0 + rng.uniform(-2e-3, 2e-3, size=v.shape[0])
    return v, u


@pytest.mark.parametrize("nd,dx", [(2, 0.01), (3, 0.02)])
def test_mixed_kick_matches_oracle(tp, orc, nd, dx):
    """One kick with random velocities and perturbed densities against the FP64 oracle."""
    import torch
    db = tp.dam_break(nd, dx)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system, precision="mixed")
    v, u = perturbed_state(semi, nd, 23)
    P = orc.params_from_semi(semi)
    dv_ref, *_ = orc.kick(P, v, u, db.fluid_system.mass, db.boundary_system.coordinates.T,
                          db.boundary_system.boundary_model.hydrodynamic_mass)
    vd, ud = torch.from_numpy(v).cuda(), torch.from_numpy(u).cuda()
    dv = torch.full_like(vd, float("nan"))
    tp.kick(dv, vd, ud, {"semi": semi}, 0.0)
    semi.synchronize()
    got = dv.cpu().numpy()
    errs = {f"dv{d}": rel_err(got[:, d], dv_ref[:, d]) for d in range(nd)}
    errs["drho"] = rel_err(got[:, nd], dv_ref[:, nd])
    report(f"mixed_kick_{nd}d", errs)
    # FP32 pair terms: relative errors of a few FP32 roundings times the cancellation in the
    # sums; dv is summed plainly, drho with compensation
    assert errs["drho"] <= 1e-5
    assert max(errs[f"dv{d}"] for d in range(nd)) <= 1e-4


def run_pair(tp, nd, dx, steps, dt=1e-5):
    """Same seeded state, `steps` CK2N54 steps in FP64 and in mixed precision."""
    db = tp.dam_break(nd, dx)
    out = {}
    for precision in ("fp64", "mixed"):
        semi = tp.Semidiscretization(db.fluid_system, db.boundary_system, precision=precision)
        ode = tp.semidiscretize(semi, (0.0, steps * dt))
        v, u = tp.solve(ode, dt=dt, nsteps=steps)
        semi.synchronize()
        ints0 = semi.integrals(ode.v0, ode.u0)
        ints = semi.integrals(v, u)
        out[precision] = (v.cpu().numpy(), u.cpu().numpy(), ints0, ints)
        semi.close()
        del ode
    return db, out


def compare_runs(db, out, nd):
    v64, u64, i0, i64 = out["fp64"]
    vm, um, _, im = out["mixed"]
    se = db.state_equation
    rho0 = se.reference_density
    rho64, rhom = v64[:, nd], vm[:, nd]

    def pressure(rho):
        return se.sound_speed ** 2 * rho0 / se.exponent * ((rho / rho0) ** se.exponent - 1.0)

    return {
        "n_fluid": int(v64.shape[0]),
        "density": float(np.abs(rhom - rho64).max() / rho0),
        "density_change_rel": float(np.abs(rhom - rho64).max() /
                                    max(np.abs(rho64 - rho0).max(), 1e-300)),
        "pressure": rel_err(pressure(rhom), pressure(rho64)),
        "velocity": rel_err(vm[:, :nd], v64[:, :nd]),
        "position_over_dx": float(np.abs(um - u64).max() / db.particle_spacing),
        "volume_drift_fp64": float((i64["volume"] - i0["volume"]) / i0["volume"]),
        "volume_drift_mixed": float((im["volume"] - i0["volume"]) / i0["volume"]),
        "mass_drift_mixed_vs_fp64": float((im["volume"] - i64["volume"]) / i64["volume"]),
        "kinetic_energy_rel": float(abs(im["kinetic_energy"] - i64["kinetic_energy"]) /
                                    max(i64["kinetic_energy"], 1e-300)),
    }


def test_mixed_trajectory_matches_oracle_2d(tp, orc):
    """200 CK2N54 steps of the 2D dam break (dt = 1e-5) in mixed precision vs the oracle."""
    db = tp.dam_break(2, 0.01)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system, precision="mixed")
    ode = tp.semidiscretize(semi, (0.0, 200e-5))
    v, u = tp.solve(ode, dt=1e-5, nsteps=200)
    P = orc.params_from_semi(semi)
    v0, u0 = ode.v0.cpu().numpy(), ode.u0.cpu().numpy()
    vr, ur = orc.solve_ck2n54(P, v0, u0, db.fluid_system.mass, db.boundary_system.coordinates.T,
                              db.boundary_system.boundary_model.hydrodynamic_mass, 1e-5, 200)
    vm, um = v.cpu().numpy(), u.cpu().numpy()
    rho0 = db.state_equation.reference_density
    errs = {"density": float(np.abs(vm[:, 2] - vr[:, 2]).max() / rho0),
            "velocity": rel_err(vm[:, :2], vr[:, :2]),
            "position_over_dx": float(np.abs(um - ur).max() / 0.01)}
    report("mixed_trajectory_2d_vs_oracle", errs)
    assert errs["density"] <= DENSITY_RTOL


@pytest.mark.parametrize("name,dx,steps", [("dam_break_3d_1M", 6.25e-3, 500),
                                           ("dam_break_3d_8M", 3.125e-3, 1000)])
def test_mixed_trajectory_matches_fp64_at_scale(tp, name, dx, steps):
    """The 3D workloads at full size: 1.02M fluid particles for 500 steps and 8.2M for 1000
    steps (dt = 1e-5), mixed precision against the FP64 path from the same state."""
    db, out = run_pair(tp, 3, dx, steps)
    errs = compare_runs(db, out, 3)
    errs["steps"] = steps
    report(f"mixed_vs_fp64_{name}", errs)
    assert errs["density"] <= DENSITY_RTOL
