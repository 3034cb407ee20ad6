"""Host-side setups against the reference's own tank fixtures and the workload sizes.

The expected coordinates come from test/setups/rectangular_tank.jl (extracted into
tests/golden/rectangular_tank.json by the generator beside it); the comparison uses Julia's
default ``isapprox`` for arrays: ||a - b|| <= sqrt(eps) * max(||a||, ||b||)."""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "rectangular_tank.json")
CASES = json.load(open(GOLDEN))["cases"]


def julia_isapprox(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    return np.linalg.norm(a - b) <= math.sqrt(np.finfo(float).eps) * max(
        np.linalg.norm(a), np.linalg.norm(b))


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_rectangular_tank_matches_reference_fixture(tp, case):
    kw = dict(n_layers=case["n_layers"], spacing_ratio=case["spacing_ratio"])
    if case["min_coordinates"] is not None:
        kw["min_coordinates"] = tuple(case["min_coordinates"])
    if case["faces"] is not None:
        kw["faces"] = tuple(case["faces"])
    tank = tp.RectangularTank(case["particle_spacing"], tuple(case["fluid_size"]),
                              tuple(case["tank_size"]), case["fluid_density"], **kw)
    assert julia_isapprox(tank.boundary.coordinates, case["boundary"])
    if case["fluid"] is not None:
        assert julia_isapprox(tank.fluid.coordinates, case["fluid"])


def test_fixture_covers_every_face_combination():
    faces2 = {tuple(c["faces"]) for c in CASES if c["faces"] and len(c["faces"]) == 4}
    assert len(faces2) == 14


@pytest.mark.parametrize("nd,dx,nf,nw_open,nw_closed", [
    (2, 0.01, 7200, 1578, 2556),
    (3, 6.25e-3, 1_024_000, 908_844, 1_166_808),
])
def test_workload_particle_counts(tp, nd, dx, nf, nw_open, nw_closed):
    """Dam-break workload sizes: fluid 120 x 60 / 160 x 80 x 80 on the lattice
    x = min + dx (i - 1/2), walls from RectangularTank with 3 layers."""
    fluid = (1.2, 0.6) if nd == 2 else (1.0, 0.5, 0.5)
    tank = (3.2, 1.0) if nd == 2 else (3.2, 1.0, 1.0)
    for open_top, nw in ((True, nw_open), (False, nw_closed)):
        faces = (True, True, True, not open_top) + ((True, True) if nd == 3 else ())
        t = tp.RectangularTank(dx, fluid, tank, 1000.0, n_layers=3, faces=faces)
        assert t.fluid.n_particles == nf
        assert t.boundary.n_particles == nw


def test_dam_break_lattice_and_hydrostatics(tp):
    """Fluid coordinates sit within the jitter of dx (i - 1/2); the hydrostatic density
    makes p = rho0 g depth (first order in dx) and masses are dx^D rho."""
    dx = 0.02
    db = tp.dam_break(2, dx)
    x = db.tank.fluid.coordinates
    frac = x / dx + 0.5
    assert np.abs(frac - np.rint(frac)).max() <= 0.01 + 1e-9
    se = db.state_equation
    p = np.array([se(r) for r in db.tank.fluid.density])
    depth = 0.6 - x[1]
    assert np.abs(p - 1000.0 * 9.81 * depth).max() < 1000.0 * 9.81 * dx
    assert np.allclose(db.fluid_system.mass, dx ** 2 * db.tank.fluid.density, rtol=1e-15)


def test_rectangular_tank_validates_dimensions(tp):
    with pytest.raises(ValueError, match="must be of length 2 for a 2D problem"):
        tp.RectangularTank(0.2, (0.5, 0.4), (0.5, 0.5, 0.0), 1000.0, spacing_ratio=3)
    with pytest.raises(ValueError, match="must be of length 3 for a 3D problem"):
        tp.RectangularTank(0.2, (0.5, 0.4, 0.5), (0.5, 0.5), 1000.0, spacing_ratio=3)
