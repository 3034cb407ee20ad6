"""The C-ABI library loads, exports every symbol include/tpsph.h declares, and rejects
invalid parameters at create time (validation runs before any HIP call, so this needs no GPU)."""
import ctypes as C
import os
import re

import pytest

from conftest import REPO


def declared_symbols(header="tpsph.h"):
    text = open(os.path.join(REPO, "include", header)).read()
    return sorted(set(re.findall(r"\b(tpsph_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(tp):
    lib = tp.native.load()
    syms = declared_symbols()
    assert "tpsph_kick" in syms and "tpsph_create" in syms
    assert {"tpsph_kick_vjp", "tpsph_adjoint_forward", "tpsph_adjoint_reverse"} <= set(syms)
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in include/tpsph.h but not exported"
    assert set(tp.native.EXPORTED_SYMBOLS) == set(syms)
    assert lib.tpsph_abi_version() == 6


def test_library_exports_every_symbol_of_every_header(tp):
    lib = tp.native.load()
    headers = sorted(f for f in os.listdir(os.path.join(REPO, "include")) if f.endswith(".h"))
    assert "tpsph_batch.h" in headers
    for header in headers:
        for s in declared_symbols(header):
            assert hasattr(lib, s), f"{s} declared in include/{header} but not exported"


def test_reverse_mode_entry_points_reject_a_null_handle(tp):
    """The adjoint entry points validate before touching the device."""
    lib = tp.native.load()
    pb = (C.c_double * 3)()
    assert lib.tpsph_kick_vjp(None, None, None, None, 0.0, None, None, pb) == \
        tp.native.TPSPH_ERR_ARGUMENT
    assert lib.tpsph_adjoint_forward(None, None, None, 0.0, 1e-5, 1, 1, None, None) == \
        tp.native.TPSPH_ERR_ARGUMENT
    assert lib.tpsph_adjoint_reverse(None, None, None, None, None, pb) == \
        tp.native.TPSPH_ERR_ARGUMENT
    assert "NULL" in lib.tpsph_last_error(None).decode()


def test_sort_entry_point_rejects_a_null_handle(tp):
    lib = tp.native.load()
    assert lib.tpsph_sort_particles(None, None, None, None) == tp.native.TPSPH_ERR_ARGUMENT
    assert "NULL" in lib.tpsph_last_error(None).decode()


def test_params_struct_layout_matches_header(tp):
    # 10 int32 + 9 doubles + acceleration[3] + offset + domain_min[3] + domain_max[3] + skin
    # + adaptive flag and its padding + 3 adaptive doubles
    # + wall viscosity kind and its padding + 4 wall viscosity doubles
    assert C.sizeof(tp.native.TpsphParams) == \
        10 * 4 + (9 + 3 + 1 + 3 + 3 + 1) * 8 + 2 * 4 + 3 * 8 + 2 * 4 + 4 * 8


def _create(tp, **overrides):
    lib = tp.native.load()
    p = tp.native.TpsphParams()
    p.ndims = 2
    p.memory = tp.native.TPSPH_MEM_HOST
    p.smoothing_length = 0.012
    p.reference_density = 1000.0
    p.sound_speed = 48.5
    p.exponent = 7.0
    for k, v in overrides.items():
        setattr(p, k, v)
    out = C.c_void_p()
    mass = (C.c_double * 1)(1.0)
    rc = lib.tpsph_create(C.byref(p), 1, 0, C.cast(mass, C.c_void_p), None, None, None,
                          C.byref(out))
    return rc, lib.tpsph_last_error(None).decode()


@pytest.mark.parametrize("field,value,msg", [
    ("ndims", 4, "ndims"),
    ("smoothing_length", 0.0, "smoothing_length"),
    ("smoothing_length", float("nan"), "smoothing_length"),
    ("reference_density", 0.0, "reference_density"),
    ("sound_speed", -1.0, "sound_speed"),
    ("exponent", 0.0, "exponent"),
    ("kernel", 3, "WendlandC2"),
    ("memory", 7, "memory"),
    ("pair_search", 4, "pair_search"),
    ("pair_search", -1, "pair_search"),
    ("skin", 1.0, "skin"),
    ("cell_divisions", 5, "cell_divisions"),
    ("use_density_diffusion", 3, "use_density_diffusion"),
    ("use_density_diffusion", -1, "use_density_diffusion"),
    ("density_calculator", 2, "density_calculator"),
    ("density_calculator", -1, "density_calculator"),
])
def test_create_rejects_invalid_parameters(tp, field, value, msg):
    rc, err = _create(tp, **{field: value})
    assert rc == tp.native.TPSPH_ERR_ARGUMENT
    assert msg in err


@pytest.mark.parametrize("fields,msg", [
    ({"mach_number_target": 0.0, "min_sound_speed": 10.0, "max_sound_speed": 100.0},
     "mach_number_target"),
    ({"mach_number_target": 0.1, "min_sound_speed": 0.0, "max_sound_speed": 100.0},
     "min_sound_speed"),
    ({"mach_number_target": 0.1, "min_sound_speed": 10.0, "max_sound_speed": 5.0},
     "max_sound_speed"),
])
def test_create_rejects_invalid_adaptive_state_equation(tp, fields, msg):
    rc, err = _create(tp, adaptive_sound_speed=1, **fields)
    assert rc == tp.native.TPSPH_ERR_ARGUMENT
    assert msg in err


def test_create_rejects_mixed_precision_with_a_quadratic_viscosity(tp):
    """The FP32 pair terms carry ArtificialViscosityMonaghan's alpha term only."""
    rc, err = _create(tp, precision=1, use_viscosity=1, alpha=0.02, beta=0.5)
    assert rc == tp.native.TPSPH_ERR_ARGUMENT
    assert "beta = 0" in err


def test_create_rejects_summation_density_in_mixed_precision(tp):
    """SummationDensity runs in FP64 (on the cell walk or the brick kernel)."""
    rc, err = _create(tp, density_calculator=1, precision=1)
    assert rc == tp.native.TPSPH_ERR_ARGUMENT
    assert "SummationDensity" in err


def test_semidiscretization_rejects_unknown_pair_search(tp):
    db = tp.dam_break(2, 0.05)
    with pytest.raises(ValueError, match="pair_search"):
        tp.Semidiscretization(db.fluid_system, db.boundary_system, pair_search="octree")


def test_create_rejects_null_pointers(tp):
    lib = tp.native.load()
    p = tp.native.TpsphParams()
    p.ndims = 3
    p.memory = tp.native.TPSPH_MEM_HOST
    p.smoothing_length = p.reference_density = p.sound_speed = p.exponent = 1.0
    out = C.c_void_p()
    assert lib.tpsph_create(C.byref(p), 10, 0, None, None, None, None, C.byref(out)) == \
        tp.native.TPSPH_ERR_ARGUMENT
    assert lib.tpsph_create(None, 10, 0, None, None, None, None, C.byref(out)) == \
        tp.native.TPSPH_ERR_ARGUMENT
    assert lib.tpsph_kick(None, None, None, None, 0.0) == tp.native.TPSPH_ERR_ARGUMENT


def test_system_constructors_reject_out_of_scope_options(tp):
    ic = tp.RectangularShape(0.1, (3, 3), (0.0, 0.0), density=1000.0)
    se = tp.StateEquationCole(sound_speed=10.0, reference_density=1000.0, exponent=7)
    k = tp.WendlandC2Kernel(2)
    with pytest.raises(NotImplementedError):
        tp.WeaklyCompressibleSPHSystem(ic, tp.ContinuityDensity(), se, k, 0.12,
                                       surface_tension=object())
    with pytest.raises(ValueError):
        tp.WeaklyCompressibleSPHSystem(ic, tp.ContinuityDensity(), se, tp.WendlandC2Kernel(3),
                                       0.12)
    with pytest.raises(NotImplementedError):
        tp.BoundaryModelDummyParticles([1000.0], [1.0], tp.AdamiPressureExtrapolation(), k,
                                       0.12, state_equation=se, viscosity=object())


def test_noalloc_kick_rejects_sizes_beyond_its_scratch(tp):
    """tpsph_kick_noalloc measures its arrays before touching the device: 2^25 fluid particles
    need far more than its fixed 64 MiB and are rejected with the limit in the message."""
    lib = tp.native.load()
    p = tp.native.TpsphParams()
    p.ndims = 3
    p.memory = tp.native.TPSPH_MEM_DEVICE
    p.smoothing_length = 0.012
    p.reference_density = 1000.0
    p.sound_speed = 48.5
    p.exponent = 7.0
    for d in range(3):
        p.domain_min[d], p.domain_max[d] = 0.0, 1.0
    dummy = C.c_void_p(8)  # never dereferenced: the size check comes first
    rc = lib.tpsph_kick_noalloc(C.byref(p), 1 << 25, 0, dummy, None, None, dummy, dummy, dummy,
                                0.0, None)
    assert rc == tp.native.TPSPH_ERR_ARGUMENT
    msg = lib.tpsph_last_error(None).decode()
    assert "limit is 67108864 bytes" in msg, msg
    p.domain_max[0] = 0.0  # no grid box
    rc = lib.tpsph_kick_noalloc(C.byref(p), 10, 0, dummy, None, None, dummy, dummy, dummy, 0.0,
                                None)
    assert rc == tp.native.TPSPH_ERR_ARGUMENT and "grid box" in lib.tpsph_last_error(None).decode()
