"""tpsph_kick_noalloc (the kick without device allocation, its arrays in the library's fixed
scratch array) against tpsph_kick on a handle with the same parameters and the cell search.

Both run the same kernels on the same grid, so the results are compared bit for bit.  Two
calls in a row, the second with a smaller input: the scratch array keeps the first call's
contents, so the second call shows that everything that must start at zero is cleared."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_available


def _case(tp, nd, dx, seed):
    db = tp.dam_break(nd, dx)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system, pair_search="cells")
    v0, u0 = semi.initial_state()
    v = v0.clone()
    gen = np.random.default_rng(seed)
    import torch
    v[:, :nd] = torch.from_numpy(gen.normal(0.0, 0.4, size=(v.shape[0], nd))).cuda()
    return db, semi, v, u0


def _noalloc_kick(tp, semi, db, v, u):
    import torch
    lib = tp.native.load()
    p = semi.params()
    fm = torch.from_numpy(np.ascontiguousarray(db.fluid_system.mass)).cuda()
    wx = torch.from_numpy(np.ascontiguousarray(db.boundary_system.coordinates.T)).cuda()
    wm = torch.from_numpy(
        np.ascontiguousarray(db.boundary_system.boundary_model.hydrodynamic_mass)).cuda()
    dv = torch.full_like(v, float("nan"))
    torch.cuda.synchronize()
    rc = lib.tpsph_kick_noalloc(C.byref(p), db.n_fluid, db.n_wall, fm.data_ptr(),
                                wx.data_ptr(), wm.data_ptr(), dv.data_ptr(), v.data_ptr(),
                                u.data_ptr(), 0.0, None)
    assert rc == 0, lib.tpsph_last_error(None).decode()
    torch.cuda.synchronize()
    return dv.cpu().numpy()


def _handle_kick(tp, semi, v, u):
    import torch
    dv = torch.full_like(v, float("nan"))
    tp.kick(dv, v, u, {"semi": semi}, 0.0)
    semi.synchronize()
    return dv.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_noalloc_kick_equals_handle_kick_twice():
    from conftest import load_package
    tp = load_package()
    for nd, dx, seed in ((3, 0.025, 1), (2, 0.02, 2)):  # the second input is smaller
        print(f"tpsph_kick_noalloc: {nd}D dam break, dx = {dx}", flush=True)
        db, semi, v, u = _case(tp, nd, dx, seed)
        got = _noalloc_kick(tp, semi, db, v, u)
        ref = _handle_kick(tp, semi, v, u)
        assert np.isfinite(ref).all()
        assert np.array_equal(got, ref), f"{nd}D: max |diff| {np.abs(got - ref).max():.3e}"
        semi.close()
