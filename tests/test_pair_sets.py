"""The pairs the kick's kernels summed (tpsph_kick_pairs_csr) against brute force.

tpsph_neighbors_csr runs a search of its own; this export instead reads the lists the kick
evaluated -- the 16-bit brick lists of k_fluid_brick mapped back through the brick row
tables, the int Verlet lists, the wall and Adami lists, and the cell walk of an overflowed
list -- and applies the kernels' exact radius test.  The workload demands bit-exact sets:
a pair dropped near q = 2 changes dv by far less than the 1e-12 tolerance of the kick
tests (grad W vanishes there as (1 - q/2)^3), so only a set comparison shows it.

Every case kicks twice: once at positions that build the lists, once after a displacement
below half the skin, so the second kick reuses the lists of the first.  Pairs are planted
at squared distance exactly R^2 (which belongs to the relation: d^2 <= R^2) and at the
nearest representable squared distances on either side."""
import numpy as np
import pytest

from conftest import gpu_available

RTOL = 1e-12


def dist2(nd, a, b):
    """The kernels' evaluation order: (dx*dx + dy*dy) + dz*dz."""
    d = a - b
    s = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
    if nd == 3:
        s = s + d[..., 2] * d[..., 2]
    return s


def boundary_offsets(x_a, R2, n=4000):
    """Coordinates x_b below x_a with (x_a - x_b)^2 == R2, and the neighbours of that value
    on either side, among the doubles near x_a - sqrt(R2).  Returns (exact, inside, outside)
    or None when no x_b gives R2 exactly."""
    base = x_a - np.sqrt(R2)
    cand = base + np.arange(-n, n + 1) * np.spacing(base)
    d = x_a - cand
    d2 = d * d
    exact = cand[d2 == R2]
    if exact.size == 0:
        return None
    inside = cand[d2 < R2]
    outside = cand[d2 > R2]
    x_in = inside[np.argmax(d2[d2 < R2])]
    x_out = outside[np.argmin(d2[d2 > R2])]
    return exact[0], x_in, x_out


def plant_pairs(nd, u, targets, movers, R2, axis=0):
    """Move movers[k] next to targets[k] along `axis` so that the squared distances are R2,
    just below and just above (three movers per target).  Returns the planted (a, b, d2)."""
    planted = []
    mi = 0
    for a in targets:
        found = boundary_offsets(u[a, axis], R2)
        if found is None:
            continue
        for xb in found:
            b = movers[mi]
            mi += 1
            u[b] = u[a]
            u[b, axis] = xb
            planted.append((a, b, float(dist2(nd, u[a], u[b]))))
        if len(planted) >= 6:
            break
    return planted


def test_plant_pairs_hits_the_radius_exactly():
    """The construction itself (CPU): R^2 exactly, and one representable value either side."""
    R = 2 * 1.2 * 0.02
    R2 = R * R
    rng = np.random.default_rng(5)
    u = rng.uniform(0.1, 0.9, size=(400, 3))
    planted = plant_pairs(3, u, range(0, 200), range(200, 400), R2)
    assert len(planted) >= 3
    d2s = sorted({p[2] for p in planted[:3]})
    assert d2s[1] == R2 and d2s[0] < R2 < d2s[2]


def csr_pairs(off, idx):
    rows = np.repeat(np.arange(off.size - 1), np.diff(off))
    return rows, idx.astype(np.int64)


def reference_sets(orc, nd, R, xa, xb, az2):
    """Brute force d^2 <= R^2 (the oracle, the kernels' evaluation order), without the pairs
    below az2 when az2 > 0 (the pair terms skip them; the Adami sums do not)."""
    off, idx = orc.neighbors(nd, R, xa, xb, brute=True)
    if az2 <= 0:
        return off, idx
    rows, cols = csr_pairs(off, idx)
    keep = dist2(nd, xa[rows], xb[cols]) >= az2
    counts = np.bincount(rows[keep], minlength=xa.shape[0])
    new_off = np.zeros(xa.shape[0] + 1, dtype=np.int64)
    np.cumsum(counts, out=new_off[1:])
    return new_off, cols[keep]


def assert_same_sets(got, ref, what):
    goff, gidx = got
    roff, ridx = ref
    assert np.array_equal(goff, roff), f"{what}: pair counts differ"
    assert np.array_equal(gidx.astype(np.int64), ridx), f"{what}: pair sets differ"


def run_case(tp, orc, nd, dx, search, mass_factor=1.0):
    import torch
    db = tp.dam_break(nd, dx)
    if mass_factor != 1.0:
        db.fluid_system.mass = db.fluid_system.mass * mass_factor
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system, pair_search=search)
    v0, u0 = semi.initial_state()
    u = u0.cpu().numpy().copy()
    v = v0.cpu().numpy().copy()
    rng = np.random.default_rng(13)
    v[:, :nd] = rng.normal(0.0, 0.3, size=(v.shape[0], nd))
    h = db.fluid_system.smoothing_length
    R = 2 * h
    R2 = R * R
    az2 = np.nextafter(R2, np.inf) - R2
    xw = db.boundary_system.coordinates.T.copy()
    nf = u.shape[0]
    # fluid-fluid pairs on the radius, interior particles moved next to others
    order = rng.permutation(nf)
    planted = plant_pairs(nd, u, order[:400], order[400:430], R2, axis=0)
    assert any(p[2] == R2 for p in planted), "no exact R^2 pair was planted"
    # fluid-wall pairs on the radius: fluid particles moved next to wall particles along y
    wall_planted = []
    for w in rng.permutation(xw.shape[0])[:300]:
        found = boundary_offsets(-xw[w, 1], R2)  # fluid above: y_f = y_w + offset
        if found is None:
            continue
        for yneg in found:
            f = order[430 + len(wall_planted)]
            u[f] = xw[w]
            u[f, 1] = -yneg
            wall_planted.append((f, w, float(dist2(nd, u[f], xw[w]))))
        if len(wall_planted) >= 3:
            break
    dv = torch.zeros_like(v0)
    vd = torch.from_numpy(v).cuda()
    tp.kick(dv, vd, torch.from_numpy(u).cuda(), {"semi": semi}, 0.0)  # builds the lists
    # a displacement below half the skin (default 0.04 R): the next kick reuses the lists
    moved = np.ones(nf, dtype=bool)
    moved[[p[1] for p in planted] + [p[0] for p in planted] + [p[0] for p in wall_planted]] = False
    u2 = u.copy()
    amp = 0.004 * R
    u2[moved] += rng.uniform(-amp, amp, size=(int(moved.sum()), nd))
    tp.kick(dv, vd, torch.from_numpy(u2).cuda(), {"semi": semi}, 0.0)
    semi.synchronize()
    got = {p: tp.kick_pairs(semi, p) for p in (tp.native.PAIR_FLUID_FLUID,
                                               tp.native.PAIR_FLUID_WALL,
                                               tp.native.PAIR_WALL_FLUID)}
    assert_same_sets(got[tp.native.PAIR_FLUID_FLUID],
                     reference_sets(orc, nd, R, u2, u2, az2), "fluid-fluid")
    assert_same_sets(got[tp.native.PAIR_FLUID_WALL],
                     reference_sets(orc, nd, R, u2, xw, az2), "fluid-wall")
    assert_same_sets(got[tp.native.PAIR_WALL_FLUID],
                     reference_sets(orc, nd, R, xw, u2, -1.0), "wall-fluid")
    # the planted pairs: in the relation exactly when d^2 <= R^2
    off, idx = got[tp.native.PAIR_FLUID_FLUID]
    for a, b, d2 in planted:
        assert (b in idx[off[a]:off[a + 1]]) == (d2 <= R2), (a, b, d2, R2)
    off, idx = got[tp.native.PAIR_FLUID_WALL]
    for f, w, d2 in wall_planted:
        assert (w in idx[off[f]:off[f + 1]]) == (d2 <= R2), (f, w, d2, R2)
    # and the kick itself at the reused lists matches the oracle
    P = orc.params_from_semi(semi)
    dv_ref, *_ = orc.kick(P, v, u2, db.fluid_system.mass, xw,
                          db.boundary_system.boundary_model.hydrodynamic_mass)
    got_dv = dv.cpu().numpy()
    for col in range(dv_ref.shape[1]):
        scale = np.abs(dv_ref[:, col]).max()
        err = np.abs(got_dv[:, col] - dv_ref[:, col]).max() / scale
        assert err <= RTOL, f"dv column {col}: {err:.3e}"
    return semi


@pytest.mark.parametrize("search", ["lists", "cells"])
@pytest.mark.parametrize("nd,dx", [(2, 0.01), (3, 0.02)])
@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_kick_pairs_match_bruteforce_after_list_reuse(tp, orc, nd, dx, search):
    run_case(tp, orc, nd, dx, search)


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")
def test_kick_pairs_with_overflowed_lists(tp, orc):
    """40x the masses: every list overflows and each target walks the cells."""
    run_case(tp, orc, 3, 0.025, "lists", mass_factor=40.0)
