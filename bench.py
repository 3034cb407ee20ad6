"""Benchmark of the MI355X WCSPH right-hand side.  Prints one JSON line.

Headline (`value`): particle-steps/s of the 3D dam break at 1.02M fluid particles
(160 x 80 x 80 at dx = 6.25 mm, 908,844 wall particles), FP64.  One step = one
CarpenterKennedy2N54 step (5 x kick! + drift! + stage update) run on the device
(tpsph_solve_ck2n54) with the state resident in HBM.

Beside it, in the same line:
  roofline       the fused fluid interaction's algorithmic bytes per launch over its average
                 launch time (HIP events on the library's stream), against 8 TB/s; its FP64
                 VALU rate from the committed counter profile, against 78.6 TFLOP/s
  moving_flow    the same case in motion: after an untimed run from rest, CFL-limited steps
                 (dt from the step-size rule, which makes the cell order rebuild often)
  precision      the 8.2M-particle refinement (dx = 3.125 mm): FP64 vs mixed precision (FP32
                 pair arithmetic), speed-up and the density error of the mixed run
  ensemble       256 uneven 2D dam breaks (about 2M fluid particles) advanced together in
                 graph-captured steps vs the same cases one at a time
  adjoint        the 250k-particle 3D case: forward run vs the reverse sweep (checkpoints
                 every 10 steps, recomputation), time ratio
  sorted_order   the headline case after one SortingCallback sort (tpsph_sort_particles):
                 the caller's order becomes the library's cell order
  cpu_baseline   the FP64 C oracle with OpenMP on the host cores: median of 5 right-hand
                 sides after 2 warm-ups and one full CK2N54 step of the headline case

  python bench.py [--gpus 1] [--steps K] [--warmup W]
  python bench.py --only headline|moving|precision|ensemble|adjoint|sorted   (one section)
  python bench.py --rebuild-kicks N    # cost of a rebuild (separate JSON line, no bench)
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6
SECTIONS = ("headline", "moving", "precision", "ensemble", "adjoint", "sorted")


def profile_text(kernel):
    import glob
    for path in sorted(glob.glob(os.path.join(REPO, "profiles", f"{kernel}_*.md"))):
        text = open(path).read()
        if f"`{kernel}" in text:
            yield path, text


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from its committed rocprofv3 summary (FETCH_SIZE +
    WRITE_SIZE, calibrated against a copy of known size), or None."""
    import re
    for path, text in profile_text(kernel):
        m = re.search(r"`%s`: fetch [\d.]+ MB \+ write [\d.]+ MB = \*\*([\d.]+) MB\*\* per launch"
                      % re.escape(kernel), text)
        if m:
            return float(m.group(1)) * 1e6, os.path.relpath(path, REPO)
    return None, None


def profiled_fp64_flops(kernel):
    """FP64 operations per launch of `kernel` from the SQ_INSTS_VALU_*_F64 counters of its
    committed summary (wavefront instructions x 64 lanes; an FMA counts 2), or None."""
    import re
    for path, text in profile_text(kernel):
        m = re.search(r"- `%s<3[^`]*>` \(\d+ dispatches, per dispatch[^)]*\): "
                      r"([^\n]*SQ_INSTS_VALU_FMA_F64[^\n]*)" % re.escape(kernel), text)
        if not m:
            continue
        vals = dict(re.findall(r"(SQ_INSTS_VALU_\w+_F64) = ([\d.e+]+)", m.group(1)))
        ops = (float(vals.get("SQ_INSTS_VALU_ADD_F64", 0)) +
               float(vals.get("SQ_INSTS_VALU_MUL_F64", 0)) +
               2.0 * float(vals.get("SQ_INSTS_VALU_FMA_F64", 0)) +
               float(vals.get("SQ_INSTS_VALU_TRANS_F64", 0)))
        return 64.0 * ops, os.path.relpath(path, REPO)
    return None, None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--dx", type=float, default=6.25e-3)
    ap.add_argument("--dt", type=float, default=1e-5)
    ap.add_argument("--only", choices=SECTIONS, default=None,
                    help="run one section (its JSON object is printed under its name)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--pair-search", default="auto", choices=["auto", "lists", "cells"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="arithmetic of the headline run")
    ap.add_argument("--rebuild-kicks", type=int, default=0, metavar="N",
                    help="instead of the bench, time N kicks that each rebuild the cell list "
                         "and the Verlet lists")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the run, copy MIB MiB device to device once (a plain copy "
                         "of known size that calibrates FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def fluid_bytes(nf, nw, nd=3):
    """Algorithmic HBM bytes of one fluid interaction: fluid x, v (8 nd each), rho, p, m read,
    dv (8 (nd+1)) written; wall x (8 nd), m, p, rho read."""
    return nf * (16 * nd + 24 + 8 * (nd + 1)) + nw * (8 * nd + 24)


def kernel_times(lib, native, h):
    times = {}
    for name, kid in native.TIMED_KERNELS.items():
        n, ms = C.c_int64(), C.c_double()
        native.check(lib.tpsph_kernel_time(h, kid, C.byref(n), C.byref(ms)), h)
        times[name] = (n.value, ms.value)
    return times


def run_solve(tp, lib, native, semi, v, u, dt, warmup, steps):
    """warmup + timed steps of tpsph_solve_ck2n54 with kernel timers; returns seconds."""
    import torch
    h = semi.handle

    def run(n):
        native.check(lib.tpsph_solve_ck2n54(h, v.data_ptr(), u.data_ptr(), 0.0, dt, n), h)

    if warmup:
        run(warmup)
    torch.cuda.synchronize()
    native.check(lib.tpsph_enable_timing(h, 1), h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    times = kernel_times(lib, native, h)
    native.check(lib.tpsph_enable_timing(h, 0), h)
    return elapsed, times


def rebuilds(lib, h):
    n = C.c_int64()
    lib.tpsph_rebuild_count(h, C.byref(n))
    return n.value


def section_headline(tp, lib, native, args):
    import numpy as np
    db = tp.dam_break(3, args.dx)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 pair_search=args.pair_search, precision=args.precision)
    ode = tp.semidiscretize(semi, (0.0, args.dt * (args.warmup + args.steps)))
    v, u = ode.v0.clone(), ode.u0.clone()
    nf, nw = db.n_fluid, db.n_wall
    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    r0 = rebuilds(lib, semi.handle)
    elapsed, times = run_solve(tp, lib, native, semi, v, u, args.dt, args.warmup, args.steps)
    r1 = rebuilds(lib, semi.handle)
    finite = bool(np.isfinite(v.cpu().numpy()).all())
    search = args.pair_search
    if search == "auto":
        search = "lists" if semi.skin != 0.0 else "cells"
    fluid_kernel = {"lists": "k_fluid_brick", "cells": "k_fluid"}[search]
    alg = fluid_bytes(nf, nw)
    traffic, traffic_src = profiled_traffic(fluid_kernel)
    flops, flops_src = profiled_fp64_flops(fluid_kernel)
    n_launch, fluid_ms = times["fluid"]
    avg_s = fluid_ms / 1e3 / max(n_launch, 1)
    achieved = alg / avg_s / 1e9
    out = {
        "value": nf * args.steps / elapsed,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "n_fluid": nf, "n_wall": nw, "finite": finite,
        "rebuilds_in_timed_steps": r1 - r0 if args.warmup == 0 else None,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": f"{fluid_kernel} + k_fluid_walls (fused fluid-fluid + fluid-wall "
                      "interaction)",
            "avg_launch_us": avg_s * 1e6,
            "algorithmic_bytes_per_launch": alg,
            "fp64_valu_tflops": flops / avg_s / 1e12 if flops else None,
            "fp64_valu_frac": flops / avg_s / 1e12 / FP64_VALU_PEAK_TFLOPS if flops else None,
            "fp64_valu_source": flops_src,
        },
        "kernel_ms_per_step": {k: ms / args.steps for k, (n, ms) in times.items()},
    }
    semi.close()
    return out, (db, ode)


def section_moving(tp, lib, native, args, spin_steps=1500, steps=100):
    """The 3D dam break in motion: `spin_steps` untimed CFL-limited steps from rest, then
    `steps` timed ones; dt from the step-size rule (CFL 0.9, StepsizeCallback)."""
    db = tp.dam_break(3, args.dx)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system)
    dt = tp.calculate_dt(None, None, 0.9, semi)
    ode = tp.semidiscretize(semi, (0.0, dt * (spin_steps + steps)))
    v, u = ode.v0.clone(), ode.u0.clone()
    native.check(lib.tpsph_solve_ck2n54(semi.handle, v.data_ptr(), u.data_ptr(), 0.0, dt,
                                        spin_steps), semi.handle)
    r0 = rebuilds(lib, semi.handle)
    elapsed, times = run_solve(tp, lib, native, semi, v, u, dt, 0, steps)
    r1 = rebuilds(lib, semi.handle)
    vmax = float(v[:, :3].norm(dim=1).max())
    semi.close()
    return {
        "value": db.n_fluid * steps / elapsed,
        "unit": "particle-steps/s",
        "ms_per_step": elapsed * 1e3 / steps,
        "dt": dt,
        "start_time": dt * spin_steps,
        "steps": steps,
        "rebuilds_per_step": (r1 - r0) / steps,
        "max_speed": vmax,
        "kernel_ms_per_step": {k: ms / steps for k, (n, ms) in times.items()},
    }


def section_sorted(tp, lib, native, args, warmup=5, steps=200):
    """The headline case with its fluid rows put into the library's cell order once (the
    sort of SortingCallback, tpsph_sort_particles) before the timed steps."""
    import torch
    db = tp.dam_break(3, args.dx)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system)
    ode = tp.semidiscretize(semi, (0.0, args.dt * (warmup + steps)))
    v, u = ode.v0.clone(), ode.u0.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    semi.sort_particles(v, u)
    torch.cuda.synchronize()
    sort_s = time.perf_counter() - t0
    elapsed, times = run_solve(tp, lib, native, semi, v, u, args.dt, warmup, steps)
    semi.close()
    return {
        "value": db.n_fluid * steps / elapsed,
        "unit": "particle-steps/s",
        "ms_per_step": elapsed * 1e3 / steps,
        "steps": steps,
        "sort_ms": sort_s * 1e3,
        "kernel_ms_per_step": {k: ms / steps for k, (n, ms) in times.items()},
    }


def section_precision(tp, lib, native, args, warmup=3, steps=20, dx=3.125e-3):
    """FP64 and mixed precision on the 8.2M-particle refinement, same initial state."""
    import numpy as np
    db = tp.dam_break(3, dx)
    res, states = {}, {}
    for precision in ("fp64", "mixed"):
        semi = tp.Semidiscretization(db.fluid_system, db.boundary_system, precision=precision)
        ode = tp.semidiscretize(semi, (0.0, args.dt * (warmup + steps)))
        v, u = ode.v0.clone(), ode.u0.clone()
        elapsed, times = run_solve(tp, lib, native, semi, v, u, args.dt, warmup, steps)
        n_launch, fluid_ms = times["fluid"]
        avg_s = fluid_ms / 1e3 / max(n_launch, 1)
        res[precision] = {"ms_per_step": elapsed * 1e3 / steps,
                          "value": db.n_fluid * steps / elapsed,
                          "fluid_pass_us": avg_s * 1e6,
                          "fluid_pass_gbs": fluid_bytes(db.n_fluid, db.n_wall) / avg_s / 1e9,
                          "kernel_ms_per_step": {k: ms / steps for k, (n, ms) in times.items()}}
        states[precision] = v.cpu().numpy()
        semi.close()
        del ode, v, u
    rho0 = db.state_equation.reference_density
    err = float(np.abs(states["mixed"][:, 3] - states["fp64"][:, 3]).max() / rho0)
    return {
        "workload": f"3D dam break at dx = {dx}: {db.n_fluid} fluid + {db.n_wall} wall "
                    f"particles, {warmup} + {steps} CK2N54 steps at dt = {args.dt}",
        "fp64": res["fp64"], "mixed": res["mixed"],
        "speedup": res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"],
        "density_error_vs_fp64": err,
        "unit": "particle-steps/s",
    }


def section_ensemble(tp, lib, native, args, n_cases=256, steps=300, alone_steps=30, dt=1e-5):
    """256 uneven 2D dam breaks: the ensemble for `steps` graph-captured steps, and every
    case alone for `alone_steps` steps -- as an ensemble of one (graph-captured) and on its
    own handle (tpsph_solve_ck2n54) -- summed over the cases."""
    import torch
    from trixiparticles_jl_amd import ensemble as en
    cases = en.dam_break_ensemble(n_cases)
    ens = en.Ensemble(cases)
    v, u = ens.v0.clone(), ens.u0.clone()
    ens.solve(dt, 2, v=v, u=u)  # capture + warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.solve(dt, steps, v=v, u=u)
    t_batch = (time.perf_counter() - t0) / steps
    # the same steps as the one-at-a-time runs, for the ratio
    v2, u2 = ens.v0.clone(), ens.u0.clone()
    ens.solve(dt, 1, v=v2, u=u2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.solve(dt, alone_steps, v=v2, u=u2)
    t_batch_short = (time.perf_counter() - t0) / alone_steps
    n_fluid, n_wall = ens.n_fluid, ens.n_wall
    tuning = en.load_tuning()
    ens.close()
    t_one_graph = 0.0
    t_one_handle = 0.0
    for c in cases:
        one = en.Ensemble([c])
        vo, uo = one.v0.clone(), one.u0.clone()
        one.solve(dt, 1, v=vo, u=uo)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        one.solve(dt, alone_steps, v=vo, u=uo)
        t_one_graph += time.perf_counter() - t0
        one.close()
        semi = tp.Semidiscretization(c.dam_break.fluid_system, c.dam_break.boundary_system)
        ode = tp.semidiscretize(semi, (0.0, dt * alone_steps))
        vh, uh = ode.v0.clone(), ode.u0.clone()
        native.check(lib.tpsph_solve_ck2n54(semi.handle, vh.data_ptr(), uh.data_ptr(), 0.0, dt,
                                            1), semi.handle)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        native.check(lib.tpsph_solve_ck2n54(semi.handle, vh.data_ptr(), uh.data_ptr(), 0.0, dt,
                                            alone_steps), semi.handle)
        torch.cuda.synchronize()
        t_one_handle += time.perf_counter() - t0
        semi.close()
    t_one_graph /= alone_steps
    t_one_handle /= alone_steps
    best_alone = min(t_one_graph, t_one_handle)
    sizes = [c.n_fluid for c in cases]
    return {
        "workload": f"{n_cases} 2D dam breaks (seed 1): {n_fluid} fluid + {n_wall} wall "
                    f"particles, {min(sizes)} to {max(sizes)} fluid per case, dt = {dt}",
        "value": n_fluid / t_batch,
        "unit": "particle-steps/s",
        "ms_per_step": t_batch * 1e3,
        "steps": steps,
        "one_at_a_time_ms_per_step": {"ensemble_of_one_graph": t_one_graph * 1e3,
                                      "own_handle": t_one_handle * 1e3,
                                      "steps_per_case": alone_steps},
        "speedup_vs_one_at_a_time": best_alone / t_batch_short,
        "launch_shapes": {k: tuning.get(k) for k in ("cell_divisions", "x_cell_divisions",
                                                     "fluid_block", "wall_block")}
        if tuning else None,
    }


def section_adjoint(tp, lib, native, args):
    """Forward run and reverse sweep of the 250k-particle 3D case (tpsph_adjoint)."""
    from trixiparticles_jl_amd import adjoint as ad
    return ad.benchmark()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(db, ode, semi_params, dt):
    """The FP64 C oracle with OpenMP on the host cores: 2 warm-up right-hand sides, the
    median of 5, then one full CK2N54 step (5 right-hand sides + stage updates)."""
    from oracle import oracle as orc
    v0, u0 = ode.v0.cpu().numpy(), ode.u0.cpu().numpy()
    xw = db.boundary_system.coordinates.T.copy()
    mw = db.boundary_system.boundary_model.hydrodynamic_mass
    m = db.fluid_system.mass
    ts = []
    for k in range(7):
        tc = time.perf_counter()
        orc.kick(semi_params, v0, u0, m, xw, mw)
        if k >= 2:
            ts.append(time.perf_counter() - tc)
    t_rhs = statistics.median(ts)
    tc = time.perf_counter()
    orc.solve_ck2n54(semi_params, v0, u0, m, xw, mw, dt, 1)
    t_step = time.perf_counter() - tc
    nf = db.n_fluid
    return {"value": nf / t_step, "unit": "particle-steps/s", "cores": orc.num_threads(),
            "kind": "port", "cpu_model": cpu_model(),
            "rhs_median_s": t_rhs, "step_s": t_step,
            "sample": f"headline case ({nf} fluid + {db.n_wall} wall particles): 2 warm-up "
                      f"right-hand sides, median of 5 = {t_rhs:.3f} s, one full CK2N54 step "
                      f"= {t_step:.3f} s (value = fluid particles / step time)"}


def rebuild_kicks(tp, lib, native, args):
    """Kicks that all rebuild: the positions alternate between u and u shifted by 0.06 dx
    along x, beyond half the default skin, so every kick rebuilds the cell-sorted order, the
    Verlet lists and the brick tables.  Prints one JSON line."""
    import torch
    db = tp.dam_break(3, args.dx)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 pair_search=args.pair_search, precision=args.precision)
    ode = tp.semidiscretize(semi, (0.0, args.dt))
    v, u = ode.v0.clone(), ode.u0.clone()
    u2 = u.clone()
    u2.view(-1, 3)[:, 0] += 0.06 * args.dx
    dv = torch.empty_like(v)
    p = {"semi": semi}
    for k in range(4):
        tp.kick(dv, v, u if k % 2 == 0 else u2, p, 0.0)
    torch.cuda.synchronize()
    lib.tpsph_enable_timing(semi.handle, 1)
    t0 = time.perf_counter()
    for k in range(args.rebuild_kicks):
        tp.kick(dv, v, u if k % 2 == 0 else u2, p, 0.0)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n, ms = C.c_int64(), C.c_double()
    lib.tpsph_kernel_time(semi.handle, 0, C.byref(n), C.byref(ms))  # TPSPH_TIME_CELL_LIST
    print(json.dumps({
        "rebuilding_kicks": args.rebuild_kicks,
        "ms_per_kick": elapsed * 1e3 / args.rebuild_kicks,
        "rebuild_chain_ms": ms.value / max(n.value, 1),
        "n_fluid": db.n_fluid, "n_wall": db.n_wall}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not an implementation)",
              file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, REPO)
    from __graft_entry__ import load_package

    tp = load_package()
    native = tp.native
    lib = native.load()
    if args.rebuild_kicks > 0:
        rebuild_kicks(tp, lib, native, args)
        return
    if args.only and args.only != "headline":
        fn = {"moving": section_moving, "precision": section_precision,
              "ensemble": section_ensemble, "adjoint": section_adjoint,
              "sorted": section_sorted}[args.only]
        print(json.dumps({args.only: fn(tp, lib, native, args)}), flush=True)
        return

    def progress(msg):
        print(f"bench: {msg}", file=sys.stderr, flush=True)

    progress("headline")
    head, (db, ode) = section_headline(tp, lib, native, args)
    extra = {}
    if not args.only:
        for name, fn in (("moving_flow", section_moving), ("precision", section_precision),
                         ("ensemble", section_ensemble), ("adjoint", section_adjoint),
                         ("sorted_order", section_sorted)):
            progress(name)
            try:
                extra[name] = fn(tp, lib, native, args)
            except Exception as e:  # a section that cannot run is reported, not hidden
                extra[name] = {"error": f"{type(e).__name__}: {e}"}
    cpu = None
    if not args.no_cpu:
        progress("cpu baseline")
        from oracle import oracle as orc
        semi = tp.Semidiscretization(db.fluid_system, db.boundary_system)
        cpu = cpu_baseline(db, ode, orc.params_from_semi(semi), args.dt)
    out = {
        "metric": "particle-steps/s, 3D WCSPH 1M",
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 pairs / f64 state",
        "data": "synthetic (seeded lattice + uniform jitter of 0.01 dx, hydrostatic density)",
        "config": {
            "workload": "3D WCSPH dam break: 1.0 x 0.5 x 0.5 m water at dx = "
                        f"{args.dx} m in a 3.2 x 1.0 x 1.0 m open tank with 3 dummy wall "
                        "layers; WendlandC2 h = 1.2 dx, Cole EOS (rho0 1000, c = 20 sqrt(g 0.6),"
                        " gamma 7), Monaghan alpha 0.02, Molteni-Colagrossi delta 0.1, Adami "
                        f"walls, CarpenterKennedy2N54 at dt = {args.dt}",
            "n_fluid": head["n_fluid"],
            "n_wall": head["n_wall"],
            "parallelism": "single GPU",
            "state_finite_after_run": head["finite"],
        },
        "roofline": head["roofline"],
        "kernel_ms_per_step": head["kernel_ms_per_step"],
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
