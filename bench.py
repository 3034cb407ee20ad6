"""Benchmark: particle-steps/s of the 3D WCSPH dam break at 1.02M fluid particles.

One step = one CarpenterKennedy2N54 step (5 x kick! + drift! + stage update), run entirely
on the device (tpsph_solve_ck2n54) with the state resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--dx 6.25e-3] [--cpu-steps S]
  python bench.py --rebuild-kicks N    # cost of a rebuild (separate JSON line, no bench)

Roofline of the dominant kernel (the fused fluid interaction): its algorithmic HBM bytes per
launch are the compulsory reads and writes of one evaluation,
  fluid: x, v (8*ndims each), rho, p, m (8 each) read, dv (8*(ndims+1)) written
  wall:  x_w (8*ndims), m_w, p_w, rho_w (8 each) read
divided by the kernel's average duration from HIP events on the library's stream.
The CPU baseline is the FP64 C oracle (OpenMP) on the host cores, on a bounded sample of the
same workload (a few steps), reported as particle-steps/s.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from its committed rocprofv3 summary (FETCH_SIZE +
    WRITE_SIZE, calibrated against a copy of known size), or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(REPO, "profiles", f"{kernel}_*.md"))):
        m = re.search(r"`%s`: fetch [\d.]+ MB \+ write [\d.]+ MB = \*\*([\d.]+) MB\*\* per launch"
                      % re.escape(kernel), open(path).read())
        if m:
            return float(m.group(1)) * 1e6, os.path.relpath(path, REPO)
    return None, None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--dx", type=float, default=6.25e-3)
    ap.add_argument("--dt", type=float, default=1e-5)
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="CK2N54 steps of the oracle in the CPU-baseline sample (0: skip)")
    ap.add_argument("--pair-search", default="auto", choices=["auto", "lists", "cells"])
    ap.add_argument("--rebuild-kicks", type=int, default=0, metavar="N",
                    help="instead of the bench, time N kicks that each rebuild the cell list "
                         "and the Verlet lists")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the run, copy MIB MiB device to device once (a plain copy "
                         "of known size that calibrates FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def rebuild_kicks(tp, lib, semi, v, u, args, nf, nw):
    """Kicks that all rebuild: the positions alternate between u and u shifted by 0.06 dx
    along x, beyond half the default skin, so every kick rebuilds the cell-sorted order, the
    Verlet lists and the brick tables.  Prints one JSON line."""
    import ctypes as C
    import torch
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
        "n_fluid": nf, "n_wall": nw}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is a design, not an implementation)",
              file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, REPO)
    import numpy as np
    import torch
    from __graft_entry__ import load_package

    tp = load_package()
    native = tp.native
    lib = native.load()
    db = tp.dam_break(3, args.dx)
    semi = tp.Semidiscretization(db.fluid_system, db.boundary_system,
                                 pair_search=args.pair_search)
    ode = tp.semidiscretize(semi, (0.0, args.dt * (args.warmup + args.steps)))
    v, u = ode.v0.clone(), ode.u0.clone()
    h = semi.handle
    nf, nw = db.n_fluid, db.n_wall
    if args.rebuild_kicks > 0:
        rebuild_kicks(tp, lib, semi, v, u, args, nf, nw)
        return

    def run(nsteps):
        native.check(lib.tpsph_solve_ck2n54(h, v.data_ptr(), u.data_ptr(), 0.0, args.dt,
                                            nsteps), h)

    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    run(args.warmup)
    torch.cuda.synchronize()
    native.check(lib.tpsph_enable_timing(h, 1), h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    times = {}
    import ctypes as C
    for name, kid in native.TIMED_KERNELS.items():
        n, ms = C.c_int64(), C.c_double()
        native.check(lib.tpsph_kernel_time(h, kid, C.byref(n), C.byref(ms)), h)
        times[name] = (n.value, ms.value)
    native.check(lib.tpsph_enable_timing(h, 0), h)
    vf = v.cpu().numpy()
    finite = bool(np.isfinite(vf).all())

    search = args.pair_search
    if search == "auto":
        search = "lists" if semi.skin != 0.0 else "cells"
    # with Verlet lists and the default cell divisions the fluid kernel stages bricks in LDS
    fluid_kernel = {"lists": "k_fluid_brick", "cells": "k_fluid"}[search]
    nd = 3
    fluid_bytes = nf * (8 * nd + 8 * nd + 8 + 8 + 8 + 8 * (nd + 1))
    wall_bytes = nw * (8 * nd + 8 + 8 + 8)
    traffic, traffic_src = profiled_traffic(fluid_kernel)
    n_launch, fluid_ms = times["fluid"]
    avg_s = fluid_ms / 1e3 / max(n_launch, 1)
    achieved = (fluid_bytes + wall_bytes) / avg_s / 1e9

    cpu = None
    if args.cpu_steps > 0:
        from oracle import oracle as orc
        P = orc.params_from_semi(semi)
        v0, u0 = ode.v0.cpu().numpy(), ode.u0.cpu().numpy()
        xw = db.boundary_system.coordinates.T.copy()
        mw = db.boundary_system.boundary_model.hydrodynamic_mass
        tc = time.perf_counter()
        orc.solve_ck2n54(P, v0, u0, db.fluid_system.mass, xw, mw, args.dt, args.cpu_steps)
        cpu_s = time.perf_counter() - tc
        cpu = {"value": nf * args.cpu_steps / cpu_s, "unit": "particle-steps/s",
               "cores": orc.num_threads(), "kind": "port",
               "sample": f"{args.cpu_steps} CK2N54 steps of the same 3D dam break "
                         f"({nf} fluid + {nw} wall particles), FP64 C oracle with OpenMP, "
                         f"{cpu_s:.2f} s"}

    ms_per_step = elapsed * 1e3 / args.steps
    out = {
        "metric": "particle-steps/s, 3D WCSPH 1M",
        "value": nf * args.steps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lattice + uniform jitter of 0.01 dx, hydrostatic density)",
        "config": {
            "workload": "3D WCSPH dam break: 1.0 x 0.5 x 0.5 m water at dx = "
                        f"{args.dx} m in a 3.2 x 1.0 x 1.0 m open tank with 3 dummy wall "
                        "layers; WendlandC2 h = 1.2 dx, Cole EOS (rho0 1000, c = 20 sqrt(g 0.6),"
                        " gamma 7), Monaghan alpha 0.02, Molteni-Colagrossi delta 0.1, Adami "
                        f"walls, CarpenterKennedy2N54 at dt = {args.dt}",
            "n_fluid": nf,
            "n_wall": nw,
            "parallelism": "single GPU",
            "state_finite_after_run": finite,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": f"{fluid_kernel} (fused fluid-fluid + fluid-wall interaction)",
            "avg_launch_us": avg_s * 1e6,
            "algorithmic_bytes_per_launch": fluid_bytes + wall_bytes,
        },
        "kernel_ms_per_step": {k: ms / args.steps for k, (n, ms) in times.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
