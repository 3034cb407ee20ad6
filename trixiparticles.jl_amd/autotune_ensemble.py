"""Pick the launch shapes of the ensemble runner on this GPU (see ensemble.autotune).

  python trixiparticles.jl_amd/autotune_ensemble.py [--cases 256] [--out PATH]

Times every cell tiling (cells per compact support of the cases' grids) and, for each, every
block shape of the packed fluid and Adami passes (ensemble.autotune) on the seeded 256-case
dam-break ensemble, and writes the fastest as JSON to PATH (default:
tuning/ensemble_gfx950.json beside this script), where dam_break_ensemble and Ensemble pick
it up by default.  No choice changes which pairs interact; the tiling changes the order of
each case's sums (each case still equals itself run alone on the same tiling)."""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

# (cell_divisions, x_cell_divisions): cells per compact support along y and along x
TILINGS = ((1, 2), (2, 2), (2, 4), (3, 3))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", type=int, default=256)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning", "ensemble_gfx950.json"))
    args = ap.parse_args()
    from __graft_entry__ import load_package
    tp = load_package()
    from trixiparticles_jl_amd import ensemble as en
    runs = []
    for cd, xd in TILINGS:
        ens = en.Ensemble(en.dam_break_ensemble(args.cases, cell_divisions=cd,
                                                x_cell_divisions=xd), tuning=None)
        res = en.autotune(ens, write=False)
        res.update({"cell_divisions": cd, "x_cell_divisions": xd})
        print(json.dumps({k: res[k] for k in ("cell_divisions", "x_cell_divisions",
                                               "fluid_block", "wall_block", "ms_per_step")}),
              flush=True)
        runs.append(res)
        ens.close()
    choice = min(runs, key=lambda r: r["ms_per_step"])
    out = {k: choice[k] for k in ("arch", "cell_divisions", "x_cell_divisions", "fluid_block",
                                  "wall_block", "ms_per_step", "workload")}
    out["tilings"] = [{k: r[k] for k in ("cell_divisions", "x_cell_divisions", "fluid_block",
                                         "wall_block", "ms_per_step", "candidates")}
                      for r in runs]
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("cell_divisions", "x_cell_divisions", "fluid_block",
                                          "wall_block", "ms_per_step")}))
    del tp


if __name__ == "__main__":
    main()
