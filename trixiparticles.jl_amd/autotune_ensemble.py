"""Pick the launch shapes of the ensemble runner on this GPU (see ensemble.autotune).

  python trixiparticles.jl_amd/autotune_ensemble.py [--cases 256] [--out PATH]

Times every shape on the seeded 256-case dam-break ensemble and writes the choice as JSON
to PATH (default: tuning/ensemble_gfx950.json beside this script), where Ensemble picks it
up by default."""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", type=int, default=256)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning", "ensemble_gfx950.json"))
    args = ap.parse_args()
    from __graft_entry__ import load_package
    tp = load_package()
    from trixiparticles_jl_amd import ensemble as en
    ens = en.Ensemble(en.dam_break_ensemble(args.cases), tuning=None)
    choice = en.autotune(ens, write=False)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(choice, f, indent=1)
    print(json.dumps({k: choice[k] for k in ("fluid_block", "wall_block", "ms_per_step")}))
    ens.close()
    del tp


if __name__ == "__main__":
    main()
