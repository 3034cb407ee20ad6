"""Copy the surge-front series of the reference's 2D dam-break validation run into
``dam_break_2d_wcsph_40.json``.

The source, ``validation/dam_break_2d/validation_reference_wcsph_40.json``, is a data file the
reference's validation test compares against (``test/validation/validation.jl:48-68``): the
``max_x_coord`` of the fluid (the surge front) and the four wall pressure sensors P1-P4
(``interpolate_line`` averages, ``validation/dam_break_2d/sensors.jl``) every 0.01 / sqrt(g /
H) s of the setup in ``validation/dam_break_2d/setup_marrone_2011.jl`` with
``examples/fluid/dam_break_2d.jl`` (40 particles per height).  The run parameters recorded in its ``meta`` block are kept with
the series.  Nothing of the reference is executed.

    python tests/golden/make_dam_break_validation_fixture.py [path/to/validation_reference_wcsph_40.json]
"""
import json
import os
import sys

SRC = "/root/reference/validation/dam_break_2d/validation_reference_wcsph_40.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dam_break_2d_wcsph_40.json")
N_SAMPLES = 701  # the whole series, t up to 1.73 s (the front reaches the far wall at 0.6 s)


def main(src=SRC):
    d = json.load(open(src))
    front = d["max_x_coord_fluid_1"]
    out = {
        "source": "validation/dam_break_2d/validation_reference_wcsph_40.json",
        "series": "max_x_coord_fluid_1",
        "time": front["time"][:N_SAMPLES],
        "values": front["values"][:N_SAMPLES],
        "wall_max_x": d["max_x_coord_boundary_1"]["values"][0],
        "pressure": {k: d[f"pressure_{k}_fluid_1"]["values"][:N_SAMPLES]
                     for k in ("P1", "P2", "P3", "P4")},
        "meta": {
            "time_integrator": d["meta"]["simulation_info"]["time_integrator"],
            "fluid": d["meta"]["system_data"]["fluid_1"],
            "boundary": d["meta"]["system_data"]["boundary_1"],
        },
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT, len(out["time"]), "samples")


if __name__ == "__main__":
    main(*sys.argv[1:])
