"""Extract the expected tank coordinates held by the reference's RectangularTank tests
(test/setups/rectangular_tank.jl) into ``rectangular_tank.json``.

The reference test file is read as text: the matrix literals it holds are data (expected
coordinates), and the setup parameters of each case are restated below from the same file
(line numbers cited per case).  Nothing of the reference is executed.  Run this script in a
checkout where the reference sits at /root/reference; the JSON it writes is committed.

    python tests/golden/make_rectangular_tank_fixtures.py [path/to/rectangular_tank.jl]
"""
import json
import os
import sys

SRC = "/root/reference/test/setups/rectangular_tank.jl"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rectangular_tank.json")


def strip_comments(text):
    return "\n".join(line.split("#")[0] for line in text.splitlines())


def _skip_ws(s, i):
    while i < len(s) and s[i] in " \t\r\n":
        i += 1
    return i


def parse_value(s, i):
    """Parse a Julia literal at s[i]: a matrix ``[a b; c d]`` (rows split by ';' or newline)
    or a vector of such values ``[v1, v2, ...]``.  Returns (value, index after it)."""
    i = _skip_ws(s, i)
    assert s[i] == "[", s[i:i + 40]
    j = _skip_ws(s, i + 1)
    if s[j] == "[":  # vector of nested values
        out = []
        i = j
        while True:
            v, i = parse_value(s, i)
            out.append(v)
            i = _skip_ws(s, i)
            if s[i] == ",":
                i = _skip_ws(s, i + 1)
                if s[i] == "]":
                    return out, i + 1
                continue
            assert s[i] == "]", s[i:i + 40]
            return out, i + 1
    end = s.index("]", i)
    body = s[i + 1:end]
    rows = []
    for chunk in body.replace(";", "\n").splitlines():
        toks = chunk.split()
        if toks:
            rows.append([float(t) for t in toks])
    assert rows and all(len(r) == len(rows[0]) for r in rows), body[:80]
    return rows, end + 1


def extract(s, name, start):
    """Value assigned to ``name`` at its first assignment after ``start``."""
    key = name + " = "
    k = s.index(key, start)
    v, end = parse_value(s, k + len(key))
    return v, end


def main(path=SRC):
    text = strip_comments(open(path).read())
    cut = text.index('"Rectangular Tank 3D"')
    blocks = {2: text[:cut], 3: text[cut:]}
    cases = []

    for nd, s in blocks.items():
        # "One Boundary Layer" / "Coordinates" (2D :2-60, 3D :254-338)
        p = s.index('"One Boundary Layer"')
        fluid, q = extract(s, "expected_fluid_coords", p)
        bound, _ = extract(s, "expected_bound_coords", q)
        if nd == 2:
            water, tank, mins = (0.5, 0.4), (0.5, 0.5), [(0.0, 0.0), (-0.3, 2.0)]
        else:
            water, tank = (0.4, 0.4, 0.4), (0.4, 0.4, 0.4)
            mins = [(0.0, 0.0, 0.0), (-0.3, 2.0, -0.5)]
        for m in mins:
            for j, spacing in enumerate((0.1, 0.2)):
                for k, ratio in enumerate((1, 3)):
                    cases.append(dict(
                        name=f"{nd}D one layer spacing={spacing} ratio={ratio} min={m}",
                        particle_spacing=spacing, fluid_size=water, tank_size=tank,
                        fluid_density=1000.0, n_layers=1, spacing_ratio=ratio,
                        min_coordinates=m, faces=None,
                        fluid=[[x + m[d] for x in row] for d, row in enumerate(fluid[j])],
                        boundary=[[x + m[d] for x in row]
                                  for d, row in enumerate(bound[j][k])]))

        # "Multiple Boundary Layers" / "Coordinates" (2D :122-141, 3D :370-386)
        p = s.index('"Multiple Boundary Layers"')
        bound, _ = extract(s, "expected_bound_coords", p)
        cases.append(dict(
            name=f"{nd}D three layers spacing=0.5", particle_spacing=0.5,
            fluid_size=(0.5,) * nd, tank_size=(1.0,) * nd, fluid_density=500.0, n_layers=3,
            spacing_ratio=1, min_coordinates=None, faces=None, fluid=None, boundary=bound))

        # "Boundary Faces" (2D :180-251, 3D :416-467)
        p = s.index('"Boundary Faces"')
        if nd == 2:
            faces = [(False, False, False, True), (False, False, True, False),
                     (False, False, True, True), (False, True, False, False),
                     (False, True, False, True), (False, True, True, False),
                     (False, True, True, True), (True, False, False, False),
                     (True, False, False, True), (True, False, True, False),
                     (True, False, True, True), (True, True, False, False),
                     (True, True, False, True), (True, True, True, False)]
        else:
            faces = [(False, False, True, False, False, True),
                     (True, True, True, True, False, False)]
        b1, q = extract(s, "expected_bound_coords1", p)
        b2, _ = extract(s, "expected_bound_coords2", q)
        for nl, bl in ((1, b1), (3, b2)):
            assert len(bl) == len(faces)
            for f, b in zip(faces, bl):
                cases.append(dict(
                    name=f"{nd}D faces={f} layers={nl}", particle_spacing=0.2,
                    fluid_size=(0.6,) * nd, tank_size=(0.8,) * nd, fluid_density=750.5,
                    n_layers=nl, spacing_ratio=1, min_coordinates=None, faces=f, fluid=None,
                    boundary=b))

    with open(OUT, "w") as fh:
        json.dump({"source": "test/setups/rectangular_tank.jl", "cases": cases}, fh)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main(*sys.argv[1:])
