"""Write tests/golden/*.json from the golden vectors held by the reference's
own unit tests (read as text; nothing of the reference is built or run).

Run in a container where the reference tree is mounted:
    python tests/golden/make_fixtures.py /root/reference
The JSON files it writes are committed; the tests read only them.
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

NUM = r"[-+]?(?:\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?)"


def _numbers(text: str):
    return [float(x) for x in re.findall(NUM, text)]


def stencil_cases(ref: str):
    """(scheme, alpha, h, bc, psi, right, expected) from `st.nbs(...)` calls
    followed by `Approx(T{...})` / `Approx(std::vector{...})` matchers."""
    out = []
    specs = [("E4_1.t.cpp", "E4_1"), ("E4u_1.t.cpp", "E4u_1"), ("E2_2.t.cpp", "E2_2")]
    for fname, scheme in specs:
        path = os.path.join(ref, "src", "stencils", fname)
        src = open(path).read()
        m = re.search(r"alpha\s*=\s*\{([^}]*)\}", src)
        alpha = _numbers(m.group(1)) if m else [0.0, 0.0]
        pat = re.compile(
            r"st\.nbs\(\s*(" + NUM + r")\s*,\s*bcs::(\w+)\s*,\s*(" + NUM + r")\s*,\s*(true|false)"
            r"[^;]*;\s*REQUIRE_THAT\(\s*\w+\s*,\s*Approx\((?:T|std::vector)\s*\{([^}]*)\}")
        for mm in pat.finditer(src):
            line = src[: mm.start()].count("\n") + 1
            out.append(dict(scheme=scheme, alpha=alpha, h=float(mm.group(1)), bc=mm.group(2),
                            psi=float(mm.group(3)), right=mm.group(4) == "true",
                            expected=_numbers(mm.group(5)),
                            source=f"src/stencils/{fname}:{line}"))
    return out


def _lua_numbers(text: str):
    """Numbers of a Lua table body whose entries may be fractions (3/25)."""
    out = []
    for tok in text.split(","):
        tok = tok.strip()
        if not tok:
            continue
        if "/" in tok:
            a, b = tok.split("/")
            out.append(float(a) / float(b))
        else:
            out.append(float(tok))
    return out


def polye2_cases(ref: str):
    """The E2-poly Dirichlet closure vector of src/stencils/polyE2_1.t.cpp
    ("dirichlet coeffs"): parameters, h, psi and the expected 2 x 4 block."""
    fname = "polyE2_1.t.cpp"
    src = open(os.path.join(ref, "src", "stencils", fname)).read()
    start = src.index('TEST_CASE("dirichlet coeffs")')
    block = src[start:src.index("TEST_CASE", start + 10)]
    da = _lua_numbers(re.search(r"dirichlet_alpha\s*=\s*\{([^}]*)\}", block).group(1))
    fa = _lua_numbers(re.search(r"floating_alpha\s*=\s*\{([^}]*)\}", block).group(1))
    h = float(re.search(r"real h\s*=\s*(" + NUM + ")", block).group(1))
    psi = float(re.search(r"real psi\s*=\s*(" + NUM + ")", block).group(1))
    exact = _numbers(re.search(r"T exact\{([^}]*)\}", block).group(1))
    line = src[:start].count("\n") + 1
    alpha = (fa + [0.0] * 6)[:6] + (da + [0.0] * 3)[:3] + [0.0] * 4
    return [dict(scheme="E2P_1", alpha=alpha, h=h, bc="Dirichlet", psi=psi, right=False, expected=exact,
                 source=f"src/stencils/{fname}:{line}")]


def mesh_cases():
    """Line extents, R indices and psi of the reference's mesh test
    (src/mesh/mesh.t.cpp:67-230): extents {21, 22, 23}, bounds
    {-1, -1, 0}..{1, 2, 2.2}, sphere centre {0.01, -0.01, 0.5} radius 0.25."""
    return dict(
        extents=[21, 22, 23], lo=[-1, -1, 0], hi=[1, 2, 2.2], sphere=[0.01, -0.01, 0.5, 0.25],
        source="src/mesh/mesh.t.cpp:67-230",
        n_intersections=[26, 42, 28],
        lines=[
            # dir, line index, start mesh coord, start obj (R index or -1), start psi,
            # end mesh coord, end obj, end psi
            dict(dir=0, index=6 * 23 + 3, start=[0, 6, 3], start_obj=-1, end=[10, 6, 3], end_obj=0,
                 end_psi=0.40365385103120377),
            dict(dir=0, index=6 * 23 + 4, start=[10, 6, 3], start_obj=1,
                 start_psi=0.2036538510312047, end=[20, 6, 3], end_obj=-1),
            dict(dir=1, index=8 * 23 + 4, start=[8, 0, 4], start_obj=-1, end=[8, 7, 4], end_obj=0,
                 end_psi=0.2884394027061823),
            dict(dir=1, index=8 * 23 + 5, start=[8, 7, 4], start_obj=1,
                 start_psi=0.42843940270618086, end=[8, 21, 4], end_obj=-1),
            dict(dir=2, index=12 * 22 + 8 + 13, start=[12, 8, 0], start_obj=-1, end=[12, 8, 5],
                 end_obj=26, end_psi=0.4491194432954615),
            dict(dir=0, index=0, start=[0, 0, 0], start_obj=-1, end=[20, 0, 0], end_obj=-1),
            dict(dir=1, index=1, start=[0, 0, 1], start_obj=-1, end=[0, 21, 1], end_obj=-1),
            dict(dir=2, index=-1, start=[20, 21, 0], start_obj=-1, end=[20, 21, 22], end_obj=-1),
        ],
        strides=[22 * 23, 23, 1],
    )


def sphere_geometry(ref: str):
    """object_geometry.t.cpp "sphere intersections": per direction the cut
    points' positions, psi and solid coordinates (in the reference's R
    order) and the solid points of that direction's lines (its S order),
    from the `exact`, `sc`, `psi` and `sx`/`sy`/`sz` initializers of each
    SECTION.  The mesh and sphere are those of the test's configuration."""
    src = open(os.path.join(ref, "src", "mesh", "object_geometry.t.cpp")).read()
    body = src[src.index('TEST_CASE("sphere intersections")'):src.index('TEST_CASE("rect_intersections")')]
    out = dict(n=[21, 22, 23], lo=[-1.0, -1.0, 0.0], hi=[1.0, 2.0, 2.2], center=[0.01, -0.01, 0.5],
               radius=0.25, dirs=[])
    names = ["X", "y", "z"]
    for d, name in enumerate(names):
        start = body.index(f'SECTION("{name}")')
        end = body.index(f'SECTION("{names[d + 1]}")') if d < 2 else len(body)
        sec = body[start:end]

        def init(var):
            m = re.search(r"std::vector<\w+>\s+" + var + r"\s*=\s*\{(.*?)\};", sec, re.S)
            return _numbers(m.group(1))

        pos = init("exact")
        sc = init("sc")
        psi = init("psi")
        solid = init("s" + "xyz"[d])
        out["dirs"].append(dict(
            pos=[pos[i:i + 3] for i in range(0, len(pos), 3)],
            solid_coord=[[int(v) for v in sc[i:i + 3]] for i in range(0, len(sc), 3)],
            psi=psi,
            solid=[[int(v) for v in solid[i:i + 3]] for i in range(0, len(solid), 3)]))
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    stencils = stencil_cases(ref) + polye2_cases(ref)
    with open(os.path.join(HERE, "stencil_closures.json"), "w") as f:
        json.dump(stencils, f, indent=1)
    with open(os.path.join(HERE, "mesh_lines.json"), "w") as f:
        json.dump(mesh_cases(), f, indent=1)
    geo = sphere_geometry(ref)
    with open(os.path.join(HERE, "object_geometry_sphere.json"), "w") as f:
        json.dump(geo, f)
    print(f"{len(stencils)} stencil cases written; sphere geometry: "
          + ", ".join(f"{len(d['pos'])} cut points / {len(d['solid'])} solid points" for d in geo["dirs"]))


if __name__ == "__main__":
    main()
