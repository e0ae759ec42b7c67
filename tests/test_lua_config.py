"""The Lua `simulation` configuration reader (shoccs_config_parse): the Lua
subset, the keys of the reference's from_lua readers, and the reference's own
configuration files when the reference tree is present (CPU only)."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

WAVE = """
-- constant-velocity wave through a sphere (this library's extension)
local a = 1 / math.sqrt(3)
n = 2^3^2 // 16          -- 512 // 16 = 32 (^ is right associative)
simulation = {
    mesh = { index_extents = {n, n, n}, domain_bounds = {min = {0, 0, 0}, max = {1, 1, 1}} },
    domain_boundaries = { xmin = "dirichlet", ymin = 'dirichlet', ["zmin"] = "dirichlet" },
    shapes = {
        { type = "sphere", center = {0.5, 0.5, 0.5}, radius = 0.3,
          boundary_condition = "dirichlet" },
    },
    --[[ a long comment
         spanning lines, with "quotes" and {braces} ]]
    scheme = { order = 1, type = "E4u", alpha = {-0.7733323791884821, 0.1623961700641681}; },
    system = { type = "scalar wave", velocity = {a, a, a},
               initial_condition = { center = {0.25, 0.25, 0.25}, width = 0.05, amplitude = 1 },
               max_error = 1e3 },
    integrator = { type = "rk4", form = "fused" },
    step_controller = { max_step = 10, cfl = { hyperbolic = 0.5 } },
    unused = { f = function(t, loc) if t > 0 then return 1 elseif t < 0 then return -1 end return 0 end },
    io = { write_every_step = 1 },
}
simulation.step_controller.min_dt = -2^2 * 1e-7 + 1e-6   -- -(2^2): unary minus binds looser than ^
"""


def _parse(text):
    import shoccs_amd as S

    return S.Config(text)


def test_lua_subset_and_wave_keys():
    import shoccs_amd as S

    c = _parse(WAVE)
    assert c.kind == S.SYS_WAVE_LINEAR
    assert c.n == (32, 32, 32)
    assert c.lo == (0.0, 0.0, 0.0) and c.hi == (1.0, 1.0, 1.0)
    assert c.grid_bc == (S.BC_DIRICHLET, S.BC_FLOATING) * 3
    assert c.spheres == [[0.5, 0.5, 0.5, 0.3]] and c.shape_bc == [S.BC_DIRICHLET]
    assert c.scheme == S.E4U_1 and c.alpha == (-0.7733323791884821, 0.1623961700641681)
    a = 1 / math.sqrt(3)
    assert c.params == [a, a, a, 0.25, 0.25, 0.25, 0.05, 1.0, 1000.0]
    assert c.integrator == S.RK4_FUSED
    assert c.max_step == 10 and c.max_time is None
    assert c.h_cfl == 0.5 and c.p_cfl == 1.0
    assert c.min_dt == pytest.approx(-4e-7 + 1e-6, rel=0, abs=1e-22)


def test_heat_keys_and_defaults():
    import shoccs_amd as S

    c = _parse("""
        dx = 2 / 50
        simulation = {
            mesh = { index_extents = {51, 51}, domain_bounds = {2, 2} },
            domain_boundaries = { xmin = "dirichlet", xmax = "dirichlet", ymin = "dirichlet",
                                  ymax = "dirichlet", zmin = "dirichlet" },
            scheme = { order = 2, type = "E4" },
            system = { type = "heat", diffusivity = 1 / 30 },
            integrator = { type = "euler" },
            step_controller = { max_time = 0.3, cfl = { parabolic = 0.5 } },
            manufactured_solution = { type = "Gaussian",
                { center = {1, 1}, variance = {1, 1}, amplitude = 2, frequency = 0.5 },
                { center = {0.2, 2}, variance = {1 / 3, 0.8}, amplitude = -dx, frequency = 1 } },
        }""")
    assert c.kind == S.SYS_HEAT and c.n == (51, 51, 1)
    assert c.hi == (2.0, 2.0, 0.0)
    # zmin on an axis of extent 1 is taken as floating (boundaries.cpp:17-39)
    assert c.grid_bc == (S.BC_DIRICHLET,) * 4 + (S.BC_FLOATING,) * 2
    assert c.scheme == S.E4_2 and c.integrator == S.EULER
    assert c.max_step is None and c.max_time == 0.3 and c.p_cfl == 0.5 and c.min_dt == 1e-6
    assert c.params[:3] == [1 / 30, S.MMS_GAUSS, 2]
    g = np.array(c.params[3:]).reshape(2, 8)
    np.testing.assert_array_equal(g[1], [0.2, 2, 0, 1 / 3, 0.8, 0, -2 / 50, 1])


@pytest.mark.parametrize("text,msg", [
    ("simulation = { mesh = { index_extents = {8} } }", "domain_bounds"),
    ("simulation = {", "expected"),
    ("x = 1 +", "unexpected"),
    ("simulation = { mesh = { index_extents = {8}, domain_bounds = {1} }, scheme = { order = 1, type = 'E6u' },"
     " system = { type = 'heat' } }", "not built"),
    ("simulation = { mesh = { index_extents = {8}, domain_bounds = {1} }, scheme = { order = 2, type = 'E2' },"
     " system = { type = 'heat' }, manufactured_solution = { type = 'lua', call = function(t, x) return 0 end } }",
     "lua"),
    ("simulation = { mesh = { index_extents = {8}, domain_bounds = {1} }, scheme = { order = 2, type = 'E2' },"
     " system = { type = 'heat' }, integrator = { type = 'rk3' } }", "integrator.type"),
    ("simulation = { mesh = { index_extents = {8}, domain_bounds = {1} }, domain_boundaries = { xmin = 'periodic' },"
     " scheme = { order = 2, type = 'E2' }, system = { type = 'heat' } }", "boundary condition"),
    ("y = math.gamma(2)", "math.gamma"),
])
def test_errors_are_reported(text, msg):
    import shoccs_amd as S

    with pytest.raises(S.ShoccsError, match=msg):
        _parse(text)


def test_syntax_error_names_the_line():
    import shoccs_amd as S

    with pytest.raises(S.ShoccsError, match="line 4: .}. expected"):
        _parse("a = 1\nb = 2\nc = {1, 2\nd = 3")


def test_cli_check(tmp_path):
    path = tmp_path / "wave.lua"
    path.write_text(WAVE)
    r = subprocess.run([sys.executable, "-m", "shoccs_amd", str(path), "--check"], cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "configuration ok" in r.stdout and "32768 points" in r.stdout


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "lua-configs")), reason="reference tree absent")
def test_reference_configs():
    """The reference's own configuration files: the 2-D radial-wave configs
    and heat.lua (floating circle, gaussian solution) read completely; the
    others are refused with the reason (a template placeholder); scalar_wave.lua
    (two circles, Dirichlet and floating, E2-poly) reads completely."""
    import shoccs_amd as S

    c = S.Config.from_file(os.path.join(REF, "lua-configs", "brady_livescu_4_3_n61.lua"))
    assert c.kind == S.SYS_WAVE_RADIAL and c.n == (61, 61, 1)
    assert c.hi == (math.sqrt(2), math.sqrt(2), 0.0)
    assert c.params == [-0.25, -0.25, 0.0, 0.0, 10.0]
    assert c.scheme == S.E4U_1 and c.alpha == (-0.7733323791884821, 0.1623961700641681)
    assert c.grid_bc == (S.BC_DIRICHLET, S.BC_FLOATING, S.BC_DIRICHLET, S.BC_FLOATING, S.BC_FLOATING,
                         S.BC_FLOATING)
    assert c.max_time == 10.0 and c.h_cfl == 0.8 and c.integrator == S.RK4
    S.Config.from_file(os.path.join(REF, "lua-configs", "brady_livescu_4_3_long.lua"))
    with pytest.raises(S.ShoccsError):  # template placeholders are comments: `{, }` has no values
        c = S.Config.from_file(os.path.join(REF, "lua-configs", "brady_livescu_4_3.lua"))
        c.problem(None)
    w = S.Config.from_file(os.path.join(REF, "scalar_wave.lua"))
    assert w.kind == S.SYS_WAVE_RADIAL and w.n == (71, 71, 1) and w.scheme == S.E2P_1
    assert list(w.alpha) == H.E2P_LUA  # floating, Dirichlet, interpolant lists in order
    assert w.shape_bc == [S.BC_DIRICHLET, S.BC_FLOATING]
    assert w.params[:4] == [-1.0, -1.0, 0.0, 0.0]
    p = w.problem(None)  # the floating circle carries cut-point rows
    assert all(len(p.cut_rows(d)["rw"]) > 0 for d in range(2))
    p.close()
    e = S.Config.from_file(os.path.join(REF, "eigenvalues.lua"))  # two yz_rect walls, E2-poly
    assert e.eigenvalues and e.n == (21, 1, 1) and e.scheme == S.E2P_1
    assert [int(sh[0]) for sh in e.shapes] == [S.SHAPE_RECT_X, S.SHAPE_RECT_X]
    assert e.shape_bc == [S.BC_DIRICHLET, S.BC_FLOATING]
    p = e.problem(None)
    assert p.nR == (2, 0, 0)
    p.close()
    with pytest.raises(S.ShoccsError, match="E2"):  # scheme order 1 "E2" (E2_1) is not built
        S.Config.from_file(os.path.join(REF, "scalar_wave_1d.lua"))
    h = S.Config.from_file(os.path.join(REF, "heat.lua"))
    assert h.kind == S.SYS_HEAT and h.shape_bc == [S.BC_FLOATING]
    assert h.params[:3] == [1 / 30, S.MMS_GAUSS, 1]  # the commented-out gaussians are skipped
    p = h.problem(None)  # a floating circle: cut-point rows in every direction
    assert sum(p.nR) > 0 and all(len(p.cut_rows(d)["rw"]) > 0 for d in range(2))
    p.close()
