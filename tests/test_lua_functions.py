"""Lua functions in the configuration reader (CPU): the manufactured solutions
of the reference's own tests, written in Lua there (heat.t.cpp:79-101,
:182-204, :294-312; rk4_v2.t.cpp:69-93; simulation_cycle.t.cpp:62-83), read
by shoccs_config_parse and called through the shx_mms_fns it hands to the
problem.  Each value equals the same formula evaluated in Python with the
same operations in the same order, bit for bit."""
import ctypes as C
import math

import numpy as np
import pytest

import shoccs_amd as S
import test_lua_mms as T

HEAT_E2 = """
simulation = {
    mesh = { index_extents = {21, 22, 23}, domain_bounds = { min = {1, 1.1, 0.3}, max = {3, 3.3, 2.2} } },
    domain_boundaries = { xmin = "dirichlet", ymin = "neumann", ymax = "neumann", zmax = "dirichlet" },
    shapes = { { type = "sphere", center = {2.0001, 2.5656565, 1.313131311}, radius = 0.25,
                 boundary_condition = "%s" } },
    scheme = { order = 2, type = "E2" },
    system = { type = "heat", diffusivity = 0.1 },
    manufactured_solution = {
        type = "lua",
        %s
    }
}
"""

CUBIC = """
        call = function(time, loc)
            local x, y, z = loc[1], loc[2], loc[3]
            return (math.sin(time) +
                x * x * (y + z) + y * y * (x + z) + z * z * (x + y) +
                3 * x * y * z + x + y + z)
        end,
        ddt = function(time, loc)
            return math.cos(time)
        end,
        grad = function(time, loc)
            local x, y, z = loc[1], loc[2], loc[3]
            return 2. * x * (y + z) + y * y + z * z + 3. * y * z + 1,
                    x * x + 2. * y * (x + z) + z * z + 3. * x * z + 1,
                    x * x + y * y + 2. * z * (x + y) + 3. * x * y + 1
        end,
        lap = function(time, loc)
            local x, y, z = loc[1], loc[2], loc[3]
            return 2. * (y + z) + 2. * (x + z) + 2. * (x + y)
        end,
        div = function(time, loc)
            return 0.0
        end
"""

BILINEAR = """
        call = function(time, loc)
            local x, y, z = loc[1], loc[2], loc[3]
            return (math.sin(time) +
                x * (y + z) + y * (x + z) + z * (x + y) +
                x + y + z)
        end,
        ddt = function(time, loc)
            return math.cos(time)
        end,
        grad = function(time, loc)
            local x, y, z = loc[1], loc[2], loc[3]
            return 2. * (y + z) + 1,
                   2. * (x + z) + 1,
                   2. * (x + y) + 1
        end,
        lap = function(time, loc)
            local x, y, z = loc[1], loc[2], loc[3]
            return 0.0
        end,
        div = function(time, loc)
            return 0.0
        end
"""


def _fns(cfg):
    assert cfg.mms is not None
    return cfg.mms.contents


def _values(f, t, loc):
    p = (C.c_double * 3)(*loc)
    g = (C.c_double * 3)()
    f.grad(f.ctx, t, p, g)
    return f.call(f.ctx, t, p), f.ddt(f.ctx, t, p), tuple(g), f.lap(f.ctx, t, p)


@pytest.mark.parametrize("body,bc,py", [(CUBIC, "dirichlet", T.cubic(math.sin, math.cos)),
                                        (BILINEAR, "floating", T.bilinear3())])
def test_lua_functions_match_their_formulas(body, bc, py):
    cfg = S.Config(HEAT_E2 % (bc, body))
    assert cfg.params[:2] == [0.1, S.MMS_HOST]
    f = _fns(cfg)
    rng = np.random.default_rng(3)
    for _ in range(50):
        t = float(rng.uniform(0, 2))
        loc = tuple(float(v) for v in rng.uniform(0.5, 3.5, 3))
        got = _values(f, t, loc)
        want = (py[0](t, loc), py[1](t, loc), tuple(float(v) for v in py[2](t, loc)), py[3](t, loc))
        assert got == want
    assert cfg.mms_error() == ""


def test_lua_function_statements():
    """Locals, if/elseif/else, nested calls, upvalues and function statements."""
    text = """
local scale = 2
local function sq(v) return v * v end
function pick(v)
    if v < 0 then
        return -1
    elseif v == 0 then
        return 0
    else
        local r = sq(v) * scale
        return r
    end
end
simulation = {
    mesh = { index_extents = {21, 22}, domain_bounds = { min = {1, 1.1}, max = {3, 3.3} } },
    scheme = { order = 2, type = "E2" },
    system = { type = "heat", diffusivity = 1.0 },
    manufactured_solution = {
        type = "lua",
        call = function(t, loc) return pick(loc[1] - 2) + t end,
        ddt = function(t, loc) return 1 end,
        grad = function(t, loc) return pick(loc[2] - 2), sq(3), math.max(loc[1], loc[2]) end,
        lap = function(t, loc) if loc[3] > 0 then return 1 end return 2 end,
        div = function(t, loc) return 0 end
    }
}
"""
    cfg = S.Config(text)  # the functions live in the configuration
    f = _fns(cfg)
    c, d, g, lap = _values(f, 0.5, (3.5, 1.5, 0.0))
    assert (c, d, g, lap) == (0.5 + 2 * 1.5 * 1.5, 1.0, (-1.0, 9.0, 3.5), 2.0)
    c, _, g, lap = _values(f, 0.0, (2.0, 2.0, 1.0))
    assert (c, g, lap) == (0.0, (0.0, 9.0, 2.0), 1.0)
    c, *_ = _values(f, 0.0, (1.0, 2.0, 1.0))
    assert c == -1.0


def test_lua_function_errors_give_nan_and_a_message():
    text = (HEAT_E2 % ("dirichlet", CUBIC)).replace("return math.cos(time)", "return nothing.field")
    cfg = S.Config(text)
    f = _fns(cfg)
    _, d, _, _ = _values(f, 0.0, (1.0, 2.0, 1.0))
    assert math.isnan(d)
    assert "non-table" in cfg.mms_error()


def test_lua_mms_requires_every_function():
    text = (HEAT_E2 % ("dirichlet", CUBIC)).replace("div = function", "xdiv = function")
    with pytest.raises(S.ShoccsError, match="requires call, ddt, grad, div, lap"):
        S.Config(text)
