"""Field output (shoccs_io_*): the dump interval (src/io/interval.t.cpp's
cases), the raw data and geometry files (src/io/field_data.cpp), the XDMF3
document (src/io/xdmf.cpp, the reference's `header` test setup) and the
Error field (scalar_system_utils.hpp:224-301).  Host only."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import helpers as H
import oracle_ffi as O


def _prob(cfg):
    # the mesh and cut points are all the output needs (extents as small as
    # 3 leave no room for the closures of any scheme)
    return H.product_problem(None, cfg, mesh_only=True)


def _fires(io, prob, seq):
    """write a zero field at each (step, time, dt); which ones were written"""
    z = np.zeros(prob.flat_size)
    return [io.write({"U": z}, s, t, dt) for s, t, dt in seq]


def test_interval_never_fires(tmp_path):
    """interval.t.cpp "Never fire": no interval, nothing written (step 0
    included)."""
    import shoccs_amd as S

    prob = _prob(H.heat3d_cfg(n=(6, 7, 8)))
    io = S.FieldIO(prob, dir=str(tmp_path / "io"))
    assert not any(_fires(io, prob, [(i, 0.1 * i, 0.1) for i in range(20)]))
    assert not (tmp_path / "io").exists()


def test_interval_time_no_rollover(tmp_path):
    """interval.t.cpp "No rollover": distance 1.5, values i/2 with tolerance
    1e-6 fire at every third i (and at step 0)."""
    import shoccs_amd as S

    prob = _prob(H.heat3d_cfg(n=(6, 7, 8)))
    io = S.FieldIO(prob, dir=str(tmp_path / "io"), write_every_time=1.5)
    got = _fires(io, prob, [(i, 0.5 * i, 1e-6) for i in range(50)])
    assert got == [i == 0 or i % 3 == 0 for i in range(50)]


def test_interval_time_rollover(tmp_path):
    """interval.t.cpp "Rollover": distance 4.5, integer values with tolerance
    1e-6 fire first at 5, last at 45, ten times."""
    import shoccs_amd as S

    prob = _prob(H.heat3d_cfg(n=(6, 7, 8)))
    io = S.FieldIO(prob, dir=str(tmp_path / "io"), write_every_time=4.5)
    got = _fires(io, prob, [(i + 1, float(i), 1e-6) for i in range(46)])
    fired = [i for i, g in enumerate(got) if g]
    assert fired[0] == 5 and fired[-1] == 45 and len(fired) == 10


def test_interval_steps(tmp_path):
    import shoccs_amd as S

    prob = _prob(H.heat3d_cfg(n=(6, 7, 8)))
    io = S.FieldIO(prob, dir=str(tmp_path / "io"), write_every_step=3, suffix_length=4)
    got = _fires(io, prob, [(i, 0.01 * i, 0.01) for i in range(11)])
    assert got == [i % 3 == 0 for i in range(11)]
    assert sorted(os.listdir(tmp_path / "io")) == ["U.0000", "U.0001", "U.0002", "U.0003", "rx", "ry", "rz",
                                                  "view.xmf"]


@pytest.mark.parametrize("cfg", [H.sphere_wave_cfg(24, scheme=O.E4U_1), H.heat2d_cfg(31)],
                         ids=["3d", "2d"])
def test_data_and_geometry_files(tmp_path, cfg):
    """field_data::write: D then Rx, Ry, Rz, native FP64, one file per
    variable and dump; write_geom: the cut-point positions, (z, y, x) on a
    mesh with one z point."""
    import shoccs_amd as S

    prob = _prob(cfg)
    io = S.FieldIO(prob, dir=str(tmp_path), write_every_step=1)
    u = np.arange(prob.flat_size, dtype=np.float64) * 0.25
    e = -u
    assert io.write({"U": u, "Error": e}, 0, 0.0, 0.0)
    assert io.write({"U": 2 * u, "Error": e}, 1, 0.1, 0.1)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "U.000000"), u)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "U.000001"), 2 * u)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "Error.000000"), e)
    for d, name in enumerate(("rx", "ry", "rz")):
        pos = prob.r_arrays(d)["pos"]
        if prob.n[2] == 1:
            pos = pos[:, ::-1]
        np.testing.assert_array_equal(np.fromfile(tmp_path / name).reshape(-1, 3), pos)


def _short(x: float) -> str:
    """the shortest round-trip form, as fmt prints a double"""
    r = repr(float(x))
    return r[:-2] if r.endswith(".0") else r


def test_xdmf_document(tmp_path):
    """xdmf.cpp: the reference's `header` test setup (extents 3 4 5, bounds
    (0.1, 0.2, 0.3)-(1.1, 1.2, 1.3), variables U and V, five-digit suffixes,
    two writes), then a mesh with cut points (seek offsets of RX, RY, RZ)."""
    import shoccs_amd as S

    cfg = dict(H.heat3d_cfg(n=(3, 4, 5)), lo=(0.1, 0.2, 0.3), hi=(1.1, 1.2, 1.3), spheres=[], obj_bc=[])
    prob = _prob(cfg)
    io = S.FieldIO(prob, dir=str(tmp_path), write_every_step=1, suffix_length=5, xdmf_filename="headertest.xmf")
    z = np.zeros(prob.flat_size)
    io.write({"U": z, "V": z}, 0, 0.0, 0.0)
    io.write({"U": z, "V": z}, 1, 0.1, 0.1)
    text = (tmp_path / "headertest.xmf").read_text()
    assert text.startswith('<?xml version="1.0" encoding="utf-8"?>\n<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>\n')
    root = ET.fromstring(text.split("\n", 2)[2])
    assert root.tag == "Xdmf" and root.get("Version") == "3.0"
    dom = root.find("Domain")
    topo, geom = dom.findall("Topology"), dom.findall("Geometry")
    assert topo[0].get("TopologyType") == "3DCoRectMesh" and topo[0].get("Dimensions") == "3 4 5"
    d = [(1.1 - 0.1) / 2, (1.2 - 0.2) / 3, (1.3 - 0.3) / 4]
    items = geom[0].findall("DataItem")
    assert items[0].text == "0.1 0.2 0.3" and items[1].text == " ".join(_short(v) for v in d)
    for a in range(3):
        assert topo[1 + a].get("TopologyType") == "Polyvertex" and topo[1 + a].get("NumberOfElements") == "0"
        assert geom[1 + a].find("DataItem").text == ("rx", "ry", "rz")[a]
    series = dom.find("Grid")
    assert series.get("Name") == "TimeSeries" and series.get("CollectionType") == "Temporal"
    grids = series.findall("Grid")
    assert [g.get("Name") for g in grids] == ["0", "1"]
    assert grids[1].find("Time").get("Value") == "0.10000000000000001"  # %.17g
    sub = grids[1].findall("Grid")
    assert [g.get("Name") for g in sub] == ["F"]  # no cut points: D only
    attrs = sub[0].findall("Attribute")
    assert [a.get("Name") for a in attrs] == ["U", "V"]
    assert [a.find("DataItem").text for a in attrs] == ["U.00001", "V.00001"]
    assert sub[0].find("Topology").get("Reference") == "/Xdmf/Domain/Topology[1]"

    # with cut points: RX, RY, RZ grids seek past D and the sets before them
    cfg = H.sphere_wave_cfg(20, scheme=O.E4U_1)
    prob = _prob(cfg)
    io = S.FieldIO(prob, dir=str(tmp_path / "b"), write_every_step=1)
    io.write({"U": np.zeros(prob.flat_size)}, 0, 0.0, 0.0)
    root = ET.fromstring((tmp_path / "b" / "view.xmf").read_text().split("\n", 2)[2])
    sub = root.find("Domain").find("Grid").find("Grid").findall("Grid")
    assert [g.get("Name") for g in sub] == ["F", "RX", "RY", "RZ"]
    seeks = [g.find("Attribute").find("DataItem").get("Seek") for g in sub]
    N, nR = prob.N, prob.nR
    assert seeks == [None, str(8 * N), str(8 * (N + nR[0])), str(8 * (N + nR[0] + nR[1]))]
    assert sub[2].find("Attribute").find("DataItem").get("Dimensions") == f"{nR[1]} 1"
    assert sub[3].find("Geometry").get("Reference") == "/Xdmf/Domain/Geometry[4]"
    counts = [t.get("NumberOfElements") for t in root.find("Domain").findall("Topology")[1:]]
    assert counts == [str(v) for v in nR]


def test_error_field():
    """write_scalar_error: |u - sol| at fluid points and at the cut points of
    non-Dirichlet bodies; zero elsewhere, on Dirichlet grid faces and at the
    cut points of Dirichlet bodies."""
    cfg = H.cut_cfgs()["two_spheres_E2"]  # one Dirichlet, one floating sphere; Dirichlet y faces
    prob = _prob(cfg)
    rng = np.random.default_rng(0)
    u, sol = rng.standard_normal(prob.flat_size), rng.standard_normal(prob.flat_size)
    err = prob.error_field(u, sol)
    want = np.zeros(prob.flat_size)
    f = prob.fluid()
    want[f] = np.abs(u[f] - sol[f])
    n = prob.n
    D = want[:prob.N].reshape(n)
    D[:, 0, :] = 0.0
    D[:, -1, :] = 0.0
    off = prob.N
    for d in range(3):
        sid = prob.r_arrays(d)["shape_id"]
        for i in range(prob.nR[d]):
            if cfg["obj_bc"][sid[i]] != O.DIRICHLET:
                want[off + i] = abs(u[off + i] - sol[off + i])
        off += prob.nR[d]
    np.testing.assert_array_equal(err, want)
    assert np.count_nonzero(err[prob.N:]) > 0


def test_lua_io_keys():
    import shoccs_amd as S

    base = ("simulation = { mesh = { index_extents = {8, 8}, domain_bounds = {1, 1} }, scheme = { order = 2, "
            "type = 'E2' }, system = { type = 'heat' }, step_controller = { max_step = 3 }")
    assert S.Config(base + " }").io is None
    assert S.Config(base + ", io = { dir = 'out' } }").io is None  # no interval: no output
    c = S.Config(base + ", io = { write_every_step = 2, dir = 'out', suffix_length = 4, "
                        "xdmf_filename = 'a.xmf' } }")
    assert c.io == dict(dir="out", xdmf_filename="a.xmf", suffix_length=4, write_every_step=2,
                        write_every_time=None)
    c = S.Config(base + ", io = { write_every_time = 0.25 } }")
    assert c.io["dir"] == "io" and c.io["write_every_time"] == 0.25 and c.io["suffix_length"] == 6
