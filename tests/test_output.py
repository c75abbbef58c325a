"""VTU/PVTU/PVD output (source/PostProcessor.cc:83-227): the patches carry
the FE field at their points (a field that FE_Q(p) holds exactly must come
out exactly), zero on inactive cells, the state ratios per cell, the
reference's file names and a time series."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from adamantine_amd.output import PostProcessor, read_vtu
from adamantine_amd.problem import Problem


def field(x):
    # degree <= 2 in each direction: exact in FE_Q(2)
    return 300.0 + 5.0 * x[:, 0] ** 2 + 3.0 * x[:, 1] * x[:, 2] - x[:, 2]


@pytest.mark.parametrize("deg,refine", [(1, 0), (2, 0), (2, 3), (3, 2)])
def test_vtu_patches(tmp_path, deg, refine):
    p = Problem(n_cells=(3, 2, 2), extent=(1.5, 1.0, 0.5),
                origin=(0.1, -0.2, 0.0), fe_degree=deg)
    X = p.node_coordinates()
    T = field(X) if deg >= 2 else 300.0 + X @ np.array([2.0, -1.0, 4.0])
    active = np.ones(p.n_cells_total, np.uint8)
    active[-1] = 0
    cs = np.zeros((p.n_cells_total, 3))
    cs[:, 0] = np.linspace(0.0, 1.0, p.n_cells_total)
    cs[:, 2] = 1.0 - cs[:, 0]
    pp = PostProcessor({"filename_prefix": "out",
                        "output_dir": str(tmp_path) + "/",
                        "additional_output_refinement": refine}, p)
    pvtu = pp.write_thermal_output(7, 0.25, T, cs, active)
    assert os.path.basename(pvtu) == "out.7.pvtu"
    d = read_vtu(str(tmp_path / "out.7.0.vtu"))
    n = max(1, refine)
    npts = (n + 1) ** 3
    assert d["points"].shape == (p.n_cells_total * npts, 3)
    assert d["types"].shape == (p.n_cells_total * n ** 3,)
    assert float(d["TIME"][0]) == 0.25
    pts = d["points"]
    want = field(pts) if deg >= 2 else 300.0 + pts @ np.array([2.0, -1.0,
                                                                4.0])
    got = d["temperature"].reshape(p.n_cells_total, npts)
    want = want.reshape(p.n_cells_total, npts)
    np.testing.assert_allclose(got[:-1], want[:-1], rtol=1e-13, atol=1e-10)
    assert (got[-1] == 0.0).all()
    # hexahedra: corners in VTK order span an axis-aligned box
    conn = d["connectivity"].reshape(-1, 8)
    c = pts[conn]
    lo, hi = c.min(1), c.max(1)
    assert (hi - lo > 0).all()
    np.testing.assert_array_equal(c[:, 0], lo)
    np.testing.assert_array_equal(c[:, 6], hi)
    # every patch sits on its own cell: its centroid is the cell centre
    cen = pts.reshape(p.n_cells_total, npts, 3).mean(1)
    np.testing.assert_allclose(cen, p.cell_centers(), rtol=0, atol=1e-12)
    np.testing.assert_array_equal(d["solid"], np.repeat(cs[:, 0], n ** 3))
    np.testing.assert_array_equal(d["powder"], np.repeat(cs[:, 2], n ** 3))
    assert (d["liquid"] == 0).all() and (d["subdomain"] == 0).all()


def test_pvtu_and_pvd(tmp_path):
    p = Problem(n_cells=(2, 2, 2), fe_degree=1)
    T = np.full(p.n_dofs, 300.0)
    cs = np.ones((p.n_cells_total, 1))
    pp = PostProcessor({"filename_prefix": "thermal",
                        "output_dir": str(tmp_path) + "/"}, p, 2)
    for k, t in enumerate((0.0, 1e-4, 2e-4)):
        pp.write_thermal_output(k, t, T, cs, np.ones(p.n_cells_total))
    pvd = pp.write_pvd()
    assert os.path.basename(pvd) == "thermal_m2.pvd"
    sets = ET.parse(pvd).getroot().find("Collection").findall("DataSet")
    assert [float(s.get("timestep")) for s in sets] == [0.0, 1e-4, 2e-4]
    piece = ET.parse(str(tmp_path / "thermal_m2.1.pvtu")).getroot().find(
        "PUnstructuredGrid").find("Piece")
    assert piece.get("Source") == "thermal_m2.1.0.vtu"
    d = read_vtu(str(tmp_path / "thermal_m2.2.0.vtu"))
    assert (d["temperature"] == 300.0).all() and (d["solid"] == 1.0).all()


@pytest.mark.parametrize("deg,refine", [(1, 0), (2, 3)])
def test_vtu_patches_2d(tmp_path, deg, refine):
    """2-D: quadrilateral patches on (x, height, 0) points, the field exact
    where FE_Q(p) holds it."""
    p = Problem(n_cells=(4, 1, 3), extent=(2.0, 1.0, 1.5),
                origin=(0.5, 0.0, -0.3), fe_degree=deg, dim=2)
    X = p.node_coordinates()

    def f2(x):
        if deg >= 2:
            return 300.0 + 4.0 * x[:, 0] ** 2 - 2.0 * x[:, 0] * x[:, 1] + x[:, 1]
        return 300.0 + 2.0 * x[:, 0] - 3.0 * x[:, 1]
    T = f2(X)
    active = np.ones(p.n_cells_total, np.uint8)
    active[0] = 0
    cs = np.zeros((p.n_cells_total, 2))
    cs[:, 0] = 1.0
    pp = PostProcessor({"filename_prefix": "two",
                        "output_dir": str(tmp_path) + "/",
                        "additional_output_refinement": refine}, p)
    pp.write_thermal_output(0, 0.0, T, cs, active)
    d = read_vtu(str(tmp_path / "two.0.0.vtu"))
    n = max(1, refine)
    npts = (n + 1) ** 2
    pts = d["points"]
    assert pts.shape == (p.n_cells_total * npts, 3)
    assert (pts[:, 2] == 0.0).all()
    assert (d["types"] == 9).all() and d["types"].size == p.n_cells_total * n * n
    got = d["temperature"].reshape(p.n_cells_total, npts)
    want = f2(pts[:, :2]).reshape(p.n_cells_total, npts)
    np.testing.assert_allclose(got[1:], want[1:], rtol=1e-13, atol=1e-10)
    assert (got[0] == 0.0).all()
    conn = d["connectivity"].reshape(-1, 4)
    c = pts[conn][:, :, :2]
    lo, hi = c.min(1), c.max(1)
    np.testing.assert_array_equal(c[:, 0], lo)
    np.testing.assert_array_equal(c[:, 2], hi)
    cen = pts.reshape(p.n_cells_total, npts, 3).mean(1)[:, :2]
    np.testing.assert_allclose(cen, p.cell_centers(), rtol=0, atol=1e-12)
