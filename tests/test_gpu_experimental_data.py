"""Observation operator and value write on the device:
set_vector_with_experimental_data_point_cloud (reference
tests/test_experimental_data.cc:75-137) on the 2x2x2 Q1 unit cube. The
reference's expected vector is in deal.II's DoF order; here every node is
checked through its coordinates: the nine top nodes hold the values of the
points that sit on them, every other node is 0."""
import os

import numpy as np
import pytest

from adamantine_amd.problem import Material, Problem

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def unit_cube(n=2, p=1):
    m = Material(n_materials=1, n_states=1, use_table=False, p_order=0)
    m.set_poly(0, "density", "solid", [1.0])
    m.set_poly(0, "specific_heat", "solid", [1.0])
    for k in ("thermal_conductivity_x", "thermal_conductivity_y",
              "thermal_conductivity_z"):
        m.set_poly(0, k, "solid", [1.0])
    return Problem(n_cells=(n, n, n), extent=(1.0, 1.0, 1.0), fe_degree=p,
                   material=m)


def test_set_values_from_point_cloud():
    import torch
    from adamantine_amd.experimental_data import (PointCloud,
                                                  get_expt_to_dof_mapping,
                                                  set_with_experimental_data)
    from adamantine_amd.thermal import ThermalOperatorDevice
    p = unit_cube()
    op = ThermalOperatorDevice(p)
    pc = PointCloud(os.path.join(GOLD, "experimental_data_#camera_#frame.csv"),
                    0, 0)
    pc.read_next_frame()
    pv = pc.get_points_values()
    mapping = get_expt_to_dof_mapping(op, pv)
    T = torch.zeros(p.n_dofs, dtype=torch.float64, device="cuda")
    set_with_experimental_data(op, pv, mapping, T)
    got = T.cpu().numpy()
    X = p.node_coordinates()
    exp = np.zeros(p.n_dofs)
    for pt, v in zip(pv.points, pv.values):
        k = int(np.argmin(((X - np.array(pt)) ** 2).sum(1)))
        assert np.allclose(X[k], pt)
        exp[k] = v
    np.testing.assert_array_equal(got, exp)
    # the reference's expected values, as a multiset
    ref = [0] * 18 + [1.2, 1.5, 1.3, 1.6, 1.8, 1.9, 1.4, 1.7, 2]
    assert sorted(got.tolist()) == sorted(ref)


def test_nearest_dofs_match_full_scan():
    """The windowed search equals a scan over all active nodes (lowest index
    among equidistant ones), on a Q2 box with inactive top layers, for points
    inside, outside, on nodes and half way between nodes."""
    from adamantine_amd.experimental_data import (PointsValues,
                                                  get_expt_to_dof_mapping)
    from adamantine_amd.thermal import ThermalOperatorDevice
    p = unit_cube(n=6, p=2)
    z = p.cell_centers()[:, 2]
    p.cell_active = (z < 0.5).astype(np.uint8)
    op = ThermalOperatorDevice(p)
    X = p.node_coordinates()
    nn = p.n_nodes
    act = np.zeros(p.n_dofs, bool)
    for c in np.nonzero(p.cell_active)[0]:
        cx, cy, cz = c % 6, (c // 6) % 6, c // 36
        for k in range(3):
            for j in range(3):
                for i in range(3):
                    act[(2 * cx + i) + nn[0] * ((2 * cy + j) +
                                               nn[1] * (2 * cz + k))] = True
    rs = np.random.RandomState(3)
    pts = np.concatenate([rs.uniform(-0.3, 1.3, (200, 3)),
                          X[rs.randint(0, p.n_dofs, 20)],
                          (X[:40:2] + X[1:41:2]) / 2,
                          [[0.5, 0.5, 2.0], [0.1, 0.2, 0.9]]])
    _, dofs = get_expt_to_dof_mapping(op, PointsValues(
        [tuple(x) for x in pts], [0.0] * len(pts)))
    idx = np.nonzero(act)[0]
    for q, x in enumerate(pts):
        d = np.sqrt(((X[idx] - x) ** 2).sum(1))
        # the nearest node; for points half way between nodes the library's
        # own coordinates decide (they may differ from these in the last bit)
        near = idx[d <= d.min() * (1 + 1e-12)]
        assert dofs[q] in near
        if len(near) == 1:
            assert dofs[q] == near[0]
