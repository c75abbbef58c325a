"""Reading the reference's 2-D inputs (source/Geometry.cc:186-260 with
dim = 2: length and height, their divisions and origins; boundary ids 0..3;
tests/data/integration_2d*.info) into device problems."""
import os

import numpy as np

from adamantine_amd.problem import (BC_ADIABATIC, SOURCE_ELECTRON_BEAM,
                                    read_info)
from adamantine_amd.run import deposition_boxes, problem_from_database

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_integration_2d_inputs():
    for name in ("integration_2d.info", "integration_2d_units.info",
                 "integration_2d_ensemble.info"):
        db = read_info(os.path.join(GOLD, name))
        p = problem_from_database(db, GOLD)
        assert p.dim == 2 and p.fe_degree == 3
        assert (p.n_cells[0], p.n_cells[2]) == (20, 10)
        assert p.n_dofs == 61 * 31 and p.n_cells_total == 200
        assert p.boundary_types == [BC_ADIABATIC] * 4
        assert p.beams[0].type == SOURCE_ELECTRON_BEAM
        # the built-in box is not scaled by units.mesh (the reference
        # scales imported meshes only): 2e-2 m, or 2e1 in the units input
        assert p.extent[0] in (2e-2, 2e1)
        assert p.node_coordinates().shape == (61 * 31, 2)
        # the beam centre of the 2-D sources: scan path x and y, the latter
        # being the height (0.1 m in both inputs)
        assert abs(p.beams[0].segments[-1][2][1] - 0.1) < 1e-12


def test_2d_deposition_boxes_widened():
    """2-D deposition boxes (x, height bounds) become rows of (x, -, z)
    bounds for the activation calls."""
    db = read_info(os.path.join(GOLD, "integration_2d.info"))
    p = problem_from_database(db, GOLD)
    g = dict(db["geometry"])
    g.update({"material_deposition": "true",
              "material_deposition_method": "file",
              "material_deposition_file": "material_deposition_2d.txt"})
    boxes, t, c, s = deposition_boxes(g, p, GOLD)
    assert boxes.shape[1] == 6 and len(boxes) == len(t) > 0
    assert (boxes[:, 1] == 0.0).all() and (boxes[:, 4] == 0.0).all()
    assert (boxes[:, 3] >= boxes[:, 0]).all() and \
        (boxes[:, 5] >= boxes[:, 2]).all()
