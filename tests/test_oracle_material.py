"""Material properties of the oracle, pinned by the reference's known values
(tests/test_material_property.hh:264-468, tolerance 1e-10)."""
import numpy as np
import pytest

from adamantine_amd.problem import (STATE_PROPERTY_NAMES,
                                    material_from_database, read_info)
from oracle import oracle as orc

RHO = STATE_PROPERTY_NAMES.index("density")
KX = STATE_PROPERTY_NAMES.index("thermal_conductivity_x")
KZ = STATE_PROPERTY_NAMES.index("thermal_conductivity_z")
SOLID = [1.0, 0.0, 0.0]
POWDER = [0.0, 0.0, 1.0]


def values(fixture, p_order):
    db = read_info(fixture)
    m = material_from_database(db["materials"], n_states=3, p_order=p_order)
    md = m.to_c()
    out = []
    for mid, ratios in ((0, SOLID), (1, SOLID), (1, POWDER)):
        out.append([orc.material_property(md, prop, mid, ratios, 15.0)
                    for prop in (RHO, KX, KZ)])
    return np.array(out)


def test_table(golden):
    got = values(golden("material_property_table.info"), 0)
    want = np.array([[1., 100., 100.], [1.75, 150., 150.], [2., 162.5, 162.5]])
    np.testing.assert_allclose(got, want, rtol=1e-10)


def test_polynomial(golden):
    got = values(golden("material_property_polynomial.info"), 4)
    want = np.array([[15., 465., 465.], [706., 681001., 681001.],
                     [720., 45280., 45280.]])
    np.testing.assert_allclose(got, want, rtol=1e-10)


def test_table_edges():
    """compute_property_from_table (MaterialProperty.templates.hh:1174-1211):
    below the first temperature -> first value, above the last -> last value,
    and a table with all 12 entries returns its last value on the final
    interval."""
    from adamantine_amd.problem import Material
    m = Material(n_materials=1, n_states=1, use_table=True)
    m.set_table(0, "density", "solid", [(10.0, 1.0), (20.0, 3.0)])
    full = [(float(i), float(i * i)) for i in range(12)]
    m.set_table(0, "specific_heat", "solid", full)
    md = m.to_c()
    f = lambda prop, T: orc.lib().oracle_property_from_table(md, 0, prop, 0, T)
    assert f(RHO, 5.0) == 1.0 and f(RHO, 10.0) == 1.0
    assert f(RHO, 15.0) == 2.0 and f(RHO, 25.0) == 3.0
    assert f(1, 10.5) == 121.0           # final interval of a full table
    assert f(1, 9.5) == 81.0 + 0.5 * 19.0
