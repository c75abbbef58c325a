"""The five workloads of the MI355X heat path, as Problem builders.

Every choice the workload statements leave open is fixed here (seeded RNGs
are numpy's legacy RandomState streams, which numpy keeps stable):

1. ``plain_small``  32x16x8 Q1 box 2x1x0.5 mm, Solid, polynomial k(T),
   Goldak 200 W moving along x at 1 m/s, forward Euler 1e-6 s x 100.
2. ``plain_large``  128x64x32 Q2 box 4x2x1 mm (2,154,945 nodes), SLP with
   8-breakpoint k(T), c_p(T) tables, top 4 cell layers powder, Goldak 400 W
   on a 5-track serpentine, radiative + convective top face, RK4 5e-7 s.
3. ``ensemble``     64x32x16 Q2, same material model, 64 members.
4. ``adjoint``      64x32x16 Q2 (12 of 16 layers active, top 4 powder),
                     200 RK4 steps of 5e-7 s, checkpoints every 20 steps,
                     200 thermocouple-like points at the final time.
5. ``deposition``   96x96x40 Q2, electron beam, layer activation.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np

from .problem import (BC_ADIABATIC, BC_CONVECTIVE, BC_RADIATIVE,
                      FORWARD_EULER, RK_FOURTH_ORDER, SOURCE_ELECTRON_BEAM,
                      SOURCE_GOLDAK, Beam, Material, Problem,
                      segments_from_path)

MM = 1e-3
UM = 1e-6


def plain_small(n_cells=(32, 16, 8)) -> Problem:
    """Q1, Solid, rho = 7904, c_p = 714, k = 11.8 + 0.0138 T; Goldak 200 W,
    absorption 0.3, 50 um radius and depth, along x at mid-y on the top
    surface at 1 m/s; adiabatic; T0 = 300 K."""
    m = Material(n_materials=1, n_states=1, use_table=False, p_order=1,
                 single_material_layout=True)
    m.set_poly(0, "density", "solid", [7904.0])
    m.set_poly(0, "specific_heat", "solid", [714.0])
    for k in ("thermal_conductivity_x", "thermal_conductivity_y",
              "thermal_conductivity_z"):
        m.set_poly(0, k, "solid", [11.8, 0.0138])
    path = segments_from_path([
        {"mode": "point", "xyz": (0.0, 0.5 * MM, 0.5 * MM), "pmod": 0.0,
         "param": 1e-9},
        {"mode": "line", "xyz": (2.0 * MM, 0.5 * MM, 0.5 * MM), "pmod": 1.0,
         "param": 1.0},
    ])
    beam = Beam(type=SOURCE_GOLDAK, absorption_efficiency=0.3, max_power=200.0,
                radius=50 * UM, depth=50 * UM, segments=path)
    return Problem(n_cells=tuple(n_cells), extent=(2 * MM, 1 * MM, 0.5 * MM),
                   fe_degree=1, quadrature=0, dim=3, material=m,
                   boundary_types=[BC_ADIABATIC] * 6, beams=[beam],
                   method=FORWARD_EULER, initial_temperature=300.0)


def slp_tables(seed_breaks: int = 2, seed_k: int = 3, seed_cp: int = 4,
               k_scale: float = 1.0) -> Material:
    """SolidLiquidPowder tables: 8 breakpoints from 300 K to 3000 K (6 sorted
    uniform interior draws for k and 6 for c_p), k = 11.8 + 0.0138 T + U(-1,1),
    c_p = 714 + 0.1 (T - 300) + U(-10,10); liquid uses the solid tables;
    powder has 0.6 rho and 0.01 k; solidus 1675 K, liquidus 1708 K, latent
    heat 2.6e5 J/kg; emissivity 0.3 and h = 10 W/m^2K; T_inf = 300 K."""
    rb = np.random.RandomState(seed_breaks)
    tk = np.concatenate([[300.0], np.sort(rb.uniform(300.0, 3000.0, 6)),
                         [3000.0]])
    tc = np.concatenate([[300.0], np.sort(rb.uniform(300.0, 3000.0, 6)),
                         [3000.0]])
    kv = 11.8 + 0.0138 * tk + np.random.RandomState(seed_k).uniform(-1, 1, 8)
    cv = 714.0 + 0.1 * (tc - 300.0) + \
        np.random.RandomState(seed_cp).uniform(-10, 10, 8)
    m = Material(n_materials=1, n_states=3, use_table=True, p_order=0,
                 single_material_layout=True)
    for state, rho_f, k_f in (("solid", 1.0, 1.0), ("liquid", 1.0, 1.0),
                              ("powder", 0.6, 0.01)):
        m.set_table(0, "density", state, [(300.0, 7904.0 * rho_f)])
        m.set_table(0, "specific_heat", state, list(zip(tc, cv)))
        for kname in ("thermal_conductivity_x", "thermal_conductivity_y",
                      "thermal_conductivity_z"):
            m.set_table(0, kname, state,
                        list(zip(tk, kv * k_f * k_scale)))
        m.set_table(0, "emissivity", state, [(300.0, 0.3)])
        m.set_table(0, "convection_heat_transfer_coef", state, [(300.0, 10.0)])
    m.set_property(0, "solidus", 1675.0)
    m.set_property(0, "liquidus", 1708.0)
    m.set_property(0, "latent_heat", 2.6e5)
    m.set_property(0, "radiation_temperature_infty", 300.0)
    m.set_property(0, "convection_temperature_infty", 300.0)
    return m


def serpentine(n_tracks=5, x0=0.5 * MM, x1=3.5 * MM, y0=0.8 * MM,
               hatch=0.1 * MM, z=1.0 * MM, speed=0.8) -> List[tuple]:
    """Tracks along x at y0 + k*hatch, alternating direction; the hatch jumps
    are beam-off line segments."""
    pts = [{"mode": "point", "xyz": (x0, y0, z), "pmod": 0.0, "param": 1e-9}]
    x_end = x0
    for k in range(n_tracks):
        y = y0 + k * hatch
        if k > 0:
            pts.append({"mode": "line", "xyz": (x_end, y, z), "pmod": 0.0,
                        "param": speed})
        x_end = x1 if k % 2 == 0 else x0
        pts.append({"mode": "line", "xyz": (x_end, y, z), "pmod": 1.0,
                    "param": speed})
    return segments_from_path(pts)


def plain_large(n_cells=(128, 64, 32), extent=(4 * MM, 2 * MM, 1 * MM),
                powder_layers: int = 4, method=RK_FOURTH_ORDER) -> Problem:
    m = slp_tables()
    beam = Beam(type=SOURCE_GOLDAK, absorption_efficiency=0.3, max_power=400.0,
                radius=100 * UM, depth=60 * UM,
                segments=serpentine(z=extent[2]))
    p = Problem(n_cells=tuple(n_cells), extent=tuple(extent), fe_degree=2,
                quadrature=0, dim=3, material=m,
                boundary_types=[BC_ADIABATIC] * 5 +
                [BC_RADIATIVE | BC_CONVECTIVE],
                beams=[beam], method=method, initial_temperature=300.0)
    hz = extent[2] / n_cells[2]
    p.set_powder_layer(extent[2], powder_layers * hz)
    return p


def random_temperature(problem: Problem, seed: int = 1, low=300.0,
                       high=2000.0) -> np.ndarray:
    return np.random.RandomState(seed).uniform(low, high, problem.n_dofs)


def smooth_temperature(problem: Problem) -> np.ndarray:
    """T = 300 + 50 sin(pi x / L) cos(pi y / W) (+ a z ramp)."""
    X = problem.node_coordinates()
    L, W, H = problem.extent
    return (300.0 + 50.0 * np.sin(np.pi * X[:, 0] / L) *
            np.cos(np.pi * X[:, 1] / W) + 10.0 * X[:, 2] / H)


def ensemble_workload(n_members: int = 64, n_cells=(64, 32, 16),
                      extent=(4 * MM, 2 * MM, 1 * MM), powder_layers: int = 4,
                      method=RK_FOURTH_ORDER):
    """The ensemble workload: ``n_members`` members over one 64x32x16 Q2 box
    with the material model of ``plain_large``.

    Member j draws beam_0.absorption_efficiency ~ N(0.3, 0.05) and a
    thermal-conductivity factor ~ N(1, 0.1), in that key order, with the
    reference's scheme (default-seeded mt19937, the k-th key discarding
    k * n_members draws, negatives redrawn). Member j has the lowest
    L_j = 12 + (j mod 5) cell layers active (at most n_z; 3/4 n_z + (j mod 5)
    for other n_z), the top ``powder_layers`` of them powder. The serpentine
    scan path runs on the top of the lowest active height (12 layers), so
    deeper members see the source below their surface.

    Returns (base problem, members)."""
    from .ensemble import Member, draw_member_values
    nz = n_cells[2]
    hz = extent[2] / nz
    base_layers = (3 * nz) // 4
    p = plain_large(n_cells=n_cells, extent=extent,
                    powder_layers=powder_layers, method=method)
    top = base_layers * hz
    p.beams[0].segments = serpentine(z=top)
    p.set_material_height(top)
    p.set_powder_layer(top, powder_layers * hz)
    vals = draw_member_values(
        [("beam_0.absorption_efficiency", 0.3, 0.05),
         ("thermal_conductivity_scale", 1.0, 0.1)], n_members)
    z = p.cell_centers()[:, -1]
    members = []
    for j in range(n_members):
        layers = min(base_layers + j % 5, nz)
        height = layers * hz
        active = (z < height).astype(np.uint8)
        st = np.zeros((p.n_cells_total, 3))
        st[:, 0] = 1.0
        powder = z >= height - powder_layers * hz
        st[powder, 0] = 0.0
        st[powder, 2] = 1.0
        members.append(Member(
            absorption_efficiency=[vals["beam_0.absorption_efficiency"][j]],
            max_power=[p.beams[0].max_power],
            conductivity_scale=float(vals["thermal_conductivity_scale"][j]),
            cell_active=active, cell_state=st))
    return p, members


def meltpool_temperature(problem: Problem, t: float, peak: float = 2600.0,
                         ambient: float = 300.0, sigma_xy: float = 150e-6,
                         sigma_z: float = 80e-6) -> np.ndarray:
    """A smooth field with a melt pool under the first beam at time t:
    T = ambient + (peak - ambient) exp(-(dx^2+dy^2)/(2 sxy^2) - dz^2/(2 sz^2)).
    Solid far from the beam, mushy and liquid in the pool: the temperature
    pattern of a scan, used as the operator's benchmark input."""
    from .problem import BeamState  # noqa: F401
    X = problem.node_coordinates()
    seg = problem.beams[0].segments
    # position on the scan path (piecewise linear between segment ends)
    t0, p0 = 0.0, np.array(seg[0][2], float)
    center = np.array(seg[-1][2], float)
    for et, _, xyz in seg:
        xyz = np.array(xyz, float)
        if t <= et:
            f = 0.0 if et == t0 else (t - t0) / (et - t0)
            center = p0 + f * (xyz - p0)
            break
        t0, p0 = et, xyz
    d2 = ((X[:, 0] - center[0]) ** 2 + (X[:, 1] - center[1]) ** 2) / (
        2 * sigma_xy ** 2) + (X[:, 2] - center[2]) ** 2 / (2 * sigma_z ** 2)
    return ambient + (peak - ambient) * np.exp(-d2)


@dataclass
class AdjointWorkload:
    problem: Problem          # the model (what the gradient is taken of)
    truth: Problem            # the model with the truth conductivities
    truth_beam: list          # [absorption, max power, radius, depth]
    obs_dof: np.ndarray       # [n_obs] observed nodes
    noise: np.ndarray         # [n_obs] N(0, 1) added to the truth values
    dt: float
    n_steps: int
    checkpoint_every: int


def adjoint_workload(n_cells=(64, 32, 16), extent=(4 * MM, 2 * MM, 1 * MM),
                     powder_layers: int = 4, n_steps: int = 200,
                     dt: float = 5e-7, n_obs: int = 200,
                     checkpoint_every: int = 20) -> AdjointWorkload:
    """The adjoint workload: the base problem of ``ensemble_workload``
    (12 of 16 layers active, the top 4 of them powder, serpentine path on
    the active top), T0 = 300 K, ``n_steps`` RK4 steps of ``dt``.

    Observations: ``n_obs`` points uniform (numpy RandomState(6)) in the
    first 0.4 mm of the first track, within 0.1 mm of it in y, on the
    active top surface, each observed at its nearest node; the data are the
    final temperatures of a truth run with absorption efficiency 0.32, beam
    radius 1.1 x the model's and every conductivity x 1.05, plus N(0, 1)
    noise drawn next from the same stream."""
    import copy
    p, _ = ensemble_workload(n_members=1, n_cells=n_cells, extent=extent,
                             powder_layers=powder_layers)
    nz = n_cells[2]
    top = ((3 * nz) // 4) * extent[2] / nz
    rs = np.random.RandomState(6)
    seg = p.beams[0].segments
    x0, y0 = seg[0][2][0], seg[0][2][1]
    x = rs.uniform(x0, x0 + 0.4 * MM, n_obs)
    y = rs.uniform(y0 - 0.1 * MM, y0 + 0.1 * MM, n_obs)
    noise = rs.standard_normal(n_obs)
    nn = p.n_nodes
    h = [extent[d] / (n_cells[d] * p.fe_degree) for d in range(3)]
    ix = np.rint((x - p.origin[0]) / h[0]).astype(np.int64)
    iy = np.rint((y - p.origin[1]) / h[1]).astype(np.int64)
    iz = int(round((top - p.origin[2]) / h[2]))
    obs_dof = ix + nn[0] * (iy + nn[1] * iz)
    truth = copy.deepcopy(p)
    truth.material = slp_tables(k_scale=1.05)
    b = p.beams[0]
    return AdjointWorkload(p, truth, [0.32, b.max_power, 1.1 * b.radius,
                                      b.depth], obs_dof, noise, dt, n_steps,
                           checkpoint_every)


@dataclass
class DepositionWorkload:
    base: Problem
    members: list
    activation_step: np.ndarray   # [n_members, n_layers] step of each layer
    layer_boxes: np.ndarray       # [n_layers, 6] deposition box per layer
    dt: float
    n_steps: int


def deposition_workload(n_members: int = 32, n_cells=(96, 96, 40),
                        extent=(3 * MM, 3 * MM, 1.25 * MM), n_layers: int = 10,
                        steps_per_layer: int = 200, dt: float = 2e-7,
                        n_steps: int = 2000) -> DepositionWorkload:
    """The deposition workload: ``n_members`` members of a 96x96x40 Q2 box
    (31.25 um cells) built layer by layer, the material model of
    ``plain_large``, an electron beam of 1500 W, depth 150 um, radius 150 um.

    Member j draws beam_0.absorption_efficiency ~ N(0.85, 0.03) with the
    reference's scheme. Layer k (k = 0 .. n_layers-1) holds cell layers
    4k .. 4k+3; layer 0 is active at t = 0 and layer k >= 1 is deposited at
    step steps_per_layer * k + delta_jk with delta ~ U{0..100}
    (numpy RandomState(7), drawn member-major), by a deposition box covering
    the layer. The beam rasters along x at 2 m/s at mid-y on the nominal top
    of the last deposited layer (layer k during steps [200k, 200(k+1)))."""
    from .ensemble import Member, draw_member_values
    from .problem import SOURCE_ELECTRON_BEAM as EB
    nz = n_cells[2]
    hz = extent[2] / nz
    cells_per_layer = nz // n_layers
    m = slp_tables()
    # raster: one track along x per layer window, z on that layer's top
    speed = 2.0
    x0, y0 = 0.25 * MM, 0.5 * extent[1]
    pts = [{"mode": "point", "xyz": (x0, y0, cells_per_layer * hz),
            "pmod": 0.0, "param": 1e-9}]
    x = x0
    for k in range(n_layers):
        z = (k + 1) * cells_per_layer * hz
        if k > 0:
            pts.append({"mode": "point", "xyz": (x, y0, z), "pmod": 0.0,
                        "param": 1e-9})
        x = x + speed * steps_per_layer * dt
        pts.append({"mode": "line", "xyz": (x, y0, z), "pmod": 1.0,
                    "param": speed})
    beam = Beam(type=EB, absorption_efficiency=0.85, max_power=1500.0,
                radius=150 * UM, depth=150 * UM,
                segments=segments_from_path(pts))
    p = Problem(n_cells=tuple(n_cells), extent=tuple(extent), fe_degree=2,
                quadrature=0, dim=3, material=m,
                boundary_types=[BC_ADIABATIC] * 5 +
                [BC_RADIATIVE | BC_CONVECTIVE],
                beams=[beam], method=FORWARD_EULER, initial_temperature=300.0)
    z = p.cell_centers()[:, -1]
    active0 = (z < cells_per_layer * hz).astype(np.uint8)
    vals = draw_member_values([("beam_0.absorption_efficiency", 0.85, 0.03)],
                              n_members)
    members = [Member(absorption_efficiency=[vals[
        "beam_0.absorption_efficiency"][j]], max_power=[1500.0],
        cell_active=active0.copy()) for j in range(n_members)]
    delta = np.random.RandomState(7).randint(0, 101, (n_members, n_layers))
    steps = steps_per_layer * np.arange(n_layers)[None, :] + delta
    steps[:, 0] = 0
    boxes = np.array([[0.0, 0.0, (k * cells_per_layer + 0.25) * hz,
                       extent[0], extent[1],
                       ((k + 1) * cells_per_layer - 0.25) * hz]
                      for k in range(n_layers)])
    return DepositionWorkload(p, members, steps, boxes, dt, n_steps)
