"""Ensemble simulation with data assimilation from an adamantine input file.

`run_ensemble` is the time loop of the reference's run_ensemble
(application/adamantine.hh:1961-2433) for the thermal physics on one
structured box, on the device:
  - the members are drawn by create_database_ensemble
    (source/ensemble_management.cc:166-265);
  - the whole ensemble is stepped by one batched call per time step (the
    reference steps member after member, :2137-2142);
  - experimental frames are read at their time stamps and assimilated with
    the EnKF, with the reference's frame cadence (:2150-2361);
  - material is deposited in every member with the activation windows of
    the loop (:2054-2128).
`run` is the single-simulation loop of the reference's run
(application/adamantine.hh:1063-1400) with material deposition: the boxes
of create_material_deposition_boxes are activated window by window
(:1134-1207) before each step.
Mesh refinement, checkpoints, output and mechanics of the reference's loops
are not part of this path.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .problem import (FORWARD_EULER, RK_FOURTH_ORDER, RK_THIRD_ORDER,
                      STATE_NAMES, STATE_PROPERTY_NAMES, Problem, _get,
                      beams_from_database, boundary_from_database,
                      material_from_database, read_info)

_METHODS = {"forward_euler": FORWARD_EULER, "rk_third_order": RK_THIRD_ORDER,
            "rk_fourth_order": RK_FOURTH_ORDER}


def material_template_parameters(materials: Dict):
    """get_material_template_parameters (application/adamantine.cc:
    155-241): number of material states (highest state with a property, plus
    one) and polynomial order (most coefficients minus one; 0 for tables)."""
    use_table = _get(materials, "property_format") == "table"
    n_states = 0
    p_order = 0
    for mid in range(int(_get(materials, "n_materials"))):
        mdb = materials.get("material_%d" % mid)
        if not isinstance(mdb, dict):
            continue
        for s, sname in enumerate(STATE_NAMES):
            sdb = mdb.get(sname)
            if not isinstance(sdb, dict):
                continue
            for pname in STATE_PROPERTY_NAMES:
                if pname in sdb:
                    n_states = max(n_states, s + 1)
                    if not use_table:
                        n = len("".join(str(sdb[pname]).split()).split(","))
                        p_order = max(p_order, n - 1)
    if not (0 < n_states < 4) or p_order > 4:
        raise ValueError("Error when computing the material parameters")
    return n_states, p_order


def problem_from_database(db: Dict, base_dir: str = ".",
                          fe_degree: Optional[int] = None) -> Problem:
    """The structured-box problem that ThermalPhysics::setup builds from an
    input database (geometry source/Geometry.cc:186-352, boundary
    source/Boundary.cc:12-70, materials, sources, active cells
    source/ThermalPhysics.templates.hh:472-493)."""
    g = db["geometry"]
    if str(_get(g, "import_mesh", "false")) == "true":
        raise NotImplementedError("imported meshes are not on the device path")
    dim = int(_get(g, "dim", 3))
    if dim not in (2, 3):
        raise ValueError("geometry.dim must be 2 or 3")
    # axis x = length, y = width, z = height (Geometry.cc:188-243); a 2-D
    # box has no width (the descriptor's y entries are placeholders)
    n_cells = (int(_get(g, "length_divisions", 10)),
               int(_get(g, "width_divisions", 10)) if dim == 3 else 1,
               int(_get(g, "height_divisions", 10)))
    origin = (float(_get(g, "length_origin", 0.0)),
              float(_get(g, "width_origin", 0.0)) if dim == 3 else 0.0,
              float(_get(g, "height_origin", 0.0)))
    extent = (float(g["length"]), float(g["width"]) if dim == 3 else 1.0,
              float(g["height"]))
    mats = db["materials"]
    n_states, p_order = material_template_parameters(mats)
    m = material_from_database(mats, n_states, p_order)
    disc = _get(db, "discretization.thermal", {})
    deg = int(fe_degree if fe_degree is not None else disc.get("fe_degree", 1))
    quad = 1 if str(disc.get("quadrature", "gauss")) == "lobatto" else 0
    ts = db["time_stepping"]
    p = Problem(n_cells=n_cells, extent=extent, origin=origin, fe_degree=deg,
                quadrature=quad, dim=dim, material=m,
                boundary_types=boundary_from_database(db.get("boundary", {}),
                                                      2 * dim),
                beams=beams_from_database(db["sources"], base_dir,
                                          db.get("units")),
                method=_METHODS[str(ts["method"])],
                initial_temperature=float(_get(mats, "initial_temperature",
                                               300.0)))
    height = float(_get(g, "material_height", 1e9))
    if height < origin[2] + extent[2]:
        p.set_material_height(height)
    if str(_get(g, "use_powder", "false")) == "true":
        p.set_powder_layer(height, float(g["powder_layer"]))
    elif n_states == 3 or n_states == 2:
        st = np.zeros((p.n_cells_total, n_states))
        st[:, 0] = 1.0
        p.cell_state = st
    return p


def deposition_boxes(g: Dict, p: Problem, base_dir: str):
    """create_material_deposition_boxes with the boxes as rows of (min x,
    y, z, max x, y, z), the layout of the activation calls; a 2-D box's
    (x, height) bounds go to x and z."""
    from .deposition import create_material_deposition_boxes
    boxes, t, c, s = create_material_deposition_boxes(g, p.beams, base_dir,
                                                      p.dim)
    if p.dim == 2:
        b6 = np.zeros((len(boxes), 6))
        b6[:, [0, 2, 3, 5]] = np.asarray(boxes).reshape(-1, 4)
        boxes = b6
    return boxes, t, c, s


def _stddev_keys(tree: Dict, path: str = "") -> List[str]:
    """Leaves of the ensemble database in input order
    (source/ensemble_management.cc:166-211), skipping the two switches."""
    out = []
    for k, v in tree.items():
        full = k if not path else path + "." + k
        if isinstance(v, dict) and v:
            out += _stddev_keys(v, full)
        elif full not in ("ensemble_simulation", "ensemble_size"):
            out.append(full)
    return out


def simulation_duration(ts: Dict, beams) -> float:
    """time_stepping.duration, or with scan_path_for_duration the earliest
    end of the heat sources' scan paths once every path file is finished
    (application/adamantine.hh:998-1004, 1041-1058). A path that is still
    being written (no SCAN_PATH_END) would make the reference wait for the
    file; that live mode is not on the device path."""
    if str(_get(ts, "scan_path_for_duration", "false")) != "true":
        return float(ts["duration"])
    paths = [b for b in beams if b.segments]
    if not all(b.scan_path_end for b in paths):
        raise NotImplementedError(
            "scan_path_for_duration with a scan path that is not finished")
    return min((b.segments[-1][0] for b in paths), default=math.inf)


@dataclass
class Run:
    temperature: object                 # [n_dofs] device tensor
    physics: object                     # ThermalPhysicsDevice
    n_steps: int = 0
    time: float = 0.0
    n_activations: int = 0


def run(db: Dict, base_dir: str = ".", fe_degree: Optional[int] = None,
        verbose: bool = False, output: bool = False,
        autotune: bool = True) -> Run:
    """run for the thermal physics on the device, with material deposition
    from the scan paths or a deposition file, and the checkpoint and restart
    sections of the input (files <prefix>.npz and <prefix>_time.txt).
    With output=True the post_processor section's VTU/PVTU/PVD files are
    written at the reference's cadence (the initial state, then every
    time_steps_between_output steps). With autotune the operator's launch
    configuration is picked by timing on the initial temperature (the
    phase state is left as it was)."""
    from bisect import bisect_left
    from .thermal import ThermalPhysicsDevice
    p = problem_from_database(db, base_dir, fe_degree)
    ph = ThermalPhysicsDevice(p)
    ph.setup()
    # checkpoint / restart (application/adamantine.hh:879-911, 964-973)
    ck = db.get("checkpoint")
    every = int(ck["time_steps_between_checkpoint"]) if ck else 0
    time, n_time_step = 0.0, 0
    if db.get("restart"):
        prefix = str(db["restart"]["filename_prefix"])
        T = ph.load_checkpoint(prefix)
        with open(prefix + "_time.txt") as f:
            tok = f.read().split()
        time, n_time_step = float(tok[0]), int(tok[1])
    else:
        T = ph.initialize_dof_vector(p.initial_temperature)
    if autotune:
        ph.operator.autotune(T, reps=3)
    pp = None
    if output:
        # output_pvtu (application/adamantine.hh:69-129, 978-981)
        from .output import PostProcessor
        post = db.get("post_processor", {})
        every_out = int(post.get("time_steps_between_output", 1))
        pp = PostProcessor(post, ph.problem)

        def write(step, t):
            active, _, _ = ph.operator.cell_data()
            pp.write_thermal_output(step, t, T,
                                    ph.set_state_to_material_properties(),
                                    active)
        write(n_time_step, time)
    n_time_step += 1
    g = db["geometry"]
    boxes, dep_t, dep_cos, dep_sin = deposition_boxes(g, p, base_dir)
    dep_t = list(dep_t)
    ts = db["time_stepping"]
    dt = float(ts["time_step"])
    duration = simulation_duration(ts, p.beams)
    activation_time = float(_get(g, "deposition_time", 0.0))
    new_temperature = float(_get(db, "materials.new_material_temperature",
                                 300.0))
    out = Run(T, ph)
    activation_time_end = -1.0
    tol = dt * 1e-6
    while time < duration:
        if time + dt > duration:
            dt = duration - time     # application/adamantine.hh:1111-1112
        if dep_t and time - activation_time_end > -tol:
            start = bisect_left(dep_t, time - tol)
            activation_time_end = min(time + max(activation_time, dt),
                                      duration)
            end = bisect_left(dep_t, activation_time_end)
            if start < end:
                b = ph.activate(boxes[start:end], dep_cos[start:end],
                                dep_sin[start:end], new_temperature, T)
                out.n_activations += 1
                if verbose:
                    print("t = %g: boxes %d..%d, %d cells" %
                          (time, start, end - 1, int((b >= 0).sum())),
                          flush=True)
        time = ph.evolve(time, dt, 1, T)
        out.n_steps += 1
        if every and n_time_step % every == 0:
            # application/adamantine.hh:1349-1372
            prefix = str(ck["filename_prefix"])
            if str(ck.get("overwrite_files", "true")) != "true":
                prefix += "_%d" % n_time_step
            ph.save_checkpoint(prefix, T)
            with open(prefix + "_time.txt", "w") as f:
                f.write("%r %d\n" % (time, n_time_step))
        if pp is not None and n_time_step % every_out == 0:
            write(n_time_step, time)
        n_time_step += 1
    if pp is not None:
        pp.write_pvd()
    out.time = time
    return out


@dataclass
class EnsembleRun:
    solution: object                    # [n_members, n_dofs] device tensor
    params: np.ndarray                  # [n_members, n_augmented]
    param_names: List[str]
    member_values: Dict[str, np.ndarray]
    times: List[float] = field(default_factory=list)  # DA times
    n_steps: int = 0
    time: float = 0.0
    n_activations: int = 0
    ensemble: object = None             # EnsembleDevice


def _cell_state(liquid: np.ndarray, powder: np.ndarray,
                n_states: int) -> np.ndarray:
    """MaterialProperty::set_state (source/MaterialProperty.templates.hh:
    537-620): cell averages of the quadrature-point ratios."""
    l = liquid.mean(axis=1)
    p = powder.mean(axis=1)
    out = np.zeros((len(l), n_states))
    if n_states == 1:
        out[:, 0] = 1.0
    elif n_states == 2:
        out[:, 1] = l
        out[:, 0] = 1.0 - l
    else:
        out[:, 1] = l
        out[:, 2] = p
        out[:, 0] = 1.0 - l - p
    return out


def run_ensemble(db: Dict, base_dir: str = ".",
                 fe_degree: Optional[int] = None,
                 verbose: bool = False, output: bool = False,
                 autotune: bool = True) -> EnsembleRun:
    """run_ensemble for the thermal physics on the device. With output=True
    every member writes the post_processor section's VTU/PVTU/PVD files
    (prefix _m<member>) at the reference's cadence: the initial state, every
    time_steps_between_output steps, and around each assimilation frame
    unless output_on_data_assimilation is false
    (application/adamantine.hh:1872-1885, 2416-2441). With autotune the
    batched launch configuration is picked by timing."""
    import torch
    from .ensemble import (DataAssimilator, EnsembleDevice, Member,
                           draw_member_values)
    from .experimental_data import (PointCloud, RayTracing,
                                    read_frame_timestamps)
    p = problem_from_database(db, base_dir, fe_degree)
    ens_db = db.get("ensemble", {})
    n = int(ens_db.get("ensemble_size", 5))
    keys = []
    for k in _stddev_keys(ens_db):
        if not k.endswith("_stddev"):
            raise ValueError("ensemble key %s is not a standard deviation" % k)
        base = k[:-len("_stddev")]
        keys.append((base, float(_get(db, base)), float(_get(ens_db, k))))
    vals = draw_member_values(keys, n)
    n_beams = len(p.beams)
    T0 = np.full(n, p.initial_temperature)
    members = []
    for j in range(n):
        eta = [b.absorption_efficiency for b in p.beams]
        power = [b.max_power for b in p.beams]
        for key, v in vals.items():
            parts = key.split(".")
            if key == "materials.initial_temperature":
                T0[j] = v[j]
            elif (len(parts) == 3 and parts[0] == "sources" and
                  parts[1].startswith("beam_") and
                  parts[2] in ("absorption_efficiency", "max_power")):
                b = int(parts[1][5:])
                if b >= n_beams:
                    raise ValueError("no heat source %s" % parts[1])
                (eta if parts[2] == "absorption_efficiency" else power)[b] = \
                    float(v[j])
            else:
                raise NotImplementedError(
                    "perturbation of %s is not on the device path" % key)
        members.append(Member(absorption_efficiency=eta, max_power=power))
    # checkpoint / restart: one file set per member, <prefix>_<member>, and
    # <prefix>_time.txt (application/adamantine.hh:1630-1661, 1725-1795,
    # 2380-2400)
    ck = db.get("checkpoint")
    every_ck = int(ck["time_steps_between_checkpoint"]) if ck else 0
    restart = None
    nc = p.n_cells_total
    dcos = [np.asarray(p.deposition_cos if p.deposition_cos is not None
                       else np.ones(nc), float).copy() for _ in range(n)]
    dsin = [np.asarray(p.deposition_sin if p.deposition_sin is not None
                       else np.zeros(nc), float).copy() for _ in range(n)]
    if db.get("restart"):
        import copy
        prefix = str(db["restart"]["filename_prefix"])
        saved = []
        for j in range(n):
            with np.load("%s_%d.npz" % (prefix, j), allow_pickle=False) as f:
                saved.append({k: f[k] for k in f.files})
        p = copy.deepcopy(p)
        ucos, usin = np.ones(nc), np.zeros(nc)
        for j, d in enumerate(saved):
            on = d["active"].astype(bool)
            members[j].cell_active = d["active"].astype(np.uint8)
            members[j].cell_state = d["cell_state"]
            dcos[j] = np.where(on, d["deposition_cos"], 1.0)
            dsin[j] = np.where(on, d["deposition_sin"], 0.0)
            ucos[on], usin[on] = dcos[j][on], dsin[j][on]
        # a cell's orientation comes from its deposition box, the same in
        # every member that has activated it
        p.deposition_cos, p.deposition_sin = ucos, usin
        with open(prefix + "_time.txt") as f:
            tok = f.read().split()
        restart = (saved, float(tok[0]), int(tok[1]))
    ens = EnsembleDevice(p, members)
    ens.compute_inverse_mass_matrix()
    if restart:
        # the saved point histories (files without them: every point took
        # its cell's state above)
        for j, d in enumerate(restart[0]):
            if "powder_qp" in d:
                ens.set_qp_state(j, d["liquid_qp"], d["powder_qp"])
                ens.set_face_state(j, d["face_powder"])
    # deposition boxes of member 0's sources, shared by all members
    # (application/adamantine.hh:1917-1925)
    from bisect import bisect_left
    g = db["geometry"]
    boxes, dep_t, dep_cos, dep_sin = deposition_boxes(g, p, base_dir)
    dep_t = list(dep_t)
    activation_time = float(_get(g, "deposition_time", 0.0))
    new_temperature = float(_get(db, "materials.new_material_temperature",
                                 300.0))
    sol = ens.initialize_dof_vector(0.0)
    if restart:
        sol.copy_(torch.tensor(np.stack([d["temperature"]
                                         for d in restart[0]]),
                               dtype=torch.float64, device=sol.device))
    else:
        sol += torch.tensor(T0, dtype=torch.float64,
                            device=sol.device)[:, None]
    if autotune:
        ens.autotune(sol, reps=2)

    # augmented state (application/adamantine.hh:1600-1700)
    da_db = db.get("data_assimilation", {})
    names = []
    if str(da_db.get("augment_with_beam_0_absorption", "false")) == "true":
        names.append("beam_0_absorption")
    if str(da_db.get("augment_with_beam_0_max_power", "false")) == "true":
        names.append("beam_0_max_power")
    params = np.zeros((n, len(names)))
    for j, m in enumerate(members):
        for i, nm in enumerate(names):
            params[j, i] = (m.absorption_efficiency[0]
                            if nm == "beam_0_absorption" else m.max_power[0])

    # experimental data
    ex = db.get("experiment", {})
    stamps, cloud = None, None
    # every member's active cells (member 0's are what the ray tracing sees,
    # application/adamantine.hh:1834-1851)
    base_active = (np.ones(p.n_cells_total, np.uint8) if p.cell_active is None
                   else np.asarray(p.cell_active, np.uint8))
    active = [(np.asarray(m.cell_active, np.uint8) if m.cell_active is not None
               else base_active).copy() for m in members]
    active0 = active[0]
    if str(ex.get("read_in_experimental_data", "false")) == "true":
        stamps = read_frame_timestamps(
            os.path.join(base_dir, str(ex["log_filename"])),
            int(ex["first_camera_id"]), int(ex["last_camera_id"]),
            int(ex["last_frame"]), int(ex.get("first_frame", 0)),
            float(ex.get("first_frame_temporal_offset", 0.0)))
        if not stamps or not stamps[0]:
            raise ValueError("the experimental log shows no frame")
        if str(ex.get("format", "point_cloud")).lower() == "point_cloud":
            cloud = PointCloud(os.path.join(base_dir, str(ex["file"])),
                               int(ex["first_camera_id"]),
                               int(ex["last_camera_id"]),
                               int(ex.get("first_frame", 0)))
        else:
            cloud = RayTracing(os.path.join(base_dir, str(ex["file"])),
                               int(ex["first_camera_id"]),
                               int(ex["last_camera_id"]), p,
                               lambda: active0, int(ex.get("first_frame", 0)))
    assimilate = (str(da_db.get("assimilate_data", "false")) == "true" and
                  cloud is not None)
    da = None
    if assimilate:
        da = DataAssimilator(
            str(da_db.get("localization_cutoff_function", "none")),
            float(da_db.get("localization_cutoff_distance", 1e100)))
    sigma = float(ex.get("estimated_uncertainty", 0.0))

    ts = db["time_stepping"]
    dt = float(ts["time_step"])
    duration = simulation_duration(ts, p.beams)
    half_window = 1.01 * dt
    time, da_time = 0.0, -1.0
    frame_index = -1
    n_step = 1
    if restart:
        time, n_step = restart[1], restart[2] + 1
        # the frame counter skips the frames before the restart time, as
        # the reference's does (application/adamantine.hh:1854-1868)
        if stamps:
            while frame_index + 1 < len(stamps[0]) and \
                    stamps[0][frame_index + 1] < time:
                frame_index += 1
    activation_time_end = -1.0
    tol = dt * 1e-6
    run = EnsembleRun(sol, params, names, vals, ensemble=ens)
    pps = []
    if output:
        from .output import PostProcessor
        post = db.get("post_processor", {})
        every_out = int(post.get("time_steps_between_output", 1))
        out_on_da = str(post.get("output_on_data_assimilation",
                                 "true")) == "true"
        pps = [PostProcessor(post, p, j) for j in range(n)]

        def write(step, t):
            for j, pp in enumerate(pps):
                lq, pw = ens.qp_state(j)
                pp.write_thermal_output(
                    step, t, sol[j], _cell_state(lq, pw, p.material.n_states),
                    active[j])
        write(0, time)
    while time < duration:
        if time + dt > duration:
            dt = duration - time     # application/adamantine.hh:2016-2017
        # material deposition, every member (:2054-2128)
        if dep_t and time - activation_time_end > -tol:
            start = bisect_left(dep_t, time - tol)
            activation_time_end = min(time + max(activation_time, dt),
                                      duration)
            end = bisect_left(dep_t, activation_time_end)
            if start < end:
                for j in range(n):
                    b = ens.activate(j, boxes[start:end], dep_cos[start:end],
                                     dep_sin[start:end], new_temperature, sol)
                    new = b >= 0
                    active[j][new] = 1
                    dcos[j][new] = np.asarray(dep_cos[start:end])[b[new]]
                    dsin[j][new] = np.asarray(dep_sin[start:end])[b[new]]
                run.n_activations += 1
        old_time = time
        time = ens.evolve(old_time, dt, 1, sol)
        if assimilate:
            frame_time = math.inf
            if frame_index + 1 < len(stamps[0]):
                if time > da_time + half_window:
                    da_time = stamps[0][frame_index + 1]
                if stamps[0][frame_index + 1] <= time:
                    frame_index = cloud.read_next_frame()
                    frame_time = stamps[0][frame_index]
            if frame_time <= time:
                if not (frame_time > old_time or n_step == 1):
                    raise RuntimeError("Unexpectedly missed a data "
                                       "assimilation frame.")
                pv = cloud.get_points_values()
                if len(pv.values) > 0:
                    da.update_dof_mapping(ens, np.array(pv.points, float))
                    R = np.eye(len(pv.values)) * sigma * sigma
                    da.update_ensemble(ens, sol, pv.values, R,
                                       params=params if names else None)
                    run.times.append(time)
                    if verbose:
                        print("data assimilation at t = %g" % time,
                              flush=True)
                    # new parameters into the members' sources
                    for j in range(n):
                        eta = members[j].absorption_efficiency
                        power = members[j].max_power
                        for i, nm in enumerate(names):
                            if nm == "beam_0_absorption":
                                eta[0] = float(params[j, i])
                            else:
                                power[0] = float(params[j, i])
                        if names:
                            ens.set_beam_parameters(j, 0, eta[0], power[0])
        if every_ck and n_step % every_ck == 0:
            prefix = str(ck["filename_prefix"])
            if str(ck.get("overwrite_files", "true")) != "true":
                prefix += "_%d" % n_step
            for j in range(n):
                lq, pw = ens.qp_state(j)
                on = active[j].astype(bool)
                np.savez("%s_%d.npz" % (prefix, j),
                         temperature=sol[j].cpu().numpy(),
                         cell_state=_cell_state(lq, pw, p.material.n_states),
                         active=active[j],
                         deposition_cos=np.where(on, dcos[j], 10.0),
                         deposition_sin=np.where(on, dsin[j], 10.0),
                         has_melted=np.zeros(nc, np.uint8),
                         liquid_qp=lq, powder_qp=pw,
                         face_powder=ens.face_state(j))
            with open(prefix + "_time.txt", "w") as f:
                f.write("%r %d\n" % (time, n_step))
        if pps and (n_step % every_out == 0 or
                    (out_on_da and da_time - half_window < time <
                     da_time + half_window)):
            write(n_step, time)
        n_step += 1
    for pp in pps:
        pp.write_pvd()
    run.n_steps = n_step - 1
    run.time = time
    return run
