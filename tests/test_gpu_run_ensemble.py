"""The reference's data-assimilation integration tests on the device path,
with their own input files (tests/golden): an ensemble of three members run
by run_ensemble with point-cloud frames assimilated at the logged times.

integration_data_assimilation (tests/test_integration_da.cc:19-84): the
members' initial temperatures are drawn around 300 K; one observation of
200 K per frame pulls the ensemble, and the average of the members' minimum
temperatures lies in [200, 300).

integration_3D_data_assimilation_augmented
(tests/test_integration_da_augmented.cc:19-61): beam 0's absorption
efficiency is in the augmented state, drawn around 0.4 while the data were
made with 0.3; the analysed average is compared with 0.3 at the reference's
tolerance (5, relative)."""
import os

import numpy as np
import pytest

from adamantine_amd.problem import read_info

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_data_assimilation_bare_plate():
    from adamantine_amd.run import run_ensemble
    print("run_ensemble: bare_plate_L_da.info (4x4x1 Q3, 3 members)",
          flush=True)
    db = read_info(os.path.join(GOLD, "bare_plate_L_da.info"))
    run = run_ensemble(db, GOLD)
    assert run.solution.shape[0] == 3
    assert len(run.times) == 2            # both frames assimilated
    assert abs(run.time - 5e-2) < 1e-12 and run.n_steps == 500
    T = run.solution.cpu().numpy()
    avg_min = T.min(axis=1).mean()
    print("  initial temperatures", run.member_values[
        "materials.initial_temperature"], "average minimum", avg_min,
        flush=True)
    assert 200.0 <= avg_min < 300.0


def test_data_assimilation_augmented_absorption():
    from adamantine_amd.run import run_ensemble
    print("run_ensemble: bare_plate_L_da_augmented.info (8x8x1 Q3)",
          flush=True)
    db = read_info(os.path.join(GOLD, "bare_plate_L_da_augmented.info"))
    run = run_ensemble(db, GOLD)
    assert run.param_names == ["beam_0_absorption"]
    prior = run.member_values["sources.beam_0.absorption_efficiency"]
    post = run.params[:, 0]
    print("  prior", prior, "analysed", post, flush=True)
    # frames at 0.02 + 0.02 and 0.04 + 0.02 s: only the first falls in the
    # 0.05 s run
    assert len(run.times) == 1
    assert abs(post.mean() - 0.3) <= 5.0 * 0.3
    # the analysis moves every member away from its prior towards the data
    assert (post < prior).all()


def test_data_assimilation_while_adding_material():
    """integration_3D_da_point_cloud_add_material
    (tests/test_integration_da.cc:89-161): material deposited along the scan
    path in every member while a point-cloud frame is assimilated; the
    average member minimum lies in [200, 300)."""
    from adamantine_amd.run import run_ensemble
    print("run_ensemble: integration_da_add_material.info (3x3x3 Q2)",
          flush=True)
    db = read_info(os.path.join(GOLD, "integration_da_add_material.info"))
    run = run_ensemble(db, GOLD)
    assert run.solution.shape[0] == 3
    assert run.n_activations > 0 and len(run.times) == 1
    T = run.solution.cpu().numpy()
    # over each member's active nodes, as the reference's vectors hold
    active = run.ensemble.get_inverse_mass_matrix().cpu().numpy() > 0
    avg_min = np.mean([T[j][active[j]].min() for j in range(T.shape[0])])
    print("  activations", run.n_activations, "average minimum", avg_min,
          flush=True)
    assert 200.0 <= avg_min < 300.0


def test_data_assimilation_ray_while_adding_material():
    """integration_3D_da_ray_add_material (tests/test_integration_da.cc:
    164-240): the same input with the frame given as a ray
    ("experiment.format ray"); the ray's hit on member 0's active cells is
    assimilated, and the average member minimum lies in [200, 300)."""
    from adamantine_amd.run import run_ensemble
    print("run_ensemble: integration_da_add_material.info, ray frames",
          flush=True)
    db = read_info(os.path.join(GOLD, "integration_da_add_material.info"))
    db["experiment"]["file"] = \
        "integration_da_add_material_expt_ray_#camera_#frame.csv"
    db["experiment"]["format"] = "ray"
    run = run_ensemble(db, GOLD)
    assert run.solution.shape[0] == 3
    assert run.n_activations > 0 and len(run.times) == 1
    T = run.solution.cpu().numpy()
    active = run.ensemble.get_inverse_mass_matrix().cpu().numpy() > 0
    avg_min = np.mean([T[j][active[j]].min() for j in range(T.shape[0])])
    print("  activations", run.n_activations, "average minimum", avg_min,
          flush=True)
    assert 200.0 <= avg_min < 300.0


def test_run_ensemble_writes_vtu(tmp_path):
    """per-member output of the ensemble loop (application/adamantine.hh:
    1872-1885, 2416-2441): with a long output period, the initial state and
    the steps around the assimilation frame; with a period of one, every
    step, the last one holding each member's own nodal temperatures at the
    written corners; a time series per member"""
    from adamantine_amd.output import read_vtu
    from adamantine_amd.run import run_ensemble

    def go(sub, period):
        out = tmp_path / sub
        out.mkdir()
        db = read_info(os.path.join(GOLD, "integration_da_add_material.info"))
        db["post_processor"] = {"filename_prefix": "ens",
                                "output_dir": str(out) + "/",
                                "time_steps_between_output": str(period)}
        run = run_ensemble(db, GOLD, output=True)
        names = sorted(os.listdir(out))
        steps = sorted({int(x.split(".")[1]) for x in names
                        if x.startswith("ens_m0.") and x.endswith(".vtu")})
        return run, out, names, steps

    run, out, names, steps = go("da", 100000)
    print("  member 0 wrote steps", steps, "of", run.n_steps, "DA at",
          run.times, flush=True)
    assert steps[0] == 0 and len(steps) >= 2 and len(run.times) == 1
    run, out, names, steps = go("all", 1)
    n = run.solution.shape[0]
    for j in range(n):
        assert "ens_m%d.pvd" % j in names
    assert steps == list(range(run.n_steps + 1))
    p = run.ensemble.base
    T = run.solution.cpu().numpy()
    h = np.array(p.extent) / np.array(p.n_cells) / p.fe_degree
    nn = np.array(p.n_nodes)
    for j in range(n):
        d = read_vtu(str(out / ("ens_m%d.%d.0.vtu" % (j, run.n_steps))))
        idx = np.rint((d["points"] - np.array(p.origin)) / h).astype(int)
        node = idx[:, 0] + nn[0] * (idx[:, 1] + nn[1] * idx[:, 2])
        nz = d["temperature"] != 0.0
        assert nz.any()
        np.testing.assert_array_equal(d["temperature"][nz], T[j][node[nz]])


def test_run_ensemble_checkpoint_restart(tmp_path):
    """the ensemble loop's checkpoint and restart (application/adamantine.hh:
    1630-1795, 2380-2400) on the DA-while-adding-material input: a run that
    checkpoints every 5 steps (files kept per step), and the same input
    restarted from its step-5 checkpoint (an activation window and the
    assimilation frame still ahead), end on the same member temperatures
    (the reference's restart bar, 1e-10)"""
    from adamantine_amd.run import run_ensemble
    db = read_info(os.path.join(GOLD, "integration_da_add_material.info"))
    prefix = str(tmp_path / "ens_ck")
    db["checkpoint"] = {"filename_prefix": prefix, "overwrite_files": "false",
                        "time_steps_between_checkpoint": "5"}
    r1 = run_ensemble(db, GOLD)
    T1 = r1.solution.cpu().numpy()
    for j in range(T1.shape[0]):
        assert os.path.exists("%s_5_%d.npz" % (prefix, j))
    with open(prefix + "_5_time.txt") as f:
        t_ck, n_ck = f.read().split()
    assert int(n_ck) == 5
    db2 = read_info(os.path.join(GOLD, "integration_da_add_material.info"))
    db2["restart"] = {"filename_prefix": prefix + "_5"}
    r2 = run_ensemble(db2, GOLD)
    T2 = r2.solution.cpu().numpy()
    print("  first run %d steps, DA at %s; restart at step 5 (t = %s): "
          "to step %d, DA at %s" % (r1.n_steps, r1.times, t_ck,
                                       r2.n_steps, r2.times), flush=True)
    assert r2.time == r1.time and r2.times == r1.times
    d = np.abs(T1 - T2)
    print("  largest difference %.3e K" % d.max(), flush=True)
    assert (d <= 1e-10 * np.abs(T1)).all()
