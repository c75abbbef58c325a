"""Benchmark of the MI355X matrix-free heat path.

Workload (metric "FP64 vmult DoF/s & % HBM peak"): the 128x64x32-cell Q2 box
4x2x1 mm (2,154,945 nodes), solid/liquid/powder with 8-breakpoint k(T) and
c_p(T) tables, Goldak source on, radiative + convective top face, FP64.
A step is one application of the thermal operator (vmult: cell kernel +
boundary-face kernel) to a resident temperature field; value = nodes
processed per second. The explicit RK4 step time of the same workload is
reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6
# names of the cell-kernel launch modes (adm_set_launch_config)
LAUNCH_MODE_NAMES = {0: "line_atomic", 1: "line_colored", 2: "plane"}
# HBM bytes of one operator application of this workload per launch mode
# (cell kernel + face kernel), from rocprofv3 FETCH_SIZE and WRITE_SIZE runs
# scaled by a 256 MiB copy; the profile named beside each value holds the
# command and the counter values. None: not profiled.
MEASURED_TRAFFIC_BYTES = {
    "line_colored": 355.6e6,
    "line_atomic": 176.0e6,
    "plane": 127.1e6,
}
TRAFFIC_PROFILE = {
    "line_colored": "profiles/hbm_traffic_plain_large.md",
    "line_atomic": "profiles/hbm_traffic_plain_large.md",
    "plane": "profiles/hbm_traffic_plane_plain_large.md",
}


def algorithmic_bytes(problem, n_powder_qp: int, n_active_cells: int) -> int:
    """Bytes one operator application must move: read src and write dst
    (8 + 8 per node), the int32 cell-to-node table of a general mesh
    (4 per cell node; this structured-box implementation derives node
    indices arithmetically but the figure is kept as the reference's
    data-structure cost), read + write of the powder ratio at powder
    quadrature points (16 each) and one activity/material byte per cell."""
    n = problem.n_dofs
    nodes_per_cell = (problem.fe_degree + 1) ** 3
    return (16 * n + 4 * n_active_cells * nodes_per_cell + 16 * n_powder_qp +
            problem.n_cells_total)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_cpu_vmults(problem, threads: int, seconds: float):
    """Whole operator applications of the CPU oracle on `threads` host
    threads for about `seconds`: (applications, elapsed s)."""
    from oracle import oracle as orc
    from adamantine_amd.workloads import random_temperature
    orc.set_threads(threads)
    o = orc.Oracle(problem)
    T = random_temperature(problem, low=300.0, high=1600.0)
    o.set_time(1e-4)
    o.vmult(T)  # warm-up
    n = 0
    t0 = time.perf_counter()
    while True:
        o.vmult(T)
        n += 1
        el = time.perf_counter() - t0
        if el > seconds:
            return n, el


def cpu_baseline(problem, seconds: float = 12.0):
    """The CPU oracle (C restatement, OpenMP) on the host cores, timed on
    whole operator applications of the same workload, once on every
    thread the process may use and once on one core."""
    threads = len(os.sched_getaffinity(0)) if hasattr(
        os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(threads, int(os.environ.get("OMP_NUM_THREADS",
                                                     threads))))
    n, el = time_cpu_vmults(problem, threads, seconds)
    n1, el1 = time_cpu_vmults(problem, 1, seconds / 2)
    return {"value": problem.n_dofs * n / el, "unit": "DoF/s",
            "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "one_core": {"value": problem.n_dofs * n1 / el1, "unit": "DoF/s",
                         "cores": 1, "sample": "%d full vmults, %.1f s"
                                               % (n1, el1)},
            "sample": "%d full vmults of the 2,154,945-node Q2 workload on "
                      "%d host threads (%.1f s)" % (n, threads, el)}


def copy_calibration(steps: int):
    """A device-to-device copy of 256 MiB (known bytes: 256 MiB read +
    256 MiB written per launch), the reference point for the HBM counters
    and the measured copy bandwidth."""
    import torch
    n = 256 * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    return {"copy_bytes": 2 * n * 8, "copy_ms": ms,
            "copy_GBs": 2 * n * 8 / (ms * 1e-3) / 1e9}


def ensemble_bench(n_members: int, n_steps: int):
    """The ensemble workload (n_members members of the 64x32x16 Q2 box):
    batched RK4 steps of all members against the serial loop of one
    single-member solver per member (the reference's run_ensemble
    schedule), both on the device."""
    import torch
    from adamantine_amd.ensemble import EnsembleDevice, member_problem
    from adamantine_amd.thermal import ThermalPhysicsDevice
    from adamantine_amd.workloads import ensemble_workload
    base, members = ensemble_workload(n_members=n_members)
    stream = torch.cuda.current_stream()
    dt = 5e-7
    ens = EnsembleDevice(base, members)
    ens.compute_inverse_mass_matrix()
    sol = ens.initialize_dof_vector(300.0)
    t = ens.evolve(0.0, dt, 3, sol)
    ens_launch = ens.autotune(sol, reps=3)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    ens.evolve(t, dt, n_steps, sol)
    e1.record(stream)
    torch.cuda.synchronize()
    batched_ms = e0.elapsed_time(e1) / n_steps
    analysis = analysis_bench(ens, sol, base, members, stream)
    analysis["cycle_ms"] = 50 * batched_ms + analysis["analysis_ms"]
    analysis["cycle"] = "50 batched RK4 steps + 1 analysis"
    del ens, sol
    # serial loop: every member's own solver, one step each, member by member
    n_loop = min(n_members, 16)
    phs, sols = [], []
    for m in members[:n_loop]:
        ph = ThermalPhysicsDevice(member_problem(base, m))
        ph.setup()
        phs.append(ph)
        sols.append(ph.initialize_dof_vector(300.0))
    for ph, s in zip(phs, sols):
        ph.evolve(0.0, dt, 3, s)
    torch.cuda.synchronize()
    k_loop = max(2, n_steps // 2)

    def time_loop(t_start):
        e0.record(stream)
        for k in range(k_loop):
            for ph, s in zip(phs, sols):
                ph.evolve(t_start + k * dt, dt, 1, s)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / (k_loop * n_loop)
    # the single-member solvers with the library's default launches, then
    # with their own best launch configuration
    default_ms = time_loop(3 * dt)
    loop_launch = phs[0].operator.autotune(sols[0], reps=5)
    for ph in phs[1:]:
        ph.operator.set_launch_config(loop_launch["mode"],
                                      loop_launch["waves_per_block"])
    loop_ms = time_loop((3 + k_loop) * dt)
    batched = n_members / (batched_ms * 1e-3)
    loop = 1.0 / (loop_ms * 1e-3)
    return {"members": n_members, "n_dofs_per_member": base.n_dofs,
            "method": "rk4", "batched_step_ms": batched_ms,
            "enkf": analysis,
            "member_steps_per_s": batched,
            "loop_member_steps_per_s": loop,
            "speedup_vs_loop": batched / loop,
            "loop_default_launch_member_steps_per_s": 1.0 / (default_ms *
                                                             1e-3),
            "speedup_vs_default_launch_loop": batched * default_ms * 1e-3,
            "loop_sample": "%d members x %d steps" % (n_loop, k_loop),
            "launch": {"mode": LAUNCH_MODE_NAMES[ens_launch["mode"]],
                       "waves_per_block": ens_launch["waves_per_block"]},
            "loop_launch": {"mode": LAUNCH_MODE_NAMES[loop_launch["mode"]],
                            "waves_per_block":
                                loop_launch["waves_per_block"]}}


def analysis_bench(ens, sol, base, members, stream, reps: int = 5):
    """The EnKF analysis of the ensemble workload: 500 observations on the
    scanned top surface, R = 4 I, absorption efficiency and conductivity
    scale as augmented parameters, no localisation (update_ensemble,
    source/DataAssimilator.cc:95-189), timed whole (host solve included);
    and the member transform X <- X Tm by itself (ensemble_transform_kernel,
    2 N^2 n flops over 2 N n 8 bytes: 8 flop/B at N = 64), with HIP events
    on the stream it runs on."""
    import numpy as np
    import torch
    from adamantine_amd.ensemble import DataAssimilator, ensemble_transform
    N, n = ens.n_members, ens.n_dofs
    rng = np.random.RandomState(5)
    zt = base.node_coordinates()[:, 2].max() * 0.75
    pts = np.stack([rng.uniform(0.2e-3, 3.8e-3, 500),
                    rng.uniform(0.7e-3, 1.3e-3, 500), np.full(500, zt)],
                   axis=1)
    da = DataAssimilator()
    obs = da.update_dof_mapping(ens, pts)
    X = sol.clone()
    y = X[:, torch.tensor(obs, device=X.device)].mean(dim=0).cpu().numpy()
    y = y + rng.normal(scale=2.0, size=obs.size)
    R = 4.0 * np.eye(obs.size)
    params0 = np.ascontiguousarray(np.stack(
        [[m.absorption_efficiency[0] for m in members],
         [m.conductivity_scale for m in members]], axis=1))
    noise = da.fill_noise_vectors(N, R)
    times = []
    for _ in range(reps + 1):
        X.copy_(sol)
        p = params0.copy()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        da.update_ensemble(ens, X, y, R, params=p, noise=noise, stream=stream)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    analysis_ms = 1e3 * float(np.median(times[1:]))
    Tm = np.eye(N) + 1e-3 * rng.standard_normal((N, N))
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    ensemble_transform(ens, X, Tm, stream=stream)
    torch.cuda.synchronize()
    k = 20
    e0.record(stream)
    for _ in range(k):
        ensemble_transform(ens, X, Tm, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    tr_ms = e0.elapsed_time(e1) / k
    flops = 2.0 * N * N * n
    bytes_ = 2.0 * N * n * 8
    return {"observations": int(obs.size), "augmented_parameters": 2,
            "analysis_ms": analysis_ms,
            "transform_ms": tr_ms,
            "transform_TFLOPs": flops / (tr_ms * 1e-3) / 1e12,
            "transform_GBs": bytes_ / (tr_ms * 1e-3) / 1e9,
            "transform_flop_per_byte": flops / bytes_,
            "transform_frac_fp64_mfma": flops / (tr_ms * 1e-3) / 1e12 /
            FP64_PEAK_TFLOPS,
            "transform_frac_hbm": bytes_ / (tr_ms * 1e-3) / 1e9 /
            HBM_PEAK_GBS}


def adjoint_bench():
    """The adjoint workload (64x32x16 Q2, 200 RK4 steps, checkpoints every
    20): one gradient (forward sweep, then the reverse sweep recomputing each
    segment) against one forward run of the same steps."""
    import numpy as np
    import torch
    from adamantine_amd.adjoint import adjoint_gradient
    from adamantine_amd.thermal import ThermalPhysicsDevice
    from adamantine_amd.workloads import adjoint_workload
    w = adjoint_workload()
    ph = ThermalPhysicsDevice(w.truth)
    ph.setup()
    ph.update_physics_parameters(0, *w.truth_beam)
    sol = ph.initialize_dof_vector(300.0)
    ph.evolve(0.0, w.dt, w.n_steps, sol)
    obs = sol.cpu().numpy()[w.obs_dof] + w.noise
    del ph, sol
    # forward alone
    ph = ThermalPhysicsDevice(w.problem)
    ph.setup()
    sol = ph.initialize_dof_vector(300.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ph.evolve(0.0, w.dt, w.n_steps, sol)
    torch.cuda.synchronize()
    fwd_s = time.perf_counter() - t0
    del ph, sol
    # gradient (fresh state)
    ph = ThermalPhysicsDevice(w.problem)
    ph.setup()
    T0 = ph.initialize_dof_vector(300.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = adjoint_gradient(ph, 0.0, w.dt, w.n_steps, T0, w.obs_dof, obs,
                           checkpoint_every=w.checkpoint_every)
    torch.cuda.synchronize()
    adj_s = time.perf_counter() - t0
    return {"workload": "64x32x16 Q2 (12 layers active, 4 powder), %d RK4 "
                        "steps of %g s, checkpoint every %d, %d observations"
                        % (w.n_steps, w.dt, w.checkpoint_every,
                           w.obs_dof.size),
            "n_dofs": w.problem.n_dofs, "forward_ms": fwd_s * 1e3,
            "gradient_ms": adj_s * 1e3, "gradient_over_forward": adj_s / fwd_s,
            "J": res.J, "grad": [float(g) for g in res.grad],
            "grad_names": ["absorption_efficiency", "conductivity_scale",
                           "beam_radius"]}


def ensemble_adjoint_bench(n_members: int = 32, start: int = 180,
                           n: int = 50, every: int = 10, loop_sample: int = 4):
    """The deposition workload's gradients (32 members of the 96x96x40 Q2
    box, forward Euler 2e-7 s): steps start .. start+n-1 with the members'
    second-layer depositions that fall in them, every member's misfit
    gradient from one batched adjoint run (adm_ensemble_adjoint_gradient,
    checkpoints every `every` steps and at each deposition), against the
    batched forward run of the same steps and against one-member adjoint
    runs of `loop_sample` members (the member loop, scaled to all)."""
    import numpy as np
    import torch
    from adamantine_amd.adjoint import ensemble_adjoint_gradient
    from adamantine_amd.ensemble import EnsembleDevice
    from adamantine_amd.workloads import deposition_workload
    w = deposition_workload(n_members=n_members)
    acts = w.activation_step[:, 1]
    events = [(j, int(acts[j]) - start, w.layer_boxes[1:2], [1.0], [0.0],
               300.0) for j in range(n_members) if start <= acts[j] < start + n]

    def at_start(members):
        ens = EnsembleDevice(w.base, members)
        ens.compute_inverse_mass_matrix()
        sol = ens.initialize_dof_vector(300.0)
        ens.evolve(0.0, w.dt, start, sol)
        return ens, sol
    t_start = 0.0
    for _ in range(start):
        t_start = t_start + w.dt
    # observed: 300 nodes of the layer-0 top (3 rows of 100 along the
    # track), the thermocouple count of the deposition workload
    nn = w.base.n_nodes
    iz = 2 * 4  # top of cell layer 3, Q2
    ix = np.linspace(10, nn[0] - 11, 100).astype(np.int64)
    obs_dof = np.concatenate([ix + nn[0] * (nn[1] // 2 + dy + nn[1] * iz)
                              for dy in (-4, 0, 4)])
    ens, sol = at_start(w.members)
    obs = sol[:, torch.tensor(obs_dof, device=sol.device)].cpu().numpy() + 5.0
    # batched forward of the same steps
    T = sol.clone()
    torch.cuda.synchronize()
    f0 = time.perf_counter()
    t = t_start
    for k in range(n):
        for (j, step, boxes, cs, sn, nt) in events:
            if step == k:
                ens.activate(j, boxes, cs, sn, nt, T)
        t = ens.evolve(t, w.dt, 1, T)
    torch.cuda.synchronize()
    fwd_s = time.perf_counter() - f0
    del ens, T
    torch.cuda.empty_cache()
    ens, sol = at_start(w.members)
    torch.cuda.synchronize()
    g0 = time.perf_counter()
    res = ensemble_adjoint_gradient(ens, t_start, w.dt, n, sol, obs_dof, obs,
                                    events=events, checkpoint_every=every)
    torch.cuda.synchronize()
    grad_s = time.perf_counter() - g0
    del ens, sol, res
    torch.cuda.empty_cache()
    # the member loop, on a sample of members
    loop_s = 0.0
    picks = [j for j in range(n_members) if any(e[0] == j for e in events)]
    picks = (picks + [j for j in range(n_members) if j not in picks])
    picks = picks[:loop_sample]
    for j in picks:
        e1, s1 = at_start([w.members[j]])
        ev1 = [(0,) + e[1:] for e in events if e[0] == j]
        torch.cuda.synchronize()
        l0 = time.perf_counter()
        ensemble_adjoint_gradient(e1, t_start, w.dt, n, s1, obs_dof,
                                  obs[j:j + 1], events=ev1,
                                  checkpoint_every=every)
        torch.cuda.synchronize()
        loop_s += time.perf_counter() - l0
        del e1, s1
    loop_all = loop_s / len(picks) * n_members
    return {"workload": "32 members of the 96x96x40 Q2 box, forward Euler "
                        "2e-07 s, steps %d..%d (%d second-layer depositions "
                        "inside), %d observed nodes, checkpoints every %d"
                        % (start, start + n - 1, len(events), obs_dof.size,
                           every),
            "forward_ms": fwd_s * 1e3, "gradient_ms": grad_s * 1e3,
            "gradient_over_forward": grad_s / fwd_s,
            "member_loop_gradient_ms": loop_all * 1e3,
            "speedup_vs_member_loop": loop_all / grad_s,
            "loop_sample": "%d members, scaled to %d" % (len(picks),
                                                         n_members)}


def deposition_bench(n_members: int = 32, start: int = 180, n: int = 125):
    """The deposition workload (32 members of the 96x96x40 Q2 box, forward
    Euler 2e-7 s): steps start .. start+n-1, in which every member deposits
    its second layer (at step 200 + U{0..100}), timed with the activations
    (cell query, state transfer, face lists and inverse mass on the device:
    adm_ensemble_activate)."""
    import numpy as np
    import torch
    from adamantine_amd.ensemble import EnsembleDevice
    from adamantine_amd.workloads import deposition_workload
    w = deposition_workload(n_members=n_members)
    ens = EnsembleDevice(w.base, w.members)
    ens.compute_inverse_mass_matrix()
    sol = ens.initialize_dof_vector(300.0)
    t = ens.evolve(0.0, w.dt, start, sol)
    torch.cuda.synchronize()
    n_act, act_s = 0, 0.0
    t0 = time.perf_counter()
    for step in range(start, start + n):
        for j in np.nonzero(w.activation_step[:, 1] == step)[0]:
            torch.cuda.synchronize()  # steps so far are not activation time
            a0 = time.perf_counter()
            ens.activate(int(j), w.layer_boxes[1:2], [1.0], [0.0], 300.0,
                         sol)
            torch.cuda.synchronize()
            act_s += time.perf_counter() - a0
            n_act += 1
        t = ens.evolve(t, w.dt, 1, sol)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"workload": "%d members of the 96x96x40 Q2 box (3,017,169 nodes "
                        "each), electron beam, forward Euler %g s, steps "
                        "%d..%d with each member's second-layer deposition"
                        % (n_members, w.dt, start, start + n - 1),
            "member_steps_per_s": n_members * n / el,
            "step_ms": (el - act_s) / n * 1e3,
            "activations": n_act,
            "activation_ms": act_s / max(n_act, 1) * 1e3,
            "wall_s": el}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--launch", default=None,
                    help="fixed launch configuration MODE,WAVES of the "
                         "operator (mode 0 line kernel with atomics, 1 line "
                         "kernel in colour launches, 2 plane kernel; e.g. "
                         "2,4) instead of the autotuner's")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--only", choices=["all", "vmult", "ensemble", "copy",
                                       "adjoint", "deposition",
                                       "ensemble_adjoint"],
                    default="all",
                    help="profiling aid: run one part and print its timing")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--assign", action="store_true",
                    help="time vmult (dst cleared, then dst = A src) instead "
                         "of vmult_add (dst += A src)")
    ap.add_argument("--field", choices=["meltpool", "random"],
                    default="meltpool",
                    help="operator input: a melt pool under the beam (solid "
                         "far away, liquid in the pool) or U(300, 1600) noise")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the operator is not sharded "
              "across GPUs; --gpus must be 1.", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from adamantine_amd.thermal import ThermalOperatorDevice, ThermalPhysicsDevice
    from adamantine_amd.workloads import plain_large, random_temperature

    if args.only == "copy":
        print(json.dumps(copy_calibration(args.steps)), flush=True)
        return 0
    if args.only == "ensemble":
        print(json.dumps(ensemble_bench(args.members, args.steps)), flush=True)
        return 0
    if args.only == "adjoint":
        print(json.dumps(adjoint_bench()), flush=True)
        return 0
    if args.only == "deposition":
        print(json.dumps(deposition_bench()), flush=True)
        return 0
    if args.only == "ensemble_adjoint":
        print(json.dumps(ensemble_adjoint_bench(loop_sample=1)), flush=True)
        return 0

    problem = plain_large()
    op = ThermalOperatorDevice(problem)
    stream = torch.cuda.current_stream()
    from adamantine_amd.workloads import meltpool_temperature
    t_eval = 1e-4  # beam on (its first track)
    field = (meltpool_temperature(problem, t_eval) if args.field == "meltpool"
             else random_temperature(problem, low=300.0, high=1600.0))
    T = torch.tensor(field, device="cuda")
    dst = op.initialize_dof_vector()
    op.set_time(t_eval)
    # launch configuration picked by the library's autotuner (colour
    # launches or atomics, wavefronts per block); phase state left as is
    if args.launch:
        col, waves = (int(v) for v in args.launch.split(","))
        op.set_launch_config(col, waves)
        tuned = {"mode": col, "waves_per_block": waves}
    else:
        tuned = op.autotune(T, reps=10)

    apply = op.vmult if args.assign else op.vmult_add
    for _ in range(args.warmup):
        apply(dst, T)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        apply(dst, T)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(end) / args.steps

    launch = {"mode": LAUNCH_MODE_NAMES[tuned["mode"]],
              "waves_per_block": tuned["waves_per_block"]}
    if args.only == "vmult":
        print(json.dumps({"vmult_ms": ms, "steps": args.steps,
                          "field": args.field, "launch": launch}), flush=True)
        return 0

    # quadrature points of cells that start as powder
    n_powder_qp = int((problem.cell_state[:, 2] > 0).sum()) * 27
    n_active = int(problem.n_cells_total)
    bytes_per_vmult = algorithmic_bytes(problem, n_powder_qp, n_active)
    achieved_gbs = bytes_per_vmult / (ms * 1e-3) / 1e9
    dofs_per_s = problem.n_dofs / (ms * 1e-3)

    # explicit RK4 steps of the same workload (context of the metric)
    ph = ThermalPhysicsDevice(problem)
    ph.setup()
    ph.operator.set_launch_config(tuned["mode"], tuned["waves_per_block"])
    sol = ph.initialize_dof_vector(300.0)
    ph.evolve(0.0, 5e-7, 5, sol)
    torch.cuda.synchronize()
    n_rk = 20
    s0 = torch.cuda.Event(enable_timing=True)
    s1 = torch.cuda.Event(enable_timing=True)
    s0.record(stream)
    ph.evolve(5 * 5e-7, 5e-7, n_rk, sol)
    s1.record(stream)
    torch.cuda.synchronize()
    rk4_ms = s0.elapsed_time(s1) / n_rk

    result = {
        "metric": "FP64 vmult DoF/s & % HBM peak; ensemble member-steps/s "
                  "vs plain-solver loop",
        "value": dofs_per_s,
        "unit": "DoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "128x64x32 Q2 box 4x2x1 mm, solid/liquid/"
                               "powder tables, Goldak 400 W, radiative+"
                               "convective top face (2,154,945 nodes)",
                   "n_dofs": problem.n_dofs, "fe_degree": 2,
                   "step": "one vmult (cell + boundary-face kernels)",
                   "launch": launch},
        "roofline": {"bound": "hbm", "achieved": achieved_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_gbs / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC_BYTES.get(launch["mode"]),
                     "traffic_source": TRAFFIC_PROFILE.get(launch["mode"]),
                     "bytes_per_launch": bytes_per_vmult},
        "rk4_step_ms": rk4_ms,
        "wall_s": wall,
    }
    del ph, sol, op, T, dst
    torch.cuda.empty_cache()
    result["ensemble"] = ensemble_bench(args.members, 10)
    torch.cuda.empty_cache()
    result["adjoint"] = adjoint_bench()
    torch.cuda.empty_cache()
    result["deposition"] = deposition_bench()
    torch.cuda.empty_cache()
    result["ensemble_adjoint"] = ensemble_adjoint_bench()
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(problem, args.cpu_seconds)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
