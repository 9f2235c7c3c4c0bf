"""Benchmark of the MI355X VBD solver: vertex-steps/s at a fixed force residual, HBM roofline, CPU baseline.

Prints ONE JSON line.  Its ``value`` is the headline metric: vertex-steps/s of the batched +
learned-warm-start solver at a fixed relative force residual, on the 128x128 wind cloth (top edge
pinned, 8-mode Fourier wind; 16,384 vertices) run as 64 worlds with their own winds (1,048,576
vertices).  One step = one substep (dt 1/600 s) of every world, each world iterated until its
residual is within ``rel_tol`` (0.3) of its value at the inertial guess; vertex-steps/s = vertices x
substeps / wall time (the drivers read the residual on the host, so time is wall-clock with the
stream synchronised).  The warm-start network (BF16 MFMA MLP) is trained on the GPU in the same run
from plain-VBD solves.  The same line carries
* ``comparison``: from identical states per substep, plain VBD solving the worlds one at a time (the
  plain GPU VBD of one cloth), plain VBD on the batch, and the batch with the learned warm start;
* ``tolerance_sweep``: the same comparison at tighter tolerances (0.1, 0.03; iteration cap 400);
* ``fp32_floor``: where plain VBD's relative residual stops falling on one world (3,000 iterations);
* ``roofline``: the color-pass kernel on the headline model, in-frame (iterations captured in a HIP
  graph, HIP events) against 8 TB/s, plus HBM traffic from two rocprofv3 --pmc child runs
  (FETCH_SIZE, WRITE_SIZE) scaled by a 256 MiB copy;
* ``single_world``: the single 128x128 cloth with each warm-start protocol, and the training record;
* ``fixed_iterations``: the reference's configured loop (10 iterations) on the 256x256
  cloth-on-sphere workload, timed while falling and after landing on the sphere;
* ``batched`` / ``batched_tets`` / ``springs`` / ``self_contact``: the 4096-world cloth, the 512 tet
  cubes, springs on every edge and hinge, and two stacked sheets with self-contact;
* ``cpu_baseline``: the FP32 CPU oracle (OpenMP, and single-threaded) on a bounded sample of the
  headline workload, and the FP64 residual of the FP32 floor iterate.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
MLP_FLOPS_PER_ROW = 2 * (48 * 128 + 2 * 128 * 128 + 128 * 3)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30,
                    help="timed substeps of the headline (every world of the batch driven to --rel-tol)")
    ap.add_argument("--warmup", type=int, default=20, help="untimed substeps of the headline batch before timing")
    ap.add_argument("--rel-tol", type=float, default=0.3, help="relative force residual of the headline")
    ap.add_argument("--tol-sweep", type=lambda t: [float(x) for x in t.split(",") if x], default=[0.1, 0.03],
                    help="further relative residuals the three solvers are compared at (comma separated)")
    ap.add_argument("--sweep-substeps", type=int, default=10, help="substeps per tolerance of the sweep")
    ap.add_argument("--max-iterations", type=int, default=400, help="iteration cap of every residual-driven solve")
    ap.add_argument("--check-every", type=int, default=1,
                    help="iterations between residual checks of every residual-driven solve (a per-world residual "
                         "evaluation costs about one iteration; a solve stops at the first check within tolerance)")
    ap.add_argument("--floor-iterations", type=int, default=3000,
                    help="plain-VBD iterations of the FP32-floor measurement (one substep of one world)")
    ap.add_argument("--residual-worlds", type=int, default=64, help="worlds of the headline batch")
    ap.add_argument("--single-world-stride", type=int, default=4,
                    help="headline: every k-th world is also solved alone (the plain single-world baseline)")
    ap.add_argument("--trial-iterations", type=int, default=1,
                    help="iterations each start runs before the residual judges it (the trial-judged variants)")
    ap.add_argument("--delta-sweeps", type=int, default=32,
                    help="1-ring low-pass sweeps of the network's change of the carried correction (deployed)")
    ap.add_argument("--smooth-sweeps", type=int, default=64,
                    help="1-ring low-pass sweeps of the deployed warm-start candidates")
    ap.add_argument("--blend", type=float, default=1.0, help="weight of the network's correction (deployed)")
    ap.add_argument("--residual-frames", type=int, default=3, help="evaluation frames of the single-world leg")
    ap.add_argument("--train-frames", type=int, default=40)
    ap.add_argument("--train-iterations", type=int, default=400, help="plain-VBD iterations per training substep")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--train-smooth", type=float, default=100.0,
                    help="weight of the 1-ring smoothness term of the warm-start training loss")
    ap.add_argument("--train-target", choices=("converged", "tolerance"), default="converged",
                    help="training target of the cloth warm start: the correction after train_iterations "
                         "iterations, or the one the tolerance solve stops at")
    ap.add_argument("--train-protocol", choices=("deployed", "converged"), default="deployed",
                    help="trajectory of the training data: the deployed tolerance solve, or converged solves")
    ap.add_argument("--verts", type=int, default=256, help="cloth vertices per side of the fixed-iteration leg")
    ap.add_argument("--iterations", type=int, default=10, help="iterations of the fixed-iteration leg")
    ap.add_argument("--fixed-steps", type=int, default=20, help="timed frames per window of the fixed-iteration leg")
    ap.add_argument("--fixed-warmup", type=int, default=5, help="first timed frame of its falling window")
    ap.add_argument("--contact-frame", type=int, default=60,
                    help="first timed frame of its on-sphere window (the cloth lands at about frame 27)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU-baseline time budget (0 disables)")
    ap.add_argument("--roofline-reps", type=int, default=20)
    ap.add_argument("--no-traffic", dest="traffic", action="store_false",
                    help="skip the rocprofv3 --pmc child runs that measure HBM traffic per color pass")
    ap.add_argument("--worlds", type=int, default=4096, help="worlds of the batched cloth leg (0 disables the leg)")
    ap.add_argument("--batched-steps", type=int, default=4)
    ap.add_argument("--pmc-mode", action="store_true",
                    help="counter runs: color passes of one prepared substep plus a plain 256 MiB device copy "
                         "(the calibration kernel for FETCH_SIZE / WRITE_SIZE), then exit")
    ap.add_argument("--pmc-workload", choices=("drop", "batched", "cubes", "springs", "wind"), default="drop")
    ap.add_argument("--roofline-only", choices=("wind", "drop", "batched", "cubes"), default=None,
                    help="profiling runs: only the in-frame color-pass roofline of the headline model ('wind', all "
                         "worlds live, after --warmup substeps of 10 iterations), of the 256x256 cloth after "
                         "--contact-frame frames ('drop'), or of the 4096-world cloth / 512 tet cubes after "
                         "--fixed-warmup frames, at a fixed --pass-shape; prints its JSON and exits")
    ap.add_argument("--pass-shape", default="auto",
                    help="color-pass shape of the fixed-iteration and --pmc-mode runs: 'auto' (autotuner) or "
                         "lanes:waves[:aligned], e.g. 256:7:1")
    ap.add_argument("--moving-steps", type=int, default=20,
                    help="timed frames of the moving-body, full-surface contact leg (0 disables)")
    ap.add_argument("--moving-frame", type=int, default=45, help="first timed frame of that leg (cloth landed)")
    ap.add_argument("--only-leg", choices=("moving", "self"), default=None,
                    help="run one secondary leg alone and print its record (for kernel-trace profiles)")
    ap.add_argument("--cube-worlds", type=int, default=512, help="worlds of the batched tet-cube legs (0 disables)")
    ap.add_argument("--cube-train-frames", type=int, default=6,
                    help="frames of plain-VBD tet-cube solves that train the cube warm start (0 disables the leg)")
    ap.add_argument("--cube-train-steps", type=int, default=300)
    ap.add_argument("--cube-warm-frames", type=int, default=30,
                    help="frames of the configured loop before the cube warm start's training data")
    ap.add_argument("--cube-eval-substeps", type=int, default=30)
    ap.add_argument("--cube-trial-iterations", type=int, default=0,
                    help="tet-cube warm start: iterations each start runs before the residual judges it")
    ap.add_argument("--cube-rel-tol", type=float, default=1e-3,
                    help="relative force residual of the tet-cube warm-start leg (mass-dominated blocks converge fast)")
    ap.add_argument("--self-contact-steps", type=int, default=10,
                    help="timed frames of the stacked-sheets self-contact drape (0 disables)")
    ap.add_argument("--self-contact-warm", type=int, default=20, help="frames before timing (sheets land)")
    ap.add_argument("--self-contact-verts", type=int, default=128, help="vertices per side of each sheet")
    ap.add_argument("--spring-steps", type=int, default=10,
                    help="timed frames of the 256x256 cloth with springs on every edge and hinge (0 disables)")
    return ap.parse_args()


# ----------------------------------------------------------------------------- helpers
def _progress(msg: str) -> None:
    """One progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _scene(name, args, dev):
    from newton_amd import scenes

    if name == "drop":
        return scenes.cloth_drop_on_sphere(dev, verts_per_side=args.verts, iterations=args.iterations)
    if name == "cubes":
        return scenes.soft_cubes_batched(dev, worlds=args.cube_worlds)
    if name == "springs":
        return scenes.cloth_drop_on_sphere(dev, verts_per_side=args.verts, iterations=args.iterations, springs=True)
    if name == "wind":
        return scenes.cloth_wind_batched(dev, worlds=args.residual_worlds)
    return scenes.cloth_batched(dev, worlds=args.worlds)


def _contact_counts(model, contacts):
    import numpy as np

    if not contacts.soft_contact_max:
        return None
    c = int(contacts.soft_contact_count.item())
    return np.bincount(contacts.soft_contact_particle.cpu().numpy()[:c], minlength=model.particle_count)


def time_frames(model, state0, info, solver, pipe, contacts, steps, warmup):
    """Graph-captured frames (10 x collide + step); returns (gpu_ms total, wall s, finite, final state)."""
    import torch

    sub, dt = info["substeps"], info["dt"]
    states = [state0, model.state()]

    def frame():
        a, b = states
        for _ in range(sub):
            a.clear_forces()
            pipe.collide(a, contacts)
            solver.step(a, b, None, contacts, dt)
            a, b = b, a
        # an even substep count leaves the result in states[0]

    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        for _ in range(max(warmup, 1)):
            frame()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            frame()
        torch.cuda.synchronize()
        for _ in range(2):
            graph.replay()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            graph.replay()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_ms = ev0.elapsed_time(ev1)
    finite = bool(torch.isfinite(states[0].particle_q).all().item())
    del graph
    return gpu_ms, wall, finite, states[0]


def inframe_roofline(model, state, info, solver, pipe, contacts, iters, reps):
    """Color-pass roofline from the passes as a substep runs them: ``iters`` iterations (every color in
    order, repeated) captured in one HIP graph after contact preparation and the forward step, replayed
    ``reps`` times; HIP events on the graph's stream.  The average pass duration is the elapsed time
    over reps x iters x colors launches (the graph's gaps between kernels included), the algorithmic
    bytes are the exact per-color counts of ``roofline.color_pass_bytes`` for the live contacts."""
    import numpy as np
    import torch

    from newton_amd.roofline import color_pass_bytes

    dt = info["dt"]
    ncol = len(model.particle_color_groups)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        state.clear_forces()
        if pipe is not None:
            pipe.collide(state, contacts)
        solver.begin_step(state, contacts, dt)
        solver.iterate(state, dt, 1)
        torch.cuda.synchronize()
        per_color_bytes = color_pass_bytes(model, contacts_per_particle=_contact_counts(model, contacts)
                                           if contacts is not None else None)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            solver.iterate(state, dt, iters)
        graph.replay()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            graph.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
    del graph
    launches = reps * iters * ncol
    total_bytes = float(per_color_bytes.sum()) * iters * reps
    achieved = total_bytes / (ms / 1e3) / 1e9
    return {
        "kernel": "k_solve_fast",
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "traffic_detail": None,
        "bytes_per_launch": round(float(per_color_bytes.mean()), 1),
        "avg_launch_us": round(ms * 1e3 / launches, 3),
        "launches_timed": launches,
        "timing": f"{iters} iterations x {ncol} colors captured in one HIP graph, replayed {reps} times "
                  "(HIP events; inter-kernel gaps of the graph included)",
    }


def pmc_mode(args):
    """Minimal workload for rocprofv3 --pmc: `reps` launches of each color pass and one known-size copy."""
    import torch

    from newton_amd import CollisionPipeline
    from newton_amd.roofline import color_pass_bytes
    from newton_amd.solvers import SolverVBD

    dev = "cuda:0"
    model, state, info = _scene(args.pmc_workload, args, dev)
    solver = SolverVBD(model, iterations=info["iterations"])
    pipe = CollisionPipeline(model, soft_contact_margin=0.01)
    contacts = pipe.contacts()
    state.clear_forces()
    pipe.collide(state, contacts)
    if args.pass_shape == "auto":
        tune = solver.autotune(state, contacts, info["dt"])
    else:
        parts = [int(x) for x in args.pass_shape.split(":")]
        lanes, waves, aligned = parts[0], parts[1], (parts[2] if len(parts) > 2 else 0)
        solver.set_block_shape(lanes)
        solver.set_pass_waves(waves)
        solver.set_aligned_items(bool(aligned))
        tune = {"block": lanes, "waves": waves, "aligned": aligned, "fixed": True}
    solver.begin_step(state, contacts, info["dt"])
    ncol = len(model.particle_color_groups)
    for _ in range(args.roofline_reps):
        for c in range(ncol):
            solver.color_pass(state, c, info["dt"])
    nbytes = 256 << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    b = color_pass_bytes(model, contacts_per_particle=_contact_counts(model, contacts))
    print(json.dumps({"pmc_mode": True, "workload": info["workload"], "autotune": tune,
                      "color_pass_launches": args.roofline_reps * ncol,
                      "algorithmic_bytes_per_color_pass": [int(x) for x in b], "copy_bytes": nbytes}))


def roofline_only(args):
    """The in-frame roofline alone, for kernel-trace profiles: no autotuning (``--pass-shape`` is
    required), so the only color-pass launches are the measured ones and one prepared substep's."""
    import torch

    from newton_amd import CollisionPipeline
    from newton_amd.solvers import SolverVBD

    if args.pass_shape == "auto":
        raise SystemExit("--roofline-only needs a fixed --pass-shape lanes:waves:aligned")
    dev = "cuda:0"
    model, state, info = _scene(args.roofline_only, args, dev)
    parts = [int(x) for x in args.pass_shape.split(":")]
    iters = {"wind": 10, "drop": args.iterations}.get(args.roofline_only, info["iterations"])
    solver = SolverVBD(model, iterations=iters)
    solver.set_pass_shape(parts[0], parts[1], parts[2] if len(parts) > 2 else 0)
    pipe = CollisionPipeline(model, soft_contact_margin=0.01)
    contacts = pipe.contacts() if pipe.soft_rigid_contact_pair_count else None
    s0, s1 = state, model.state()
    if args.roofline_only == "wind":
        wind = info["wind"]
        for _ in range(args.warmup):
            s0.particle_f.copy_(wind.force(s0.particle_q, model.particle_mass))
            solver.step(s0, s1, None, None, info["dt"])
            s0, s1 = s1, s0
        s0.particle_f.copy_(wind.force(s0.particle_q, model.particle_mass))
    else:
        frames = args.contact_frame if args.roofline_only == "drop" else args.fixed_warmup
        for _ in range(frames * info["substeps"]):
            s0.clear_forces()
            pipe.collide(s0, contacts)
            solver.step(s0, s1, None, contacts, info["dt"])
            s0, s1 = s1, s0
    torch.cuda.synchronize()
    roof = inframe_roofline(model, s0, info, solver, pipe if contacts is not None else None, contacts, iters,
                            args.roofline_reps)
    kernel = f"k_solve_fast<{parts[0]}, {parts[1]}>"
    print(json.dumps({"roofline_only": args.roofline_only, "workload": info["workload"], "kernel": kernel,
                      "pass_shape": args.pass_shape, **roof}))


def measure_traffic(args, workload="drop", timeout_s: float = 240.0, shape: str | None = None):
    """HBM bytes per color-pass launch from two rocprofv3 --pmc child runs (FETCH_SIZE, then WRITE_SIZE).

    Each counter is scaled by what it reads for the 256 MiB device copy made in
    the same run (bytes copied / counter bytes), so the figure is in real bytes.
    ``shape`` ("lanes:waves:aligned") fixes the pass shape of the child run (default: the
    fixed-iteration leg's ``--pass-shape`` for "drop", the autotuner's choice otherwise).
    Returns (bytes_per_launch, details) or (None, reason).
    """
    import csv
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None, "rocprofv3 not found"
    out = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter)
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "pmc", "--", sys.executable,
                   str(ROOT / "bench.py"), "--pmc-mode", "--pmc-workload", workload, "--verts", str(args.verts),
                   "--iterations", str(args.iterations), "--worlds", str(args.worlds), "--cube-worlds",
                   str(args.cube_worlds), "--residual-worlds", str(args.residual_worlds), "--roofline-reps", "5",
                   "--pass-shape", shape or (args.pass_shape if workload == "drop" else "auto")]
            try:
                proc = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s, cwd="/tmp",
                                      env={**os.environ, "TMPDIR": "/tmp"})
            except subprocess.TimeoutExpired:
                return None, f"rocprofv3 --pmc {counter} timed out"
            if proc.returncode != 0:
                return None, f"rocprofv3 --pmc {counter} failed ({proc.returncode})"
            path = os.path.join(d, "pmc_counter_collection.csv")
            if not os.path.exists(path):
                return None, f"no counter output for {counter}"
            rows = list(csv.DictReader(open(path)))
            solve = [float(r["Counter_Value"]) for r in rows if "k_solve_fast" in r["Kernel_Name"]]
            copies = [float(r["Counter_Value"]) for r in rows if r["Kernel_Name"] == "__amd_rocclr_copyBuffer"]
            if not solve or not copies:
                return None, f"counter rows missing for {counter}"
            scale = (256 << 20) / (max(copies) * 1024.0)
            out[counter] = (sum(solve) / len(solve) * 1024.0, scale)
    fetch, fs = out["FETCH_SIZE"]
    write, ws = out["WRITE_SIZE"]
    total = fetch * fs + write * ws
    return total, {"fetch_bytes": round(fetch * fs), "write_bytes": round(write * ws), "fetch_scale": round(fs, 4),
                   "write_scale": round(ws, 4)}


# ----------------------------------------------------------------------------- legs
def fixed_iterations_leg(args):
    """The reference's configured loop on the 256x256 cloth-on-sphere workload (65,536 vertices; one
    frame = 10 x (clear_forces -> collide -> SolverVBD.step with 10 iterations), captured in a HIP
    graph): timed while the cloth falls and again after it has landed on the sphere, each window with
    its in-frame color-pass roofline; HBM traffic per pass from two rocprofv3 --pmc child runs."""
    import torch

    from newton_amd import CollisionPipeline
    from newton_amd.solvers import SolverVBD

    dev = "cuda:0"
    model, state0, info = _scene("drop", args, dev)
    n = model.particle_count
    sub = info["substeps"]
    solver = SolverVBD(model, iterations=args.iterations)
    pipe = CollisionPipeline(model, soft_contact_margin=0.01)
    contacts = pipe.contacts()
    state0.clear_forces()
    pipe.collide(state0, contacts)
    if args.pass_shape == "auto":
        tune = solver.autotune(state0, contacts, info["dt"])
        args.pass_shape = f"{tune['block']}:{tune['waves']}:{tune['aligned']}"  # the counter runs use the timed shape
    else:
        parts = [int(x) for x in args.pass_shape.split(":")]
        solver.set_pass_shape(parts[0], parts[1], parts[2] if len(parts) > 2 else 0)
        tune = {"block": parts[0], "waves": parts[1], "aligned": parts[2] if len(parts) > 2 else 0, "fixed": True}
    windows = {}
    state = state0
    frame = 0
    for name, start in (("falling", args.fixed_warmup), ("on_sphere", args.contact_frame)):
        warm = max(start - frame - 2, 1)  # time_frames runs max(warmup, 1) eager frames and 2 replays first
        _progress(f"fixed-iteration cloth, {name}: frames {frame + warm + 2}..{frame + warm + 2 + args.fixed_steps}")
        gpu_ms, wall, finite, state = time_frames(model, state, info, solver, pipe, contacts, args.fixed_steps, warm)
        first = frame + warm + 2
        frame = first + args.fixed_steps
        pipe.collide(state, contacts)
        torch.cuda.synchronize()
        n_contacts = int(contacts.soft_contact_count.item())
        probe = model.state()
        probe.particle_q.copy_(state.particle_q)
        probe.particle_qd.copy_(state.particle_qd)
        roof = inframe_roofline(model, probe, info, solver, pipe, contacts, args.iterations, args.roofline_reps)
        windows[name] = {"frames": [first, frame], "ms_per_step": round(gpu_ms / args.fixed_steps, 4),
                         "vertex_steps_per_s": round(n * sub * args.fixed_steps / (gpu_ms / 1e3), 1),
                         "soft_contacts": n_contacts, "finite": finite, "wall_s": round(wall, 4), "roofline": roof}
    out = {"workload": info["workload"], "vertices": n, "triangles": model.tri_count,
           "bending_edges": model.edge_count, "colors": len(model.particle_color_groups), "substeps_per_step": sub,
           "iterations": args.iterations, "solver": "plain VBD, the reference's fixed iteration count",
           "hip_graph": True, "steps": args.fixed_steps, "pass_shape": tune, **windows}
    if args.traffic:
        _progress("HBM traffic: two rocprofv3 --pmc child runs")
        traffic, detail = measure_traffic(args, "drop")
        out["falling"]["roofline"]["traffic"] = None if traffic is None else round(traffic, 1)
        out["falling"]["roofline"]["traffic_detail"] = detail
    return out, model, info


def moving_body_leg(args):
    """The 256x256 cloth falling onto a sphere that an external rigid solver moves (sweeping, rising,
    spinning; SolverVBD(integrate_with_external_rigid_solver=True) reads the poses of the step's end and
    start), with the full-surface edge/face contacts on: 10 x (pose update, collide, step with 10
    iterations) per frame, the pose update computed on the device from a device time so the frame is
    captured in one HIP graph; timed after the cloth has landed."""
    import torch

    from newton_amd import CollisionPipeline, scenes
    from newton_amd.solvers import SolverVBD

    dev = "cuda:0"
    model, state0, info = scenes.cloth_on_moving_sphere(dev, verts_per_side=args.verts, iterations=args.iterations)
    n, sub, dt = model.particle_count, info["substeps"], info["dt"]
    pose = info["body_pose"]
    solver = SolverVBD(model, iterations=args.iterations, integrate_with_external_rigid_solver=True)
    pipe = CollisionPipeline(model, soft_contact_margin=0.01, enable_rigid_soft_full_surface_contact=True)
    contacts = pipe.contacts()
    states = [state0, model.state()]
    t = torch.zeros((), dtype=torch.float64, device=dev)

    def frame():
        a, b = states
        for _ in range(sub):
            a.body_q.copy_(pose(t))
            t.add_(dt)
            b.body_q.copy_(pose(t))
            a.clear_forces()
            pipe.collide(a, contacts)
            solver.step(a, b, None, contacts, dt)
            a, b = b, a

    warm = max(args.moving_frame - 2, 1)
    _progress(f"moving body, full surface: frames {warm + 2}..{warm + 2 + args.moving_steps}")
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        for _ in range(warm):
            frame()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            frame()
        for _ in range(2):
            graph.replay()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(args.moving_steps):
            graph.replay()
        ev1.record(stream)
        torch.cuda.synchronize()
        gpu_ms = ev0.elapsed_time(ev1)
    final = states[0]
    pipe.collide(final, contacts)
    torch.cuda.synchronize()
    cnt = min(int(contacts.soft_contact_count.item()), contacts.soft_contact_max)
    idx = contacts.soft_contact_indices[:cnt].cpu().numpy()
    kinds = {"particle": int((idx[:, 1] < 0).sum()), "edge": int(((idx[:, 1] >= 0) & (idx[:, 2] < 0)).sum()),
             "face": int((idx[:, 2] >= 0).sum())}
    q = final.particle_q
    centre = final.body_q[0, :3]
    gap = float((torch.linalg.norm(q - centre, dim=1) - 0.4).min().item())
    del graph
    return {"workload": info["workload"], "vertices": n, "triangles": model.tri_count, "edges": model.edge_count,
            "bodies": model.body_count, "iterations": args.iterations, "substeps_per_step": sub,
            "full_surface": True, "candidate_pairs": {"particle": pipe.soft_rigid_contact_pair_count,
                                                      "edge": len(pipe.soft_edge_rigid_pairs_host),
                                                      "face": len(pipe.soft_face_rigid_pairs_host)},
            "hip_graph": True, "frames": [warm + 2, warm + 2 + args.moving_steps],
            "ms_per_step": round(gpu_ms / args.moving_steps, 3),
            "vertex_steps_per_s": round(n * sub * args.moving_steps / (gpu_ms / 1e3), 1),
            "records_at_end": kinds, "min_vertex_gap_to_sphere_m": round(gap, 5),
            "finite": bool(torch.isfinite(q).all().item())}


def batched_leg(args, name="batched", steps=None):
    """Many worlds in one launch per color pass (4096 cloths or 512 tet cubes), or the headline cloth
    with a spring on every edge and hinge; graph-captured frames; color-pass roofline; the work-group
    shape picked by the autotuner."""
    steps = args.batched_steps if steps is None else steps
    from newton_amd import CollisionPipeline
    from newton_amd.solvers import SolverVBD

    dev = "cuda:0"
    t0 = time.perf_counter()
    model, state, info = _scene(name, args, dev)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    solver = SolverVBD(model, iterations=info["iterations"])
    t_solver = time.perf_counter() - t0
    pipe = CollisionPipeline(model, soft_contact_margin=0.01)
    contacts = pipe.contacts()
    state.clear_forces()
    pipe.collide(state, contacts)
    tune = solver.autotune(state, contacts, info["dt"])
    gpu_ms, wall, finite, last = time_frames(model, state, info, solver, pipe, contacts, steps, 1)
    n = model.particle_count
    roof = inframe_roofline(model, last, info, solver, pipe, contacts, info["iterations"], max(3, args.roofline_reps // 4))
    return {
        "workload": info["workload"],
        "vertices": n,
        "worlds": model.world_count,
        "triangles": model.tri_count,
        "tets": model.tet_count,
        "springs": model.spring_count,
        "autotune": tune,
        "colors": len(model.particle_color_groups),
        "iterations": info["iterations"],
        "substeps_per_step": info["substeps"],
        "steps": steps,
        "ms_per_step": round(gpu_ms / steps, 3),
        "vertex_steps_per_s": round(n * info["substeps"] * steps / (gpu_ms / 1e3), 1),
        "hip_graph": True,
        "finite": finite,
        "host_build_s": round(t_build, 2),
        "solver_create_s": round(t_solver, 2),
        "layout": solver.layout_stats(),
        "roofline": roof,
    }


def self_contact_leg(args):
    """Two stacked cloth sheets draping a sphere with self-contact on: warm-up frames until the sheets
    touch, then graph-captured frames; detection and color-pass (rows + solve + truncation) times from
    HIP events, and the live contact counts."""
    import numpy as np
    import torch

    from newton_amd import CollisionPipeline, scenes
    from newton_amd.solvers import SolverVBD

    dev = "cuda:0"
    model, state, info = scenes.cloth_layers_on_sphere(dev, verts_per_side=args.self_contact_verts)
    solver = SolverVBD(model, iterations=info["iterations"], **info["solver_kwargs"])
    pipe = CollisionPipeline(model, soft_contact_margin=0.01)
    contacts = pipe.contacts()
    n = model.particle_count
    gpu_ms, wall, finite, last = time_frames(model, state, info, solver, pipe, contacts, args.self_contact_steps,
                                             args.self_contact_warm)
    ci = solver.collision_info()
    vt = int(ci["vertex_colliding_triangles_count"].sum().item())
    ee = int(ci["edge_colliding_edges_count"].sum().item())
    dt, reps = info["dt"], 10
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        last.clear_forces()
        pipe.collide(last, contacts)
        solver.begin_step(last, contacts, dt)
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        solver.collision_detection(last)
        e[0].record(stream)
        for _ in range(reps):
            solver.collision_detection(last)
        e[1].record(stream)
        for _ in range(reps):
            for c in range(len(model.particle_color_groups)):
                solver.color_pass(last, c, dt)
        e[2].record(stream)
        torch.cuda.synchronize()
    return {
        "workload": info["workload"],
        "vertices": n,
        "triangles": model.tri_count,
        "edges": model.edge_count,
        "colors": len(model.particle_color_groups),
        "iterations": info["iterations"],
        "substeps_per_step": info["substeps"],
        "self_contact": {k: v for k, v in info["solver_kwargs"].items() if k != "particle_enable_self_contact"},
        "warmup_frames": args.self_contact_warm,
        "steps": args.self_contact_steps,
        "ms_per_step": round(gpu_ms / args.self_contact_steps, 3),
        "vertex_steps_per_s": round(n * info["substeps"] * args.self_contact_steps / (gpu_ms / 1e3), 1),
        "vertex_triangle_contacts": vt,
        "edge_edge_contacts": ee,
        "detection_ms": round(e[0].elapsed_time(e[1]) / reps, 4),
        "iteration_ms": round(e[1].elapsed_time(e[2]) / reps, 4),
        "grid_overflow": solver.self_contact_overflow(),
        "hip_graph": True,
        "finite": finite,
    }


def residual_leg(args):
    """Wind cloth driven to a relative force residual: plain VBD vs the learned warm start.

    The warm-start network is trained in this run, on the GPU, from plain-VBD
    substeps of the same scene (``train_frames`` frames, ``train_iterations``
    iterations each), then both methods solve the next ``residual_frames``
    frames from identical states; each substep stops at ``rel_tol`` (or
    ``max_iterations``).  Times are wall-clock per substep with the stream
    synchronised (the residual checks are host round trips).
    """
    import numpy as np
    import torch

    from newton_amd import scenes
    from newton_amd.solvers import SolverVBD
    from newton_amd.warmstart import LearnedWarmStart

    dev = "cuda:0"
    model, state, info = scenes.cloth_wind(dev)
    wind, dt, sub = info["wind"], info["dt"], info["substeps"]
    n = model.particle_count
    solver = SolverVBD(model, iterations=10)
    ws = LearnedWarmStart(model, dt, seed=0)
    shape = solver.autotune(state, None, dt)  # pass shape of this cloth (results do not depend on it)
    mass = model.particle_mass
    s0, s1 = state, model.state()
    # Training data.  "deployed": along the trajectory the deployed solver follows -- each substep
    # starts from the network-free warm start, iterates to the tolerance (that iterate continues the
    # trajectory and its correction is the next substep's c_prev), then keeps iterating to
    # train_iterations in total; the converged correction is the target.  "converged": the
    # trajectory itself follows the train_iterations solves (c_prev is a converged correction).
    deployed = args.train_protocol == "deployed"
    feats, targs = [], []
    t0 = time.perf_counter()
    for frame in range(args.train_frames):
        wind.redraw(frame)
        if frame % 10 == 0:
            _progress(f"residual leg: training data frame {frame}/{args.train_frames}")
        for _ in range(sub):
            s0.particle_f.copy_(wind.force(s0.particle_q, mass))
            solver.begin_step(s0, None, dt)
            feats.append(ws.features(solver, s0).clone())
            r0 = float(solver.residual(s0, dt).item())
            guess = ws.reuse(solver, s0, dt)
            trial = solver.set_iterate(s0, guess, commit=False)
            if float(solver.residual(_trial(trial), dt).item()) < r0:
                solver.set_iterate(s0, guess, commit=True)
            its = 0
            tol_target = args.train_target == "tolerance"
            if deployed or tol_target:
                its, _ = solver.iterate_to_residual(s0, dt, args.rel_tol * args.rel_tol * r0, args.max_iterations)
                x_tol = s0.particle_q.clone()
            if tol_target:
                targs.append(ws.target(solver, s0))
            if not (tol_target and deployed):
                solver.iterate(s0, dt, max(args.train_iterations - its, 0))
            if not tol_target:
                targs.append(ws.target(solver, s0))
            if deployed:
                solver.set_iterate(s0, x_tol, commit=True)
            ws.observe(solver, s0)
            solver.end_step(s0, s1, dt)
            s0, s1 = s1, s0
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    rows = n * len(feats)
    gen = torch.Generator(device="cpu")
    gen.manual_seed(0)
    batch = n  # one substep's particles per step: the smoothness term couples neighbours
    t0 = time.perf_counter()
    first = None
    loss = None
    for step in range(args.train_steps):
        k = int(torch.randint(0, len(feats), (1,), generator=gen))
        lr = max(1e-3 * 0.5 * (1.0 + math.cos(math.pi * step / args.train_steps)), 1e-5)
        loss = ws.train_substep(feats[k], targs[k], lr=lr, smooth=args.train_smooth)
        if step == 0:
            first = float(loss.item())
        if step % 500 == 0:
            _progress(f"residual leg: training step {step}/{args.train_steps}")
    last = float(loss.item()) if loss is not None else None
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    act = ws.active
    target_ms = float(torch.stack([(t[act] ** 2).mean() for t in targs]).mean().item())
    train_flops = 4 * MLP_FLOPS_PER_ROW * batch * args.train_steps  # forward + training forward/backward
    del feats, targs

    # MFMA roofline of the inference kernel on one substep's rows (the last features gathered) and on
    # the 64-world batch's 1,048,576 rows
    mlp_rates = []
    for rows_mult in (1, 64):
        feat_b = ws.feat if rows_mult == 1 else ws.feat.repeat(rows_mult, 1)
        pred = torch.empty((feat_b.shape[0], 3), dtype=torch.float32, device=dev)
        ws.mlp.forward(feat_b, pred)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            ws.mlp.forward(feat_b, pred)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 20
        mlp_rates.append((feat_b.shape[0], ms, MLP_FLOPS_PER_ROW * feat_b.shape[0] / (ms * 1e-3) / 1e12))
        del feat_b, pred
    fwd_ms, fwd_tflops = mlp_rates[0][1], mlp_rates[0][2]

    variants = ("plain", "reuse", "learned_trial", "learned_unjudged", "parallel_starts", "learned_smoothed",
                "learned")
    # the deployed batch's candidates on the single world: low-passed, judged at the start
    ws_s = LearnedWarmStart(model, dt, mlp=ws.mlp, blend=args.blend, smooth_sweeps=args.smooth_sweeps)
    ws_s.corr = ws.corr
    accepted_smoothed = 0
    # the world's substep solved from three starts at once (inertial, y + c, learned) in a 3-copy model
    from newton_amd.solvers.parallel_starts import ParallelStarts

    model_k, _, _ = scenes.cloth_wind_batched(dev, worlds=2)
    solver_k = SolverVBD(model_k, iterations=10)
    solver_k.autotune(model_k.state(), None, dt)
    ps = ParallelStarts(model_k, solver_k, LearnedWarmStart(model_k, dt, mlp=ws.mlp), copies=2, candidates=(1,))
    ps_wins = [0, 0]
    res = {k: {"iters": [], "s": 0.0, "reached": 0} for k in variants}
    accepted = 0
    accepted_trial = 0
    r_ws_rel = []
    for frame in range(args.residual_frames):
        wind.redraw(args.train_frames + frame)
        for _ in range(sub):
            q0, qd0, corr0 = s0.particle_q.clone(), s0.particle_qd.clone(), ws.corr.clone()
            f_ext = wind.force(s0.particle_q, mass)
            for name in variants:
                s0.particle_q.copy_(q0)
                s0.particle_qd.copy_(qd0)
                s0.particle_f.copy_(f_ext)
                ws.corr.copy_(corr0)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                if name == "parallel_starts":
                    pr = ps.step(s0, s1, ws.corr, dt, args.rel_tol, args.max_iterations)
                    ps_wins[pr["winner"]] += 1
                    r = {"iterations": pr["iterations"], "r": pr["r"], "r0": pr["r0"], "warm_start_index": None,
                         "r_warm_start": None}
                elif name == "learned_unjudged":
                    # ablation outside the judged protocol: the learned start is taken without the veto
                    solver.begin_step(s0, None, dt)
                    r0_sq = float(solver.residual(s0, dt).item())
                    solver.set_iterate(s0, ws(solver, s0, dt)[-1], commit=True)
                    its, r_sq = solver.iterate_to_residual(s0, dt, args.rel_tol * args.rel_tol * r0_sq,
                                                           args.max_iterations, args.check_every)
                    solver.end_step(s0, s1, dt)
                    r = {"iterations": its, "r": math.sqrt(r_sq), "r0": math.sqrt(r0_sq), "warm_start_index": 1,
                         "r_warm_start": None}
                else:
                    start = {"plain": None, "reuse": ws.reuse, "learned": ws, "learned_trial": ws,
                             "learned_smoothed": ws_s}[name]
                    r = solver.step_to_residual(
                        s0, s1, None, None, dt, rel_tol=args.rel_tol, max_iterations=args.max_iterations,
                        check_every=args.check_every, warm_start=start,
                        trial_iterations=args.trial_iterations if name == "learned_trial" else 0)
                torch.cuda.synchronize()
                res[name]["s"] += time.perf_counter() - t0
                res[name]["iters"].append(r["iterations"])
                res[name]["reached"] += int(r["r"] <= args.rel_tol * r["r0"])
                if name == "learned_trial":
                    accepted_trial += int(r["warm_start_index"] == 1)
                if name == "learned_smoothed":
                    accepted_smoothed += int(r["warm_start_index"] == 1)
                if name == "learned":
                    accepted += int(r["warm_start_index"] == 1)
                    if r["r_warm_start"] is not None and r["r0"] > 0:
                        r_ws_rel.append(r["r_warm_start"] / r["r0"])
            ws.observe(solver, s0)  # the learned run's iterate carries on
            s0, s1 = s1, s0
    substeps = args.residual_frames * sub
    out = {"workload": info["workload"], "vertices": n, "rel_tol": args.rel_tol, "max_iterations": args.max_iterations,
           "substeps": substeps, "check_every": 1,
           "pass_shape": {k: shape[k] for k in ("block", "waves", "aligned")}}
    for name in variants:
        d = res[name]
        out[name] = {"mean_iterations": round(float(np.mean(d["iters"])), 2),
                     "reached_tol": f"{d['reached']}/{substeps}",
                     "vertex_steps_per_s": round(n * substeps / d["s"], 1)}
    out["speedup_vertex_steps"] = round(res["plain"]["s"] / res["learned"]["s"], 3)
    out["speedup_vs_reuse"] = round(res["reuse"]["s"] / res["learned"]["s"], 3)
    out["iteration_ratio"] = round(float(np.mean(res["plain"]["iters"]) /
                                         max(float(np.mean(res["learned"]["iters"])), 1e-9)), 3)
    out["learned_candidate_chosen"] = f"{accepted}/{substeps}"
    out["speedup_trial_vertex_steps"] = round(res["plain"]["s"] / res["learned_trial"]["s"], 3)
    out["trial_iterations"] = args.trial_iterations
    out["learned_candidate_chosen_after_trial"] = f"{accepted_trial}/{substeps}"
    out["speedup_smoothed_vertex_steps"] = round(res["plain"]["s"] / res["learned_smoothed"]["s"], 3)
    out["learned_candidate_chosen_smoothed"] = f"{accepted_smoothed}/{substeps}"
    out["smooth_sweeps"] = args.smooth_sweeps
    out["speedup_unjudged_vertex_steps"] = round(res["plain"]["s"] / res["learned_unjudged"]["s"], 3)
    out["speedup_parallel_starts_vertex_steps"] = round(res["plain"]["s"] / res["parallel_starts"]["s"], 3)
    out["parallel_starts_winner"] = {"inertial": ps_wins[0], "network": ps_wins[1]}
    out["learned_blend"] = ws.blend
    out["warm_start_rel_residual_mean"] = round(float(np.mean(r_ws_rel)), 4) if r_ws_rel else None
    out["training"] = {"protocol": args.train_protocol, "frames": args.train_frames, "substeps": args.train_frames * sub,
                       "iterations_per_substep": args.train_iterations, "rows": rows, "steps": args.train_steps,
                       "loss": "mse + smooth * |L e|^2 over one substep per step", "smooth": args.train_smooth,
                       "batch": batch, "loss_first": None if first is None else round(first, 5),
                       "loss_last": None if last is None else round(last, 6),
                       "target_mean_square": round(target_ms, 5), "data_s": round(t_data, 2),
                       "train_s": round(t_train, 2), "train_tflops": round(train_flops / t_train / 1e12, 2)}
    big_rows, big_ms, big_tflops = mlp_rates[1]
    out["mlp_roofline"] = {"kernel": "k_infer (libvbd_ws)", "bound": "mfma", "achieved": round(big_tflops, 2),
                           "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": round(big_tflops / BF16_PEAK_TFLOPS, 5), "rows": big_rows,
                           "avg_launch_us": round(big_ms * 1e3, 2),
                           "flops_per_row": MLP_FLOPS_PER_ROW,
                           "single_world": {"rows": n, "avg_launch_us": round(fwd_ms * 1e3, 2),
                                            "achieved": round(fwd_tflops, 2)}}
    return out, s0, ws


def _solve_variants(args, tol, substeps, frame0, ctx, with_extra=False):
    """``substeps`` substeps of the batched wind cloth, every world driven to relative residual ``tol``
    (cap ``max_iterations``) by each solver on its own trajectory from one common start (the headline
    batch's current state): plain VBD one world at a time (every ``single_world_stride``-th world on
    a single-world solver, its rate taken as the rate for all worlds), plain VBD on the batch, and
    the batch with the learned warm start (``ctx["variants"]`` lists the batched solvers; each keeps
    its own positions, velocities and carried correction).  The winds are the same functions of
    position for every solver.  Times are wall-clock around each solve, stream synchronised.  The
    deployed solver's final state becomes the common start of the next call."""
    import numpy as np
    import torch

    model, wind, info, solver = (ctx[k] for k in ("model", "wind", "info", "solver"))
    one_solver, per, W = (ctx[k] for k in ("one_solver", "per", "W"))
    dt, sub, mass = info["dt"], info["substeps"], model.particle_mass
    one_mass = one_solver.model.particle_mass
    cap = args.max_iterations
    sample = list(range(0, W, max(args.single_world_stride, 1)))
    batched = ctx["variants"] if with_extra else ("plain_batched", ctx["deployed"])
    base = ctx["state"]
    # per-solver trajectories from the common start
    traj = {}
    for name in batched:
        a, b = model.state(), model.state()
        a.particle_q.copy_(base.particle_q)
        a.particle_qd.copy_(base.particle_qd)
        ws = ctx["protocol"][name].get("warm_start")
        if ws is not None:
            ws.corr.copy_(ctx["corr"])
        traj[name] = [a, b]
    single = {w: (base.particle_q[w * per:(w + 1) * per].clone(), base.particle_qd[w * per:(w + 1) * per].clone())
              for w in sample}
    one_in, one_out = ctx["one_state"], ctx["one_out"]
    res = {k: {"iters": [], "witers": [], "s": 0.0, "reached": 0, "chosen": None}
           for k in ("plain_single_world",) + tuple(batched)}
    for i in range(substeps):
        if i % sub == 0:
            wind.redraw(frame0 + i // sub)
        d = res["plain_single_world"]
        for w in sample:
            q, qd = single[w]
            one_in.particle_q.copy_(q)
            one_in.particle_qd.copy_(qd)
            one_in.particle_f.copy_(wind.winds[w].force(q, one_mass))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = one_solver.step_to_residual(one_in, one_out, None, None, dt, rel_tol=tol, max_iterations=cap,
                                            check_every=args.check_every)
            torch.cuda.synchronize()
            d["s"] += time.perf_counter() - t0
            d["iters"].append(r["iterations"])
            d["reached"] += int(r["r"] <= tol * r["r0"])
            q.copy_(one_out.particle_q)
            qd.copy_(one_out.particle_qd)
        for name in batched:
            a, b = traj[name]
            a.particle_f.copy_(wind.force(a.particle_q, mass))
            prot = ctx["protocol"][name]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = solver.step_to_residual_worlds(a, b, None, None, dt, rel_tol=tol, max_iterations=cap,
                                               check_every=args.check_every, **prot)
            torch.cuda.synchronize()
            d = res[name]
            d["s"] += time.perf_counter() - t0
            d["iters"].append(r["iterations"])
            d["witers"].append(r["world_iterations"])
            d["reached"] += r["reached"]
            if r["candidate_chosen"]:
                d["chosen"] = r["candidate_chosen"] if d["chosen"] is None else \
                    [x + y for x, y in zip(d["chosen"], r["candidate_chosen"])]
            if prot.get("warm_start") is not None:
                prot["warm_start"].observe(solver, a)
            traj[name] = [b, a]
    # the deployed solver's state carries on
    dep = ctx["deployed"]
    ctx["state"].particle_q.copy_(traj[dep][0].particle_q)
    ctx["state"].particle_qd.copy_(traj[dep][0].particle_qd)
    ctx["corr"].copy_(ctx["protocol"][dep]["warm_start"].corr)
    n = model.particle_count
    out = {"rel_tol": tol, "substeps": substeps, "max_iterations": cap, "trajectories": "each solver its own"}
    d = res["plain_single_world"]
    single_vs = len(sample) * per * substeps / d["s"]
    out["plain_single_world"] = {"worlds_timed": len(sample), "mean_iterations": round(float(np.mean(d["iters"])), 2),
                                 "reached_tol": f"{d['reached']}/{len(sample) * substeps} world-substeps",
                                 "vertex_steps_per_s": round(single_vs, 1), "seconds": round(d["s"], 4)}
    for name in batched:
        d = res[name]
        wit = np.stack(d["witers"])
        out[name] = {"mean_iterations_launched": round(float(np.mean(d["iters"])), 2),
                     "mean_world_iterations": round(float(wit.mean()), 2),
                     "mean_slowest_world_iterations": round(float(wit.max(axis=1).mean()), 2),
                     "world_substeps_at_cap": int((wit >= cap).sum()),
                     "reached_tol": f"{d['reached']}/{substeps * W} world-substeps",
                     "vertex_steps_per_s": round(n * substeps / d["s"], 1), "seconds": round(d["s"], 4)}
        if d["chosen"] is not None:
            out[name]["world_substeps_choosing"] = dict(zip(ctx["candidate_names"], d["chosen"]))
    for name in batched:
        if name != "plain_batched":
            out[f"speedup_{name}_vs_plain_single_world"] = round(out[name]["vertex_steps_per_s"] / single_vs, 3)
            out[f"speedup_{name}_vs_plain_batched"] = round(res["plain_batched"]["s"] / res[name]["s"], 3)
    out["speedup_batching_only"] = round(out["plain_batched"]["vertex_steps_per_s"] / single_vs, 3)
    return out


def headline_leg(args, s_single, ws):
    """The headline: the 128x128 wind cloth (pinned top edge, Fourier wind) as ``residual_worlds`` worlds
    in one model, every world driven by its own wind and judged by its own force residual.  Every world
    starts from the single-world leg's last state; ``warmup`` substeps of plain VBD at ``rel_tol`` let the
    worlds drift apart (a common state that does not depend on the protocols compared); then each
    solver -- plain VBD one world at a time, plain VBD batched, batched + learned warm start (the
    deployed protocol) and the same protocol without the network -- runs ``steps`` substeps at
    ``rel_tol`` on its own trajectory from that common state, and again at each tolerance of
    ``tol_sweep``; the color-pass roofline of this model is measured in-frame; and one substep shows
    where plain VBD's relative residual stops falling (the FP32 floor)."""
    import numpy as np
    import torch

    from newton_amd import scenes
    from newton_amd.solvers import SolverVBD
    from newton_amd.warmstart import LearnedWarmStart

    dev = "cuda:0"
    W = args.residual_worlds
    model, s0, info = scenes.cloth_wind_batched(dev, worlds=W)
    wind, dt, sub = info["wind"], info["dt"], info["substeps"]
    n = model.particle_count
    s1 = model.state()
    s0.particle_q.copy_(s_single.particle_q.repeat(W, 1))
    s0.particle_qd.copy_(s_single.particle_qd.repeat(W, 1))
    solver = SolverVBD(model, iterations=10)
    tune = solver.autotune(s0, None, dt)
    # deployed: y + c_prev and y + c_prev + S(blend * MLP) (the network's change low-passed by
    # delta_sweeps 1-ring sweeps), judged against the inertial start after trial_iterations VBD
    # iterations each
    ws_dep = LearnedWarmStart(model, dt, mlp=ws.mlp, seed=0, blend=0.5, smooth_sweeps=0,
                              delta_sweeps=args.delta_sweeps)
    # ablation: the same protocol without the network (y + c_prev against the inertial start)
    ws_reuse = LearnedWarmStart(model, dt, mlp=ws.mlp, seed=0, blend=0.0, offer_reuse=False)
    one_model, one_state, _ = scenes.cloth_wind(dev)
    one_solver = SolverVBD(one_model, iterations=10)
    one_solver.autotune(one_state, None, dt)
    corr = ws.corr.repeat(W, 1).contiguous()
    ctx = {"model": model, "solver": solver, "wind": wind, "info": info, "state": s0, "corr": corr,
           "one_state": one_state, "one_out": one_model.state(), "one_solver": one_solver,
           "per": one_model.particle_count, "W": W, "candidate_names": ("reuse", "network"),
           "deployed": "batched_learned",
           "variants": ("plain_batched", "batched_learned", "batched_reuse_only"),
           "protocol": {"plain_batched": {},
                        "batched_learned": {"warm_start": ws_dep, "trial_iterations": args.trial_iterations},
                        "batched_reuse_only": {"warm_start": ws_reuse, "trial_iterations": args.trial_iterations}}}
    frame = args.train_frames + args.residual_frames
    mass = model.particle_mass
    _progress(f"headline: {W} worlds x {one_model.particle_count} vertices, {args.warmup} warm-up substeps")
    ws_dep.corr.copy_(corr)
    for i in range(args.warmup):
        if i % sub == 0:
            wind.redraw(frame + i // sub)
        s0.particle_f.copy_(wind.force(s0.particle_q, mass))
        # plain VBD to the tolerance: the common state does not depend on the protocols compared
        solver.step_to_residual_worlds(s0, s1, None, None, dt, rel_tol=args.rel_tol,
                                       max_iterations=args.max_iterations, check_every=args.check_every)
        ws_dep.observe(solver, s0)
        s0, s1 = s1, s0
    ctx["state"] = s0
    corr.copy_(ws_dep.corr)
    frame += -(-args.warmup // sub)
    _progress(f"headline: {args.steps} timed substeps at rel_tol {args.rel_tol}, each solver on its own trajectory")
    main = _solve_variants(args, args.rel_tol, args.steps, frame, ctx, with_extra=True)
    frame += -(-args.steps // sub)
    sweep = []
    for tol in args.tol_sweep:
        _progress(f"headline: tolerance sweep, rel_tol {tol}, {args.sweep_substeps} substeps")
        sweep.append(_solve_variants(args, tol, args.sweep_substeps, frame, ctx))
        frame += -(-args.sweep_substeps // sub)
    ctx["s0"] = ctx["state"]
    # in-frame color-pass roofline of this model (all worlds live, 10 iterations of one substep)
    probe = model.state()
    probe.particle_q.copy_(ctx["s0"].particle_q)
    probe.particle_qd.copy_(ctx["s0"].particle_qd)
    probe.particle_f.copy_(wind.force(probe.particle_q, mass))
    roof = inframe_roofline(model, probe, info, solver, None, None, 10, max(2, args.roofline_reps // 4))
    roof["pass_shape"] = {k: tune[k] for k in ("block", "waves", "aligned")}
    # where plain VBD's relative residual stops falling: one world's substep iterated far past the tolerances
    st = ctx["one_state"]
    st.particle_q.copy_(ctx["s0"].particle_q[:ctx["per"]])
    st.particle_qd.copy_(ctx["s0"].particle_qd[:ctx["per"]])
    st.particle_f.copy_(wind.winds[0].force(st.particle_q, mass[:ctx["per"]]))
    one_solver.begin_step(st, None, dt)
    r0 = float(one_solver.residual(st, dt).item())
    hist = []
    for _ in range(args.floor_iterations // 50):
        one_solver.iterate(st, dt, 50)
        hist.append(float(one_solver.residual(st, dt).item()))
    rel = np.sqrt(np.asarray(hist) / r0)
    floor = {"workload": "one world of the batch, one substep of plain VBD", "iterations": len(hist) * 50,
             "rel_residual_every_50": [round(float(x), 5) for x in rel],
             "min_rel_residual": round(float(rel.min()), 5),
             "iterations_at_min": int(50 * (int(rel.argmin()) + 1))}
    ctx["floor_state"] = (st.particle_q.clone(), one_solver.scratch("q_prev"), one_solver.scratch("inertia"),
                          st.particle_f.clone())
    return {"workload": info["workload"], "worlds": W, "vertices": n, "stop": "every world at rel_tol",
            "warmup_substeps": args.warmup, "pass_shape": {k: tune[k] for k in ("block", "waves", "aligned")},
            "main": main, "tolerance_sweep": sweep, "roofline": roof, "fp32_floor": floor}, ctx


def cubes_residual_leg(args, ws_cloth):
    """Batched tet cubes with the learned warm start: the 512 cubes (6..12 cells, ground contact) as one
    model, every world judged by its own force residual (relative ``cube_rel_tol``, floored at 1 % of
    the world's weight norm, since a free-falling unstressed cube starts at round-off).  The MLP starts
    from the cloth network and is retrained on the GPU on tet-mesh 1-ring features from
    ``cube_train_frames`` frames of plain-VBD solves (targets: ``train_iterations`` iterations); then
    plain VBD batched and batched + learned warm start (the deployed protocol) each run
    ``cube_eval_substeps`` substeps on their own trajectory from one common state.  Training starts
    after ``cube_warm_frames`` frames of the configured loop, when the cubes reach the ground."""
    import numpy as np
    import torch

    from newton_amd import CollisionPipeline, scenes
    from newton_amd.solvers import SolverVBD
    from newton_amd.warmstart import LearnedWarmStart, WarmStartMLP

    dev = "cuda:0"
    model, s0, info = scenes.soft_cubes_batched(dev, worlds=args.cube_worlds)
    dt, sub, W = info["dt"], info["substeps"], model.world_count
    n = model.particle_count
    solver = SolverVBD(model, iterations=info["iterations"])
    pipe = CollisionPipeline(model, soft_contact_margin=0.01)
    contacts = pipe.contacts()
    s0.clear_forces()
    pipe.collide(s0, contacts)
    tune = solver.autotune(s0, contacts, dt)
    mlp = WarmStartMLP(dev, seed=0)
    mlp.load_params(ws_cloth.mlp.params.cpu().numpy())
    ws = LearnedWarmStart(model, dt, mlp=mlp, seed=0, blend=0.5)
    ps = info["world_particle_start"]
    w_of = np.repeat(np.arange(W), np.diff(ps))
    m = model.particle_mass.cpu().numpy().astype(np.float64)
    abs_tol = 1e-2 * 9.81 * np.sqrt(np.bincount(w_of, weights=m * m, minlength=W))
    tol = args.cube_rel_tol
    s1 = model.state()
    # the configured loop (8 iterations) until the cubes have reached the ground (the drops take up
    # to ~33 frames): training and evaluation then see landings, bounces and resting contact
    for _ in range(args.cube_warm_frames * sub):
        s0.clear_forces()
        pipe.collide(s0, contacts)
        solver.step(s0, s1, None, contacts, dt)
        s0, s1 = s1, s0
    feats, targs = [], []
    t0 = time.perf_counter()
    for frame in range(args.cube_train_frames):
        for _ in range(sub):
            s0.clear_forces()
            pipe.collide(s0, contacts)
            solver.begin_step(s0, contacts, dt)
            feats.append(ws.features(solver, s0).clone())
            r0 = solver.residual_worlds(s0, dt).cpu().numpy()
            target = np.maximum(tol * tol * r0, abs_tol * abs_tol)
            solver.iterate_to_residual_worlds(s0, dt, target, args.max_iterations)
            x_tol, d_tol = s0.particle_q.clone(), solver.scratch("displacement")
            solver.iterate(s0, dt, args.train_iterations // 2)
            targs.append(ws.target(solver, s0))
            s0.particle_q.copy_(x_tol)
            solver.set_displacement(d_tol)
            ws.observe(solver, s0)
            solver.end_step(s0, s1, dt)
            s0, s1 = s1, s0
        _progress(f"cubes residual: training data frame {frame + 1}/{args.cube_train_frames}")
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    gen = torch.Generator(device="cpu")
    gen.manual_seed(1)
    t0 = time.perf_counter()
    first = last = None
    for step in range(args.cube_train_steps):
        k = int(torch.randint(0, len(feats), (1,), generator=gen))
        lr = max(1e-3 * 0.5 * (1.0 + math.cos(math.pi * step / args.cube_train_steps)), 1e-5)
        loss = ws.train_substep(feats[k], targs[k], lr=lr, smooth=args.train_smooth)
        if step == 0:
            first = float(loss.item())
    last = float(loss.item()) if args.cube_train_steps else None
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del feats, targs
    # evaluation: each solver on its own trajectory from the common state
    variants = {"plain_batched": {},
                "batched_learned": {"warm_start": ws, "trial_iterations": args.cube_trial_iterations}}
    base_q, base_qd, base_corr = s0.particle_q.clone(), s0.particle_qd.clone(), ws.corr.clone()
    out = {"workload": info["workload"], "worlds": W, "vertices": n, "tets": model.tet_count, "rel_tol": tol,
           "abs_floor": "1 % of each world's weight norm", "max_iterations": args.max_iterations,
           "substeps": args.cube_eval_substeps, "first_frame": args.cube_warm_frames + args.cube_train_frames,
           "pass_shape": {k: tune[k] for k in ("block", "waves", "aligned")}, "trajectories": "each solver its own"}
    for name, prot in variants.items():
        a, b = model.state(), model.state()
        a.particle_q.copy_(base_q)
        a.particle_qd.copy_(base_qd)
        ws.corr.copy_(base_corr)
        c = pipe.contacts()
        secs, witers, reached, chosen = 0.0, [], 0, None
        for _ in range(args.cube_eval_substeps):
            a.clear_forces()
            pipe.collide(a, c)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = solver.step_to_residual_worlds(a, b, None, c, dt, rel_tol=tol, max_iterations=args.max_iterations,
                                               check_every=args.check_every,
                                               abs_tol=abs_tol, **prot)
            torch.cuda.synchronize()
            secs += time.perf_counter() - t0
            witers.append(r["world_iterations"])
            reached += r["reached"]
            if r["candidate_chosen"]:
                chosen = r["candidate_chosen"] if chosen is None else [x + y for x, y in zip(chosen,
                                                                                             r["candidate_chosen"])]
            if prot:
                ws.observe(solver, a)
            a, b = b, a
        wit = np.stack(witers)
        out[name] = {"mean_world_iterations": round(float(wit.mean()), 2),
                     "mean_slowest_world_iterations": round(float(wit.max(axis=1).mean()), 2),
                     "world_substeps_at_cap": int((wit >= args.max_iterations).sum()),
                     "reached_tol": f"{reached}/{W * args.cube_eval_substeps} world-substeps",
                     "vertex_steps_per_s": round(n * args.cube_eval_substeps / secs, 1), "seconds": round(secs, 4),
                     "finite": bool(torch.isfinite(a.particle_q).all().item())}
        if chosen is not None:
            out[name]["world_substeps_choosing"] = dict(zip(("reuse", "network"), chosen))
    out["speedup_batched_learned_vs_plain_batched"] = round(out["plain_batched"]["seconds"] /
                                                           out["batched_learned"]["seconds"], 3)
    out["training"] = {"frames": args.cube_train_frames, "substeps": args.cube_train_frames * sub,
                       "target_iterations": args.train_iterations // 2, "steps": args.cube_train_steps,
                       "init": "the cloth network", "loss_first": first, "loss_last": last,
                       "data_s": round(t_data, 2), "train_s": round(t_train, 2)}
    return out


class _trial:  # noqa: N801  (positions-only state for a residual evaluation)
    def __init__(self, q):
        self.particle_q = q


def cpu_baseline_leg(args, ctx, fixed_model, fixed_q, fixed_qd, fixed_dt):
    """CPU baseline: the FP32 C oracle on the GPU host's cores.

    Headline workload: plain VBD on one 128x128 wind-cloth world from the batch's last state, each
    substep iterated to ``rel_tol`` (the residual evaluated after every iteration, as the GPU driver
    does), OpenMP over each color's vertices, for about ``cpu_seconds``; also single-threaded (Warp's
    CPU device runs a launch serially) on a shorter sample.  Beside it: the fixed-iteration 256x256
    cloth, and the FP64 residual of the FP32 iterate the FP32 floor measurement ended at."""
    import numpy as np
    import torch

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle  # noqa: PLC0415  (the checker, timed as the CPU baseline)

    cores = min(len(os.sched_getaffinity(0)), 16)
    one_model, per, info = ctx["one_solver"].model, ctx["per"], ctx["info"]
    wind, dt, sub = info["wind"], info["dt"], info["substeps"]
    om = oracle.OracleModel(one_model)
    q = ctx["s0"].particle_q[:per].cpu().numpy().copy()
    qd = ctx["s0"].particle_qd[:per].cpu().numpy().copy()
    mass = one_model.particle_mass

    def run(threads, budget, limit):
        nonlocal q, qd
        prm = oracle.params(args.max_iterations, num_threads=threads, residual_stop=args.rel_tol)
        done, spent, its = 0, 0.0, []
        while spent < budget and done < limit:
            # the world's wind (world 0 of the batch) at the current positions; not timed
            f = (wind.winds[0].acceleration(torch.as_tensor(q, device=mass.device)) * mass[:, None]).cpu().numpy()
            t0 = time.perf_counter()
            _, q, qd, hist = oracle.step(om, prm, q, qd, f, None, dt, residual_history=True)
            spent += time.perf_counter() - t0
            its.append(int(np.count_nonzero(hist[1:])))
            done += 1
        return done, spent, its

    done, spent, its = run(cores, args.cpu_seconds, 200)
    done1, spent1, its1 = run(1, 0.3 * args.cpu_seconds, 50)
    out = {
        "value": round(per * done / spent, 1),
        "unit": "vertex-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{done} substeps of one {info['workload'].split('_')[-1]} wind-cloth world (the headline's "
                  f"world 0, from its last state), plain VBD iterated to relative force residual {args.rel_tol} "
                  f"(cap {args.max_iterations}, mean {np.mean(its):.1f} iterations), FP32 C oracle, "
                  f"OpenMP {cores} threads",
        "single_thread": {"value": round(per * done1 / spent1, 1), "cores": 1,
                          "sample": f"{done1} substeps of the same trajectory, mean {np.mean(its1):.1f} "
                                    "iterations, FP32 C oracle, 1 thread"},
    }
    # the reference's fixed-iteration loop on the 256x256 cloth (10 iterations)
    omf = oracle.OracleModel(fixed_model)
    prm = oracle.params(args.iterations, num_threads=cores)
    qf, qdf = fixed_q, fixed_qd
    nf, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 0.3 * args.cpu_seconds and nf < 100:
        qf, qdf, _ = oracle.simulate(omf, prm, qf, qdf, 1, fixed_dt, 0.01, len(omf.pairs) // 2)
        nf += 1
    out["fixed_iterations"] = {"value": round(fixed_model.particle_count * nf / (time.perf_counter() - t0), 1),
                               "cores": cores, "sample": f"{nf} substeps (collide + step, {args.iterations} "
                               "iterations) of the 256x256 cloth from its initial state"}
    # FP32 floor: the FP32 and FP64 residuals of the FP32 iterate the floor measurement ended at
    if "floor_state" in ctx:
        qe, q_prev, inertia, f_ext = (t.cpu().numpy() for t in ctx["floor_state"])
        prm64 = oracle.params(1, num_threads=cores)
        _, r_end64 = oracle.residual(om, prm64, None, qe, q_prev, inertia, dt, precision="f64")
        _, r0_64 = oracle.residual(om, prm64, None, inertia, q_prev, inertia, dt, precision="f64")
        _, r_end32 = oracle.residual(om, prm64, None, qe, q_prev, inertia, dt, precision="f32")
        _, r0_32 = oracle.residual(om, prm64, None, inertia, q_prev, inertia, dt, precision="f32")
        out["fp32_floor_check"] = {
            "what": "residual of the FP32 iterate after the floor measurement's iterations, evaluated by the "
                    "oracle in FP32 and in FP64 (relative to each precision's residual at the inertial guess)",
            "rel_fp32": round(r_end32 / r0_32, 6), "rel_fp64": round(r_end64 / r0_64, 6)}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU only; --gpus must be 1 (worlds would shard "
                          "across GPUs with no per-step exchange)"}))
        sys.exit(2)
    if args.pmc_mode:
        pmc_mode(args)
        return
    if args.roofline_only:
        roofline_only(args)
        return
    if args.only_leg == "moving":
        import torch

        torch.cuda.set_device(0)
        print(json.dumps({"only_leg": "moving_body_full_surface", **moving_body_leg(args)}))
        return
    if args.only_leg == "self":
        import torch

        torch.cuda.set_device(0)
        print(json.dumps({"only_leg": "self_contact", **self_contact_leg(args)}))
        return
    import torch

    torch.cuda.set_device(0)
    _progress("residual leg: training data, GPU training, single-world comparisons")
    single, s_single, ws = residual_leg(args)
    _progress(f"single world: learned after trial sweeps {single['speedup_trial_vertex_steps']}x plain")
    head, ctx = headline_leg(args, s_single, ws)
    main_rec = head["main"]
    value = main_rec["batched_learned"]["vertex_steps_per_s"]
    ms_per_step = 1e3 * main_rec["batched_learned"]["seconds"] / args.steps
    _progress(f"headline: {value:.4g} vertex-steps/s ({main_rec['speedup_batched_learned_vs_plain_single_world']}x "
              f"plain single world, {main_rec['speedup_batched_learned_vs_plain_batched']}x plain batched)")
    torch.cuda.empty_cache()

    fixed, fixed_model, fixed_info = fixed_iterations_leg(args)
    _progress(f"fixed iterations: {fixed['falling']['vertex_steps_per_s']:.4g} / "
              f"{fixed['on_sphere']['vertex_steps_per_s']:.4g} vertex-steps/s")
    torch.cuda.empty_cache()
    roofline = dict(head["roofline"])
    if args.traffic:
        _progress("HBM traffic of the headline model: two rocprofv3 --pmc child runs")
        traffic, detail = measure_traffic(args, "wind")
        roofline["traffic"] = None if traffic is None else round(traffic, 1)
        roofline["traffic_detail"] = detail

    legs = {}
    for name, enabled, steps in (("batched", args.worlds > 0, None), ("batched_tets", args.cube_worlds > 0, None),
                                 ("springs", args.spring_steps > 0, args.spring_steps)):
        if enabled:
            _progress(f"{name} leg")
            legs[name] = batched_leg(args, {"batched_tets": "cubes"}.get(name, name), steps)
            _progress(f"{name}: {legs[name]['vertex_steps_per_s']:.4g} vertex-steps/s")
            torch.cuda.empty_cache()
            if args.traffic and name in ("batched", "batched_tets"):
                # HBM bytes per pass at the pass shape the leg ran (two rocprofv3 --pmc child runs)
                _progress(f"HBM traffic of the {name} leg: two rocprofv3 --pmc child runs")
                tune = legs[name]["autotune"]
                shape = f"{tune['block']}:{tune['waves']}:{int(tune['aligned'])}"
                traffic, detail = measure_traffic(args, {"batched_tets": "cubes"}.get(name, name), shape=shape)
                legs[name]["roofline"]["traffic"] = None if traffic is None else round(traffic, 1)
                legs[name]["roofline"]["traffic_detail"] = detail
                legs[name]["roofline"]["traffic_vs_algorithmic"] = (
                    None if traffic is None else round(traffic / legs[name]["roofline"]["bytes_per_launch"], 3))
    if args.moving_steps > 0:
        legs["moving_body_full_surface"] = moving_body_leg(args)
        _progress(f"moving body, full surface: {legs['moving_body_full_surface']['vertex_steps_per_s']:.4g} "
                  "vertex-steps/s")
        torch.cuda.empty_cache()
    if args.cube_worlds > 0 and args.cube_train_frames > 0:
        _progress("tet cubes with the learned warm start")
        legs["batched_tets_residual"] = cubes_residual_leg(args, ws)
        _progress(f"tet cubes: learned {legs['batched_tets_residual']['speedup_batched_learned_vs_plain_batched']}x "
                  "plain batched")
        torch.cuda.empty_cache()
    if args.self_contact_steps > 0:
        _progress("self-contact leg: stacked sheets on a sphere")
        legs["self_contact"] = self_contact_leg(args)
        torch.cuda.empty_cache()
    cpu = None
    if args.cpu_seconds > 0:
        _progress(f"CPU baseline: about {1.6 * args.cpu_seconds:.0f} s")
        fq = fixed_model.particle_q.cpu().numpy().copy()
        fqd = fixed_model.particle_qd.cpu().numpy().copy()
        cpu = cpu_baseline_leg(args, ctx, fixed_model, fq, fqd, fixed_info["dt"])

    per = ctx["per"]
    rec = {
        "metric": "vertex-steps/s at fixed force residual; plain VBD vs batched+learned warm-start",
        "value": round(value, 1),
        "unit": "vertex-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{head['workload']}: the 128x128 wind cloth (top edge pinned, 8-mode Fourier wind, "
                        f"16,384 vertices) as {head['worlds']} worlds with their own winds, every world driven "
                        f"to relative force residual {args.rel_tol}",
            "worlds": head["worlds"],
            "vertices": head["vertices"],
            "vertices_per_world": per,
            "rel_tol": args.rel_tol,
            "max_iterations": args.max_iterations,
            "step": "one substep (dt 1/600 s) of every world, each world iterated until r_w <= rel_tol * r0_w",
            "check_every": args.check_every,
            "solver": "batched + learned warm start: per world the inertial start y, y + c_prev and "
                      f"y + c_prev + S(0.5 MLP) (the BF16 network trained in this run; S = {args.delta_sweeps} "
                      "sweeps of the 1-ring mean on the network's change), judged by their force residuals after "
                      f"{args.trial_iterations} VBD iteration(s) each; every solver on its own trajectory from one "
                      "common state; batched_reuse_only in the comparison is the same protocol without the network",
            "pass_shape": head["pass_shape"],
        },
        "comparison": main_rec,
        "tolerance_sweep": head["tolerance_sweep"],
        "fp32_floor": head["fp32_floor"],
        "roofline": roofline,
        "single_world": single,
        "fixed_iterations": fixed,
        **legs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
