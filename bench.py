"""Benchmark: vertex-steps/s of the MI355X VBD solver on the 256x256 cloth-on-sphere workload.

One step = one frame of the reference's usage loop: 10 substeps, each
clear_forces -> CollisionPipeline.collide -> SolverVBD.step (10 iterations),
captured once in a HIP graph and replayed.  Prints ONE JSON line with the
throughput, a roofline record for the dominant kernel (the color pass) and a
CPU baseline (the FP32 CPU oracle, OpenMP over each color's vertices, on a
bounded sample of the same workload).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed frames")
    ap.add_argument("--warmup", type=int, default=3, help="untimed frames")
    ap.add_argument("--verts", type=int, default=256, help="cloth vertices per side")
    ap.add_argument("--iterations", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU-baseline time budget (0 disables)")
    ap.add_argument("--roofline-reps", type=int, default=20)
    ap.add_argument("--no-traffic", dest="traffic", action="store_false",
                    help="skip the two rocprofv3 --pmc child runs that measure HBM traffic per color pass")
    ap.add_argument("--pmc-mode", action="store_true",
                    help="counter runs: color passes of one prepared substep plus a plain 256 MiB device copy "
                         "(the calibration kernel for FETCH_SIZE / WRITE_SIZE), then exit")
    return ap.parse_args()


def pmc_mode(args):
    """Minimal workload for rocprofv3 --pmc: `reps` launches of each color pass and one known-size copy."""
    import numpy as np
    import torch

    from newton_amd import CollisionPipeline, scenes
    from newton_amd.roofline import color_pass_bytes
    from newton_amd.solvers import SolverVBD

    dev = "cuda:0"
    model, state, info = scenes.cloth_drop_on_sphere(dev, verts_per_side=args.verts, iterations=args.iterations)
    solver = SolverVBD(model, iterations=args.iterations)
    pipe = CollisionPipeline(model, soft_contact_margin=0.01)
    contacts = pipe.contacts()
    state.clear_forces()
    pipe.collide(state, contacts)
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
    counts = np.bincount(contacts.soft_contact_particle.cpu().numpy()[: int(contacts.soft_contact_count.item())],
                         minlength=model.particle_count)
    b = color_pass_bytes(model, contacts_per_particle=counts)
    print(json.dumps({"pmc_mode": True, "color_pass_launches": args.roofline_reps * ncol,
                      "algorithmic_bytes_per_color_pass": [int(x) for x in b], "copy_bytes": nbytes}))


def measure_traffic(args, timeout_s: float = 240.0):
    """HBM bytes per color-pass launch from two rocprofv3 --pmc child runs (FETCH_SIZE, then WRITE_SIZE).

    Each counter is scaled by what it reads for the 256 MiB device copy made in
    the same run (bytes copied / counter bytes), so the figure is in real bytes.
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
                   str(ROOT / "bench.py"), "--pmc-mode", "--verts", str(args.verts), "--iterations",
                   str(args.iterations), "--roofline-reps", "5"]
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


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (worlds would shard "
                          "across GPUs with no per-step exchange, see DESIGN.md)"}))
        sys.exit(2)
    if args.pmc_mode:
        pmc_mode(args)
        return
    import numpy as np
    import torch

    from newton_amd import CollisionPipeline, scenes
    from newton_amd.roofline import color_pass_bytes
    from newton_amd.solvers import SolverVBD

    dev = "cuda:0"
    torch.cuda.set_device(0)
    model, state0, info = scenes.cloth_drop_on_sphere(dev, verts_per_side=args.verts, iterations=args.iterations)
    n = model.particle_count
    dt = info["dt"]
    sub = info["substeps"]
    solver = SolverVBD(model, iterations=args.iterations)
    pipe = CollisionPipeline(model, soft_contact_margin=0.01)
    contacts = pipe.contacts()
    state1 = model.state()
    q_init = state0.particle_q.cpu().numpy().copy()
    qd_init = state0.particle_qd.cpu().numpy().copy()
    states = [state0, state1]

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
        for _ in range(max(args.warmup, 1)):
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
        for _ in range(args.steps):
            graph.replay()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_ms = ev0.elapsed_time(ev1)
    q_now = states[0].particle_q
    finite = bool(torch.isfinite(q_now).all().item())
    elapsed = gpu_ms / 1e3
    value = n * sub * args.steps / elapsed

    # ---- roofline of the color-pass kernel: back-to-back launches timed with events on its stream
    with torch.cuda.stream(stream):
        s = states[0]
        s.clear_forces()
        pipe.collide(s, contacts)
        solver.begin_step(s, contacts, dt)
        torch.cuda.synchronize()
        counts = np.bincount(contacts.soft_contact_particle.cpu().numpy()[: int(contacts.soft_contact_count.item())],
                             minlength=n) if contacts.soft_contact_max else None
        per_color_bytes = color_pass_bytes(model, contacts_per_particle=counts)
        ncol = len(model.particle_color_groups)
        color_ms = np.zeros(ncol)
        for c in range(ncol):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            solver.color_pass(s, c, dt)  # warm
            e0.record(stream)
            for _ in range(args.roofline_reps):
                solver.color_pass(s, c, dt)
            e1.record(stream)
            torch.cuda.synchronize()
            color_ms[c] = e0.elapsed_time(e1) / args.roofline_reps
    total_bytes = float(per_color_bytes.sum())
    total_s = float(color_ms.sum()) / 1e3
    achieved = total_bytes / total_s / 1e9
    roofline = {
        "kernel": "k_solve_fast",
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "traffic_detail": None,
        "bytes_per_launch": round(total_bytes / ncol, 1),
        "avg_launch_us": round(float(color_ms.mean()) * 1e3, 3),
        "launches_per_substep": ncol * args.iterations,
    }

    if args.traffic:
        traffic, detail = measure_traffic(args)
        roofline["traffic"] = None if traffic is None else round(traffic, 1)
        roofline["traffic_detail"] = detail

    # ---- CPU baseline: the FP32 CPU oracle on a bounded sample of the same workload
    cpu = None
    if args.cpu_seconds > 0:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle  # noqa: PLC0415  (the checker, timed as the CPU baseline)

        cores = min(len(os.sched_getaffinity(0)), 16)
        om = oracle.OracleModel(model)
        prm = oracle.params(args.iterations, num_threads=cores)
        q = q_init
        qd = qd_init
        done = 0
        t0 = time.perf_counter()
        while True:
            q, qd, _ = oracle.simulate(om, prm, q, qd, 1, dt, 0.01, len(om.pairs) // 2)
            done += 1
            if time.perf_counter() - t0 >= args.cpu_seconds or done >= 200:
                break
        cpu_t = time.perf_counter() - t0
        cpu = {
            "value": round(n * done / cpu_t, 1),
            "unit": "vertex-steps/s",
            "cores": cores,
            "kind": "port",
            "sample": f"{done} substeps (collide + step, {args.iterations} iterations) of the same "
                      f"{args.verts}x{args.verts} cloth from its initial state, FP32 C oracle, OpenMP {cores} threads",
        }

    rec = {
        "metric": "vertex-steps/s at fixed force residual; plain VBD vs batched+learned warm-start",
        "value": round(value, 1),
        "unit": "vertex-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(gpu_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"cloth_drop_on_sphere_{args.verts}x{args.verts}",
            "vertices": n,
            "triangles": model.tri_count,
            "bending_edges": model.edge_count,
            "colors": len(model.particle_color_groups),
            "substeps_per_step": sub,
            "iterations": args.iterations,
            "solver": "plain VBD, fixed iterations",
            "hip_graph": True,
            "wall_s": round(wall, 4),
            "finite": finite,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
