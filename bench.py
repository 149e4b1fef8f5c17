"""Benchmark: particle-steps/s of the DEM time step on one MI355X.

Workload (default, --workload hopper): 1,000,000 polydisperse spheres
(r ~ U[0.5, 1.5] mm, rho 2500) poured above a conical STL hopper (half-angle
30 deg, outlet radius 8 mm, 4,096 triangles), Hertz normal + Mindlin
tangential, E=5e7, nu=0.25, e=0.6, mu=0.4, dt=2e-6 s, FP64, gravity on.
The particles start from a seeded random sequential packing (solid fraction
0.25) in a 0.1 x 0.1 m column above the hopper.

A "step" is one full DEM time step over all particles (broad phase rebuild
when the Verlet skin demands it, contact gather, walls, integration); the
inputs are resident in HBM before timing starts.

Printed: one JSON line (rank 0) with value = particle-steps/s, the roofline
of the dominant kernel (the fused contact + integration kernel, timed with
HIP events on the engine's stream), and a CPU baseline: the oracle (a C++
restatement of the reference's step_cpu, OpenMP over the host cores) timed
on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle-step of the fused contact + integration
# kernel (FP64 structure of arrays): read x, v, w and their Kahan residuals
# (6 x 24 B), radius and mass (16 B), id/flags (8 B) = 168 B; write x, v, w and
# residuals = 144 B.  Neighbour reads are reuse (cache), not counted.
FORCE_BYTES_PER_PARTICLE = 312.0
# Per directed contact carrying tangential history (Mindlin/LSC with a
# non-zero shear modulus): 32 B read + 32 B write.  The reference's shear
# modulus is 0 (material.rs:49), so no history is carried and this term is 0.
HIST_BYTES_PER_CONTACT = 64.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name: str, n: int | None):
    from rusty_particles_amd import scenes
    if name == "hopper":
        return scenes.hopper_workload(n or 1_000_000)
    if name == "box":
        return scenes.box_settle_workload(n or 10_000)
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(p, x, r, m, meshes, steps: int, threads: int, cell: float):
    """The oracle (port of step_cpu) on the same inputs, reference grid cell =
    largest contact cutoff (the cell-size heuristic of simulation.rs:64)."""
    import numpy as np

    import oracle
    from rusty_particles_amd import abi
    lib = oracle.lib()
    lib.rpo_set_threads(threads)
    o = oracle.OracleSimulation(p.dt, None, None, params=abi.RpParams.from_buffer_copy(p))
    o.set_state(x, r, m)
    for v9, n3 in meshes:
        o.add_mesh_arrays(v9, n3)
    o.set_cell_size(cell)
    o.set_record(False)
    o.step(1)  # warm-up (allocations, first touch)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return x.shape[0] * steps / dt, dt


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="hopper")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=100)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run the library's 256 MiB copy kernel so a counter profile of this run can "
                         "scale FETCH_SIZE/WRITE_SIZE by a kernel of known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one domain on one GPU; multi-GPU runs are not part of "
                                   "this build (use --gpus 1)"}))
        return 2

    import numpy as np

    from rusty_particles_amd import abi, native
    if native.device_count() < 1:
        raise SystemExit("bench.py: no HIP device visible")

    t0 = time.perf_counter()
    p, x, r, m, meshes = build_workload(args.workload, args.n)
    n = x.shape[0]
    ntri = sum(v9.shape[0] for v9, _ in meshes)
    log(f"workload {args.workload}: {n} particles, {ntri} triangles, built in {time.perf_counter() - t0:.1f}s")

    eng = native.Engine(abi.RpParams.from_buffer_copy(p))
    eng.upload(x, r, m)
    for v9, n3 in meshes:
        eng.add_mesh(v9, n3)
    eng.step(args.warmup)
    eng.synchronize()
    log(f"warmup {args.warmup} steps done")

    # timed region: whole job, K steps, one enqueue, synchronised on both sides
    eng.synchronize()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.synchronize()
    wall = time.perf_counter() - t0
    st = eng.stats()
    value = n * args.steps / wall
    log(f"timed {args.steps} steps: {wall:.3f}s, {value:.4g} particle-steps/s, rebuilds {st['rebuilds']}, "
        f"max list length {st['max_neighbours']}, list capacity {st['neighbour_capacity']}")

    # kernel roofline: HIP events around each launch on the engine stream
    eng.set_timing(True)
    eng.step(args.timing_steps)
    tm = eng.timing()
    eng.set_timing(False)
    f_ms, f_cnt = tm["force"]
    avg_s = f_ms / max(f_cnt, 1) * 1e-3
    bytes_per_launch = FORCE_BYTES_PER_PARTICLE * n
    achieved = bytes_per_launch / avg_s / 1e9
    log(f"force kernel: {f_cnt} launches, avg {avg_s * 1e6:.1f} us, {achieved:.1f} GB/s algorithmic; "
        f"rebuild chain avg {tm['rebuild'][0] / max(tm['rebuild'][1], 1) * 1e3:.1f} us")
    if args.calibrate_copy:
        ms = native.probe_copy(256 * 1024 * 1024)
        log(f"copy calibration: k_copy_probe read 268435456 B + wrote 268435456 B in {ms:.3f} ms "
            f"({2 * 268435456 / ms / 1e6:.0f} GB/s)")
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"{args.workload}_force_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        cell = 2.0 * float(r.max()) * (1.5 if p.normal_model in (abi.RP_NORMAL_JKR, abi.RP_NORMAL_SJKR) else 1.0)
        log(f"cpu baseline: oracle, {threads} threads, {args.cpu_steps} step(s) of the same {n}-particle workload")
        cval, csec = cpu_baseline(p, x, r, m, meshes, args.cpu_steps, threads, cell)
        cpu = {"value": cval, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} step(s) of the full {n}-particle workload after 1 warm-up step, "
                         f"reference grid cell {cell * 1e3:.2f} mm (largest contact cutoff), brute-force "
                         f"{ntri}-triangle wall loop as in simulation.rs:465-523; {csec:.1f}s"}
        log(f"cpu baseline {cval:.4g} particle-steps/s ({csec:.1f}s)")

    out = {
        "metric": "particle-steps/s on 1 GPU, plain vs surrogate+batched at 1e-3 force err; HBM GB/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} spheres, {ntri}-triangle STL, "
                               f"{'Hertz+Mindlin' if args.workload == 'hopper' else 'Hertz+Coulomb'}, "
                               f"dt={p.dt:g}, FP64", "particles": n, "triangles": ntri, "form": "plain"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "contacts_per_step": st["contacts"],
        "wall_contacts_per_step": st["wall_contacts"],
        "force_kernel_avg_us": avg_s * 1e6,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
