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

--workload jkr: 2,000,000 spheres (r lognormal, median 50 um, sigma 0.3,
rho 1500) in a periodic box at solid fraction 0.35, JKR + Coulomb (E=1e8,
nu=0.3, gamma=0.05, mu=0.3), dt=1e-8 s, velocities ~N(0, 0.05) m/s.  The
step runs once per JKR solver on the same initial state: "exact" (the
reference's Newton solve), "shortcut" (identical results; no iterations for
contacts that cannot have a root) and "surrogate" (the BF16-MFMA network,
trained on the GPU before timing, plus FP64 polish and exact fallback);
value = the surrogate form, the others and the surrogate's force errors
(check mode, exact solve alongside) are in the same line.

--workload batch: an ensemble of 256 independent simulations (N_k ~
U{2000..50000}, 6.7M spheres in total, r ~ U[1, 3] mm, per-instance friction,
restitution and surface energy; simplified JKR + Coulomb, dt = 5e-6 s,
jittered lattices in a 0.2 m box STL column) stepped by one batched engine
(one set of launches per step over all instances, HIP-graph step loop);
value = total particle-steps/s.  The same instances stepped one engine at a
time (each with its own launches) are timed beside it.

--workload jkr-batch: 64 periodic JKR packings (N_k ~ U{5000..40000}, 1.5M
spheres, r lognormal median 50 um, surface energy ~U[0.02, 0.08] J/m^2) in
one batched engine, once per JKR solver (exact, shortcut, surrogate); value
= the surrogate+batched form.  With --physics-fix (batch workloads): the
physics-fix mode with Mindlin springs and their tangential history, and a
friction coefficient drawn per instance on top of the sweep; on the single
workloads it switches the mode only (the hopper's Hertz + Mindlin then keeps
a real tangential history).

--workload drum: the capacity run, 16,000,000 spheres (r ~ U[0.4, 0.8] mm,
rho 2500) in a closed 16,384-triangle STL drum of radius 0.25 m turning at
30 rpm (the surface velocity of the moving-wall extension), JKR (surface energy 0.01 J/m^2) +
Coulomb (mu 0.5), E=5e7, nu=0.3, e=0.5, dt=2e-6 s, 1,000 steps per timed
region; a jittered-lattice bed filling 30 % of the drum.  Forms: shortcut,
then surrogate (the value).  CPU baseline sample: the bed's first 20,000
spheres in the same drum.  --workload drum-fix: the capacity run as
specified, JKR + Mindlin with tangential history, in the physics-fix mode
(real shear modulus, history table carried through re-sorts), with the bed
in contact (scenes.drum_chain_workload: chains of touching spheres along the
axis filling 30 % of a 0.99 m drum); its roofline counts 64 B per stored
history entry on top of the 312 B per particle.

Timing: --steps steps per timed region, --repeats regions (default 3), each
on a fresh engine from the same initial state after the same warm-up (so
each times the same simulated window), value = the median region;
contacts/s = the contacts of the last step times steps/s.

--workload hopper-bed: the hopper configuration's 1M spheres resting in a
cone of the same half-angle, outlet and triangle count (top radius 0.3 m so
the bed fits): vertical columns in static Hertz equilibrium, each standing on
a fixed foot sphere that touches the cone (scenes.hopper_bed_workload), ~2
contacts per sphere: a contact-bearing window of the same workload.

Printed: one JSON line (rank 0) with value = particle-steps/s, the roofline
of the dominant kernel (the fused contact + integration kernel, timed with
HIP events on the engine's stream), and a CPU baseline: the oracle (a C++
restatement of the reference's step_cpu, OpenMP over the host cores) timed
on a bounded sample of the same workload.  The default hopper line also
carries, in the same JSON object: "contact_window" (the hopper-bed workload,
same steps and warm-up: contacts/s and its force kernel's roofline) and
"surrogate_batched" (the jkr-batch workload with the exact, shortcut and
surrogate+batched forms and the surrogate's force errors); --no-extras
leaves them out.
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
# Per directed contact carrying tangential history (the physics-fix mode's
# Mindlin/LSC with a real shear modulus): partner id and displacement, 32 B
# read + 32 B write.  In the parity mode the reference's shear modulus is 0
# (material.rs:49), no history is carried and this term is 0.
HIST_BYTES_PER_CONTACT = 64.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name: str, n: int | None):
    """(params, x, r, m, v or None, meshes, description, fixed flags or None)."""
    from rusty_particles_amd import scenes
    if name == "hopper":
        p, x, r, m, meshes = scenes.hopper_workload(n or 1_000_000)
        return p, x, r, m, None, meshes, "Hertz+Mindlin hopper", None
    if name == "hopper-bed":
        p, x, r, m, meshes, fixed = scenes.hopper_bed_workload(n or 1_000_000)
        return p, x, r, m, None, meshes, "Hertz+Mindlin hopper, bed of columns standing in the cone", fixed
    if name == "box":
        p, x, r, m, meshes = scenes.box_settle_workload(n or 10_000)
        return p, x, r, m, None, meshes, "Hertz+Coulomb box", None
    if name == "jkr":
        p, x, r, m, v = scenes.jkr_periodic_workload(n or 2_000_000)
        return p, x, r, m, v, [], "JKR+Coulomb periodic packing", None
    if name == "jkr-dense":
        p, x, r, m, v = scenes.jkr_dense_workload(n or 2_000_000)
        return p, x, r, m, v, [], "JKR+Coulomb compressed FCC crystal", None
    if name == "drum":
        p, x, r, m, meshes = scenes.drum_workload(n or 16_000_000)
        return p, x, r, m, None, meshes, "JKR+Coulomb drum turning at 30 rpm", None
    if name == "drum-fix":
        from rusty_particles_amd import abi
        p, x, r, m, meshes = scenes.drum_chain_workload(n or 16_000_000)
        p.tangential_model = abi.RP_TANGENTIAL_MINDLIN
        p.mode = abi.RP_MODE_PHYSICS_FIX
        return (p, x, r, m, None, meshes, "JKR+Mindlin with tangential history (physics-fix mode) drum turning at 30 rpm",
                None)
    raise SystemExit(f"unknown workload {name}")


def traffic_of(workload: str):
    """(HBM bytes per launch of the workload's force kernel, source): from its
    committed counter profile (rocprofv3 FETCH_SIZE / WRITE_SIZE of an earlier
    run, scaled by the copy probe in the same run), not measured live; None
    when there is no profile."""
    name = f"{workload}_force_traffic.json"
    tfile = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(tfile):
        return None, None
    with open(tfile) as fh:
        return json.load(fh).get("bytes_per_launch"), f"profiles/{name} (rocprofv3 counters, earlier run)"


def cpu_info() -> dict:
    """Host CPU model and logical CPU count, and the threads the baseline may
    use: the GPU box allots 16 host cores per GPU (its OMP_NUM_THREADS)."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    ncpu = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or 16
    return {"cpu_model": model, "host_cpus": ncpu, "threads": max(1, min(ncpu, share))}


# The surrogate network as its MFMA tiles execute it: 4 inputs padded to 16,
# two 64-wide hidden layers, 2 outputs padded to 16 (multiply + add = 2 flops).
MLP_FLOPS_PER_ROW = 2 * (16 * 64 + 64 * 64 + 64 * 16)
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def surrogate_kernels(tm: dict, rows_per_step: float = None) -> dict:
    """Average durations of the surrogate's collect and network launches from
    an engine timing dict; with the rows the network ran on per step (the
    contacts routed to the surrogate), its achieved BF16 TFLOP/s and the
    fraction of the dense MFMA peak."""
    (c_ms, c_n), (n_ms, n_n) = tm["collect"], tm["network"]
    if not n_n:
        return {}
    out = {"collect_avg_us": c_ms / max(c_n, 1) * 1e3, "network_avg_us": n_ms / n_n * 1e3}
    out["surrogate_kernels_avg_us"] = out["collect_avg_us"] + out["network_avg_us"]
    if rows_per_step:
        tflops = rows_per_step * MLP_FLOPS_PER_ROW / (out["network_avg_us"] * 1e-6) / 1e12
        out.update({"network_rows_per_step": rows_per_step, "network_tflops": tflops,
                    "network_mfma_frac": tflops / BF16_PEAK_TFLOPS})
    return out


def run_batch(args) -> int:
    """Batched ensembles: --workload batch (sJKR material sweep, plain form)
    and --workload jkr-batch (JKR surface-energy sweep, one line per JKR
    solver, the surrogate last); each against one engine per instance."""
    import numpy as np

    from rusty_particles_amd import abi, native, scenes
    t0 = time.perf_counter()
    jkr = args.workload == "jkr-batch"
    if jkr:
        p, counts, mats, x, r, m, v = scenes.jkr_ensemble_workload(args.instances or 64)
        meshes = []
        forms = args.jkr_forms.split(",")
        desc = "JKR+Coulomb periodic packings, surface-energy sweep"
    else:
        p, counts, mats, x, r, m, meshes = scenes.batch_ensemble_workload(args.instances or 256)
        v = None
        forms = ["plain"]
        desc = "sJKR+Coulomb, material sweep"
    if args.physics_fix:
        # Mindlin springs with their history; the sweep adds friction, which
        # the parity mode's inert spring (shear modulus 0) would ignore
        p.mode = abi.RP_MODE_PHYSICS_FIX
        p.tangential_model = abi.RP_TANGENTIAL_MINDLIN
        rng = np.random.default_rng(0x5EED)
        mats = [abi.RpMaterial(q.youngs_modulus, q.poissons_ratio, q.density, float(rng.uniform(0.1, 0.9)),
                               q.restitution_coefficient, q.surface_energy) for q in mats]
        desc = desc.replace("+Coulomb", "+Mindlin with history (physics-fix mode)") + ", friction sweep"
    n = int(sum(counts))
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    log(f"workload {args.workload}: {len(counts)} instances, {n} particles, built in {time.perf_counter() - t0:.1f}s")

    net = None
    surr_info = {}
    if "surrogate" in forms:
        from rusty_particles_amd import surrogate
        rs = r[:, None] * r[None, :200] / (r[:, None] + r[None, :200])
        gam = [q.surface_energy for q in mats]
        e_star = 1.0 / (2 * (1 - p.particle.poissons_ratio ** 2) / p.particle.youngs_modulus)
        env = surrogate.envelope(-0.40, 30.0, (0.9 * rs.min(), 1.1 * rs.max()), e_star, (min(gam), max(gam)))
        t0 = time.perf_counter()
        net = surrogate.train(env, nsamples=args.train_samples, epochs=args.train_epochs, batch=4096, lr=3e-3)
        surr_info["train_s"] = time.perf_counter() - t0
        log(f"surrogate trained in {surr_info['train_s']:.1f}s")

    def engine(lo, hi, form, batch=True):
        q = abi.RpParams.from_buffer_copy(p)
        sl = slice(int(offs[lo]), int(offs[hi]))
        vv = None if v is None else v[sl]
        if batch:
            e = native.Engine(q)
            e.upload_batch(counts[lo:hi], mats[lo:hi], x[sl], r[sl], m[sl], v=vv)
        else:
            q.particle = mats[lo]
            e = native.Engine(q)
            e.upload(x[sl], r[sl], m[sl], v=vv)
        for v9, n3 in meshes:
            e.add_mesh(v9, n3)
        if form != "plain":
            e.set_jkr_solver(form, net)
        if args.no_graph:
            e.set_graph(False)
        return e

    def timed(make, steps):
        """Median wall time of --repeats timed regions of `steps` steps, each
        on a fresh engine from make() (the same initial state), warmed up and
        autotuned first; returns (median, all regions, the last engine)."""
        walls, e = [], None
        for _ in range(max(args.repeats, 1)):
            if e is not None:
                e.close()
            e = make()
            e.step(args.warmup)
            e.synchronize()
            if args.autotune:
                sk, ce = e.autotune(args.autotune)
                log(f"autotune: skin {sk:.3g} m, cell {ce:.3g} m, tiling {e.tiling()}")
            e.synchronize()
            t0 = time.perf_counter()
            e.step(steps)
            e.synchronize()
            walls.append(time.perf_counter() - t0)
        return float(np.median(walls)), walls, e

    results = {}
    for form in forms:
        wall, walls, e = timed(lambda: engine(0, len(counts), form), args.steps)
        log(f"[{form}] {len(walls)} timed regions of {args.steps} steps: " + ", ".join(f"{w:.4f}s" for w in walls))
        value = n * args.steps / wall
        st = e.stats()
        e.set_timing(True)
        e.step(args.timing_steps)
        tm = e.timing()
        e.set_timing(False)
        avg_s = tm["force"][0] / max(tm["force"][1], 1) * 1e-3
        alg_bytes = FORCE_BYTES_PER_PARTICLE * n + HIST_BYTES_PER_CONTACT * st.get("history_contacts", 0)
        res = {"value": value, "ms_per_step": wall / args.steps * 1e3, "force_kernel_avg_us": avg_s * 1e6,
               "achieved_gbs": alg_bytes / avg_s / 1e9, "contacts_per_step": st["contacts"],
               "history_contacts_per_step": st.get("history_contacts", 0),
               "contacts_per_s": st["contacts"] * args.steps / wall,
               "repeat_values": [n * args.steps / w for w in walls], "wall": wall, "stats": st}
        res.update(surrogate_kernels(tm))
        log(f"[{form}] batched: {args.steps} steps in {wall:.3f}s, {value:.4g} particle-steps/s, force kernel "
            f"{avg_s * 1e6:.1f} us")
        if form == "surrogate":
            e.set_jkr_solver("surrogate", net, check=True)
            errs = []
            for _ in range(5):
                e.step(1)
                errs.append(e.jkr_stats())
            surr_info.update({
                "served_per_step": float(np.mean([q["served"] for q in errs])),
                "fallback_per_step": float(np.mean([q["fallback"] for q in errs])),
                "force_rel_err_max": max(q["err_max"] for q in errs),
                "force_rel_err_mean": float(np.mean([q["err_mean"] for q in errs])),
            })
            res.update(surrogate_kernels(tm, surr_info["served_per_step"] + surr_info["fallback_per_step"]))
            log(f"[surrogate] check: {surr_info}")
        e.close()
        results[form] = res
    main = forms[-1]
    R = results[main]

    graph_off = None
    if not jkr:
        def no_graph_engine():
            q = engine(0, len(counts), main)
            q.set_graph(False)
            return q
        wall_off, _, e = timed(no_graph_engine, args.steps)
        graph_off = n * args.steps / wall_off
        log(f"batched, graph off: {graph_off:.4g} particle-steps/s")
        e.close()

    # the same instances, one engine each, run back to back (graph on)
    sep_steps = max(args.steps // 4, 20)
    total = 0.0
    for k in range(0 if args.no_compare else len(counts)):
        e = engine(k, k + 1, main, batch=False)
        e.step(args.warmup)
        e.synchronize()
        t0 = time.perf_counter()
        e.step(sep_steps)
        e.synchronize()
        total += time.perf_counter() - t0
        e.close()
    sep_value = n * sep_steps / total if total > 0.0 else None
    if sep_value:
        log(f"one engine per instance: {sep_steps} steps each, {total:.3f}s in all, {sep_value:.4g} particle-steps/s")

    # what a plain streaming copy reaches on this GPU (with --calibrate-copy
    # also the reference point of a counter profile)
    nbytes = (1 << 30) if not args.calibrate_copy else 256 * 1024 * 1024
    ms = native.probe_copy(nbytes)
    copy_gbs = 2 * nbytes / ms / 1e6
    log(f"copy probe: k_copy_probe read {nbytes} B + wrote {nbytes} B in {ms:.3f} ms ({copy_gbs:.0f} GB/s)")

    wname = args.workload + ("-fix" if args.physics_fix else "")
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or cpu_info()["threads"]
        sample = list(range(0, len(counts), max(len(counts) // 4, 1)))[:4]
        csec, cwork, csteps = 0.0, 0, []
        for k in sample:
            sl = slice(int(offs[k]), int(offs[k + 1]))
            q = abi.RpParams.from_buffer_copy(p)
            q.particle = mats[k]
            cell = 2.0 * float(r.max()) * 1.5
            val, sec, ns = cpu_baseline(q, x[sl], r[sl], m[sl], None if v is None else v[sl], meshes,
                                        args.cpu_steps, threads, cell, args.cpu_seconds / len(sample))
            csec += sec
            cwork += counts[k] * ns
            csteps.append(ns)
        cpu = {"value": cwork / csec, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"instances {sample} ({sum(counts[k] for k in sample)} particles), {csteps} "
                         f"steps after 1 warm-up step each, reference's JKR solve; {csec:.1f}s"}
        log(f"cpu baseline {cwork / csec:.4g} particle-steps/s")
    out = {
        "metric": "particle-steps/s on 1 GPU, plain vs surrogate+batched at 1e-3 force err; HBM GB/s",
        "value": R["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": R["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{wname}: {len(counts)} instances, {n} spheres in all, {desc}, dt={p.dt:g}, "
                               f"FP64", "particles": n, "instances": len(counts),
                   "form": "batched" if main == "plain" else f"{main}+batched"},
        "roofline": {"bound": "hbm", "achieved": R["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": R["achieved_gbs"] / HBM_PEAK_GBS, "traffic": traffic_of(wname)[0],
                     "traffic_source": traffic_of(wname)[1]},
        "stream_copy_gbs": copy_gbs,
        "cpu_baseline": cpu,
        "repeats": len(R["repeat_values"]), "repeat_values": R["repeat_values"],
        "contacts_per_step": R["contacts_per_step"], "contacts_per_s": R["contacts_per_s"],
        "history_contacts_per_step": R["history_contacts_per_step"],
        "force_kernel_avg_us": R["force_kernel_avg_us"],
        "one_engine_per_instance_value": sep_value,
    }
    if graph_off is not None:
        out["graph_off_value"] = graph_off
    if len(results) > 1:
        out["forms"] = {f: {k: q for k, q in res.items() if k not in ("stats", "wall")} for f, res in results.items()}
    if surr_info:
        out["surrogate"] = surr_info
    return out


def mesh_motions(workload: str):
    """(mesh, origin, omega) of the workload's moving meshes."""
    from rusty_particles_amd import scenes
    return [(0, scenes.DRUM_ORIGIN, scenes.DRUM_OMEGA)] if workload in ("drum", "drum-fix") else []


def cpu_baseline(p, x, r, m, v, meshes, steps: int, threads: int, cell: float, seconds: float = 0.0,
                 motions=(), bits: int = 64, fixed=None):
    """The oracle (port of step_cpu) on the same inputs, reference grid cell =
    largest contact cutoff (the cell-size heuristic of simulation.rs:64).
    Times at least `steps` steps, more when the warm-up step says that
    `seconds` of CPU time hold more (at most 10,000).  Returns
    (particle-steps/s, seconds, steps timed)."""
    import numpy as np

    import oracle
    from rusty_particles_amd import abi
    lib = oracle.lib()
    lib.rpo_set_threads(threads)
    o = oracle.OracleSimulation(p.dt, None, None, bits=bits, params=abi.RpParams.from_buffer_copy(p))
    o.set_state(x, r, m, fixed, v=v)
    for v9, n3 in meshes:
        o.add_mesh_arrays(v9, n3)
    for k, org, om in motions:
        o.set_mesh_motion(k, org, om)
    o.set_cell_size(cell)
    o.set_record(False)
    t0 = time.perf_counter()
    o.step(1)  # warm-up (allocations, first touch)
    warm = time.perf_counter() - t0
    if seconds > 0.0:
        steps = max(steps, min(10_000, int(seconds / max(warm, 1e-6))))
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    return x.shape[0] * steps / dt, dt, steps


def run_single(args) -> dict:
    """One workload (single engine): the forms, the timed regions, the kernel
    roofline and the CPU baseline; returns the JSON object."""
    import numpy as np

    from rusty_particles_amd import abi, native
    t0 = time.perf_counter()
    p, x, r, m, v, meshes, desc, fixed = build_workload(args.workload, args.n)
    wname = args.workload
    if args.physics_fix and p.mode != abi.RP_MODE_PHYSICS_FIX:
        p.mode = abi.RP_MODE_PHYSICS_FIX
        wname += "-fix"
        desc += (", physics-fix mode (Mindlin with tangential history)"
                 if p.tangential_model == abi.RP_TANGENTIAL_MINDLIN else ", physics-fix mode")
    n = x.shape[0]
    ntri = sum(v9.shape[0] for v9, _ in meshes)
    log(f"workload {args.workload}: {n} particles, {ntri} triangles, built in {time.perf_counter() - t0:.1f}s")

    forms = args.jkr_forms.split(",") if p.normal_model == abi.RP_NORMAL_JKR else ["plain"]
    net = None
    surr_info = {}
    if "surrogate" in forms:
        from rusty_particles_amd import surrogate
        rr = r[::max(1, n // 2_000_000)]  # at most ~2M radii span the envelope
        rs = rr[:, None] * rr[None, :200] / (rr[:, None] + rr[None, :200])
        env = surrogate.envelope(-0.40, 30.0, (0.9 * rs.min(), 1.1 * rs.max()),
                                 1.0 / (2 * (1 - p.particle.poissons_ratio ** 2) / p.particle.youngs_modulus),
                                 p.particle.surface_energy)
        t0 = time.perf_counter()
        net = surrogate.train(env, nsamples=args.train_samples, epochs=args.train_epochs, batch=4096, lr=3e-3)
        surr_info["train_s"] = time.perf_counter() - t0
        surr_info["train_loss"] = [float(net.epoch_loss[0]), float(net.epoch_loss[-1])]
        log(f"surrogate trained in {surr_info['train_s']:.1f}s, loss {net.epoch_loss[0]:.3g} -> "
            f"{net.epoch_loss[-1]:.3g}")

    def fresh_engine(form):
        """An engine on the initial state, warmed up (and autotuned)."""
        e = native.Engine(abi.RpParams.from_buffer_copy(p))
        if args.no_graph:
            e.set_graph(False)
        e.upload(x, r, m, fixed, v=v)
        for v9, n3 in meshes:
            e.add_mesh(v9, n3)
        for k, org, om in mesh_motions(args.workload):
            e.set_mesh_motion(k, org, om)
        if form != "plain":
            e.set_jkr_solver(form, net)
        e.step(args.warmup)
        e.synchronize()
        tuned = None
        if args.autotune:
            t0 = time.perf_counter()
            tuned = e.autotune(args.autotune)
            log(f"[{form}] autotune: skin {tuned[0]:.3g} m, cell {tuned[1]:.3g} m (0 = automatic), "
                f"tiling {e.tiling()}, {time.perf_counter() - t0:.1f}s")
        return e, tuned

    results = {}
    eng = None
    for form in forms:
        # timed regions: whole job, K steps each, one enqueue, synchronised on
        # both sides, each on a fresh engine from the same initial state (so
        # every region times the same simulated window); value = the median
        walls = []
        for _ in range(max(args.repeats, 1)):
            if eng is not None:
                eng.close()
            eng, tuned = fresh_engine(form)
            eng.synchronize()
            t0 = time.perf_counter()
            eng.step(args.steps)
            eng.synchronize()
            walls.append(time.perf_counter() - t0)
        log(f"[{form}] {len(walls)} timed regions of {args.steps} steps after {args.warmup} warm-up steps: "
            + ", ".join(f"{w:.4f}s" for w in walls))
        wall = float(np.median(walls))
        st = eng.stats()
        value = n * args.steps / wall
        log(f"[{form}] timed {args.steps} steps: {wall:.3f}s, {value:.4g} particle-steps/s, rebuilds "
            f"{st['rebuilds']}, max list length {st['max_neighbours']}, list capacity {st['neighbour_capacity']}")

        # kernel roofline: HIP events around each launch on the engine stream
        eng.set_timing(True)
        eng.step(args.timing_steps)
        tm = eng.timing()
        eng.set_timing(False)
        f_ms, f_cnt = tm["force"]
        avg_s = f_ms / max(f_cnt, 1) * 1e-3
        # history tables (physics-fix mode): each stored entry read and written once
        alg_bytes = FORCE_BYTES_PER_PARTICLE * n + HIST_BYTES_PER_CONTACT * st.get("history_contacts", 0)
        achieved = alg_bytes / avg_s / 1e9
        res = {"value": value, "ms_per_step": wall / args.steps * 1e3, "force_kernel_avg_us": avg_s * 1e6,
               "achieved_gbs": achieved, "contacts_per_step": st["contacts"],
               "contacts_per_s": st["contacts"] * args.steps / wall,
               "repeat_values": [n * args.steps / w for w in walls],
               "jkr_failures_per_step": st["jkr_failures"], "history_contacts_per_step": st.get("history_contacts", 0),
               "wall": wall, "stats": st,
               "autotuned_skin_m": tuned[0] if tuned else None, "autotuned_cell_m": tuned[1] if tuned else None,
               "tiling": eng.tiling()}
        res.update(surrogate_kernels(tm))
        log(f"[{form}] force kernel: {f_cnt} launches, avg {avg_s * 1e6:.1f} us, {achieved:.1f} GB/s algorithmic; "
            f"rebuild chain avg {tm['rebuild'][0] / max(tm['rebuild'][1], 1) * 1e3:.1f} us"
            + (f"; collect avg {res['collect_avg_us']:.1f} us, network avg {res['network_avg_us']:.1f} us"
               if "network_avg_us" in res else ""))
        if form == "surrogate":
            eng.set_jkr_solver("surrogate", net, check=True)
            errs = []
            for _ in range(5):
                eng.step(1)
                errs.append(eng.jkr_stats())
            surr_info.update({
                "served_per_step": float(np.mean([e["served"] for e in errs])),
                "fallback_per_step": float(np.mean([e["fallback"] for e in errs])),
                "force_rel_err_max": max(e["err_max"] for e in errs),
                "force_rel_err_mean": float(np.mean([e["err_mean"] for e in errs])),
                "network_force_rel_err_max": max(e["raw_max"] for e in errs),
                "network_force_rel_err_mean": float(np.mean([e["raw_mean"] for e in errs])),
                "rows_capacity": errs[-1]["rows_capacity"],
            })
            res.update(surrogate_kernels(tm, surr_info["served_per_step"] + surr_info["fallback_per_step"]))
            log(f"[surrogate] check: {surr_info}")
        results[form] = res
    main_form = forms[-1]
    R = results[main_form]
    value, wall, st, avg_s, achieved = R["value"], R["wall"], R["stats"], R["force_kernel_avg_us"] * 1e-6, \
        R["achieved_gbs"]
    # what a plain streaming copy reaches on this GPU (context for the roofline;
    # with --calibrate-copy also the reference point of a counter profile)
    nbytes = (1 << 30) if not args.calibrate_copy else 256 * 1024 * 1024
    ms = native.probe_copy(nbytes)
    copy_gbs = 2 * nbytes / ms / 1e6
    log(f"copy probe: k_copy_probe read {nbytes} B + wrote {nbytes} B in {ms:.3f} ms ({copy_gbs:.0f} GB/s)")
    traffic, traffic_src = traffic_of(wname)

    cpu = None
    if not args.no_cpu:
        info = cpu_info()
        threads = args.cpu_threads or info["threads"]
        cell = 2.0 * float(r.max()) * (1.5 if p.normal_model in (abi.RP_NORMAL_JKR, abi.RP_NORMAL_SJKR) else 1.0)
        # the drum's 16M spheres x 16,384 triangles (brute-force wall loop) take minutes per CPU step:
        # its sample is the bed's first 20,000 sites (bottom layers) in the same drum
        ns = n if args.workload not in ("drum", "drum-fix") else min(n, 20_000)
        what = f"the full {n}-particle workload" if ns == n else f"the first {ns} of the {n} particles"
        log(f"cpu baseline: oracle, {threads} threads ({info['host_cpus']} host CPUs, {info['cpu_model']}), "
            f"~{args.cpu_seconds:g}s of {what}")
        fx = None if fixed is None else fixed[:ns]
        cval, csec, csteps = cpu_baseline(p, x[:ns], r[:ns], m[:ns], None if v is None else v[:ns], meshes,
                                          args.cpu_steps, threads, cell, args.cpu_seconds,
                                          mesh_motions(args.workload), fixed=fx)
        log(f"cpu baseline f64 {cval:.4g} particle-steps/s ({csec:.1f}s)")
        # the reference's own precision (f32), the same sample
        fval, fsec, fsteps = cpu_baseline(p, x[:ns], r[:ns], m[:ns], None if v is None else v[:ns], meshes,
                                          args.cpu_steps, threads, cell, 0.5 * args.cpu_seconds,
                                          mesh_motions(args.workload), bits=32, fixed=fx)
        log(f"cpu baseline f32 {fval:.4g} particle-steps/s ({fsec:.1f}s)")
        # as shipped: the reference's 1 m grid cell (simulation.rs:66), every
        # walk appending all particles of up to 27 cells -- O(N^2) per step,
        # so a sample of particles' gathers is timed and scaled to the workload
        import oracle
        oracle.lib().rpo_set_threads(threads)
        o = oracle.OracleSimulation(p.dt, None, None, params=abi.RpParams.from_buffer_copy(p))
        o.set_state(x[:ns], r[:ns], m[:ns], fx, v=None if v is None else v[:ns])
        nsamp = 2000 if ns <= 100_000 else 64
        gsec = o.time_gather(nsamp, 1.0)
        del o
        aval = nsamp / gsec if gsec > 0 else None
        log(f"cpu baseline as shipped (1 m cells): {nsamp} gathers in {gsec:.2f}s -> {aval:.4g} particle-steps/s")
        cpu = {"value": cval, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{csteps} step(s) of {what} after 1 warm-up step, FP64, "
                         f"reference grid cell {cell * 1e3:.2f} mm (largest contact cutoff, the cell-size "
                         f"heuristic of simulation.rs:64), brute-force {ntri}-triangle wall loop as in "
                         f"simulation.rs:465-523; {csec:.1f}s",
               "cpu_model": info["cpu_model"], "host_cpus": info["host_cpus"],
               "threads_note": "the GPU box allots 16 host cores per GPU (OMP_NUM_THREADS)",
               "f32_value": fval, "f32_sample": f"{fsteps} step(s), the same sample in f32 (the reference's "
                                                f"precision); {fsec:.1f}s",
               "as_shipped_value": aval,
               "as_shipped_sample": f"the gathers of {nsamp} particles spread over {what} with the reference's "
                                    f"1 m grid cell (simulation.rs:66; each walk appends every particle of up to 27 "
                                    f"cells), {gsec:.2f}s, scaled to particle-steps/s; walls and integration "
                                    f"not included"}

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
        "config": {"workload": f"{wname}: {n} spheres, {ntri}-triangle STL, {desc}, dt={p.dt:g}, FP64",
                   "particles": n, "triangles": ntri, "form": main_form},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src},
        "stream_copy_gbs": copy_gbs,
        "cpu_baseline": cpu,
        "repeats": len(R["repeat_values"]),
        "repeat_values": R["repeat_values"],
        "contacts_per_step": st["contacts"],
        "contacts_per_s": R["contacts_per_s"],
        "wall_contacts_per_step": st["wall_contacts"],
        "force_kernel_avg_us": avg_s * 1e6,
        "jkr_failures_per_step": st["jkr_failures"],
        "history_contacts_per_step": R["history_contacts_per_step"],
        "autotuned_skin_m": R["autotuned_skin_m"], "autotuned_cell_m": R["autotuned_cell_m"], "tiling": R["tiling"],
    }
    if len(results) > 1 or main_form != "plain":
        out["forms"] = {f: {k: v for k, v in res.items() if k not in ("stats", "wall")} for f, res in results.items()}
    if surr_info:
        out["surrogate"] = surr_info
    return out




def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="steps per timed region (5000; drum 1000)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="hopper")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=1, help="least CPU-baseline steps")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU-baseline sample length: as many steps as fit, going by the warm-up step")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel on its own (no HIP-graph step loop), e.g. under a profiler")
    ap.add_argument("--no-compare", action="store_true",
                    help="batch workloads: skip the one-engine-per-instance comparison")
    ap.add_argument("--timing-steps", type=int, default=100)
    ap.add_argument("--repeats", type=int, default=3,
                    help="timed regions of --steps steps, each on a fresh engine from the initial state; "
                         "value = the median")
    ap.add_argument("--jkr-forms", default=None,
                    help="JKR solvers to time on the JKR workloads, the last one is the reported value "
                         "(default exact,shortcut,surrogate; drum shortcut,surrogate)")
    ap.add_argument("--train-samples", type=int, default=1_000_000)
    ap.add_argument("--train-epochs", type=int, default=40)
    ap.add_argument("--instances", type=int, default=None, help="ensemble size (256 batch, 64 jkr-batch)")
    ap.add_argument("--physics-fix", action="store_true",
                    help="the physics-fix mode (real shear modulus, Mindlin springs with their tangential "
                         "history); batch workloads switch to Mindlin and draw a friction per instance")
    ap.add_argument("--autotune", type=int, default=None,
                    help="steps per chunk of the engine's skin / cell-size autotuner before timing (0 = off; "
                         "default 20 for the steady jkr and batch workloads, off for the hopper, whose particles "
                         "start at rest and speed up, so an early tuning picks a skin too small for later)")
    ap.add_argument("--no-extras", action="store_true",
                    help="hopper: leave out the contact window and the surrogate+batched forms")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run the library's 256 MiB copy kernel so a counter profile of this run can "
                         "scale FETCH_SIZE/WRITE_SIZE by a kernel of known traffic")
    args = ap.parse_args()
    if args.autotune is None:
        args.autotune = 0 if args.workload in ("hopper", "box", "drum", "drum-fix") else 20
    if args.steps is None:
        args.steps = 1000 if args.workload in ("drum", "drum-fix") else 5000
    if args.jkr_forms is None:
        args.jkr_forms = "shortcut,surrogate" if args.workload in ("drum", "drum-fix") else "exact,shortcut,surrogate"
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs one domain on one GPU; multi-GPU runs are not "
                                   "supported (use --gpus 1)"}))
        return 2

    import numpy as np

    from rusty_particles_amd import abi, native
    if native.device_count() < 1:
        raise SystemExit("bench.py: no HIP device visible")
    if args.workload in ("batch", "jkr-batch"):
        print(json.dumps(run_batch(args)), flush=True)
        return 0
    out = run_single(args)
    if args.workload == "hopper" and not args.no_extras:
        out["contact_window"] = contact_window(args)
        out["surrogate_batched"] = surrogate_batched(args)
    print(json.dumps(out), flush=True)
    return 0


def contact_window(args) -> dict:
    """The hopper-bed workload (the hopper's spheres packed in the cone) with
    the same steps, warm-up and repeats: a window in which contacts and wall
    contacts carry force.  Summary of its JSON object."""
    a = argparse.Namespace(**vars(args))
    a.workload, a.n, a.no_cpu, a.autotune = "hopper-bed", None, True, 0
    log("contact window: hopper-bed workload")
    r = run_single(a)
    return {"workload": r["config"]["workload"], "particles": r["config"]["particles"], "value": r["value"],
            "unit": "particle-steps/s", "ms_per_step": r["ms_per_step"], "contacts_per_step": r["contacts_per_step"],
            "wall_contacts_per_step": r["wall_contacts_per_step"], "contacts_per_s": r["contacts_per_s"],
            "force_kernel_avg_us": r["force_kernel_avg_us"], "roofline": r["roofline"],
            "repeat_values": r["repeat_values"]}


def surrogate_batched(args) -> dict:
    """The jkr-batch workload (64 periodic JKR packings, surface-energy sweep)
    in its three forms -- the reference's exact solve, the shortcut (identical
    results) and the surrogate+batched form -- with the same steps and
    warm-up, one timed region each, no autotuning (the same settings for every
    form); the surrogate's force errors from check mode."""
    a = argparse.Namespace(**vars(args))
    a.workload, a.instances, a.jkr_forms, a.autotune = "jkr-batch", 64, "exact,shortcut,surrogate", 0
    a.physics_fix = False
    a.no_compare, a.no_cpu, a.repeats, a.timing_steps = True, True, 1, 50
    log("surrogate+batched: jkr-batch workload")
    r = run_batch(a)
    f = r["forms"]
    s = r.get("surrogate", {})
    return {"workload": r["config"]["workload"], "particles": r["config"]["particles"],
            "forms": {k: {"value": v["value"], "ms_per_step": v["ms_per_step"],
                          "force_kernel_avg_us": v["force_kernel_avg_us"],
                          **{q: v[q] for q in ("collect_avg_us", "network_avg_us", "network_rows_per_step",
                                               "network_tflops", "network_mfma_frac") if q in v}}
                      for k, v in f.items()},
            "surrogate_vs_exact": f["surrogate"]["value"] / f["exact"]["value"],
            "surrogate_vs_shortcut": f["surrogate"]["value"] / f["shortcut"]["value"],
            "force_rel_err_max": s.get("force_rel_err_max"), "force_rel_err_mean": s.get("force_rel_err_mean"),
            "served_per_step": s.get("served_per_step"), "fallback_per_step": s.get("fallback_per_step"),
            "train_s": s.get("train_s")}



if __name__ == "__main__":
    sys.exit(main())
