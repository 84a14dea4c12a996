#!/usr/bin/env python3
"""Throughput of the DEM time step on one MI355X.

Workload (default): 1,000,000 polydisperse FP64 spheres, radii U[0.8, 1.2] mm,
triply periodic cube at solid fraction 0.30 (random sequential addition),
v ~ N(0, 0.1^2) m/s, no gravity, Hertz + Coulomb contact (E* 1e7 Pa,
G* = E*/4, en 0.85, mu 0.4), cell edge 2 * 1.2 * r_max, dt 5e-6 s.

One step = one kick-drift-kick time step of every particle (binning, contact
forces, integration).  value = particles * timed steps / seconds of the timed
region, state resident in HBM.

Prints one JSON line with the throughput, the roofline of the dominant kernel
(the integrator kernel k_integ of the packed sphere step, which streams the
state and sums the contact records; timed with HIP events on the context's
stream), the whole step's bandwidth, the list rebuilds inside the timed window,
and CPU baselines (the oracle restatement of the reference's CPU step on one
thread and on every host core, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 1M spheres; batched calibration gradients/s vs plain"
HBM_PEAK_GBS = 8000.0
# compulsory HBM bytes per particle-step (x, v, w, q, r read; x, v, w, q
# written; cell key and permutation) and per contact (one contact record
# written and read): the algorithmic traffic of one step
BYTES_PER_PARTICLE_STEP = 232.0
BYTES_PER_CONTACT = 112.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_integ_periodic_hertz_1M_traffic.txt")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cached_workload(n, seed):
    from pacific_amd import workloads
    cache_dir = os.path.join(os.environ.get("TMPDIR", "/tmp"), "gdem_cache")
    os.makedirs(cache_dir, exist_ok=True)
    path = os.path.join(cache_dir, f"periodic_hertz_{n}_s{seed}.npz")
    wl = None
    if os.path.exists(path):
        try:
            z = np.load(path)
            wl = workloads.periodic_hertz(16, seed)  # structure; arrays replaced below
            r = z["radius"]
            wl["particles"] = workloads._particles(z["x"], z["v"], r)
            wl["size"] = tuple(z["size"])
            wl["cell_size"] = float(z["cell_size"])
            wl["name"] = f"periodic_hertz_{n}"
        except Exception:
            wl = None
    if wl is None:
        log(f"generating {n} spheres by random sequential addition (seed {seed})")
        t = time.time()
        wl = workloads.periodic_hertz(n, seed)
        log(f"generated in {time.time() - t:.1f} s")
        p = wl["particles"]
        np.savez(path, x=p["x"], v=p["v"], radius=p["radius"], size=np.array(wl["size"]),
                 cell_size=wl["cell_size"])
    return wl


def cached_convex(n, seed):
    from pacific_amd import workloads
    cache_dir = os.path.join(os.environ.get("TMPDIR", "/tmp"), "gdem_cache")
    os.makedirs(cache_dir, exist_ok=True)
    path = os.path.join(cache_dir, f"convex_box_{n}_s{seed}_phi025.npz")
    wl = workloads.convex_box(64, seed)  # structure; arrays replaced below
    if os.path.exists(path):
        z = np.load(path)
        wl["particles"] = {k[2:]: z[k] for k in z.files if k.startswith("p_")}
        wl["size"] = tuple(z["size"])
        wl["walls"] = workloads.box_walls((0.0, 0.0, 0.0), wl["size"], 1)
    else:
        log(f"generating {n} convex particles by random sequential addition (seed {seed})")
        t = time.time()
        wl = workloads.convex_box(n, seed)
        log(f"generated in {time.time() - t:.1f} s")
        np.savez(path, size=np.array(wl["size"]),
                 **{"p_" + k: v for k, v in wl["particles"].items()})
    wl["name"] = f"convex_box_{n}"
    return wl


def read_traffic():
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(wl, steps, warm=2, threads=1):
    """The oracle (restatement of the reference CPU step): one thread, or the
    all-core variant (slabs of cell rows in two colours, OpenMP)."""
    from oracle import py_oracle
    o = py_oracle.ConvexOracle(wl) if "shapes" in wl else py_oracle.Oracle(wl)
    o.prime()
    step = o.step if threads == 1 else (lambda k, dt: o.step_parallel(k, dt, threads))
    step(warm, wl["dt"])
    t = time.perf_counter()
    step(steps, wl["dt"])
    el = time.perf_counter() - t
    n = len(wl["particles"]["x"])
    return n * steps / el, el


def host_cores():
    """Threads this process may use (the box's CPU share), and the CPU model."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return max(n, 1), model


def run_batch(args):
    """Batched calibration-sweep systems: one persistent launch for all of
    them, against looping the single-system step over a sample of them."""
    from pacific_amd import autotune, gdem, workloads
    steps = args.steps or 1000
    warm = 10 if args.warmup is None else args.warmup
    t = time.time()
    systems = workloads.batch_systems(args.systems, seed=4 if args.seed is None else args.seed)
    ntot = sum(len(w["particles"]["radius"]) for w in systems)
    nsub = sum(len(w["particles"]["radius"]) * w["substeps"] for w in systems)
    dt = systems[0]["dt"]
    log(f"{len(systems)} systems, {ntot} particles, generated in {time.time() - t:.1f} s")
    cache = os.path.join(os.environ.get("TMPDIR", "/tmp"), "gdem_autotune.json")
    tuning = autotune.tune(systems, dt, sample=min(256, len(systems)), steps=20, cache=cache, log=log)
    log(f"autotune: workgroup {tuning['block_size']}, cell factor {tuning['cell_factor']}")
    tuned = autotune.with_cell_factor(systems, tuning["cell_factor"])
    b = gdem.build_batch(tuned, block_size=tuning["block_size"])
    b.prime()
    b.step(warm, dt)
    b.synchronize()
    log(f"warmup {warm} steps done")
    b.set_kernel_timing(True)
    t0 = time.perf_counter()
    b.step(steps, dt)
    b.synchronize()
    el = time.perf_counter() - t0
    cnt = b.counters()
    value = ntot * steps / el
    log(f"timed {steps} steps in {el:.3f} s (kernel {cnt['kernel_ms']:.1f} ms): "
        f"{value:.4e} particle-steps/s")
    b.close()
    # the plain single-system step looped over a sample of the systems
    sample = autotune.sample_systems(tuned, 32)
    loop_steps = min(steps, 100)
    tot_ps, tot_t = 0, 0.0
    for wl in sample:
        sub = wl["substeps"]
        s1 = gdem.build_system(wl)
        s1.prime()
        s1.step(5 * sub, dt / sub)
        s1.synchronize()
        t1 = time.perf_counter()
        s1.step(loop_steps * sub, dt / sub)
        s1.synchronize()
        tot_t += time.perf_counter() - t1
        tot_ps += len(wl["particles"]["radius"]) * loop_steps
        s1.close()
    loop_rate = tot_ps / tot_t
    log(f"looped single-system step: {loop_rate:.4e} particle-steps/s (32 systems x {loop_steps} steps)")
    cpu = None
    if not args.no_cpu:
        from oracle import py_oracle
        cs = args.cpu_steps or 1000
        cps, cel = 0, 0.0
        for wl in sample:
            sub = wl["substeps"]
            o = py_oracle.Oracle(wl)
            o.prime()
            t2 = time.perf_counter()
            o.step(cs * sub, dt / sub)
            cel += time.perf_counter() - t2
            cps += len(wl["particles"]["radius"]) * cs
        cpu = {"value": cps / cel, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"the same 32 systems ({sum(len(w['particles']['radius']) for w in sample)} "
                         f"spheres) x {cs} steps, one after the other, single thread ({cel:.1f} s)"}
        log(f"CPU {cpu['value']:.4e} particle-steps/s ({cel:.1f} s)")
    out = {
        "metric": "particle-steps/s, 4,096 batched settling systems (secondary line)",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": el / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batch_{len(systems)}", "systems": len(systems),
                   "n_particles": ntot, "particles_per_system": [64, 2048],
                   "contact": "Hooke+Coulomb, per-system kn/en/mu", "dt": dt,
                   "kn_range": [1.0e4, 1.0e6],
                   "substeps": "per system, omega dt / substeps <= 1 (1 to 6)",
                   "block_size": tuning["block_size"], "cell_factor": tuning["cell_factor"],
                   "parallelism": "single GPU, one persistent launch"},
        "substep_particle_steps_per_s": nsub * steps / el,
        "looped_single_system": {"value": loop_rate, "sample": f"32 systems x {loop_steps} steps"},
        "speedup_vs_looped": value / loop_rate,
        "autotune": tuning["table"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_calibration(args, nsys=None):
    """Batched adjoint gradients of calibration trajectories against central
    finite differences looped over systems with the single-system step."""
    from pacific_amd import gdem, workloads
    nsys = nsys or (args.systems if args.systems != 4096 else 512)
    nsteps, K, obs, dt = 400, 50, (100, 200, 300, 400), 1.0e-5
    t = time.time()
    systems, hidden = workloads.calibration_systems(nsys, n=1000, seed=5)
    v0 = np.concatenate([w["particles"]["v"] for w in systems])
    log(f"{nsys} calibration systems of 1000 spheres generated in {time.time() - t:.1f} s")
    eng = gdem.GdemDiff(systems, nsteps, dt, K, obs)
    eng.set_parameters(hidden, v0)
    target = eng.record()
    eng.set_targets(target)
    guess = np.array([1.2, 0.9, 1.1])
    theta = hidden * guess[None, :]
    eng.set_parameters(theta, v0)
    eng.gradient()  # warm-up
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        loss, gth, gv0 = eng.gradient()
    tg = (time.perf_counter() - t0) / reps
    rate = nsys / tg
    log(f"adjoint: {tg:.3f} s per launch for {nsys} systems: {rate:.1f} gradients/s")
    # accuracy 1: the reverse sweep against the forward-mode (dual tangent)
    # derivative along random directions, on this configuration
    rng = np.random.default_rng(1)
    jvp_worst = 0.0
    for _ in range(2):
        dth = rng.normal(size=theta.shape) * theta
        dv = rng.normal(size=v0.shape)
        _, dl = eng.directional(dth, dv)
        off = eng.offsets
        for s in range(nsys):
            want = float((gth[s] * dth[s]).sum() + (gv0[off[s]:off[s + 1]] * dv[off[s]:off[s + 1]]).sum())
            jvp_worst = max(jvp_worst, abs(dl[s] - want) / max(abs(want), 1e-300))
    log(f"adjoint vs forward mode: max relative error {jvp_worst:.2e}")
    eng.close()
    # accuracy 2: central differences of the loss over a 48-step horizon of 4
    # of the same systems (over 400 steps the colliding packing makes the loss
    # far rougher than its exact derivative at any step differences can
    # take); the step is chosen per component (calibration.central_difference)
    from pacific_amd import calibration
    nfd = 4
    short = gdem.GdemDiff(systems[:nfd], 48, dt, 10, (16, 32, 48))
    nv = sum(len(w["particles"]["radius"]) for w in systems[:nfd])
    short.set_parameters(hidden[:nfd], v0[:nv])
    short.set_targets(short.record())
    short.set_parameters(theta[:nfd], v0[:nv])
    _, sg, _ = short.gradient()
    worst_norm, worst_comp, nres = 0.0, 0.0, 0
    for s in range(nfd):
        fd, agree = np.zeros(3), np.zeros(3)
        for c in range(3):
            def f(h):
                t = theta[:nfd].copy()
                t[s, c] += h
                short.set_parameters(t)
                return short.loss()[s]
            fd[c], agree[c] = calibration.central_difference(f, theta[s, c])
        short.set_parameters(theta[:nfd])
        nw, cw, nr = calibration.gradient_error(sg[s], fd, agree, theta[s])
        worst_norm, worst_comp, nres = max(worst_norm, nw), max(worst_comp, cw), nres + nr
    short.close()
    log(f"adjoint vs central differences (48 steps, {nfd} systems): normwise {worst_norm:.2e}, "
        f"{nres} of {3 * nfd} components resolved, worst {worst_comp:.2e}")
    # the single-system step looped over systems, gradient by central differences
    nloop = 0 if getattr(args, "no_looped", False) else 4
    t1 = time.perf_counter()
    for s in range(nloop):
        wl = systems[s]
        for c in range(3):
            for sign in (1.0, -1.0):
                th = theta[s].copy()
                th[c] += sign * 1e-6 * theta[s, c]
                w = dict(wl)
                hooke = (th[0], th[1], -1.0, th[2], 0.0)
                w["models"] = [(0, 0, gdem.MODEL_HOOKE, hooke), (0, 1, gdem.MODEL_HOOKE, hooke)]
                one = gdem.build_system(w)
                last = 0
                sub = wl["substeps"]
                for step in obs:
                    one.step((step - last) * sub, dt / sub)
                    last = step
                    one.get_state()
                one.close()
    loop_rate = None
    if nloop:
        tl = (time.perf_counter() - t1) / nloop
        loop_rate = 1.0 / tl
        log(f"looped single-system step with central differences: {tl:.3f} s per gradient "
            f"({loop_rate:.2f} gradients/s)")
    return {
        "metric": "batched calibration gradients/s (secondary line)", "value": rate,
        "unit": "gradients/s", "n_gpus": 1, "steps": nsteps, "warmup": 1,
        "ms_per_step": tg / nsteps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"calibration_{nsys}x1000", "systems": nsys, "particles_per_system": 1000,
                   "window": nsteps, "checkpoint_every": K, "observations": list(obs), "dt": dt,
                   "gradient": "d loss / d (kn, en, mu) and d loss / d v0 per system",
                   "parallelism": "single GPU, one persistent launch"},
        "accuracy": {"vs_forward_mode_max_rel": jvp_worst,
                     "vs_central_differences_max_rel": worst_norm,
                     "vs_central_differences_resolved_components_max_rel": worst_comp,
                     "central_differences_resolved_components": f"{nres} of {3 * nfd}",
                     "central_differences": f"48-step horizon, {nfd} systems; step per "
                                            "component among 1e-4..1e-8 relative (best "
                                            "agreement of consecutive steps); max_rel is "
                                            "normwise over theta_i dL/dtheta_i, resolved "
                                            "components agree to 1e-6 between steps"},
        "looped_single_system_fd": {"value": loop_rate, "unit": "gradients/s",
                                    "sample": f"{nloop} systems, 6 runs of {nsteps} steps each"},
        "speedup_vs_looped": rate / loop_rate if loop_rate else None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="500 (spheres) / 300 (convex)")
    ap.add_argument("--warmup", type=int, default=None, help="20 (spheres) / 10 (convex)")
    ap.add_argument("--workload", choices=["periodic_hertz", "convex_box", "batch", "calibration"],
                    default="periodic_hertz",
                    help="periodic_hertz: 1M-sphere throughput case (default); convex_box: "
                         "200k cubes + tetrahedra with GJK; batch: 4,096 small settling "
                         "systems in one persistent launch (secondary lines)")
    ap.add_argument("--systems", type=int, default=4096, help="batch: number of systems")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=None, help="20 (spheres) / 5 (convex)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-calibration", action="store_true",
                    help="skip the batched calibration-gradient part of the default line")
    ap.add_argument("--no-looped", action="store_true",
                    help="calibration: skip the looped single-system baseline")
    ap.add_argument("--narrow-phase", choices=["GJK", "GJK_SV"], default="GJK",
                    help="convex_box: Johnson (default) or signed-volume GJK")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the DEM step runs on one GPU: it is not sharded across "
                                   "GPUs (the multi-GPU decomposition is a design only)"}))
        sys.exit(2)

    from pacific_amd import gdem
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "calibration":
        print(json.dumps(run_calibration(args)), flush=True)
        return
    convex = args.workload == "convex_box"
    if args.steps is None:
        args.steps = 300 if convex else 500
    if args.warmup is None:
        args.warmup = 10 if convex else 20
    if args.cpu_steps is None:
        args.cpu_steps = 5 if convex else 20
    if convex:
        wl = cached_convex(args.n or 200_000, 3 if args.seed is None else args.seed)
        wl["narrow_phase"] = args.narrow_phase
    else:
        wl = cached_workload(args.n or 1_000_000, 2 if args.seed is None else args.seed)
    n = len(wl["particles"]["x"])
    dt = wl["dt"]
    sysg = gdem.build_system(wl)
    nc, edge = sysg.grid()
    log(f"grid {nc}, edge {edge[0]:.4e} m, size {tuple(round(v, 5) for v in wl['size'])} m")
    sysg.prime()
    sysg.synchronize()
    sysg.step(args.warmup, dt)
    sysg.synchronize()
    log(f"warmup {args.warmup} steps done")

    # timed region: K steps, no instrumentation
    sysg.reset_counters()
    rebuilds0 = sysg.get_skin()[1] if not convex else 0
    sysg.synchronize()
    t0 = time.perf_counter()
    sysg.step(args.steps, dt)
    sysg.synchronize()
    el = time.perf_counter() - t0
    cnt = sysg.counters()
    rebuilds = (sysg.get_skin()[1] - rebuilds0) if not convex else None
    mean_contacts = cnt["contacts"] / max(cnt["force_evals"], 1)
    value = n * args.steps / el
    log(f"timed {args.steps} steps in {el:.4f} s: {value:.4e} particle-steps/s, "
        f"mean contacts {mean_contacts:.0f}, list rebuilds {rebuilds}")

    # kernel durations with HIP events on the context's stream (the extended
    # launch records each kernel's own start and end)
    ksteps = min(args.steps, 200)
    sysg.reset_counters()
    sysg.set_kernel_timing(True)
    sysg.step(ksteps, dt)
    sysg.synchronize()
    kc = sysg.counters()
    kt = sysg.kernel_times()
    sysg.set_kernel_timing(False)
    k_ms = kc["force_kernel_ms"] / max(kc["timed_launches"], 1)
    pair_ms = kt["ms"][1] / kt["launches"][1] if kt["launches"][1] else None
    k_contacts = kc["contacts"] / max(kc["force_evals"], 1)
    alg_bytes = BYTES_PER_PARTICLE_STEP * n + BYTES_PER_CONTACT * k_contacts
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    step_bytes = BYTES_PER_PARTICLE_STEP * n + BYTES_PER_CONTACT * mean_contacts
    step_achieved = step_bytes * args.steps / el / 1e9
    log(f"force/integrator kernel {k_ms * 1e3:.1f} us/launch, {achieved:.0f} GB/s algorithmic; "
        f"pair kernel {pair_ms * 1e3 if pair_ms else 0:.1f} us/launch")
    sysg.close()

    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline: oracle, 1 thread, {args.cpu_steps} steps of the same workload")
        cv, cel = cpu_baseline(wl, args.cpu_steps)
        log(f"CPU {cv:.4e} particle-steps/s ({cel:.1f} s)")
        one = {"value": cv, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"{n} particles of the same workload, {args.cpu_steps} steps after 2 "
                         f"warm-up steps, single thread ({cel:.1f} s)"}
        cpu = one
        cores, model = host_cores()
        if not convex and cores > 1:
            asteps = 4 * args.cpu_steps
            log(f"CPU baseline: oracle, {cores} threads, {asteps} steps")
            av, ael = cpu_baseline(wl, asteps, threads=cores)
            log(f"CPU all cores {av:.4e} particle-steps/s ({ael:.1f} s)")
            cpu = {"value": av, "unit": "particle-steps/s", "cores": cores, "kind": "port",
                   "sample": f"{n} particles of the same workload, {asteps} steps after 2 "
                             f"warm-up steps, {cores} OpenMP threads: cell-row slabs swept "
                             f"in two colours ({ael:.1f} s)",
                   "cpu_model": model, "one_core": one,
                   "gpu_vs_all_cores": value / av, "gpu_vs_one_core": value / cv}

    if convex:
        out = {
            "metric": "particle-steps/s at 200k convex particles (secondary line)", "value": value,
            "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": wl["name"], "n_particles": n, "shapes": "cube 2 mm / tetrahedron 2.5 mm",
                       "solid_fraction": 0.25, "dt": dt,
                       "contact": f"Hooke+Coulomb, {wl.get('narrow_phase', 'GJK')}",
                       "grid": nc, "parallelism": "single GPU"},
            "detect_force_us": k_ms * 1e3, "mean_contacts": k_contacts,
            "cpu_baseline": cpu,
        }
        print(json.dumps(out), flush=True)
        return
    calib = None
    if not args.no_calibration:
        # second half of the metric: batched calibration gradients/s against
        # the single-system step looped over systems (central differences)
        c = run_calibration(args, nsys=512)
        calib = {k: c[k] for k in ("value", "unit", "config", "accuracy", "looped_single_system_fd",
                                   "speedup_vs_looped")}
    traffic = read_traffic()
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"periodic_hertz_{n}", "n_particles": n, "solid_fraction": 0.30,
                   "radius_mm": [0.8, 1.2], "contact": "Hertz+Coulomb", "dt": dt,
                   "cell_factor": 1.2, "grid": nc, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "counter_frac": (traffic / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
                                      if traffic else None),
                     "kernel": "k_integ<MODE_MOVE>", "kernel_us": k_ms * 1e3,
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "mean_contacts": k_contacts},
        "pair_kernel_us": pair_ms * 1e3 if pair_ms else None,
        "rebuilds_in_timed_window": rebuilds,
        "step_roofline": {"achieved": step_achieved, "unit": "GB/s",
                          "frac": step_achieved / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "calibration_gradients": calib,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
