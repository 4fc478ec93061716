#!/usr/bin/env python
"""Throughput of the DEM time step on one MI355X.

Workload: the polydisperse periodic chute -- 2,000,000 flowing spheres with
d ~ U[0.5, 1.5] (three grid levels) over a rough base of fixed spheres,
tilted 24 degrees, periodic in x and y, HertzianViscoelasticMindlin with
e = 0.8 and mu = 0.5, dt = t_c / 50, FP64.

Protocol:
  1. preparation (untimed): the spheres start on a seeded, non-overlapping
     rock-salt lattice and fall onto the rough base under |g| = 1 straight
     down until the bed is at rest -- a random loose packing 30 diameters
     deep (settle_steps);
  2. the configured run: 50,000 steps under the chute gravity (|g| = 1, 24
     degrees), timed chunk by chunk (run_average, and steady_window over the
     run's last chunks, rebuilds included);
  3. after W warm-up steps, K steps are timed: `value`.

Prints one JSON line: particle-steps/s (flowing spheres x steps / s), the
roofline of the dominant kernel (the force kernel, which also integrates the
step; HIP-event timed on the engine's stream) with the whole-step byte model
264 + 64 C per particle, and a CPU baseline (the oracle's OpenMP build on the
host's cores, on a bounded sample of the same state).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per force-kernel launch on this workload, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE passes (calibrated against a plain copy); see the
# file for the commands and the scaling
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_force_chute2M_traffic.md")


def measured_traffic():
    """(traffic_bytes_per_launch, contacts_per_particle of the profiled state)
    from the committed counter summary, or (None, None)."""
    traffic, cbar = None, None
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = int(line.split(":", 1)[1])
            if line.startswith("contacts_per_particle:"):
                cbar = float(line.split(":", 1)[1])
    except OSError:
        pass
    return traffic, cbar


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=2_000_000, help="flowing spheres")
    ap.add_argument("--settle-steps", type=int, default=80000,
                    help="preparation: steps of the lattice falling onto the base under |g| straight down")
    ap.add_argument("--run-steps", type=int, default=50000,
                    help="the configured run before the timed window: steps from the settled packing under the "
                         "chute gravity, timed chunk by chunk (run_average, steady_window)")
    ap.add_argument("--chunk", type=int, default=500, help="steps per timed chunk of the configured run")
    ap.add_argument("--steady-steps", type=int, default=2000,
                    help="the last chunks of the configured run reported as steady_window")
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="steps of the OpenMP oracle in the CPU baseline (0: skip)")
    ap.add_argument("--skin", type=float, default=0.125,
                    help="Verlet skin (0: the engine's default, 0.4 x the smallest radius = 0.1); 0.125 ran "
                         "fastest of 0.05 .. 0.2 on the configured chute state")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="contact arithmetic: fp64 (the metric) or mixed (FP32 contact geometry and forces)")
    ap.add_argument("--mixed-steps", type=int, default=500,
                    help="after the FP64 timing, time this many steps of the same state in mixed precision (0: skip)")
    ap.add_argument("--ensemble", type=int, default=0,
                    help="instead of the chute: this many calibration heaps (sizes log-uniform in [500, 20000]) "
                         "stepped as one packed ensemble, against the same heaps run one at a time")
    ap.add_argument("--ensemble-steps", type=int, default=30000, help="steps per simulation for sims/hour")
    ap.add_argument("--churn", type=int, default=0,
                    help="instead of the timed chute steps: the cost of dem_add_particles and dem_remove_particles "
                         "during a run of the chute (--n flowing spheres), appending batches of this many spheres "
                         "and removing 1 to 20,000 of them")
    ap.add_argument("--save-state", default="",
                    help="after the configured run, write the particle state and stored interactions to this .npz")
    ap.add_argument("--load-state", default="",
                    help="start from a state written by --save-state (no preparation, no configured run): lets a "
                         "profiler cover the timed window's state only")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many bytes device-to-device (calibrates HBM counters) and exit")
    ap.add_argument("--list-sweep", choices=("particle", "wavefront"), default="particle",
                    help="Verlet-list build: one sweep of the neighbour cells per particle, or the wavefront-shared "
                         "sweep staged in LDS (dem_config.list_sweep)")
    ap.add_argument("--engine-lib", default="",
                    help="path of another build of libmercurydpm_dem.so to measure (an A/B comparison of builds); "
                         "default: the in-tree library")
    return ap.parse_args()


def run_ensemble(a):
    """Batched ensemble: a.ensemble calibration heaps (Drivers/Calibration/
    CalibrationHeap.cpp), particle counts log-uniform in [500, 20000], d ~
    U[0.8, 1.2], each with its own sliding friction U[0.1, 0.8], rolling
    friction U[0, 0.1] and restitution U[0.5, 0.95], its own base and casing
    (per-species wall groups), run together in one context for
    a.ensemble_steps steps of the heap protocol (fill, casing lifted once
    settled, spheres below the base deleted, angle of repose), with the layout
    and skin picked by the autotuner.  Against: the same protocol run one heap
    at a time (each at the faster of the default and the tuned skin), timed
    on 16 heap sizes spread over the range and interpolated (log-log) to
    every heap.  One JSON line: simulations per hour both ways."""
    import numpy as np

    from mercurydpm_fork_amd import dem
    from mercurydpm_fork_amd import ensemble as E
    from mercurydpm_fork_amd import scenarios as S

    def log(msg):
        print("# " + msg, file=sys.stderr, flush=True)

    params = E.draw_heap_params(a.ensemble, a.seed)
    members = [E.calibration_heap(p) for p in params]
    sizes = np.array([p.n for p in params])
    t_tune = time.time()
    schedule, skin, tune = E.autotune(members, log=log)
    t_tune = time.time() - t_tune
    ens = E.pack_heaps(members, schedule=schedule)
    e = dem.Engine(device=0, skin=skin)
    S.apply(e, ens.scenario)
    e.compute_forces()
    e.stats()
    t0 = time.perf_counter()
    run = E.run_heap_protocol(e, ens, a.ensemble_steps, log=log)
    t_packed = time.perf_counter() - t0
    e.close()
    log(f"packed: {a.ensemble} heaps, {int(sizes.sum())} spheres, {a.ensemble_steps} steps in {t_packed:.1f}s")

    sample = np.unique(np.geomspace(500, 20000, 16).astype(int))
    t_sample = []
    for k, n in enumerate(sample):
        hp = E.HeapParams(int(n), 0.45, 0.05, 0.725, 900 + k)
        one = E.pack_heaps([E.calibration_heap(hp)])
        best = None
        # each heap alone at the engine's default skin and at the ensemble's
        # tuned one; the faster counts
        for sk in (0.0, skin):
            g = dem.Engine(device=0, skin=sk)
            S.apply(g, one.scenario)
            g.compute_forces()
            g.stats()
            t1 = time.perf_counter()
            E.run_heap_protocol(g, one, a.ensemble_steps)
            t = time.perf_counter() - t1
            g.close()
            best = t if best is None else min(best, t)
        t_sample.append(best)
        log(f"sequential: heap of {n} spheres, {a.ensemble_steps} steps in {t_sample[-1]:.2f}s")
    t_sample = np.array(t_sample)
    t_seq = float(np.sum(np.exp(np.interp(np.log(sizes), np.log(sample), np.log(t_sample)))))
    per_hour = lambda t: 3600.0 * a.ensemble / t  # noqa: E731
    ok = np.isfinite(run.angle_of_repose)
    print(json.dumps({
        "metric": "calibration heaps per hour (batched ensemble)",
        "value": per_hour(t_packed), "unit": "simulations/hour", "n_gpus": 1,
        "steps": a.ensemble_steps, "ms_per_step": 1e3 * t_packed / a.ensemble_steps, "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic (seeded loose lattice fills of each heap's casing)",
        "config": {"workload": "CalibrationHeap protocol, LinearViscoelasticFriction with rolling friction, "
                               "per-heap sliding friction, rolling friction and restitution",
                   "simulations": int(a.ensemble), "particles": int(sizes.sum()), "size_range": [500, 20000],
                   "steps_per_simulation": a.ensemble_steps, "layout": schedule, "skin": skin,
                   "autotune": tune, "autotune_seconds": round(t_tune, 1)},
        "protocol": {"casings_lifted": int(np.isfinite(run.fill_time).sum()), "angles_measured": int(ok.sum()),
                     "angle_of_repose_mean": float(np.mean(run.angle_of_repose[ok])) if ok.any() else None,
                     "angle_at_end_mean": float(np.nanmean(run.angle_at_end)),
                     "angle_vs_sliding_friction_correlation": float(np.corrcoef(
                         [p.sliding_mu for p in params], np.nan_to_num(run.angle_at_end))[0, 1]),
                     "spheres_deleted": int(run.removed), "seconds": run.seconds},
        "sequential": {"value": per_hour(t_seq), "unit": "simulations/hour", "seconds_all": round(t_seq, 1),
                       "sample_seconds": {int(n): round(float(t), 3) for n, t in zip(sample, t_sample)}},
        "speedup_vs_sequential": t_seq / t_packed}))


def run_churn(a):
    """Insertion and removal during a run (ParticleHandler::copyAndAddObject
    of the insertion boundaries, removeObject of the deletion boundaries):
    the chute of a.n flowing spheres is uploaded and stepped; then five
    batches of a.churn spheres are appended above the bed (the first through
    the re-upload that leaves room, the others in place; every second one
    between dem_step_begin and dem_step_end), and 1, 100, 2000 and 20,000
    spheres at random handler indices are removed, each followed by 20
    steps.  One JSON line: the milliseconds of each call."""
    import numpy as np

    from mercurydpm_fork_amd import dem
    from mercurydpm_fork_amd import scenarios as S

    def log(msg):
        print("# " + msg, file=sys.stderr, flush=True)

    sc = S.chute_hvm(n_flow=a.n, seed=a.seed)
    st = sc.state
    l = sc.info["footprint"]
    top = float(st.x[:, 2].max())
    rng = np.random.default_rng(a.seed)
    nc = int(l / 1.6)
    cells = rng.permutation(nc * nc)
    flow = np.flatnonzero(st.inv_mass > 0)
    e = dem.Engine(device=0, skin=a.skin)
    S.apply(e, sc)
    e.compute_forces()
    e.step(10)
    n0 = e.n
    k = a.churn
    next_id = int(st.id.max()) + 1
    appends = []
    for b in range(5):
        c = cells[(b * k) % cells.size:][:k]
        src = rng.choice(flow, c.size)
        pos = np.stack([(c % nc + 0.5) * 1.6, (c // nc + 0.5) * 1.6, np.full(c.size, top + 2.0 + 1.6 * b)], 1)
        part = dem.State(id=np.arange(next_id, next_id + c.size, dtype=np.uint32), species=st.species[src],
                         radius=st.radius[src], inv_mass=st.inv_mass[src], inv_inertia=st.inv_inertia[src],
                         x=pos, v=np.zeros((c.size, 3)), w=np.zeros((c.size, 3)),
                         q=np.tile([1.0, 0.0, 0.0, 0.0], (c.size, 1)), f=None, t=None)
        next_id += c.size
        cap0 = e.stats().capacity
        mid = b % 2 == 1
        if mid:
            e.step_begin()
        t = time.perf_counter()
        e.add_particles(part)
        ms = 1e3 * (time.perf_counter() - t)
        if mid:
            e.step_end()
        cap = e.stats().capacity
        appends.append({"k": int(c.size), "ms": round(ms, 3), "in_place": bool(cap == cap0), "mid_step": mid,
                        "n_after": int(e.n), "capacity": int(cap)})
        log(f"append {c.size}: {ms:.2f} ms ({'in place' if cap == cap0 else 're-upload'})")
        e.step(20)
    removals = []
    for kr in (1, 100, 2000, 20000):
        n = e.n
        idx = [int(rng.integers(0, n - j)) for j in range(kr)]
        t = time.perf_counter()
        e.remove_particles(idx)
        ms = 1e3 * (time.perf_counter() - t)
        removals.append({"k": kr, "ms": round(ms, 3), "n_after": int(e.n)})
        log(f"remove {kr}: {ms:.2f} ms")
        e.step(20)
    x = e.download()["x"]
    if not np.all(np.isfinite(x)):
        raise RuntimeError("non-finite positions after the insertions and removals")
    e.close()
    print(json.dumps({"metric": "milliseconds per insertion / removal call during a run",
                      "unit": "ms", "n_gpus": 1, "higher_is_better": False, "dtype": "f64",
                      "data": "synthetic (the chute's loose lattice; appended spheres copy resident radii)",
                      "config": {"workload": "polydisperse chute, HertzianViscoelasticMindlin",
                                 "spheres_at_start": int(n0), "append_batch": k},
                      "append": appends, "remove": removals}))


_roctx = None


def profile_region(on):
    """Open (on) or close the region rocprofv3 --selected-regions records
    (roctxProfilerResume / roctxProfilerPause of the ROCTx library), so that a
    kernel trace or counter run of this command covers only the timed window.
    Without rocprofv3 the calls do nothing."""
    global _roctx
    if _roctx is None:
        import ctypes
        _roctx = False
        for name in ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so"):
            try:
                _roctx = ctypes.CDLL(name)
                break
            except OSError:
                continue
    if _roctx:
        (_roctx.roctxProfilerResume if on else _roctx.roctxProfilerPause)(0)


def b_alg(cbar, mixed=False):
    """Algorithmic HBM bytes per particle-step of the step (plain FP64,
    sliding-spring laws): state read x, v, omega (72) + radius, inverse mass
    (16) + id, species, flags (8); state written x, v, omega (72); force and
    torque written by the force phase and read by the closing half-kick (96;
    48 in mixed precision, where they are FP32); per active contact (pairs
    once, plus wall contacts) an 8-byte key and a 24-byte spring, read and
    written (64)."""
    return (216.0 if mixed else 264.0) + 64.0 * cbar


def fused_bytes(cbar, lbar):
    """The force kernel's own model (it also integrates): own state in
    (posr, vel, omg, quat, disp: 160; mask, list length, species, wall mask:
    20) and out (posr, vel, omg, quat, disp: 160; masks: 12), 8 B per list
    entry (partner index and reverse slot, one 8-byte load), 64 B per pair in
    contact."""
    return 352.0 + 8.0 * lbar + 64.0 * cbar


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may use,
    capped by OMP_NUM_THREADS when set."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def with_state(S, dem, sc, d):
    """The scenario sc with the dynamic state d (a download) in place of its
    initial one."""
    return S.Scenario(sc.name, sc.species, sc.walls, sc.periodic, sc.gravity, sc.dt,
                      dem.State(id=sc.state.id, species=sc.state.species, radius=sc.state.radius,
                                inv_mass=sc.state.inv_mass, inv_inertia=sc.state.inv_inertia,
                                x=d["x"], v=d["v"], w=d["w"], q=d["q"], f=d["f"], t=d["t"]))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the engine runs a simulation on a single MI355X; "
              f"--gpus {a.gpus} is not supported (multi-GPU slab decomposition is a design only)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if a.copy_probe > 0:
        src = torch.ones(a.copy_probe // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.copy_probe}))
        return

    from mercurydpm_fork_amd import dem
    from mercurydpm_fork_amd import scenarios as S

    if a.engine_lib:
        dem.HIP_LIB_PATH = os.path.abspath(a.engine_lib)
    sweep = dem.SWEEP_WAVEFRONT if a.list_sweep == "wavefront" else dem.SWEEP_PER_PARTICLE
    if a.ensemble > 0:
        return run_ensemble(a)
    if a.churn > 0:
        return run_churn(a)

    def log(msg):
        print("# " + msg, file=sys.stderr, flush=True)

    t_gen = time.time()
    sc = S.chute_hvm(n_flow=a.n, seed=a.seed)
    n_total = sc.state.n
    n_flow = sc.info["n_flowing"]
    prec = dem.PRECISION_MIXED if a.precision == "mixed" else dem.PRECISION_FP64
    eng = dem.Engine(device=0, skin=a.skin, precision=prec, list_sweep=sweep)
    settled_cn = None
    if a.load_state:
        z = np.load(a.load_state)
        S.apply(eng, with_state(S, dem, sc, z), dem.Contacts(z["p_id"], z["i_id"], z["spring"], z["ts"]))
        eng.set_time(float(z["time"]), int(z["n_steps"]))
        eng.compute_forces()
        a.run_steps = 0
        log(f"loaded the state of step {int(z['n_steps'])} from {a.load_state}")
    else:
        S.apply(eng, sc)
        eng.compute_forces()
        log(f"generated {n_total} spheres ({n_flow} flowing, {sc.info['n_base']} base, footprint "
            f"{sc.info['footprint']:.1f}) in {time.time() - t_gen:.1f}s")
        t_settle = time.time()
        sst = S.settle_chute(eng, sc, a.settle_steps, log=log)
        settled_cn = sst.n_contacts / n_total
        log(f"settled in {time.time() - t_settle:.1f}s: C/N {settled_cn:.3f}")
    rb_start = eng.stats().n_rebuilds

    # the configuration's run: the settled packing under the chute gravity
    # (|g| = 1, 24 degrees) for run_steps steps, timed chunk by chunk
    chunks = []
    done = 0
    t_run = 0.0
    rb_prev = rb_start
    while done < a.run_steps:
        k = min(a.chunk, a.run_steps - done)
        eng.set_profiling(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.step(k)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = eng.stats()
        done += k
        t_run += dt
        chunks.append(dict(steps=k, seconds=dt, contacts=st.n_contacts, rebuilds=st.n_rebuilds - rb_prev,
                           force_ms=st.t_force_ms, kick_ms=st.t_kick_ms, rebuild_ms=st.t_rebuild_ms))
        rb_prev = st.n_rebuilds
        if done % (10 * a.chunk) == 0 or done == a.run_steps:
            log(f"run {done}/{a.run_steps} steps, {t_run:.1f}s: {dt / k * 1e3:.3f} ms/step, "
                f"C/N {st.n_contacts / n_total:.3f}, {st.n_rebuilds} rebuilds")
    run_avg = None
    if a.run_steps > 0:
        run_avg = {"value": n_flow * a.run_steps / t_run, "unit": "particle-steps/s", "steps": a.run_steps,
                   "seconds": round(t_run, 3), "ms_per_step": round(1e3 * t_run / a.run_steps, 4),
                   "rebuilds": int(sum(c["rebuilds"] for c in chunks)),
                   "note": "the whole configured run from the settled packing, rebuilds included"}
    if a.save_state:
        d = eng.download()
        ct = eng.contacts()
        st = eng.stats()
        np.savez(a.save_state, x=d["x"], v=d["v"], w=d["w"], q=d["q"], f=d["f"], t=d["t"], p_id=ct.p_id,
                 i_id=ct.i_id, spring=ct.spring, ts=ct.ts, time=st.time, n_steps=st.n_steps)
        log(f"state of step {st.n_steps} written to {a.save_state}")
    # the late window: the last chunks of the run covering >= steady_steps
    steady = None
    if chunks:
        win, s_win = [], 0
        for c in reversed(chunks):
            win.append(c)
            s_win += c["steps"]
            if s_win >= a.steady_steps:
                break
        secs = sum(c["seconds"] for c in win)
        reb_ms = sum(c["rebuild_ms"] for c in win)
        steady = {"steps": s_win, "from_step": a.run_steps - s_win, "ms_per_step": round(1e3 * secs / s_win, 4),
                  "value": n_flow * s_win / secs, "rebuilds": int(sum(c["rebuilds"] for c in win)),
                  "rebuild_share": round(reb_ms / max(1e3 * secs, 1e-12), 4),
                  "force_ms_per_step": round(sum(c["force_ms"] for c in win) / s_win, 4),
                  "contacts_per_particle": round(float(np.mean([c["contacts"] for c in win])) / n_total, 4)}

    # the driver's window: warm-up, then K steps timed at the run's end
    eng.step(a.warmup)
    c0 = eng.contact_count()
    rb0 = eng.stats().n_rebuilds
    torch.cuda.synchronize()
    eng.set_profiling(True)
    profile_region(True)
    t0 = time.perf_counter()
    eng.step(a.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    profile_region(False)
    st = eng.stats()
    c1 = st.n_contacts
    elapsed = t1 - t0
    value = n_flow * a.steps / elapsed
    # contacts per particle over the window (pairs once + wall contacts)
    cbar = 0.5 * (c0 + c1) / n_total
    lbar = st.n_list_entries / n_total
    force_ms = st.t_force_ms / max(st.n_timed_steps, 1)
    mixed_run = a.precision == "mixed"
    balg = b_alg(cbar, mixed_run)
    achieved = balg * n_total / (force_ms * 1e-3) / 1e9
    bfused = fused_bytes(cbar, lbar)
    log(f"timed {a.steps} steps in {elapsed:.3f}s: force+integration {force_ms:.3f} ms, kick "
        f"{st.t_kick_ms / max(st.n_timed_steps, 1):.3f} ms, rebuild "
        f"{st.t_rebuild_ms / max(st.n_timed_steps, 1):.3f} ms per step; {st.n_rebuilds - rb0} rebuilds in the "
        f"window; C/N {cbar:.3f}; max list {st.max_list_length}")

    # the mixed-precision variant on the same state: a second engine, started
    # from the FP64 run's particles and stored contacts (not the metric)
    mixed = None
    if a.mixed_steps > 0 and a.precision == "fp64":
        d = eng.download()
        ct = eng.contacts()
        s3 = with_state(S, dem, sc, d)
        em = dem.Engine(device=0, skin=a.skin, precision=dem.PRECISION_MIXED, list_sweep=sweep)
        S.apply(em, s3, ct)
        em.set_time(st.time, int(st.n_steps))
        em.compute_forces()
        em.step(200)
        torch.cuda.synchronize()
        em.set_profiling(True)
        tm0 = time.perf_counter()
        em.step(a.mixed_steps)
        torch.cuda.synchronize()
        tm1 = time.perf_counter()
        sm = em.stats()
        mk = sm.t_force_ms / max(sm.n_timed_steps, 1)
        mixed = {"value": n_flow * a.mixed_steps / (tm1 - tm0), "unit": "particle-steps/s",
                 "ms_per_step": 1e3 * (tm1 - tm0) / a.mixed_steps, "steps": a.mixed_steps,
                 "kernel_ms": round(mk, 4),
                 "roofline_frac": round(b_alg(cbar, True) * n_total / (mk * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                 "dtype": "f32 contact geometry and forces, f64 state and springs",
                 "speedup_vs_fp64": round((n_flow * a.mixed_steps / (tm1 - tm0)) / value, 4)}
        em.close()
        log(f"mixed precision: {mixed['ms_per_step']:.3f} ms per step, force {mixed['kernel_ms']:.3f} ms")

    cpu = None
    if a.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        d = eng.download()
        ct = eng.contacts()
        s2 = with_state(S, dem, sc, d)
        threads = cpu_threads()
        o, got = oracle_lib.engine_omp(threads)
        S.apply(o, s2, ct)
        o.set_time(st.time, int(st.n_steps))
        o.step(1)  # first touch of the oracle's pools, untimed
        tc0 = time.perf_counter()
        o.step(a.cpu_steps)
        tc1 = time.perf_counter()
        cpu = {"value": n_flow * a.cpu_steps / (tc1 - tc0), "unit": "particle-steps/s", "cores": int(got),
               "kind": "port", "seconds": round(tc1 - tc0, 2),
               "sample": f"oracle/dem_oracle.c built with OpenMP (the restated reference step: pair, wall and "
                         f"integration loops on {got} threads, as the reference's OpenMP build runs them), "
                         f"{a.cpu_steps} steps of the same {n_total}-sphere chute state the GPU timed"}
        o.close()
        log(f"cpu baseline: {cpu['value']:.3g} particle-steps/s on {got} threads ({cpu['seconds']} s)")

    traffic, traffic_cbar = measured_traffic()
    out = {
        "metric": "particle-steps/sec at 2M polydisperse spheres",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32 contacts / f64 state",
        "data": "synthetic (seeded rock-salt lattice settled onto the rough base under |g| = 1, then the "
                "configured run under the chute gravity)",
        "config": {"workload": "polydisperse periodic chute, HertzianViscoelasticMindlin e=0.8 mu=0.5, 24 deg",
                   "flowing_spheres": n_flow, "base_spheres": sc.info["n_base"],
                   "diameter": "U[0.5,1.5]", "footprint": round(sc.info["footprint"], 3),
                   "dt": sc.dt, "settle_steps": a.settle_steps, "verlet_skin": a.skin, "list_sweep": a.list_sweep,
                   "contacts_per_particle_settled": None if settled_cn is None else round(settled_cn, 4),
                   "steps_before_window": a.run_steps + a.warmup,
                   "contacts_per_particle": round(cbar, 4), "contacts_per_particle_dense_bed": "2-3",
                   "list_entries_per_particle": round(lbar, 3),
                   "rebuilds_in_window": int(st.n_rebuilds - rb0),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": traffic if a.n == 2_000_000 else None,
                     "traffic_contacts_per_particle": traffic_cbar,
                     "traffic_source": "profiles/k_force_chute2M_traffic.md (rocprofv3 FETCH_SIZE, WRITE_SIZE, "
                                       "copy-calibrated)",
                     "kernel": "k_force", "kernel_ms": round(force_ms, 4),
                     "bytes_per_particle": round(balg, 2),
                     "bytes_model": "264 + 64 C (whole step, per particle of the launch)",
                     "particles_per_launch": n_total,
                     "fused_kernel_model": {"bytes_per_particle": round(bfused, 2),
                                            "achieved": round(bfused * n_total / (force_ms * 1e-3) / 1e9, 2),
                                            "frac": round(bfused * n_total / (force_ms * 1e-3) / 1e9 /
                                                          HBM_PEAK_GBS, 5)}},
        "run_average": run_avg,
        "steady_window": steady,
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
