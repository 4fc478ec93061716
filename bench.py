#!/usr/bin/env python
"""Throughput of the DEM time step on one MI355X.

Workload: the polydisperse periodic chute -- 2,000,000 flowing spheres with
d ~ U[0.5, 1.5] (three grid levels) over a rough base of fixed spheres,
tilted 24 degrees, periodic in x and y, HertzianViscoelasticMindlin with
e = 0.8 and mu = 0.5, dt = t_c / 50, FP64.  The spheres start on a
non-overlapping lattice and are first settled into a dense bed (untimed:
strong gravity straight down, then the same with background drag), then the
chute gravity is switched on and, after the warm-up steps, K steps are timed.

Prints one JSON line: particle-steps/s (flowing spheres x steps / s), the
roofline of the dominant kernel (the force kernel, which also integrates the
step; HIP-event timed on the engine's stream) and a CPU baseline (the oracle, one thread, on a bounded
sample of the same state).
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
    """traffic_bytes_per_launch from the committed counter summary, or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--n", type=int, default=2_000_000, help="flowing spheres")
    ap.add_argument("--settle-steps", type=int, default=8000, help="untimed settling under strong gravity")
    ap.add_argument("--settle-gravity", type=float, default=100.0)
    ap.add_argument("--damp-steps", type=int, default=3000, help="untimed settling with background drag")
    ap.add_argument("--damp-drag", type=float, default=15.0)
    ap.add_argument("--save-state", default=None, help="write the settled state (npz) before timing")
    ap.add_argument("--load-state", default=None, help="start from a state written by --save-state")
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps in the CPU baseline (0: skip)")
    ap.add_argument("--skin", type=float, default=0.0)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="contact arithmetic: fp64 (the metric) or mixed (FP32 contact geometry and forces)")
    ap.add_argument("--mixed-steps", type=int, default=500,
                    help="after the FP64 timing, time this many steps of the same state in mixed precision (0: skip)")
    ap.add_argument("--ensemble", type=int, default=0,
                    help="instead of the chute: this many calibration heaps (sizes log-uniform in [500, 20000]) "
                         "stepped as one packed ensemble, against the same heaps run one at a time")
    ap.add_argument("--ensemble-steps", type=int, default=20000, help="steps per simulation for sims/hour")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many bytes device-to-device (calibrates HBM counters) and exit")
    return ap.parse_args()


def run_ensemble(a):
    """Batched ensemble: a.ensemble heaps (log-uniform sizes in [500, 20000],
    rolling friction) packed side by side in one context and stepped
    together, against the same kind of heaps stepped one at a time.  The
    sequential cost is measured on 16 sizes spread over the range and
    interpolated (log-log) to every member.  One JSON line: simulations per
    hour at a.ensemble_steps steps per simulation, both ways."""
    import numpy as np
    import torch

    from mercurydpm_fork_amd import dem
    from mercurydpm_fork_amd import ensemble as E
    from mercurydpm_fork_amd import scenarios as S

    rng = np.random.default_rng(a.seed)
    sizes = np.exp(rng.uniform(np.log(500), np.log(20000), a.ensemble)).astype(int)
    sp = E.calibration_species()

    def time_steps(sc, warm, steps):
        e = dem.Engine(device=0)
        S.apply(e, sc)
        e.compute_forces()
        e.step(warm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.step(steps)
        e.stats()  # waits for the device
        dt = (time.perf_counter() - t0) / steps
        e.close()
        return dt

    members = [E.heap(int(n), 1000 + k, sp) for k, n in enumerate(sizes)]
    packed = E.pack(members)
    t_packed = time_steps(packed.scenario, 200, a.steps)
    sample = np.unique(np.geomspace(500, 20000, 16).astype(int))
    t_sample = np.array([time_steps(E.heap(int(n), 7 + k, sp), 200, a.steps) for k, n in enumerate(sample)])
    t_seq = float(np.sum(np.exp(np.interp(np.log(sizes), np.log(sample), np.log(t_sample)))))
    per_hour = lambda t: 3600.0 / (t * a.ensemble_steps)  # noqa: E731
    print(json.dumps({
        "metric": "calibration heaps per hour (batched ensemble)",
        "value": per_hour(t_packed) * a.ensemble, "unit": "simulations/hour", "n_gpus": 1,
        "steps": a.steps, "ms_per_step": t_packed * 1e3, "higher_is_better": True, "dtype": "f64",
        "data": "synthetic (jittered lattice columns above a floor)",
        "config": {"workload": "batched calibration heaps, LinearViscoelasticFriction with rolling friction",
                   "simulations": int(a.ensemble), "particles": int(sizes.sum()),
                   "size_range": [500, 20000], "steps_per_simulation": a.ensemble_steps},
        "sequential": {"value": per_hour(t_seq) * a.ensemble, "unit": "simulations/hour",
                       "ms_per_step_all_members": t_seq * 1e3,
                       "sample": {int(n): round(t * 1e3, 4) for n, t in zip(sample, t_sample)}},
        "speedup_vs_sequential": t_seq / t_packed}))


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

    if a.ensemble > 0:
        return run_ensemble(a)

    t_gen = time.time()
    sc = S.chute_hvm(n_flow=a.n, seed=a.seed)
    n_total = sc.state.n
    prec = dem.PRECISION_MIXED if a.precision == "mixed" else dem.PRECISION_FP64
    eng = dem.Engine(device=0, skin=a.skin, precision=prec)
    chute_g = sc.gravity
    S.apply(eng, sc)
    eng.compute_forces()
    print(f"# generated {n_total} spheres ({sc.info['n_flowing']} flowing, {sc.info['n_base']} base, "
          f"footprint {sc.info['footprint']:.1f}) in {time.time() - t_gen:.1f}s", file=sys.stderr, flush=True)

    if a.load_state:
        z = np.load(a.load_state)
        s2 = S.Scenario(sc.name, sc.species, sc.walls, sc.periodic, chute_g, sc.dt,
                        dem.State(id=sc.state.id, species=sc.state.species, radius=sc.state.radius,
                                  inv_mass=sc.state.inv_mass, inv_inertia=sc.state.inv_inertia,
                                  x=z["x"], v=z["v"], w=z["w"], q=z["q"], f=z["f"], t=z["t"]))
        eng = dem.Engine(device=0, skin=a.skin, precision=prec)
        S.apply(eng, s2, dem.Contacts(z["p_id"], z["i_id"], z["spring"], z["ts"]))
        eng.set_time(float(z["time"]), int(z["nt"]))
        a.settle_steps = 0
    # settle into a dense bed (untimed): strong gravity straight down, then
    # background drag while gravity is lowered step-wise to |g| = 1, so the
    # bed is at rest and not over-compressed when the chute gravity starts
    if a.settle_steps > 0:
        t0 = time.time()
        stages = [(a.settle_gravity, 0.0, a.settle_steps)]
        ramp = [a.settle_gravity * (1.0 / a.settle_gravity) ** ((s + 1) / 5.0) for s in range(5)]
        stages += [(g, a.damp_drag, a.damp_steps // 5) for g in ramp]
        done = 0
        for g, drag, n in stages:
            eng.set_gravity((0.0, 0.0, -g))
            eng.set_background_drag(drag)
            eng.step(n)
            done += n
            print(f"# settle {done} steps (g {g:.3g}, drag {drag:g}): {eng.contact_count()} contacts, "
                  f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
        eng.set_gravity(chute_g)
        eng.set_background_drag(0.0)

    if a.save_state:
        d = eng.download()
        ct = eng.contacts()
        st0 = eng.stats()
        np.savez(a.save_state, x=d["x"], v=d["v"], w=d["w"], q=d["q"], f=d["f"], t=d["t"], p_id=ct.p_id,
                 i_id=ct.i_id, spring=ct.spring, ts=ct.ts, time=st0.time, nt=st0.n_steps)
        print(f"# saved state to {a.save_state}", file=sys.stderr, flush=True)
    eng.step(a.warmup)
    c0 = eng.contact_count()
    rb0 = eng.stats().n_rebuilds
    torch.cuda.synchronize()
    eng.set_profiling(True)
    t0 = time.perf_counter()
    eng.step(a.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = eng.stats()
    c1 = st.n_contacts
    rebuilds_window = st.n_rebuilds - rb0
    elapsed = t1 - t0
    n_flow = sc.info["n_flowing"]
    value = n_flow * a.steps / elapsed
    # contacts per particle over the window (pairs counted once + wall contacts)
    cbar = 0.5 * (c0 + c1) / n_total
    lbar = st.n_list_entries / n_total  # Verlet-list entries per particle
    force_ms = st.t_force_ms / max(st.n_timed_steps, 1)
    # algorithmic HBM bytes per particle of the force kernel (which also
    # integrates): own state in (posr, vel, omg, quat, disp: 160; mask, list
    # length, species, wall mask: 20) and out (posr, vel, omg, quat, disp:
    # 160; masks: 12), 5 B per list entry (partner index, reverse slot), and
    # 64 B per pair in contact (history record in and out)
    b_force = 352.0 + 5.0 * lbar + 64.0 * cbar
    achieved = b_force * n_total / (force_ms * 1e-3) / 1e9
    step_gbs = b_force * n_total * a.steps / elapsed / 1e9
    print(f"# timed {a.steps} steps in {elapsed:.3f}s: force+integration {force_ms:.3f} ms, kick "
          f"{st.t_kick_ms / max(st.n_timed_steps, 1):.3f} ms, rebuild {st.t_rebuild_ms / max(st.n_timed_steps, 1):.3f} ms "
          f"per step; {st.n_rebuilds} rebuilds total; C/N {cbar:.3f}; max list {st.max_list_length}",
          file=sys.stderr, flush=True)

    # the mixed-precision variant on the same state: a second engine, started
    # from the FP64 run's particles and stored contacts (not the metric)
    mixed = None
    if a.mixed_steps > 0 and a.precision == "fp64":
        d = eng.download()
        ct = eng.contacts()
        s3 = S.Scenario(sc.name, sc.species, sc.walls, sc.periodic, chute_g, sc.dt,
                        dem.State(id=sc.state.id, species=sc.state.species, radius=sc.state.radius,
                                  inv_mass=sc.state.inv_mass, inv_inertia=sc.state.inv_inertia,
                                  x=d["x"], v=d["v"], w=d["w"], q=d["q"], f=d["f"], t=d["t"]))
        em = dem.Engine(device=0, skin=a.skin, precision=dem.PRECISION_MIXED)
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
        mixed = {"value": n_flow * a.mixed_steps / (tm1 - tm0), "unit": "particle-steps/s",
                 "ms_per_step": 1e3 * (tm1 - tm0) / a.mixed_steps, "steps": a.mixed_steps,
                 "kernel_ms": round(sm.t_force_ms / max(sm.n_timed_steps, 1), 4),
                 "dtype": "f32 contact geometry and forces, f64 state and springs",
                 "speedup_vs_fp64": round((n_flow * a.mixed_steps / (tm1 - tm0)) / value, 4)}
        em.close()
        print(f"# mixed precision: {mixed['ms_per_step']:.3f} ms per step, force {mixed['kernel_ms']:.3f} ms",
              file=sys.stderr, flush=True)

    cpu = None
    if a.cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        d = eng.download()
        ct = eng.contacts()
        s2 = S.Scenario(sc.name, sc.species, sc.walls, sc.periodic, chute_g, sc.dt,
                        dem.State(id=sc.state.id, species=sc.state.species, radius=sc.state.radius,
                                  inv_mass=sc.state.inv_mass, inv_inertia=sc.state.inv_inertia,
                                  x=d["x"], v=d["v"], w=d["w"], q=d["q"], f=d["f"], t=d["t"]))
        o = oracle_lib.engine()
        S.apply(o, s2, ct)
        o.set_time(0.0, int(st.n_steps))
        tc0 = time.perf_counter()
        o.step(a.cpu_steps)
        tc1 = time.perf_counter()
        cpu = {"value": n_flow * a.cpu_steps / (tc1 - tc0), "unit": "particle-steps/s", "cores": 1,
               "kind": "port",
               "sample": f"oracle/dem_oracle.c (serial restatement of the reference step), {a.cpu_steps} steps "
                         f"of the same {n_total}-sphere chute state the GPU timed, on 1 host core"}

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
        "data": "synthetic (seeded rock-salt lattice packing settled under gravity)",
        "config": {"workload": "polydisperse periodic chute, HertzianViscoelasticMindlin e=0.8 mu=0.5, 24 deg",
                   "flowing_spheres": n_flow, "base_spheres": sc.info["n_base"],
                   "diameter": "U[0.5,1.5]", "footprint": round(sc.info["footprint"], 3),
                   "dt": sc.dt, "contacts_per_particle": round(cbar, 4),
                   "list_entries_per_particle": round(lbar, 3),
                   "rebuilds_in_window": int(rebuilds_window), "step_hbm_gbs_algorithmic": round(step_gbs, 2),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic() if a.n == 2_000_000 else None,
                     "traffic_source": "profiles/k_force_chute2M_traffic.md (rocprofv3 FETCH_SIZE x2.000 + WRITE_SIZE)",
                     "kernel": "k_force", "kernel_ms": round(force_ms, 4),
                     "bytes_per_particle": round(b_force, 2)},
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
