#!/usr/bin/env python3
"""Throughput of the dsmcFoam+ per-timestep particle engine on one MI355X.

A step is one dsmcCloud::evolve of the workload: freestream injection, move
with wall/deletion boundaries, cell sort, NTC + VHS collisions and
dsmcVolFields sampling.  Default workload: 2D Mach-10 argon flow over a
cylinder of radius 0.1 m, 400x400 cells, 25 simulators per freestream cell
(about 4M particles).  Inputs are synthetic (a uniform freestream fill) and
resident in HBM before the timed region; a fixed number of development
steps (--develop, default 4000, about 2.6 freestream flow-throughs of the
1 m domain) forms the bow shock and the wake in set-up, whatever --warmup
is.

Prints one JSON line: particle-steps/s, the HBM roofline of the dominant
kernel (algorithmic bytes / HIP-event duration), a bounded CPU baseline
(the C restatement of the reference algorithm built -O3 -march=native with
OpenMP, on all of the host's threads and on one, on the same developed
state) and the plain path's time to fields within 2% of a 10k-step long
average.  Exits non-zero, with "valid": false, if a device-side bound
(tracking, tile or capacity overflow) tripped in the timed steps.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) HBM bytes of each stage for the argon
# structure-of-arrays layout: FP64 position (24 B), FP32 velocity (12 B),
# int32 cell (4 B).  Per particle and per cell of one launch.
STAGE_BYTES = {
    # read pos+vel+cell 40, write pos+cell 28 and the histogram slot 4
    "move": dict(per_particle=72.0, per_cell=0.0),
    # read counts, write start offsets, clear counts
    "scan": dict(per_particle=0.0, per_cell=12.0),
    # read cell + slot, write the permutation entry
    "scatter": dict(per_particle=12.0, per_cell=0.0),
    # read perm + pos + vel, write sorted pos + vel + cell + sub-cell byte; per cell: start
    "gather": dict(per_particle=81.0, per_cell=4.0),
    # read sorted vel + sub-cell byte, write vel; per cell: start, sigmaTcRMax
    # and remainder (read+write), volume and the 7 sample sums of one species
    # without internal modes (nColls, collision separation, N, m U, m|U|^2), read+write
    "collide": dict(per_particle=25.0, per_cell=4 + 16 + 16 + 8 + 7 * 16),
    # per inserted particle: 40 B written (counted per particle moved: small)
    "inject": dict(per_particle=0.0, per_cell=0.0),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cylinder_case(seed=20251202):
    from hystrath_amd import abi, cases

    # dsmcVolFields sums every step and the wall's boundaryMeasurements, which
    # the reference's diffuse wall always takes (dsmcDiffuseWallPatch.C:167)
    return cases.cylinder_flow(nx=400, ny=400, ppc=25, dt=2.5e-7, seed=seed,
                               flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_WALL)


# Reacting air (workload "air"): particles carry rotational energy (4 B),
# one vibrational level (2 B) and a species byte; the collide stage is the
# reacting NTC (its sub-cell lists, 8 B read+write), the products' placement
# and the cells' sums of five species read in a second pass.
STAGE_BYTES_AIR = {
    "move": dict(per_particle=72.0, per_cell=0.0),
    "scan": dict(per_particle=0.0, per_cell=12.0),
    "scatter": dict(per_particle=12.0, per_cell=0.0),
    # read perm + pos + vel + erot + vib + type 47, write pos + vel + cell + sub + erot + vib + type 48
    "gather": dict(per_particle=95.0, per_cell=4.0),
    # read vel + erot + vib + type + sub 20, write vel + erot + vib + type 19, list 8, sums read 19;
    # per cell: start, sigmaTcRMax and remainder read+write, volume, product count and
    # offset, the 2 + 5*10 sums read+write
    "collide": dict(per_particle=66.0, per_cell=4 + 16 + 16 + 8 + 16 + 52 * 16),
    "inject": dict(per_particle=0.0, per_cell=0.0),
}


def air_case(seed=20251203):
    from hystrath_amd import abi, cases

    return cases.air_cylinder(nx=1000, ny=1000, ppc=30, seed=seed, flags=abi.HS_FLAG_SAMPLE | abi.HS_FLAG_WALL)


def build_case(workload):
    from hystrath_amd import abi, cases

    if workload == "air":
        case = air_case()
        desc = {"workload": "air5_qk_cylinder_7kms_1000x1000", "cells": case.ncells, "ppc_freestream": 30,
                "n_inf": 1e21, "T_inf": case.T_inf, "U_inf": case.U_inf, "T_wall": 1000.0, "radius": 0.01,
                "dt": case.cfg.dt, "species": "N2 O2 NO N O", "reactions": "heatBath-5species chemReactDict (12 QK)",
                "collision_model": "LarsenBorgnakkeVariableHardSphere", "partner_selection": "noTimeCounter",
                "sampling": "every step, cells and wall faces"}

        def fill(e):
            e.fill(case.n_in, case.T_inf, u=(case.U_inf, 0.0, 0.0), fill_id=0)
    elif workload == "cylinder":
        case = cylinder_case()
        desc = {"workload": "cylinder_mach10_argon_400x400", "cells": case.ncells, "ppc_freestream": 25,
                "n_inf": case.n_inf, "T_inf": case.T_inf, "U_inf": round(case.U_inf, 3), "T_wall": 500.0,
                "radius": 0.1, "dt": case.cfg.dt, "collision_model": "VariableHardSphere",
                "partner_selection": "noTimeCounter", "sampling": "every step, cells and wall faces"}

        def fill(e):
            e.fill([case.n_inf], case.T_inf, u=(case.U_inf, 0.0, 0.0), fill_id=0)
    else:
        case = cases.periodic_box(n_side=20, ppc=20, seed=20251202, flags=abi.HS_FLAG_SAMPLE)
        desc = {"workload": "periodic_box_argon_20x20x20", "cells": case.ncells, "ppc": 20, "dt": case.cfg.dt,
                "collision_model": "VariableHardSphere", "partner_selection": "noTimeCounter",
                "sampling": "every step"}

        def fill(e):
            e.fill([case.n_density / 2], 300.0, exact=True, fill_id=0)
            e.fill([case.n_density / 2], case.T_hot, exact=True, fill_id=1)
    return case, desc, fill


def run_evolve(eng, nsteps, chunk=50, what="steps"):
    """Enqueue nsteps; synchronise after each chunk of steps that is in flight
    for more than ~10 s so progress is reported while the GPU works."""
    done = 0
    t0 = time.time()
    while done < nsteps:
        k = min(chunk, nsteps - done)
        eng.evolve(k)
        done += k
        if time.time() - t0 > 10:
            eng.synchronize()
            log(f"  {what}: {done}/{nsteps}")
            t0 = time.time()


def _threaded_oracle():
    """The OpenMP build of the C oracle compiled for this host
    (-O3 -march=native) into a temporary directory; the prebuilt portable
    one (oracle/_build/libdsmc_oracle_mt.so, -march=x86-64-v2) if gcc
    cannot build it.  Returns (path, march)."""
    import subprocess
    import tempfile

    src = os.path.join(ROOT, "oracle", "dsmc_oracle.c")
    out = os.path.join(tempfile.mkdtemp(prefix="hs_cpu_"), "libdsmc_oracle_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-fopenmp", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
           "-shared", "-o", out, src, "-lm"]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        if r.returncode == 0:
            return out, "native"
        log(f"[bench] gcc -march=native failed, using the portable build: {r.stderr[-300:]}")
    except (OSError, subprocess.SubprocessError) as e:
        log(f"[bench] gcc unavailable ({e}); using the portable build")
    return os.path.join(ROOT, "oracle", "_build", "libdsmc_oracle_mt.so"), "x86-64-v2"


def _host():
    import platform
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(case, eng, budget_s):
    """Time the CPU oracle on the engine's current state (a bounded sample of
    full steps), on all threads the host gives this job (OMP_NUM_THREADS, or
    every logical CPU) and on one.  Parcels are shared among threads in the
    move, cells in the sort, the collisions and the sampling; the threaded
    build's steps equal the single-thread checker's (tests/
    test_cpu_baseline.py)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob

    path, march = _threaded_oracle()
    L = ob.lib(path)
    model, ncpu = _host()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or ncpu
    state = (eng.download(), *eng.cell_state())
    inflow = eng.inflow()
    results = {}
    for nt in (threads, 1):
        L.or_set_threads(nt)
        st = ob.OracleState(case, lib_path=path)
        st.load(*state)
        if st.inflow is not None:
            st.inflow[:] = inflow[: st.inflow.size]
        st.step = eng.step
        n0 = st.n
        t1, psteps, steps, chunk = 0.0, 0, 0, 1
        while t1 < budget_s:
            n1 = st.n
            t = time.perf_counter()
            st.evolve(chunk)
            dt = time.perf_counter() - t
            t1 += dt
            psteps += int(0.5 * (n1 + st.n) * chunk)
            steps += chunk
            chunk = max(1, min(int(0.5 * (budget_s - t1) / max(dt / chunk, 1e-6)), 50))
        results[nt] = dict(value=psteps / t1, steps=steps, seconds=round(t1, 3), particles=n0)
        log(f"[bench] CPU {nt} thread(s): {psteps / t1:.4e} particle-steps/s ({steps} steps, {t1:.1f} s)")
    a, one = results[threads], results[1]
    return {"value": a["value"], "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{a['steps']} full steps of the developed workload state ({a['particles']} particles): "
                      f"the C restatement of the reference loops (oracle/dsmc_oracle.c) built -O3 "
                      f"-march={march} with OpenMP, {threads} threads; the reference dsmcFoam+ itself needs "
                      f"OpenFOAM v1706 and MPI, which this host lacks",
            "seconds": a["seconds"], "threads": threads, "nproc": ncpu, "cpu_model": model, "march": march,
            "single_core": {"value": one["value"], "steps": one["steps"], "seconds": one["seconds"]}}


def time_to_2pct(args, case_fn, fill, ms_per_step):
    """The plain path's wall time from case start to running-average fields
    within 2% of a long average.  A long-average run (another seed) develops for
    --t2-develop steps and averages --t2-long steps; the measured run starts
    from its own fill, develops for the same number of steps without
    sampling (the reference's schedule: sampling from step 10k of 20k), then
    samples and is checked every --t2-check steps.  The time is the steps
    taken times the timed steps' ms/step; the wall time with the checks'
    read-backs is reported beside it."""
    from hystrath_amd import convergence as cv
    from hystrath_amd.engine import Engine

    dev, nj, chk = args.t2_develop, args.t2_long, args.t2_check
    case_j = case_fn(seed=20251202 + 1)
    with Engine(case_j) as ej:
        fill(ej)
        flags = case_j.cfg.flags
        ej.set_flags(flags & ~1)
        run_evolve(ej, dev, what="long-average development")
        ej.set_flags(flags)
        ej.reset_samples()
        run_evolve(ej, nj, what="long average")
        s, n = ej.samples()
        long_avg = cv.field_set(s, n, case_j)
    case = case_fn(seed=20251202)
    t0 = time.perf_counter()
    with Engine(case) as e:
        fill(e)
        flags = case.cfg.flags
        e.set_flags(flags & ~1)
        run_evolve(e, dev, what="plain development")
        e.set_flags(flags)
        e.reset_samples()
        done, err, trace = 0, None, []
        while done < args.t2_max:
            run_evolve(e, chk, what="plain sampling")
            done += chk
            s, n = e.samples()
            err = cv.errors(cv.field_set(s, n, case), long_avg)
            trace.append([done, {k: round(v, 5) for k, v in err.items()}])
            if cv.converged(err, 0.02):
                break
        wall = time.perf_counter() - t0
    ok = err is not None and cv.converged(err, 0.02)
    steps = dev + done
    log(f"[bench] time-to-2%: {'reached' if ok else 'NOT reached'} after {dev} + {done} steps; errors {err}")
    return {"reached": ok, "steps": steps, "develop_steps": dev, "sampled_steps": done,
            "seconds": round(steps * ms_per_step * 1e-3, 4), "wall_seconds_with_checks": round(wall, 3),
            "errors": err, "long_average": {"develop_steps": dev, "average_steps": nj, "seed": 20251202 + 1},
            "check_every": chk, "trace": trace[-5:],
            "definition": "relative L2 over cells with >= 1 parcel on average in the long average; velocity "
                          "components scaled by the long average's speed; argon has no internal temperature"}


def warm_start_eval(args, log_fn):
    """The learned warm start against the plain path on held-out 256 x 256
    cylinder cases (Mach 5-25, Kn 0.01-1, T_wall 300-1000 K), with the
    trained network committed in hystrath_amd/data.  Per case: the long average is
    the 10,000-step long average of one run; the plain path (fill, 3,000
    development steps, sampling) and the warm path (the network's fields
    fill the domain, 500 relaxation steps, one window denoised, the plain
    average as fallback) are independent runs timed to fields within 2% of
    the long average."""
    from hystrath_amd import warmstart as W
    from hystrath_amd.engine import Engine
    path = os.path.join(ROOT, "hystrath_amd", "data", "ws_cylinder256.npz")
    if not os.path.exists(path):
        return None
    nets = W.load_nets(path)
    cs = W.make_cases(args.ws_cases, seed=7777)
    gen = W.generate(cs, develop=3000, average=10000, log=log_fn)
    out = []
    for c, (_, target) in zip(cs, gen):
        long_avg = W.long_average_fields(c, target)
        # ms per step of this case, graph-replayed, on a developed state
        with Engine(c) as e:
            W.fill_freestream(e, c)
            e.evolve(1000)
            e.synchronize()
            t = time.perf_counter()
            e.evolve(500)
            e.synchronize()
            ms = 1e3 * (time.perf_counter() - t) / 500
        plain = W.plain_time_to_tol(c, long_avg, develop=3000)
        warm = W.warm_time_to_tol(nets, c, long_avg, relax=500)
        tp = plain["steps"] * ms * 1e-3
        tw = warm["steps"] * ms * 1e-3 + warm["inference_s"]
        out.append({"params": {k: round(v, 4) for k, v in c.ws_params.items()}, "ms_per_step": round(ms, 4),
                    "plain": {"steps": plain["steps"], "seconds": round(tp, 4), "reached": plain["reached"],
                              "errors": plain["errors"]},
                    "warm": {"steps": warm["steps"], "seconds": round(tw, 4), "reached": warm["reached"],
                             "used": warm["used"], "inference_s": round(warm["inference_s"], 4),
                             "errors": warm["errors"], "errors_denoised": warm["errors_denoised"],
                             "errors_window": warm["errors_window"], "errors_surrogate": warm["errors_surrogate"]},
                    "speedup": round(tp / max(tw, 1e-9), 3)})
        log_fn(f"[bench] warm start {c.ws_params}: plain {plain['steps']} steps, warm {warm['steps']} "
               f"({warm['used']}), x{tp / max(tw, 1e-9):.2f}")
    sp = [r["speedup"] for r in out]
    return {"cases": out, "speedup_geomean": round(float(np.exp(np.mean(np.log(sp)))), 3) if sp else None,
            "network": "surrogate and denoiser: 6-layer dilated 3x3 conv (dilations 1-2-4-8-16-1), "
                       "11-32-64-64-64-32-5 channels, bfloat16 MFMA kernels, FP32 accumulation",
            "tolerance": 0.02}


def run_batch(args):
    """64 independent cylinder cases (argon and N2 in turn, Mach 5-25, Kn
    0.01-1, 50^2 to 300^2 cells, 10k to 2M parcels each) stepped together by
    hs_dsmc_evolve_batch (each engine's launches, interleaved on the engines'
    streams), after the autotuner has measured the tilings of each population
    class; value = particle-steps of all cases per second."""
    from hystrath_amd import cases
    from hystrath_amd.engine import Engine, evolve_batch

    cs = cases.cylinder_batch(64)
    engs = []
    for c in cs:
        e = Engine(c)
        e.fill([c.n_inf], c.T_inf, u=(c.U_inf, 0.0, 0.0), fill_id=0)
        engs.append(e)
    sizes = [e.count() for e in engs]
    log(f"[bench] batch of {len(engs)} cases, {sum(sizes)} particles ({min(sizes)} to {max(sizes)} a case); "
        f"{args.develop} development steps, warm-up {args.warmup} steps")

    def steps(n, what):
        done, t0 = 0, time.time()
        while done < n:
            k = min(50, n - done)
            evolve_batch(engs, k)
            done += k
            if time.time() - t0 > 10:
                for e in engs:
                    e.synchronize()
                log(f"  {what}: {done}/{n}")
                t0 = time.time()
        for e in engs:
            e.synchronize()

    steps(args.develop, "development")
    # the autotuner: one engine per (kernel variant, population class) times the
    # candidate tilings; the other engines of its class take its choice
    tuned = {}
    for e, c in zip(engs, cs):
        key = (c.batch_params["species"], int(np.log2(max(e.count(), 1))))
        if key not in tuned:
            tuned[key] = e.autotune(steps_per_trial=4)
            log(f"[bench] autotune {key}: gather {tuned[key][0]}, collide {tuned[key][1]} "
                f"({tuned[key][2]:.4f} ms/step)")
    steps(args.warmup, "warm-up")
    for e in engs:
        e.reset_counters()
        e.synchronize()
    t0 = time.perf_counter()
    steps(args.steps, "timed")
    elapsed = time.perf_counter() - t0
    cnts = [e.counters() for e in engs]
    psteps = sum(c["particle_steps"] for c in cnts)
    bad = {k: sum(c[k] for c in cnts) for k in ("track_overflow", "tile_overflow", "capacity_overflow")}
    sizes = [e.count() for e in engs]
    value = psteps / elapsed
    ms_step = 1e3 * elapsed / max(args.steps, 1)
    log(f"[bench] {args.steps} steps of the batch in {elapsed:.3f}s: {value:.4e} particle-steps/s {bad}")
    # whole-step algorithmic bytes of every case (argon record; N2 adds its
    # internal-energy bytes, not counted here, so the figure is a lower bound)
    per_particle = sum(b["per_particle"] for b in STAGE_BYTES.values())
    step_bytes = sum(per_particle * n + sum(b["per_cell"] for b in STAGE_BYTES.values()) * c.ncells
                     for n, c in zip(sizes, cs))
    gbs = step_bytes / (ms_step * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "whole step, all cases",
                "step_algorithmic_bytes": int(step_bytes), "note": "argon record bytes for every case"}
    cpu = None
    if args.cpu_seconds > 0:
        # three cases of the batch (the smallest, a median one, an N2 one) on the host
        order = np.argsort(sizes)
        picks = [int(order[0]), int(order[len(order) // 2])]
        picks.append(next(k for k in order[::-1] if cs[k].batch_params["species"] == "N2" and k not in picks))
        rs = [cpu_baseline(cs[k], engs[k], args.cpu_seconds / len(picks)) for k in picks]
        tot_ps = sum(r["value"] * r["seconds"] for r in rs)
        tot_s = sum(r["seconds"] for r in rs)
        cpu = dict(rs[0])
        cpu.update(value=tot_ps / tot_s, seconds=round(tot_s, 3),
                   sample=f"cases {picks} of the batch ({[sizes[k] for k in picks]} particles), "
                          f"each for {args.cpu_seconds / len(picks):.1f} s per thread count: " + rs[0]["sample"])
        cpu["single_core"] = {"value": sum(r["single_core"]["value"] * r["single_core"]["seconds"] for r in rs)
                              / sum(r["single_core"]["seconds"] for r in rs)}
    for e in engs:
        e.close()
    valid = not any(bad.values())
    out = {"metric": "particle-steps/s on 1 GPU", "value": round(value, 1), "unit": "particle-steps/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms_step, 5), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64/f32",
           "data": f"synthetic: uniform freestream fills, {args.develop} development steps in set-up",
           "config": {"workload": "cylinder_batch_64", "cases": len(engs), "particles": sum(sizes),
                      "particles_min": min(sizes), "particles_max": max(sizes),
                      "cases_params": [c.batch_params for c in cs],
                      "autotuned": {f"{k[0]}/2^{k[1]}": v[:2] for k, v in tuned.items()},
                      "parallelism": "single GPU, cases interleaved on their own streams"},
           "roofline": roofline, "cpu_baseline": cpu, "overflow": bad, "valid": valid}
    print(json.dumps(out), flush=True)
    return 0 if valid else 3


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100, help="untimed steps after the development steps")
    ap.add_argument("--develop", type=int, default=4000,
                    help="development steps in set-up, before the warm-up (fixed, whatever --warmup is)")
    ap.add_argument("--workload", choices=["cylinder", "box", "batch64", "air"], default="cylinder",
                    help="batch64: 64 uneven cylinder cases stepped together (hs_dsmc_evolve_batch)")
    ap.add_argument("--profile-steps", type=int, default=200, help="instrumented steps for kernel durations")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU-baseline budget per thread count (0 disables)")
    ap.add_argument("--t2pct", type=int, default=1, help="measure the plain path's time to 2%% (0 disables)")
    ap.add_argument("--t2-develop", type=int, default=10000, help="steps before sampling starts (long average and run)")
    ap.add_argument("--t2-long", type=int, default=10000, help="steps of the long average")
    ap.add_argument("--t2-check", type=int, default=250, help="sampled steps between error checks")
    ap.add_argument("--t2-max", type=int, default=10000, help="give up after this many sampled steps")
    ap.add_argument("--ws-cases", type=int, default=2, help="held-out cases of the warm-start comparison (0 disables)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE {world})", file=sys.stderr)
        return 2

    if args.workload == "batch64":
        return run_batch(args)

    ws_weights = os.path.join(ROOT, "hystrath_amd", "data", "ws_cylinder256.npz")
    if args.ws_cases and os.path.exists(ws_weights):
        # the warm start's network keeps its memory in torch: torch's HIP
        # runtime starts before the engines'
        import torch
        torch.zeros(1, device="cuda")

    from hystrath_amd import abi
    from hystrath_amd.engine import Engine, device_copy_ms

    case, desc, fill = build_case(args.workload)
    log(f"[bench] {desc['workload']}: {case.ncells} cells, capacity {case.cfg.capacity}")
    eng = Engine(case)
    fill(eng)
    n_fill = eng.count()
    develop = {"cylinder": args.develop, "air": min(args.develop, 2000)}.get(args.workload, 0)
    log(f"[bench] filled {n_fill} particles; {develop} development steps, then warm-up {args.warmup} steps")
    run_evolve(eng, develop + args.warmup, what="development + warm-up")
    eng.synchronize()

    # ---- timed region ---------------------------------------------------------
    eng.reset_counters()
    eng.synchronize()
    t0 = time.perf_counter()
    run_evolve(eng, args.steps, what="timed")
    eng.synchronize()
    elapsed = time.perf_counter() - t0
    cnt = eng.counters()
    psteps = cnt["particle_steps"]
    value = psteps / elapsed
    ms_per_step = 1e3 * elapsed / max(args.steps, 1)
    n_now = eng.count()
    bad = {k: cnt[k] for k in ("track_overflow", "tile_overflow", "capacity_overflow") if cnt[k]}
    log(f"[bench] {args.steps} steps in {elapsed:.3f}s: {value:.4e} particle-steps/s, N={n_now} {bad or ''}")

    # ---- instrumented pass: per-kernel durations on the engine's stream ------------
    eng.set_flags(case.cfg.flags | abi.HS_FLAG_TIMING)
    eng.kernel_times(reset=True)
    eng.reset_counters()
    run_evolve(eng, args.profile_steps, what="profile")
    kt = eng.kernel_times()
    pc = eng.counters()
    eng.set_flags(case.cfg.flags & ~abi.HS_FLAG_TIMING)
    n_step = pc["particle_steps"] / max(args.profile_steps, 1)
    stages = {}
    for name, (ms, launches) in kt.items():
        if launches == 0:
            continue
        avg_ms = ms / launches
        b = (STAGE_BYTES_AIR if args.workload == "air" else STAGE_BYTES)[name]
        nbytes = b["per_particle"] * n_step + b["per_cell"] * case.ncells
        stages[name] = {"avg_ms": round(avg_ms, 5), "launches": launches,
                        "algorithmic_bytes": int(nbytes),
                        "gbs": round(nbytes / (avg_ms * 1e-3) / 1e9, 1) if avg_ms > 0 else None}
    dominant = max(stages, key=lambda k: stages[k]["avg_ms"])
    dom = stages[dominant]
    step_bytes = sum(s["algorithmic_bytes"] for s in stages.values())
    log(f"[bench] per-stage: {json.dumps(stages)}")

    # achievable bandwidth of a plain copy
    copy_bytes = 1 << 30
    copy_ms = device_copy_ms(copy_bytes, 10)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dominant}_{args.workload}.txt")
    if os.path.exists(tfile):
        for line in open(tfile):
            if line.startswith("bytes_per_launch_per_particle"):
                traffic = int(float(line.split("=")[1]) * n_step)

    roofline = {"bound": "hbm", "achieved": dom["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dom["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dominant,
                "algorithmic_bytes_per_launch": dom["algorithmic_bytes"], "avg_launch_ms": dom["avg_ms"],
                "step_algorithmic_bytes": step_bytes,
                "step_gbs": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                "achievable_copy_gbs": round(copy_gbs, 1), "stages": stages}

    cpu = None
    if args.cpu_seconds > 0:
        log(f"[bench] CPU baseline ({args.cpu_seconds}s budget per thread count)")
        cpu = cpu_baseline(case, eng, args.cpu_seconds)
    eng.close()

    t2 = None
    if args.t2pct and args.workload == "cylinder":
        t2 = time_to_2pct(args, cylinder_case, fill, ms_per_step)
    ws = None
    if args.ws_cases and args.workload == "cylinder" and os.path.exists(ws_weights):
        ws = warm_start_eval(args, log)

    desc.update({"particles": n_now, "develop_steps": develop, "warmup_steps": args.warmup,
                 "precision": "FP64 arithmetic, positions and sums; FP32 velocity storage",
                 "parallelism": "single GPU"})
    valid = not bad
    out = {"metric": "particle-steps/s on 1 GPU", "value": round(value, 1), "unit": "particle-steps/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64/f32",
           "data": (f"synthetic: uniform freestream fill, then {develop} development steps in set-up "
                    f"(bow shock and wake formed) and {args.warmup} warm-up steps, all untimed"
                    if develop else "synthetic: bimodal Maxwellian fill"),
           "config": desc, "roofline": roofline, "cpu_baseline": cpu, "time_to_2pct": t2, "warm_start": ws,
           "counters": cnt, "valid": valid}
    if not valid:
        out["invalid_reason"] = f"device-side bounds tripped in the timed steps: {bad}"
    print(json.dumps(out), flush=True)
    return 0 if valid else 3


if __name__ == "__main__":
    sys.exit(main())
