#!/usr/bin/env python3
"""Benchmark of the micSA sweep on one MI355X.

Workload (one step = one sweep = one iteration with nc = 1): plain micSA on
the L=160 periodic cubic 3D Edwards-Anderson lattice, 64 +-J samples per
uint64 lane, R = 2 replicas sharing J, one demon layer, linear anneal
T 1.8 -> 0.6 over 2^18 sweeps with the demon refresh every 1024 sweeps.
State is constructed on the device (counter-based) and stays in HBM; the
timed region covers exactly --steps sweeps, bracketed by stream syncs and
timed with HIP events on the library's stream.

Prints one JSON line.  `value` = spin updates per picosecond
(64 lanes x R x V x steps / seconds / 1e12).  `roofline` compares the
algorithmic HBM bytes of the timed sweeps, (48R+24) V per sweep and
(32R+24) V per sweep with a fused refresh (its demon read is elided), with
the MI355X HBM peak of 8 TB/s.  `cpu_baseline` times the oracle's
word-level CPU sweep (tests/oracle_lib.py) on a bounded sample of the same
lattice on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def load_package():
    import importlib.util
    if "micsa_ea_mc_amd" in sys.modules:
        return sys.modules["micsa_ea_mc_amd"]
    d = os.path.join(ROOT, "micsa-ea-mc_amd")
    spec = importlib.util.spec_from_file_location(
        "micsa_ea_mc_amd", os.path.join(d, "__init__.py"), submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["micsa_ea_mc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def alg_bytes(V, R, nsweeps, nrefresh):
    """Compulsory HBM bytes: spins R+W 16R, active demon layer R+W 32R, J read
    once 24 per site and sweep; a sweep with a fused refresh skips the demon
    read (16R less)."""
    return (nsweeps - nrefresh) * (48 * R + 24) * V + nrefresh * (32 * R + 24) * V


def cpu_baseline(L, R, nc, budget_s, threads):
    """Oracle sweeps on this host's cores for about budget_s seconds."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    orc = oracle_lib.load()
    o = oracle_lib.Lattice(L, R, nc)
    orc.init_philox(o, 0x5EED)
    rng = np.random.default_rng(7)
    o.D0[...] = rng.integers(0, 2 ** 64, size=o.D0.shape, dtype=np.uint64)
    o.D1[...] = 0
    n = 0
    t0 = time.perf_counter()
    while True:
        orc.iteration(o, nthreads=threads)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s:
            break
    return {"value": 64.0 * R * L ** 3 * n / dt / 1e12, "unit": "spin updates/ps",
            "cores": threads, "kind": "port",
            "sample": "%d sweeps of the L=%d R=%d lattice (no refresh), oracle word-level sweep, "
                      "%.1f s" % (n, L, R, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4096)
    ap.add_argument("--warmup", type=int, default=256)
    ap.add_argument("--L", type=int, default=160)
    ap.add_argument("--R", type=int, default=2)
    ap.add_argument("--nc", type=int, default=1)
    ap.add_argument("--refresh-every", type=int, default=1024)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the lattice is not sharded); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    pkg = load_package()
    L, R, nc, every = args.L, args.R, args.nc, args.refresh_every
    V = L ** 3
    nsched = 1 << 18
    lat = pkg.Lattice(L, R, nc, device=0)
    lat.init_philox(0x5EED)
    total = args.warmup + args.steps
    sched = pkg.schedule(1.8, 0.6, nsched, every)
    ctr, keys, luts, idx = pkg.schedule_tables(sched, total, R, [3003 + r for r in range(R)], 0)
    lat.set_luts(luts)
    lat.set_refresh(ctr, keys, idx)

    lat.run(0, args.warmup, every)
    lat.sync()
    w0 = time.perf_counter()
    ms = lat.run_timed(args.warmup, args.steps, every)
    lat.sync()
    wall = time.perf_counter() - w0
    m = lat.measure()
    lat.close()

    t0 = args.warmup
    nref = sum(1 for t in range(t0, t0 + args.steps) if every > 0 and t % every == 0)
    sec = ms * 1e-3
    updates = 64.0 * R * V * args.steps
    value = updates / sec / 1e12
    nbytes = alg_bytes(V, R, args.steps, nref)
    achieved = nbytes / sec / 1e9
    launches = 2 * nc * args.steps
    out = {
        "metric": "spin_updates_per_ps",
        "value": round(value, 4),
        "unit": "spin updates/ps",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "plain micSA, L=%d periodic 3D EA +-J, 64 samples per uint64 lane, "
                        "R=%d replicas, nc=%d, anneal T 1.8->0.6 over 2^18 sweeps, demon "
                        "refresh every %d sweeps" % (L, R, nc, every),
            "L": L, "R": R, "nc": nc, "lanes": 64, "refresh_every": every,
            "sweeps_timed": args.steps, "refreshes_timed": nref,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "k_halfsweep",
            "avg_launch_us": round(ms * 1e3 / launches, 3),
            "alg_bytes_per_launch": int(nbytes / launches),
        },
        "wall_s": round(wall, 4),
        "mean_energy_per_site": round(float(-m["energy"].mean()) / V, 6),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(L, R, nc, args.cpu_budget, args.cpu_threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
