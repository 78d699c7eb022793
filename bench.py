#!/usr/bin/env python3
"""Benchmark of the micSA sweep on one MI355X.

Workload (one step = one sweep = one iteration with nc = 1): plain micSA on
the L=160 periodic cubic 3D Edwards-Anderson lattice, 64 +-J samples per
uint64 lane, R = 2 replicas sharing J, one demon layer, linear anneal
T 1.8 -> 0.6 over 2^18 sweeps with the demon refresh every 1024 sweeps.
State is constructed on the device (counter-based) and stays in HBM; the
timed region covers exactly --steps sweeps and the measurements of the
integer energies, demon energies, q and q_link every 1024 sweeps (the
configuration's cadence), bracketed by stream syncs and timed with the host
clock (the sweeps alone, timed with HIP events on the library's stream, are
reported beside it).  The configuration's work mix is one refresh sweep and
one measurement per 1024 sweeps: when the window holds fewer of them than
its length calls for (a window shorter than 1024 sweeps holds none), the
extra cost of a refresh (a refresh sweep and one sweep against one sweep
pair, median of three short back-to-back blocks) and of a measurement, timed
separately after the window, is added to the window's time for the share it
lacks.

The sweep kernel and its tile are first autotuned on this lattice (every
candidate gives bit-identical results; --no-autotune keeps the default).
Prints one JSON line.  `value` = spin updates per picosecond
(64 lanes x R x V x steps / seconds / 1e12).  `roofline` is for the
dominant kernel, the sweep kernel of sweeps without a refresh: its
algorithmic HBM bytes per launch, (48R+24) V per sweep times the sweeps one
launch performs (two for the sweep-pair kernel), over its average launch time,
measured with HIP events on the library's stream over a separate run of
sweeps that all use it, against the MI355X HBM peak of 8 TB/s.  `traffic`
(HBM bytes per launch from counters) is read from the profile summary in
profiles/ when one exists for the same kernel and workload.
`cpu_baseline` times the oracle's word-level CPU sweep (tests/oracle_lib.py)
on a bounded sample of the same lattice on this host, with one refresh sweep
and one measurement folded in at the same cadence.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

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


KERNELS = {0: "k_halfsweep", 1: "k_sweep_fused<%d>", 2: "k_sweep_fused2<%d>", 3: "k_sweep_pair",
           4: "k_sweep_resident"}
RES_MAX = 64          # iterations per resident launch (mode 4)


def kernel_name(mode, R):
    name = KERNELS[mode]
    return name % R if "%d" in name else name


def read_traffic(kernel, L, R, tiling):
    """HBM bytes per launch measured by counters for this kernel, lattice and
    tiling (mode, rows, planes), from profiles/traffic_*.txt (key=value
    lines), or None."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for fn in sorted(os.listdir(d)):
        if not (fn.startswith("traffic_") and fn.endswith(".txt")):
            continue
        kv = {}
        with open(os.path.join(d, fn)) as f:
            for ln in f:
                if "=" in ln and not ln.startswith("#"):
                    k, v = ln.split("=", 1)
                    kv[k.strip()] = v.strip()
        if (kv.get("kernel") == kernel and kv.get("L") == str(L) and kv.get("R") == str(R)
                and kv.get("tiling") == "%d,%d,%d" % tuple(tiling)):
            return int(kv["hbm_bytes_per_launch"])
    return None


def alg_bytes(V, R, nsweeps, nrefresh):
    """Compulsory HBM bytes: spins R+W 16R, active demon layer R+W 32R, J read
    once 24 per site and sweep; a sweep with a fused refresh skips the demon
    read (16R less)."""
    return (nsweeps - nrefresh) * (48 * R + 24) * V + nrefresh * (32 * R + 24) * V


def host_cpu():
    """nproc, the cores this process may run on, and the CPU model string."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def cpu_baseline(L, R, nc, budget_s, threads, every, measure_every):
    """The oracle on this host's cores: refresh-free sweeps for about budget_s
    seconds, then one refresh sweep and one measurement (energies, demon
    energies, q, q_link), folded in at the configuration's cadence (one of
    each per `every` / `measure_every` sweeps)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from sched_util import oracle_schedule
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
    t_sweep = dt / n
    ctr, keys, _, luts = oracle_schedule(orc, R, [every], [1.0 / 1.8])
    r0 = time.perf_counter()
    orc.iteration(o, ctr[0], keys, luts[0], nthreads=threads)
    t_ref = time.perf_counter() - r0
    m0 = time.perf_counter()
    orc.energy(o)
    orc.demon_energy(o)
    for p in range(R // 2):
        orc.overlap(o, 2 * p, 2 * p + 1)
    t_meas = time.perf_counter() - m0
    per_sweep = t_sweep + ((t_ref - t_sweep) / every if every > 0 else 0.0) \
        + (t_meas / measure_every if measure_every > 0 else 0.0)
    cpu = host_cpu()
    return {"value": 64.0 * R * L ** 3 / per_sweep / 1e12, "unit": "spin updates/ps",
            "cores": threads, "kind": "port",
            "sample": "oracle word-level sweep of the L=%d R=%d lattice on %d threads: %d "
                      "refresh-free sweeps in %.1f s (%.1f ms each), one refresh sweep %.1f ms and "
                      "one measurement %.1f ms folded in per %d and %d sweeps"
                      % (L, R, threads, n, dt, t_sweep * 1e3, t_ref * 1e3, t_meas * 1e3, every,
                         measure_every),
            "nproc": cpu["nproc"], "cpu_affinity": cpu["affinity"], "cpu_model": cpu["model"]}


def cpu_extras(threads, budget_s=3.0):
    """The rest of the CPU-baseline plan, on this host with the oracle:
    the batched configuration's serial loop (each lattice size of the mix
    {16,24,32,48,64,96}, R=1, refresh every sweep, timed for about budget_s
    and weighted by the mix's volume: 512 lattices, lengths from
    {4096,16384,65536}, xoshiro256++ seed 4242), and the reference's own demon
    refresh (InitWalkers, ini.c:712-781: one sequential PR+LCG and
    xoshiro256++ stream per sample and replica; the reference runs one host
    thread per sample, main.cu:371-375) as demon words per second and the
    cost of one refresh of the reference's workload (L=160, nc=6, 16
    replicas per sample)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from sched_util import oracle_schedule
    orc = oracle_lib.load()
    rng = Xoshiro256pp(4242)
    sizes = (16, 24, 32, 48, 64, 96)
    lens = (4096, 16384, 65536)
    lats = [(sizes[rng.next() % 6], lens[rng.next() % 3]) for _ in range(512)]
    per = {}
    for L in sizes:
        o = oracle_lib.Lattice(L, 1, 1)
        orc.init_philox(o, 77)
        ctr, keys, _, luts = oracle_schedule(orc, 1, [0], [1.0 / 1.5])
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget_s / len(sizes) or n < 2:
            orc.iteration(o, ctr[0], keys, luts[0], nthreads=threads)
            n += 1
        per[L] = (time.perf_counter() - t0) / n
    serial_s = sum(per[L] * n for L, n in lats)
    updates = sum(64.0 * L ** 3 * n for L, n in lats)
    pr, xo, _ = orc.cpu_generators(2002)
    thr = orc.thresholds(0.90909090909090)
    nwords = 1 << 16
    t0 = time.perf_counter()
    orc.init_walkers(pr, xo, thr, nwords)
    wps = nwords / (time.perf_counter() - t0)
    ref_words = 160 ** 3 // 64 * 6 * 16
    return {"batched_serial_loop": {
                "spin_updates_per_ps": round(updates / serial_s / 1e12, 6),
                "serial_s_for_the_mix": round(serial_s, 1), "threads": threads,
                "sweep_ms_by_L": {str(L): round(per[L] * 1e3, 4) for L in sizes},
                "note": "per-size oracle refresh sweeps timed on this host and weighted by "
                        "the 512-lattice mix: the loop itself would run serial_s_for_the_mix"},
            "reference_demon_refresh": {
                "demon_words_per_s_per_thread": round(wps),
                "s_per_refresh_of_one_sample": round(ref_words / wps, 3),
                "note": "InitWalkers restatement (oracle), one stream: 64 draws per word; one "
                        "refresh of one sample of the reference's workload (L=160, nc=6, 16 "
                        "replicas) is %d words on one host thread" % ref_words}}


def nref_window(t0, steps, every):
    """Refreshes among iterations t0 .. t0+steps-1 at a uniform cadence."""
    if every <= 0:
        return 0
    return sum(1 for t in range(t0, t0 + steps) if t % every == 0)


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON result)."""
    print("bench.py: " + msg, file=sys.stderr, flush=True)


class Ticker:
    """progress() at most every `every` seconds, for long loops."""

    def __init__(self, every=30.0):
        self.every = every
        self.last = time.perf_counter()

    def __call__(self, msg):
        now = time.perf_counter()
        if now - self.last >= self.every:
            self.last = now
            progress(msg)


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x, z ^ (z >> 31)


class Xoshiro256pp:
    """xoshiro256++ (Vigna), state seeded from splitmix64(seed)."""

    def __init__(self, seed):
        self.s = []
        x = seed
        for _ in range(4):
            x, v = splitmix64(x)
            self.s.append(v)

    def next(self):
        M = 0xFFFFFFFFFFFFFFFF
        s = self.s
        rotl = lambda v, k: ((v << k) | (v >> (64 - k))) & M
        out = (rotl((s[0] + s[3]) & M, 23) + s[0]) & M
        t = (s[1] << 17) & M
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t
        s[3] = rotl(s[3], 45)
        return out


def plain_seconds(pkg, tables, reqs, jobs):
    """Device time of the plain form: each lattice alone through msa_run, one
    after the other (its setup not counted), same tables and seeds as the
    batched and served forms; the kernel, tile and refresh mode of each
    lattice size are the autotuner's (tuned once per size)."""
    total = 0.0
    tick = Ticker()
    for n, (req, job) in enumerate(zip(reqs, jobs)):
        tick("plain form: %d of %d lattices" % (n, len(reqs)))
        lat = pkg.Lattice(req.L, req.R, 1)
        lat.autotune(16)
        lat.init_philox(req.seed)
        lat.set_luts(tables.luts)
        ctr = np.ctypeslib.as_array(job.ctr, shape=(job.nref * 4 * req.R,)).copy()
        idx = np.ctypeslib.as_array(job.lut_idx, shape=(job.nref,)).copy()
        key = np.ctypeslib.as_array(job.key, shape=(req.R,)).copy()
        lat.set_refresh(ctr, key, idx)
        lat.sync()
        t0 = time.perf_counter()
        lat.run(0, req.niter, 1)
        lat.sync()
        total += time.perf_counter() - t0
        lat.close()
    return total


def batched_vs_plain(pkg, n_lat, len_div, block=None, chunk=64):
    """The batched configuration scaled down: n_lat lattices, L uniform over
    {16,24,32,48,64,96}, lengths uniform over {4096,16384,65536}/len_div,
    T 2.0->0.7, refresh every sweep, R=1; drawn with xoshiro256++ seeded 4242.
    Plain = each lattice alone through msa_run (two launches per sweep);
    batched = all lattices in one msa_batch (graph-captured launches)."""
    import importlib
    serve = importlib.import_module("micsa_ea_mc_amd.serve")
    msa = importlib.import_module("micsa_ea_mc_amd.msa")
    rng = Xoshiro256pp(4242)
    sizes = (16, 24, 32, 48, 64, 96)
    lens = (4096 // len_div, 16384 // len_div, 65536 // len_div)
    lats = [(sizes[rng.next() % 6], lens[rng.next() % 3]) for _ in range(n_lat)]
    tables = serve.Tables()
    jobs = []
    for i, (L, n) in enumerate(lats):
        req = serve.Request(i, L, n, 2.0, 0.7, 77 + i, 0.0)
        jobs.append((req, tables.job(req, i)))
    updates = sum(64 * L ** 3 * n for L, n in lats)
    t_plain = plain_seconds(pkg, tables, [req for req, _ in jobs], [job for _, job in jobs])
    # batched, with the block size picked by the autotuner on the same lattices
    tuned = None
    if block is None:
        block, tuned = serve.tune_block(tables, [req for req, _ in jobs], jobs=[j for _, j in jobs])
    b = msa.Batch(tables.luts)
    b.set_shape(block, chunk)
    for req, job in jobs:
        b.add(job)
    b.sync()
    b.run(0)
    b.sync()
    # in segments of 8192 iterations, a progress line after each (the
    # segment boundaries cost one sync each)
    t0 = time.perf_counter()
    left, tick = max(n for _, n in lats), Ticker()
    while left > 0:
        seg = min(left, 8192)
        b.run(seg)
        b.sync()
        left -= seg
        tick("batched form: %d iterations left" % left)
    t_batch = time.perf_counter() - t0
    t, te = b.progress(len(lats))
    assert (t == te).all()
    b.close()
    return {"lattices": n_lat, "length_divisor": len_div, "block": block, "graph_chunk": chunk,
            "block_tuning_s": {str(k): round(v, 5) for k, v in tuned.items()} if tuned else None,
            "spin_updates": updates,
            "plain_spin_updates_per_ps": round(updates / t_plain / 1e12, 4),
            "batched_spin_updates_per_ps": round(updates / t_batch / 1e12, 4),
            "batched_over_plain": round(t_plain / t_batch, 3),
            "plain_s": round(t_plain, 3), "batched_s": round(t_batch, 3)}


def served(pkg, n_req, load, log2_len, max_bytes, epoch=64):
    """Served configuration scaled down: L uniform over {16,32,48,64,96,128},
    length 2^U{log2_len}, T 2.0 -> U[0.6,0.9], refresh every sweep, admission
    by device footprint.  The batched capacity is measured first (all requests
    present at t = 0); the Poisson stream (seed 5151) then arrives at `load`
    times that capacity.  The plain form serves the same stream one request
    at a time in arrival order, each alone through msa_run (its finish times
    computed from each request's measured plain device time)."""
    import importlib
    serve = importlib.import_module("micsa_ea_mc_amd.serve")
    tables = serve.Tables()
    burst = serve.make_requests(n_req, 1e9, seed=5151, log2_len=log2_len)
    _, cap = serve.serve(burst, max_bytes=max_bytes, epoch=epoch, tables=tables, progress=progress)
    work = [64 * r.R * r.L ** 3 * r.niter for r in burst]
    capacity = sum(work) / cap["makespan_s"]                      # updates / s
    rate = load * capacity / (sum(work) / len(work))              # requests / s
    reqs = serve.make_requests(n_req, rate, seed=5151, log2_len=log2_len)
    _, summ = serve.serve(reqs, max_bytes=max_bytes, epoch=epoch, tables=tables, progress=progress)
    # plain: FIFO, one request alone at a time
    tp = []
    tick = Ticker()
    for r in reqs:
        tp.append(plain_seconds(pkg, tables, [r], [tables.job(r, r.rid)]))
        tick("plain form: %d of %d requests" % (len(tp), len(reqs)))
    finish, lat = 0.0, []
    for r, t in sorted(zip(reqs, tp), key=lambda x: x[0].arrival):
        finish = max(finish, r.arrival) + t
        lat.append(finish - r.arrival)
    updates = sum(64 * r.R * r.L ** 3 * r.niter for r in reqs)
    summ.update({"load_of_capacity": load, "arrival_rate_per_s": round(rate, 2),
                 "batched_capacity_spin_updates_per_ps": round(capacity / 1e12, 4),
                 "log2_length": list(log2_len), "max_bytes": max_bytes,
                 "plain_fifo_makespan_s": round(finish, 4),
                 "plain_fifo_spin_updates_per_ps": round(updates / finish / 1e12, 4),
                 "plain_fifo_latency_p50_s": round(float(np.percentile(lat, 50)), 4),
                 "plain_fifo_latency_p99_s": round(float(np.percentile(lat, 99)), 4),
                 "plain_device_spin_updates_per_ps": round(updates / sum(tp) / 1e12, 4),
                 "served_over_plain": round(finish / summ["makespan_s"], 3)})
    for k in ("makespan_s", "spin_updates_per_ps", "latency_p50_s", "latency_p99_s"):
        summ[k] = round(summ[k], 4)
    return summ


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4096)
    ap.add_argument("--warmup", type=int, default=256)
    ap.add_argument("--L", type=int, default=160)
    ap.add_argument("--R", type=int, default=2)
    ap.add_argument("--nc", type=int, default=1)
    ap.add_argument("--refresh-every", type=int, default=1024)
    ap.add_argument("--anneal", type=str, default="1.8,0.6,262144",
                    help="T0,T1,N: linear anneal T0 -> T1 over N sweeps (default: the L=160 "
                         "configuration's; the L=64 one is 2.0,0.7,65536)")
    ap.add_argument("--measure-every", type=int, default=1024,
                    help="measure energies and overlaps every this many sweeps inside the timed region (0: never)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, len(os.sched_getaffinity(0))),
                    help="oracle threads (default: the cores this process may use, at most 16, "
                         "the host share of one GPU)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-extras", action="store_true",
                    help="also time the batched mix's serial CPU loop and the reference's host "
                         "demon refresh (InitWalkers) with the oracle")
    ap.add_argument("--batched", type=int, default=512, help="lattices in the batched test (0: skip)")
    ap.add_argument("--batched-len-div", type=int, default=16,
                    help="anneal lengths {4096,16384,65536} divided by this")
    ap.add_argument("--served", type=int, default=1000, help="requests in the served test (0: skip)")
    ap.add_argument("--served-len", type=str, default="6,11",
                    help="log2 of the served requests' lengths, LO,HI (the served configuration: 12,17)")
    ap.add_argument("--served-load", type=float, default=0.8,
                    help="arrival rate as a fraction of the measured batched capacity")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default tiling instead of autotuning it on this lattice")
    ap.add_argument("--walker-refreshes", type=int, default=0,
                    help="also time this many refreshes from the reference's walker streams")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also time a streaming copy of this many MiB (bandwidth calibration)")
    ap.add_argument("--tiling", type=str, default="",
                    help="MODE,TY,TZ: use this sweep kernel and tile (no autotune)")
    ap.add_argument("--refresh-mode", type=int, default=-1,
                    help="refresh sweeps: 0 drawn inside the sweep, 1 drawn ahead on a second "
                         "stream, 2 as 1 with two refresh sweeps per pass, -1 the library's "
                         "choice for the geometry (after the autotune)")
    ap.add_argument("--thermal-iters", type=int, default=0,
                    help="also time this many Metropolis thermalization iterations (msa_run_thermal)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the lattice is not sharded); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    pkg = load_package()
    L, R, nc, every = args.L, args.R, args.nc, args.refresh_every
    V = L ** 3
    aT0, aT1, nsched = args.anneal.split(",")
    aT0, aT1, nsched = float(aT0), float(aT1), int(nsched)
    lat = pkg.Lattice(L, R, nc, device=0)
    lat.init_philox(0x5EED)
    tuned = None
    if args.tiling:
        lat.set_tiling(*[int(x) for x in args.tiling.split(",")])
    elif not args.no_autotune:
        tuned = lat.autotune(16)
    tiling = lat.tiling()
    lat.set_refresh_mode(args.refresh_mode)
    total = args.warmup + args.steps
    sched = pkg.schedule(aT0, aT1, nsched, every)
    # tables reach four refreshes past the timed window: the refresh blocks
    # timed after it (below) start at the first refreshes after the window
    ntab = total + 4 * every if every > 0 else total
    ctr, keys, luts, idx = pkg.schedule_tables(sched, ntab, R, [3003 + r for r in range(R)], 0)
    lat.set_luts(luts)
    lat.set_refresh(ctr, keys, idx)

    lat.run(0, args.warmup, every)
    lat.sync()
    # the timed region: the sweeps, and the configuration's measurements of
    # the integer energies, demon energies and overlaps every measure_every
    # sweeps (iterations that are multiples of it).  Each measurement is
    # queued behind its sweeps and collected once the next sweeps are queued
    # (msa_measure_async / msa_measure_wait), so the device does not idle
    # while the host waits for the counts.
    me = args.measure_every

    def chunks():
        t, left = args.warmup, args.steps
        while left > 0:
            n = min(left, me - t % me) if me > 0 else left
            yield t, n
            t += n
            left -= n

    w0 = time.perf_counter()
    nmeas, queued = 0, False
    for t, n in chunks():
        lat.run(t, n, every)
        if queued:
            m = lat.measure_wait()
            queued = False
        if me > 0 and (t + n) % me == 0:
            lat.measure_async()
            queued = True
            nmeas += 1
    if queued:
        m = lat.measure_wait()
    lat.sync()
    wall = time.perf_counter() - w0
    m = lat.measure()
    # the sweeps alone: the same iterations again, HIP events around each chunk
    ms_sweeps = sum(lat.run_timed(t, n, every) for t, n in chunks())
    # the configuration's work mix has one refresh per `every` and one
    # measurement per `me` sweeps; a window shorter than that holds a share
    # of them only by chance, so their cost is timed here and the share the
    # window lacks is added to its time (value = updates / (wall + missing))
    t_after = args.warmup + args.steps
    t_ref_extra = 0.0
    if every >= 4:
        # a refresh makes a block of `every` sweeps run the refresh sweep, one
        # single sweep and the rest in pairs where `every` plain sweeps run
        # all in pairs: its extra cost is (refresh sweep + one sweep) - (one
        # pair).  Timed on short back-to-back blocks, median of three, so that
        # clock drift between two long blocks does not enter the difference.
        tb = (t_after // every + 1) * every          # first refresh after the window
        diffs = []
        for j in range(3):
            t0r = tb + j * every
            ms_ref = lat.run_timed(t0r, 2, every)        # refresh sweep + one sweep
            ms_plain = lat.run_timed(t0r + 2, 2, every)  # one pair
            diffs.append(ms_ref - ms_plain)
        t_ref_extra = max(0.0, float(np.median(diffs)) * 1e-3)
    elif every > 0:
        tb = (t_after // every + 1) * every
        ms_ref = lat.run_timed(tb, every, every)     # `every` sweeps, the first one refreshing
        ms_plain = lat.run_timed(tb + every, every, 0)
        t_ref_extra = max(0.0, (ms_ref - ms_plain) * 1e-3)
    t_meas = 0.0
    if me > 0:
        q0 = time.perf_counter()
        for _ in range(3):
            lat.measure()
        t_meas = (time.perf_counter() - q0) / 3
    miss_ref = (args.steps / every - nref_window(args.warmup, args.steps, every)) if every > 0 else 0.0
    miss_meas = (args.steps / me - nmeas) if me > 0 else 0.0
    missing = miss_ref * t_ref_extra + miss_meas * t_meas
    # the dominant kernel alone: sweeps without a refresh (an even count, so
    # that the sweep-pair kernel runs alone)
    ksweeps = max(2, min(max(args.steps, 512), 1024) & ~1)
    ms_k = lat.run_timed(args.warmup + args.steps, ksweeps, 0)
    # the state across the boundary: host arrays in and out (PCIe), once
    p0 = time.perf_counter()
    S_h, D0_h, D1_h = lat.download()
    J_h = lat.couplings()
    p1 = time.perf_counter()
    lat.upload(J_h, S_h, D0_h, D1_h)
    lat.sync()
    p2 = time.perf_counter()
    thermal = None
    if args.thermal_iters > 0:
        gens = pkg.generators([3003 + r for r in range(R)])
        tctr = pkg.thermal_counters(gens, args.thermal_iters, 0)
        lut = pkg.lut_beta(1.0 / 1.8)
        lat.run_thermal(lut, tctr[:1], gens[2])             # warm-up
        h0 = time.perf_counter()
        lat.run_thermal(lut, tctr, gens[2])
        h1 = time.perf_counter()
        thermal = {"iterations": args.thermal_iters, "beta": round(1.0 / 1.8, 6),
                   "ms_per_iteration": round((h1 - h0) * 1e3 / args.thermal_iters, 4),
                   "spin_updates_per_ps": round(64.0 * R * V * args.thermal_iters / (h1 - h0) / 1e12, 4)}
    walkers = None
    if args.walker_refreshes > 0 and L % 8 == 0:
        # the reference's own demon refresh (two samples, its walker streams):
        # iterations 2^k refresh, so each timed call is one refresh + one sweep
        streams, _ = pkg.walker_streams(2002, 2, R)
        lat.set_walker_streams([0, 1], streams,
                               np.tile(pkg.thresholds(1.0 / 1.8), (args.walker_refreshes + 2, 1)))
        lat.run(0, 1, -1)
        both = [lat.run_timed(1 << k, 1, -1) for k in range(args.walker_refreshes)]
        plain = lat.run_timed(3, 1, -1)
        lat.set_walker_streams([], [], np.zeros((1, 3), np.uint64))
        walkers = {"refreshes": args.walker_refreshes, "samples": 2,
                   "ms_per_refresh": round(float(np.mean(both)) - plain, 4),
                   "ms_sweep_alone": round(plain, 4)}
    rmode = lat.refresh_mode()
    lat.close()
    copy = None
    if args.copy_probe_mb > 0:
        ms_c = pkg.copy_probe(args.copy_probe_mb << 20, 10)
        copy = {"bytes": args.copy_probe_mb << 20,
                "GBs": round(2.0 * (args.copy_probe_mb << 20) / (ms_c * 1e-3) / 1e9, 1)}

    nref = nref_window(args.warmup, args.steps, every)
    sec = wall + missing
    updates = 64.0 * R * V * args.steps
    value = updates / sec / 1e12
    if tiling[0] == 4:
        launches_k = -(-ksweeps // RES_MAX)
    else:
        launches_k = int({0: 2.0, 1: 1.0, 2: 1.0, 3: 0.5}[tiling[0]] * nc * ksweeps)
    kbytes = alg_bytes(V, R, ksweeps, 0) / launches_k
    avg_s = ms_k * 1e-3 / launches_k
    achieved = kbytes / avg_s / 1e9
    kname = kernel_name(tiling[0], R)
    traffic = read_traffic(kname, L, R, tiling)
    traffic_src = ("recorded by rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over this "
                   "kernel, lattice and tile (profiles/traffic_*.txt), not measured in this run"
                   if traffic is not None else None)
    out = {
        "metric": "spin_updates_per_ps",
        "value": round(value, 4),
        "unit": "spin updates/ps",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "plain micSA, L=%d periodic 3D EA +-J, 64 samples per uint64 lane, "
                        "R=%d replicas, nc=%d, anneal T %g->%g over %d sweeps, demon "
                        "refresh every %d sweeps" % (L, R, nc, aT0, aT1, nsched, every),
            "L": L, "R": R, "nc": nc, "lanes": 64, "refresh_every": every,
            "sweeps_timed": args.steps, "refreshes_timed": nref,
            "measure_every": me, "measurements_timed": nmeas,
            "kernel_mode": ["two-pass", "one-pass", "one-pass, single-load",
                            "two sweeps per pass",
                            "resident: up to %d sweeps per launch, state in LDS" % RES_MAX][tiling[0]],
            "tile_rows_planes": [tiling[1], tiling[2]],
            "refresh_mode": ["drawn inside the sweep", "drawn ahead on a second stream",
                             "drawn ahead, two refresh sweeps per pass"][rmode],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": kname,
            "avg_launch_us": round(avg_s * 1e6, 3),
            "launches_timed": launches_k,
            "alg_bytes_per_launch": int(kbytes),
            "timed_region_alg_GBs": round(alg_bytes(V, R, args.steps, nref) / (ms_sweeps * 1e-3) / 1e9, 1),
        },
        "wall_s": round(wall, 4),
        "work_mix": {"note": "value = updates / (wall of the timed window + the cost of the "
                             "refreshes and measurements the window lacks at the configuration's "
                             "cadence, timed separately)",
                     "refresh_every": every, "measure_every": me,
                     "refresh_extra_ms": round(t_ref_extra * 1e3, 4),
                     "measurement_ms": round(t_meas * 1e3, 4),
                     "refreshes_added": round(miss_ref, 5), "measurements_added": round(miss_meas, 5),
                     "added_ms": round(missing * 1e3, 5)},
        "sweeps_only": {"spin_updates_per_ps": round(updates / (ms_sweeps * 1e-3) / 1e12, 4),
                        "ms_per_sweep": round(ms_sweeps / args.steps, 5),
                        "note": "HIP events around the timed region's iterations run again "
                                "without the measurements"},
        "state_transfer": {"bytes_each_way": int(S_h.nbytes + D0_h.nbytes + D1_h.nbytes + J_h.nbytes),
                           "download_s": round(p1 - p0, 4), "upload_s": round(p2 - p1, 4),
                           "note": "host arrays in canonical order, layout conversion included; "
                                   "not part of value"},
        "mean_energy_per_site": round(float(-m["energy"].mean()) / V, 6),
    }
    if tiling[0] == 4:
        out["roofline"]["note"] = ("resident launches keep the state in LDS across their sweeps: "
                                   "achieved counts the per-sweep compulsory bytes of the HBM "
                                   "forms, not HBM traffic; this lattice is bound by LDS and "
                                   "barriers, not HBM")
    if tuned is not None:
        out["autotune_ms_per_sweep"] = round(tuned[3], 5)
    if thermal is not None:
        out["thermal"] = thermal
    if walkers is not None:
        out["walker_refresh"] = walkers
    if copy is not None:
        out["copy_probe"] = copy
    if args.batched > 0:
        progress("headline lattice done; batched form: %d lattices" % args.batched)
        out["batched"] = batched_vs_plain(pkg, args.batched, args.batched_len_div)
    if args.served > 0:
        progress("served form: %d requests" % args.served)
        # default lengths 2^6..2^11: the served configuration's 2^12..2^17 sweeps / 64
        lo, hi = (int(x) for x in args.served_len.split(","))
        out["served"] = served(pkg, args.served, args.served_load, (lo, hi), 8 << 30)
    if not args.no_cpu:
        progress("CPU baseline: %.0f s" % args.cpu_budget)
        out["cpu_baseline"] = cpu_baseline(L, R, nc, args.cpu_budget, args.cpu_threads, every, me)
        if args.cpu_extras:
            progress("CPU extras: batched serial loop, reference demon refresh")
            out["cpu_baseline"]["extras"] = cpu_extras(args.cpu_threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
