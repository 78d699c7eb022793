"""Served form: a request loop with continuous batching on one GPU.

Requests (annealing runs of different L and lengths) arrive over time.  A
request is admitted into the running batch (include/msa.h msa_batch_*) as
soon as it has arrived and its device footprint fits under the memory limit;
admitted lattices join the next launches while the others keep running, and
finished ones are retired, freeing their memory for the queue.  All state
(lattice, demons, counters, tables) stays on the device from admission to
retirement.  Every result is bit-identical to the same request run alone,
because each lattice's update depends only on its own state and tables.

No reference counterpart: the reference runs one sample set per process.
"""
import concurrent.futures
import dataclasses
import time

import numpy as np

from . import msa

TGRID = 1.0 / 1024          # temperature grid of the shared LUT table
T_LO, T_HI = 0.5, 2.5


@dataclasses.dataclass
class Request:
    rid: int
    L: int
    niter: int
    T0: float
    T1: float
    seed: int
    arrival: float          # seconds after the start of the stream
    R: int = 1


def make_requests(n, rate, seed=5151, sizes=(16, 32, 48, 64, 96, 128), log2_len=(12, 17),
                  T0=2.0, T1=(0.6, 0.9)):
    """Poisson stream: exponential gaps of mean 1/rate, L uniform over `sizes`,
    length 2^U{log2_len}, final temperature uniform in T1, seed = request id."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n)
    arr = np.cumsum(gaps) - gaps[0]
    out = []
    for i in range(n):
        L = int(rng.choice(sizes))
        niter = 1 << int(rng.integers(log2_len[0], log2_len[1] + 1))
        t1 = float(rng.uniform(*T1))
        out.append(Request(i, L, niter, T0, t1, 1000003 * (i + 1), float(arr[i])))
    return out


class Tables:
    """The shared temperature-grid LUT table and per-request refresh tables."""

    def __init__(self):
        self.nlo = int(round(T_LO / TGRID))
        self.nlut = int(round(T_HI / TGRID)) - self.nlo + 1
        self.luts = msa.lut_grid(TGRID, self.nlo, self.nlut)

    def refresh_tables(self, req, jobid):
        """Counters (nref, R, 4), keys (R,) and grid indices (nref,) of a request:
        seed_MC[r] = seed + r, jobid = request id mod 2^14, refresh every sweep."""
        sched = msa.schedule(req.T0, req.T1, req.niter, 1)
        return msa.schedule_tables_grid(sched, req.niter, req.R,
                                        [req.seed + r for r in range(req.R)],
                                        jobid % 16384, TGRID, self.nlo, self.nlut)

    def job(self, req, jobid):
        ctr, keys, idx = self.refresh_tables(req, jobid)
        return msa.Batch.make_job(req.L, req.R, req.seed, req.niter, 1, ctr, keys, idx)

    def subset(self, idx):
        """The LUT pairs that grid indices `idx` use, as a table of their own,
        and the indices into it."""
        used, remap = np.unique(np.asarray(idx), return_inverse=True)
        raw = np.frombuffer(bytes(self.luts), np.uint8).reshape(self.nlut, 2 * msa.LUT_BYTES)
        return np.ascontiguousarray(raw[used]).reshape(-1), remap.astype(np.int32)


BLOCKS = (256, 512, 1024)


def tune_block(tables, reqs, iters=8, candidates=BLOCKS, jobs=None):
    """Autotuner of the batched launch shape: threads per block of the batch
    kernel, timed on a scratch batch holding `reqs` (the workload or a sample
    of it; `jobs` may give their prepared msa_job) for `iters` iterations after
    two of warm-up.  Returns (best block, {block: seconds})."""
    times = {}
    jobs = jobs or [tables.job(r, i) for i, r in enumerate(reqs)]
    for block in candidates:
        b = msa.Batch(tables.luts)
        b.set_shape(block, iters)
        for j in jobs:
            b.add(j)
        b.run(2)
        b.sync()
        t0 = time.perf_counter()
        b.run(iters)
        b.sync()
        times[block] = time.perf_counter() - t0
        b.close()
    return min(times, key=times.get), times


def size_sample(sizes=(16, 32, 48, 64, 96, 128), niter=64, T0=2.0, T1=0.7):
    """One request per lattice size: the sample the served loop tunes on."""
    return [Request(i, L, niter, T0, T1, 7 + i, 0.0) for i, L in enumerate(sizes)]


def run_alone(req, tables, jobid):
    """The request by itself on the plain path: one lattice through msa_run
    (the single-lattice kernels, not the batch kernel), state built by
    msa_init_philox(seed), the same counters, keys and grid LUT pairs.  Returns
    ((spins, d0, d1), per-lane observables): the result the served form must
    reproduce bit for bit."""
    ctr, keys, idx = tables.refresh_tables(req, jobid)
    luts, sub_idx = tables.subset(idx)
    lat = msa.Lattice(req.L, req.R, 1)
    try:
        lat.init_philox(req.seed)
        lat.set_luts(luts)
        lat.set_refresh(ctr, keys, sub_idx)
        lat.run(0, req.niter, 1)
        return lat.download(), lat.measure()
    finally:
        lat.close()


def serve(requests, max_bytes, epoch=64, block=None, tables=None, keep_state=False, clock=None,
          workers=4, progress=None):
    """Run the stream; returns (per-request dicts, summary dict).  block=None
    picks the batch kernel's block size with tune_block on one request per
    lattice size of the stream.  The refresh tables of arrived requests are
    built by `workers` host threads (the host library releases the
    interpreter lock) while the device runs; a request is admitted, in
    arrival order, once its tables are ready and its footprint fits.
    progress, if given, is called with a message after every 1000th
    finished request and at least every 30 s while requests finish."""
    tables = tables or Tables()
    if block is None:
        block, _ = tune_block(tables, size_sample(sorted({r.L for r in requests})))
    batch = msa.Batch(tables.luts, max_bytes=max_bytes)
    batch.set_shape(block, epoch)
    pool = concurrent.futures.ThreadPoolExecutor(max_workers=workers)
    reqs = sorted(requests, key=lambda r: r.arrival)
    clock = clock or time.perf_counter
    t_start = clock()
    nxt = 0
    queue = []
    active = {}                 # slot -> (request, admit time)
    prepared = {}               # rid -> future of its job tables
    done = {}
    updates = 0
    max_active = 0
    max_bytes_used = 0
    finishing = []              # slots that reach their length in the run enqueued last
    last_note = [clock()]       # when progress was last called

    def retire(slot):
        nonlocal updates
        req, t_admit = active.pop(slot)
        fin = clock() - t_start
        rec = {"rid": req.rid, "L": req.L, "niter": req.niter, "arrival": req.arrival,
               "admit": t_admit, "finish": fin, "latency": fin - req.arrival,
               "energy": batch.measure(slot)["energy"]}
        if keep_state:
            rec["state"] = batch.download(slot)
        done[req.rid] = rec
        updates += 64 * req.R * req.L ** 3 * req.niter
        batch.remove(slot)
        if progress and (len(done) % 1000 == 0 or clock() - last_note[0] >= 30.0):
            last_note[0] = clock()
            progress("%d of %d requests finished" % (len(done), len(reqs)))

    try:
        while len(done) < len(reqs):
            now = clock() - t_start
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                req = reqs[nxt]
                prepared[req.rid] = pool.submit(tables.job, req, req.rid)
                queue.append(req)
                nxt += 1
            while queue:                              # FIFO admission by footprint
                req = queue[0]
                fut = prepared[req.rid]
                if not fut.done() and active:
                    break                             # admitted at a later epoch
                slot = batch.add(fut.result())
                if slot is None:
                    if not active:
                        raise RuntimeError("request %d alone exceeds the memory limit" % req.rid)
                    break
                queue.pop(0)
                del prepared[req.rid]
                active[slot] = (req, clock() - t_start)
            max_active = max(max_active, len(active))
            max_bytes_used = max(max_bytes_used, batch.bytes_in_use())
            running = [s for s in active if s not in finishing]
            if not running and not finishing:
                if nxt < len(reqs):
                    time.sleep(max(0.0, min(0.01, reqs[nxt].arrival - (clock() - t_start))))
                continue
            # one epoch stays queued on the device while the previous one's
            # finished requests are measured and retired
            new_finishing = []
            if running:
                batch.run(epoch)
                t, tend = batch.planned(max(running) + 1)
                new_finishing = [s for s in running if t[s] >= tend[s]]
                batch.wait_previous()
            else:
                batch.sync()
            for slot in finishing:
                retire(slot)
            finishing = new_finishing
        makespan = clock() - t_start
    finally:
        pool.shutdown()
        batch.close()
    lat = np.array([done[r.rid]["latency"] for r in reqs])
    summary = {"requests": len(reqs), "makespan_s": makespan,
               "spin_updates_per_ps": updates / makespan / 1e12,
               "latency_p50_s": float(np.percentile(lat, 50)),
               "latency_p99_s": float(np.percentile(lat, 99)),
               "max_concurrent": max_active, "max_bytes_in_use": max_bytes_used,
               "epoch": epoch, "block": block}
    return done, summary
