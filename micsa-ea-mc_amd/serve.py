"""Served form: a request loop with continuous batching on one GPU.

Requests (annealing runs of different L and lengths) arrive over time.  A
request is admitted into the running batch (include/msa.h msa_batch_*) as
soon as it has arrived and its device footprint fits under the memory limit;
admitted lattices join the next launches while the others keep running, and
finished ones are retired, freeing their memory for the queue.  All state
(lattice, demons, counters, tables) stays on the device from admission to
retirement.  Every result is bit-identical to the same request run alone,
because each lattice's update depends only on its own state and tables.
The loop itself is native (driver/msa_serve.c); this module writes its
request file, reads its results, and holds the plain-path reference run and
the batch block-size autotuner.

No reference counterpart: the reference runs one sample set per process.
"""
import dataclasses
import os
import time

import numpy as np

from . import msa

SERVE_BIN = os.path.join(msa.PKG, "bin", "msa_serve")

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


def serve(requests, max_bytes, epoch=64, block=None, tables=None, keep_state=False, workers=4,
          progress=None, solo_min_L=128, solo_min_iter=1024, solo_share=2.0):
    """Run the stream through the native request loop (bin/msa_serve, the
    C driver of the served form); returns (per-request dicts, summary dict).
    block=None lets the driver's autotuner pick the batch kernel's block size
    (one request per lattice size of the stream).  Requests are admitted in
    arrival order once their refresh tables (built by `workers` host threads)
    are ready and their footprint fits under max_bytes.  keep_state also
    returns each request's final spins and demons.  progress, if given, gets
    the driver's progress lines.  `tables` is accepted for callers that hold
    one; the driver builds the same temperature grid itself.  Requests with
    L >= solo_min_L and at least solo_min_iter sweeps run one at a time as
    plain lattices beside the batch (0: every request in the batch), while
    the solo lane's pending work is at most solo_share times the batch's
    (0: always)."""
    import json
    import os
    import subprocess
    import tempfile

    del tables
    reqs = sorted(requests, key=lambda r: r.arrival)
    with tempfile.TemporaryDirectory(prefix="msa_serve_") as d:
        rq = os.path.join(d, "requests.txt")
        with open(rq, "w") as f:
            for r in reqs:
                f.write("%d %d %d %d %.17g %.17g %d %.9f\n"
                        % (r.rid, r.L, r.R, r.niter, r.T0, r.T1, r.seed, r.arrival))
        res = os.path.join(d, "results.txt")
        cmd = [SERVE_BIN, rq, res, "--max-bytes", str(int(max_bytes)), "--epoch", str(epoch),
               "--block", str(block or 0), "--workers", str(workers),
               "--solo-min-L", str(int(solo_min_L)), "--solo-min-iter", str(int(solo_min_iter)),
               "--solo-share", repr(float(solo_share))]
        if keep_state:
            os.makedirs(os.path.join(d, "state"))
            cmd += ["--state-dir", os.path.join(d, "state")]
        proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        err_lines = []
        for ln in proc.stderr:
            err_lines.append(ln)
            if progress:
                progress(ln.rstrip())
        out = proc.stdout.read()
        if proc.wait() != 0:
            raise msa.MsaError("msa_serve failed: %s" % "".join(err_lines[-20:]))
        summary = json.loads(out.strip().splitlines()[-1])
        done = {}
        with open(res) as f:
            for ln in f:
                v = ln.split()
                rid, L, R, niter = (int(x) for x in v[:4])
                arrival, admit, finish, latency = (float(x) for x in v[4:8])
                rec = {"rid": rid, "L": L, "niter": niter, "arrival": arrival, "admit": admit,
                       "finish": finish, "latency": latency,
                       "energy": np.array(v[8:8 + 64 * R], np.int32).reshape(R, 64)}
                if keep_state:
                    V = L ** 3
                    st = np.fromfile(os.path.join(d, "state", "req_%d.bin" % rid), np.uint64)
                    rec["state"] = (st[:R * V].reshape(R, V).copy(),
                                    st[R * V:2 * R * V].reshape(R, 1, V).copy(),
                                    st[2 * R * V:].reshape(R, 1, V).copy())
                done[rid] = rec
    return done, summary
