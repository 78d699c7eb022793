"""Refresh schedules built with the oracle (the checker side of parity tests)."""
import numpy as np


def beta_of(t, n, t0=2.0, t1=0.7):
    return 1.0 / (t0 + (t1 - t0) * t / n)


def oracle_thermal(orc, R, nterm, seed_mc=3003, jobid=0):
    """Generators, keys and thermal counters (nterm, 2, R, 4): _fill_bits at
    iteration 2t + p, parity p (main.cu:208-229); the generators are left
    advanced for the refresh counters that follow."""
    gens = [orc.cpu_generators(seed_mc + r) for r in range(R)]
    ctr = np.zeros((nterm, 2, R, 4), np.uint32)
    for t in range(nterm):
        for p in (0, 1):
            for r in range(R):
                ctr[t, p, r] = orc.fill_counter(gens[r][0], gens[r][1], 2 * t + p, p, jobid)
    return gens, ctr


def oracle_schedule(orc, R, times, betas, seed_mc=3003, jobid=0, gens=None):
    """Counters (nref, R, 4), keys (R,), LUT bytes (2*nref tables), per-refresh LUT objects.
    gens: generator states to continue from (default: fresh from seed_mc)."""
    if gens is None:
        gens = [orc.cpu_generators(seed_mc + r) for r in range(R)]
    keys = np.array([g[2] for g in gens], np.uint32)
    ctr = np.zeros((len(times), R, 4), np.uint32)
    luts, raw = [], []
    for k, (t, b) in enumerate(zip(times, betas)):
        for r in range(R):
            ctr[k, r] = orc.fill_counter(gens[r][0], gens[r][1], t, 0, jobid)
        pair = orc.lut_demon(b)
        luts.append(pair)
        raw.append(bytes(pair[0]) + bytes(pair[1]))
    return ctr, keys, b"".join(raw), luts
