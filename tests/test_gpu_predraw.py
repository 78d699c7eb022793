"""Refresh sweeps with the demons drawn ahead (msa_set_refresh_mode(ctx, 1)):
k_refresh_draw writes the Philox + LUT demons of up to 64 coming refreshes on
a second stream, and each refresh sweep runs the refresh-free kernels
reading them.  Bit-exact against the oracle (small lattices, refresh
cadences 1, 3, 4 and powers of two, several layers, runs split at arbitrary
iterations, more refreshes than the two halves of the drawn buffer hold) and
against the fused refresh of mode 0 on larger lattices with every sweep
kernel."""
import numpy as np
import pytest

from sched_util import beta_of, oracle_schedule
from test_gpu_parity import assert_same_measures, assert_same_state, random_state

pytestmark = pytest.mark.gpu


def refresh_times(niter, every):
    if every > 0:
        return list(range(0, niter, every))
    return [0] + [1 << j for j in range(40) if (1 << j) < niter]


def oracle_run(orc, o, t0, n, every, times, ctr, keys, luts, nthreads=8):
    for t in range(t0, t0 + n):
        if t in times:
            k = times.index(t)
            orc.iteration(o, ctr[k], keys, luts[k], nthreads=nthreads)
        else:
            orc.iteration(o, nthreads=nthreads)


@pytest.mark.parametrize("L,R,nc,every,niter,splits,tiling", [
    (16, 1, 1, 1, 300, (7, 150, 143), None),      # 300 refreshes: five draws, both halves reused
    (16, 2, 2, 3, 40, (10, 30), None),
    (24, 2, 1, 4, 33, (33,), (3, 4, 4)),          # refresh sweeps paired with the next sweep
    (24, 2, 1, 3, 33, (14, 19), (3, 4, 4)),       # ... and with the sweep before (second refreshed)
    (16, 2, 1, 1, 70, (31, 39), (3, 2, 4)),       # both sweeps of every pair refreshed
    (8, 1, 6, 2, 12, (5, 7), (2, 2, 2)),
    (16, 2, 1, -1, 40, (3, 37), (0, 0, 0)),       # t = 0, 1, 2, 4, ..., two-pass sweeps
])
def test_drawn_ahead_matches_oracle(pkg, orc, L, R, nc, every, niter, splits, tiling):
    print("case L=%d R=%d nc=%d every=%d niter=%d tiling=%s" % (L, R, nc, every, niter, tiling),
          flush=True)
    o = random_state(L, R, nc, 4242 + L + nc)
    times = refresh_times(niter, every)
    betas = [beta_of(t, niter) for t in times]
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, betas, jobid=9)
    dev = pkg.Lattice(L, R, nc)
    if tiling is not None:
        dev.set_tiling(*tiling)
    dev.set_refresh_mode(1)
    assert dev.refresh_mode() == 1
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(len(times), dtype=np.int32))
    t = 0
    for n in splits:
        dev.run(t, n, every)
        oracle_run(orc, o, t, n, every, times, ctr, keys, luts)
        t += n
        assert_same_state(dev, o, "t=%d" % t)
        assert_same_measures(dev, orc, o)
    dev.close()


@pytest.mark.parametrize("L,R,every,tiling", [
    (64, 2, 1, None),          # the L = 64 configuration's cadence, autotuned kernel
    (64, 2, 3, (3, 8, 4)),     # refresh sweeps paired with the sweep after them
    (48, 4, 1, (2, 6, 12)),
    (160, 2, 5, (3, 5, 40)),   # the roofline lattice's kernel and tile
])
def test_drawn_ahead_equals_fused_refresh(pkg, L, R, every, tiling):
    print("case L=%d R=%d every=%d tiling=%s" % (L, R, every, tiling), flush=True)
    niter = 150 if L <= 64 else 12
    times = refresh_times(niter, every)
    sched = pkg.schedule(2.0, 0.7, niter, every)
    ctr, keys, luts, idx = pkg.schedule_tables(sched, niter, R, [77 + r for r in range(R)], 3)
    assert len(idx) == len(times)
    devs = []
    for mode in (0, 1):
        d = pkg.Lattice(L, R, 1)
        if tiling is None:
            d.autotune(8)
        else:
            d.set_tiling(*tiling)
        d.set_refresh_mode(mode)
        d.init_philox(0xABCD)
        d.set_luts(luts)
        d.set_refresh(ctr, keys, idx)
        devs.append(d)
    t = 0
    for n in (niter // 3, niter - niter // 3):
        for d in devs:
            d.run(t, n, every)
        t += n
        a, b = devs[0].download(), devs[1].download()
        for x, y, what in zip(a, b, ("spins", "D0", "D1")):
            assert np.array_equal(x, y), "%s differ at t=%d" % (what, t)
        ma, mb = devs[0].measure(), devs[1].measure()
        for key in ma:
            assert np.array_equal(ma[key], mb[key]), key
    for d in devs:
        d.close()


def test_mode_switch_between_runs(pkg, orc):
    """Mode 1, then mode 0, then mode 1 again on one lattice: the same bits as
    the oracle across the switches."""
    L, R, nc, every, niter = 16, 2, 1, 1, 30
    o = random_state(L, R, nc, 31337)
    times = refresh_times(niter, every)
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, [beta_of(t, niter) for t in times])
    dev = pkg.Lattice(L, R, nc)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(len(times), dtype=np.int32))
    t = 0
    for mode, n in ((1, 11), (0, 9), (1, 10)):
        dev.set_refresh_mode(mode)
        dev.run(t, n, every)
        oracle_run(orc, o, t, n, every, times, ctr, keys, luts)
        t += n
        assert_same_state(dev, o, "t=%d mode %d" % (t, mode))
    dev.close()
