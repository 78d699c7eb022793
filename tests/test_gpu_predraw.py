"""Refresh sweeps with the demons drawn ahead (msa_set_refresh_mode(ctx, 1)
and 2): k_refresh_draw writes the Philox + LUT demons of up to 64 coming
refreshes on a second stream, and each refresh sweep runs the refresh-free
kernels reading them (mode 2: two refresh sweeps in a row as one sweep
pair).  Bit-exact against the oracle (small lattices, refresh
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
@pytest.mark.parametrize("mode", [2, 1, 0])
def test_drawn_ahead_matches_oracle(pkg, orc, L, R, nc, every, niter, splits, tiling, mode):
    """Modes 2 and 1 (drawn ahead, with and without the pairs whose second
    sweep refreshes) and, as the check of the other path on the same cases,
    mode 0 (drawn inside the sweep)."""
    print("case L=%d R=%d nc=%d every=%d niter=%d tiling=%s mode %d"
          % (L, R, nc, every, niter, tiling, mode), flush=True)
    o = random_state(L, R, nc, 4242 + L + nc)
    times = refresh_times(niter, every)
    betas = [beta_of(t, niter) for t in times]
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, betas, jobid=9)
    dev = pkg.Lattice(L, R, nc)
    if tiling is not None:
        dev.set_tiling(*tiling)
    dev.set_refresh_mode(mode)
    assert dev.refresh_mode() == mode
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


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("L,R,every,tiling", [
    (64, 2, 1, None),          # the L = 64 configuration's cadence, autotuned kernel
    (64, 2, 3, (3, 8, 4)),     # refresh sweeps paired with the sweep after them
    (48, 4, 1, (2, 6, 12)),
    (160, 2, 5, (3, 5, 40)),   # the roofline lattice's kernel and tile
])
def test_drawn_ahead_equals_fused_refresh(pkg, L, R, every, tiling, mode):
    print("case L=%d R=%d every=%d tiling=%s mode %d" % (L, R, every, tiling, mode), flush=True)
    niter = 150 if L <= 64 else 12
    times = refresh_times(niter, every)
    sched = pkg.schedule(2.0, 0.7, niter, every)
    ctr, keys, luts, idx = pkg.schedule_tables(sched, niter, R, [77 + r for r in range(R)], 3)
    assert len(idx) == len(times)
    devs = []
    for m in (0, mode):
        d = pkg.Lattice(L, R, 1)
        if tiling is None:
            d.autotune(8)
        else:
            d.set_tiling(*tiling)
        d.set_refresh_mode(m)
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


@pytest.mark.parametrize("L,R,mode", [(64, 2, 2), (96, 1, 1), (16, 1, 1), (160, 2, 0), (160, 1, 0)])
def test_default_mode_by_geometry(pkg, L, R, mode):
    """The untuned default (a fresh context, no msa_autotune on it)."""
    d = pkg.Lattice(L, R, 1)
    assert d.refresh_mode() == mode
    d.set_refresh_mode(0)
    assert d.refresh_mode() == 0
    d.set_refresh_mode(-1)
    assert d.refresh_mode() == mode
    d.close()


def test_mode_switch_between_runs(pkg, orc):
    """Modes 2, 0, 1 and 2 again on one lattice: the same bits as the oracle
    across the switches."""
    L, R, nc, every, niter = 16, 2, 1, 1, 37
    o = random_state(L, R, nc, 31337)
    times = refresh_times(niter, every)
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, [beta_of(t, niter) for t in times])
    dev = pkg.Lattice(L, R, nc)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(len(times), dtype=np.int32))
    t = 0
    for mode, n in ((2, 11), (0, 9), (1, 10), (2, 7)):
        dev.set_refresh_mode(mode)
        dev.run(t, n, every)
        oracle_run(orc, o, t, n, every, times, ctr, keys, luts)
        t += n
        assert_same_state(dev, o, "t=%d mode %d" % (t, mode))
    dev.close()


def test_pipeline_across_calls_and_restarts(pkg, orc):
    """Calls that continue the drawn-ahead pipeline (each call's first refresh
    already drawn by the one before), a call that goes back to t = 0 on a
    re-uploaded state, and a call after new refresh tables: every result
    equals the oracle's."""
    L, R, nc, every, niter = 16, 2, 1, 1, 200
    o0 = random_state(L, R, nc, 2718)
    times = refresh_times(niter, every)
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, [beta_of(t, niter) for t in times], jobid=4)
    dev = pkg.Lattice(L, R, nc)
    dev.set_refresh_mode(2)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(len(times), dtype=np.int32))
    o = random_state(L, R, nc, 2718)
    dev.upload(o.J, o.S, o.D0, o.D1)
    t = 0
    for n in (13, 13, 40, 1, 60, 73):
        dev.run(t, n, every)
        oracle_run(orc, o, t, n, every, times, ctr, keys, luts)
        t += n
        assert_same_state(dev, o, "t=%d" % t)
    # back to the start on the initial state: the chunks drawn for the end of
    # the run are not the ones needed
    o = random_state(L, R, nc, 2718)
    dev.upload(o0.J, o0.S, o0.D0, o0.D1)
    dev.run(0, 50, every)
    oracle_run(orc, o, 0, 50, every, times, ctr, keys, luts)
    assert_same_state(dev, o, "rerun t=50")
    # new tables from here on (another job id): the drawn chunks are dropped
    ctr2, keys2, raw2, luts2 = oracle_schedule(orc, R, times, [beta_of(t, niter) for t in times], jobid=6)
    dev.set_luts(raw2)
    dev.set_refresh(ctr2, keys2, np.arange(len(times), dtype=np.int32))
    dev.run(50, 70, every)
    oracle_run(orc, o, 50, 70, every, times, ctr2, keys2, luts2)
    assert_same_state(dev, o, "new tables t=120")
    dev.close()
