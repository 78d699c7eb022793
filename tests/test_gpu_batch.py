"""The batched form: lattices of uneven L in one launch sequence (captured
graph), each bit-identical to the plain per-lattice run and to the oracle."""
import numpy as np
import pytest

from oracle_lib import Lattice as OLattice
from sched_util import oracle_schedule

pytestmark = pytest.mark.gpu


def lattice_case(orc, L, R, niter, every, seed, betas_from=2.0):
    rng = np.random.default_rng(seed)
    o = OLattice(L, R, 1)
    for a in (o.J, o.S, o.D0, o.D1):
        a[...] = rng.integers(0, 2 ** 64, size=a.shape, dtype=np.uint64)
    times = [t for t in range(niter) if every > 0 and t % every == 0]
    betas = [1.0 / (betas_from - 1.2 * t / niter) for t in times]
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, betas, seed_mc=500 + seed, jobid=seed % 16384)
    return o, ctr, keys, raw, luts, times


def test_batch_matches_plain_and_oracle(pkg, orc):
    cases = [(8, 1, 13, 1), (16, 2, 20, 1), (12, 1, 7, 2), (24, 1, 16, 3), (16, 1, 0, 1), (10, 2, 9, 4)]
    built = [lattice_case(orc, L, R, n, e, 11 + i) for i, (L, R, n, e) in enumerate(cases)]
    # one shared LUT table: concatenate each lattice's pairs, offset its indices
    raws, offs, off = [], [], 0
    for b in built:
        raws.append(b[3])
        offs.append(off)
        off += len(b[5])
    raws.append(built[0][3][:2 * 4096] if built[0][3] else bytes(2 * 4096))
    shared = b"".join(raws)
    batch = pkg.Batch(shared)
    batch.set_shape(128, 4)
    slots, plains = [], []
    for (L, R, n, e), b, o0 in zip(cases, built, offs):
        o, ctr, keys, raw, luts, times = b
        idx = np.arange(len(times), dtype=np.int32) + o0
        if len(times) == 0:
            ctr = np.zeros((1, R, 4), np.uint32)
            idx = np.zeros(1, np.int32)
        job = pkg.Batch.make_job(L, R, 1234, n, e, ctr, keys, idx)
        s = batch.add(job)
        assert s is not None
        batch.upload(s, o.J, o.S, o.D0, o.D1)
        slots.append(s)
        p = pkg.Lattice(L, R, 1)
        p.upload(o.J, o.S, o.D0, o.D1)
        p.set_luts(shared)
        p.set_refresh(ctr, keys, idx)
        p.run(0, n, e)
        plains.append(p)
    # graph replays (chunk 4) and a remainder, with syncs and reads in between
    for step in (4, 4, 3, 8, 5):
        batch.run(step)
        batch.sync()
        t, te = batch.progress(len(slots))
        assert all(t[i] <= te[i] for i in range(len(slots)))
    t, te = batch.progress(len(slots))
    assert list(t) == [c[2] for c in cases]
    for i, (s, p) in enumerate(zip(slots, plains)):
        got = batch.download(s)
        want = p.download()
        for g, w in zip(got, want):
            assert np.array_equal(g, w), "slot %d differs from the plain run" % s
        assert np.array_equal(batch.measure(s)["energy"], p.measure()["energy"])
    # the oracle on two of them
    for i in (1, 3):
        L, R, n, e = cases[i]
        o, ctr, keys, raw, luts, times = built[i]
        for tt in range(n):
            if tt in times:
                k = times.index(tt)
                orc.iteration(o, ctr[k], keys, luts[k])
            else:
                orc.iteration(o)
        S, D0, D1 = batch.download(slots[i])
        assert np.array_equal(S, o.S) and np.array_equal(D0, o.D0) and np.array_equal(D1, o.D1)
    for p in plains:
        p.close()
    batch.close()


@pytest.mark.parametrize("block", [128, 1024])
def test_batch_runs_span_lattices(pkg, orc, block):
    """More tiles than resident blocks: each persistent block's run crosses
    lattice boundaries and stages a different lattice's tables part-way, and
    lattices finish at different iterations (the split is rebuilt)."""
    cases = [(64, 1, 9, 1), (32, 1, 6, 1), (40, 2, 5, 2), (32, 1, 11, 3), (48, 1, 4, 1)]
    built = [lattice_case(orc, L, R, n, e, 40 + i) for i, (L, R, n, e) in enumerate(cases)]
    raws, offs, off = [], [], 0
    for b in built:
        raws.append(b[3])
        offs.append(off)
        off += len(b[5])
    shared = b"".join(raws)
    batch = pkg.Batch(shared)
    batch.set_shape(block, 2)
    slots, plains = [], []
    for (L, R, n, e), b, o0 in zip(cases, built, offs):
        o, ctr, keys, raw, luts, times = b
        idx = np.arange(len(times), dtype=np.int32) + o0
        s = batch.add(pkg.Batch.make_job(L, R, 1234, n, e, ctr, keys, idx))
        batch.upload(s, o.J, o.S, o.D0, o.D1)
        slots.append(s)
        p = pkg.Lattice(L, R, 1)
        p.upload(o.J, o.S, o.D0, o.D1)
        p.set_luts(shared)
        p.set_refresh(ctr, keys, idx)
        p.run(0, n, e)
        plains.append(p)
    for step in (3, 1, 5, 2):
        batch.run(step)
    batch.sync()
    t, _ = batch.progress(len(slots))
    assert list(t) == [c[2] for c in cases]
    for s, p in zip(slots, plains):
        for g, w in zip(batch.download(s), p.download()):
            assert np.array_equal(g, w), "slot %d differs from the plain run" % s
        assert np.array_equal(batch.measure(s)["energy"], p.measure()["energy"])
        p.close()
    batch.close()


def test_batch_planned_progress_and_early_reads(pkg, orc):
    """The host's planned counts equal the device's after a sync; a lattice
    that finished in a run can be read (on its own stream) while the next run
    is still queued, and its state is the final one."""
    a = lattice_case(orc, 16, 1, 6, 1, 21)
    b = lattice_case(orc, 24, 1, 40, 1, 22)
    raw = a[3] + b[3]
    batch = pkg.Batch(raw)
    batch.set_shape(256, 4)
    sa = batch.add(pkg.Batch.make_job(16, 1, 5, 6, 1, a[1], a[2], np.arange(len(a[5]), dtype=np.int32)))
    sb = batch.add(pkg.Batch.make_job(24, 1, 6, 40, 1, b[1], b[2],
                                      np.arange(len(b[5]), dtype=np.int32) + len(a[5])))
    batch.upload(sa, a[0].J, a[0].S, a[0].D0, a[0].D1)
    batch.upload(sb, b[0].J, b[0].S, b[0].D0, b[0].D1)
    batch.run(8)                                  # a finishes at 6, b reaches 8
    t, te = batch.planned(2)
    assert list(t) == [6, 8] and list(te) == [6, 40]
    batch.run(32)                                 # queued behind the first run
    batch.wait_previous()                         # the first run is done
    early = batch.download(sa)                    # a is final; b's run may still be going
    e_early = batch.measure(sa)["energy"]
    batch.sync()
    t2, _ = batch.progress(2)
    assert list(t2) == [6, 40] and list(batch.planned(2)[0]) == [6, 40]
    for x, y in zip(early, batch.download(sa)):
        assert np.array_equal(x, y)
    # the oracle on lattice a: six iterations with refreshes every iteration
    o, ctr, keys, _, luts, times = a
    for tt in range(6):
        orc.iteration(o, ctr[times.index(tt)], keys, luts[times.index(tt)])
    assert np.array_equal(early[0], o.S) and np.array_equal(early[1], o.D0)
    assert np.array_equal(e_early, orc.energy(o))
    batch.close()


def test_batch_admission_and_churn(pkg, orc):
    o, ctr, keys, raw, luts, times = lattice_case(orc, 16, 1, 12, 1, 3)
    job = pkg.Batch.make_job(16, 1, 77, 12, 1, ctr, keys, np.arange(len(times), dtype=np.int32))
    probe = pkg.Batch(raw)
    need = probe.job_bytes(job)
    probe.close()
    batch = pkg.Batch(raw, max_bytes=2 * need + need // 2)
    a = batch.add(job)
    b = batch.add(job)
    assert a is not None and b is not None
    assert batch.add(job) is None                 # a third does not fit
    assert batch.bytes_in_use() == 2 * need
    batch.run(5)
    batch.remove(a)                               # retire one mid-run
    c = batch.add(job)                            # and admit a new one in its place
    assert c == a
    t, _ = batch.progress(2)
    assert t[c] == 0 and t[b] == 5                # the new lattice starts from zero
    batch.run(20)
    batch.sync()
    t, te = batch.progress(2)
    assert list(t) == [12, 12]
    # b ran 5 + 7 iterations around the churn, c ran 12 from the start: same job, same result
    sb, sc = batch.download(b), batch.download(c)
    for x, y in zip(sb, sc):
        assert np.array_equal(x, y)
    batch.close()
