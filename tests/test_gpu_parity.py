"""Bit-exact parity of the HIP path (through the C ABI) with the oracle.

Same inputs on both sides: J, spins and demons, Philox counters/keys and LUT
tables.  Integer and bitwise work, so the bar is bit-exact equality of every
spin word, every demon word and every per-lane count.
"""
import numpy as np
import pytest

from oracle_lib import Lattice as OLattice
from sched_util import beta_of, oracle_schedule

pytestmark = pytest.mark.gpu


def random_state(L, R, nc, seed):
    rng = np.random.default_rng(seed)
    o = OLattice(L, R, nc)
    for a in (o.J, o.S, o.D0, o.D1):
        a[...] = rng.integers(0, 2 ** 64, size=a.shape, dtype=np.uint64)
    return o


def assert_same_state(dev, o, what=""):
    S, D0, D1 = dev.download()
    assert np.array_equal(S, o.S), "spins differ " + what
    assert np.array_equal(D0, o.D0), "D0 differs " + what
    assert np.array_equal(D1, o.D1), "D1 differs " + what


def assert_same_measures(dev, orc, o):
    m = dev.measure()
    assert np.array_equal(m["energy"], orc.energy(o))
    assert np.array_equal(m["demon"], orc.demon_energy(o))
    for p in range(o.R // 2):
        q, ql = orc.overlap(o, 2 * p, 2 * p + 1)
        assert np.array_equal(m["q"][p], q)
        assert np.array_equal(m["qlink"][p], ql)


@pytest.mark.parametrize("L,R,nc", [(16, 1, 1), (16, 2, 1), (24, 2, 3), (8, 3, 6), (32, 2, 2)])
def test_microcanonical_steps(pkg, orc, L, R, nc):
    print("case L=%d R=%d nc=%d" % (L, R, nc), flush=True)
    o = random_state(L, R, nc, 1000 + L + R + nc)
    dev = pkg.Lattice(L, R, nc)
    dev.upload(o.J, o.S, o.D0, o.D1)
    assert_same_state(dev, o, "after upload")
    assert_same_measures(dev, orc, o)
    for it in range(3):
        dev.run(it, 1, 0)
        orc.iteration(o)
        assert_same_state(dev, o, "iteration %d" % it)
    assert_same_measures(dev, orc, o)
    dev.close()


@pytest.mark.parametrize("L,R,nc,every,niter", [(16, 1, 1, 1, 24), (16, 2, 2, 3, 12),
                                                (24, 2, 1, 4, 16), (8, 1, 6, 2, 6)])
def test_annealing_with_refresh(pkg, orc, L, R, nc, every, niter):
    print("case L=%d R=%d nc=%d every=%d" % (L, R, nc, every), flush=True)
    o = random_state(L, R, nc, 77 + L)
    times = list(range(0, niter, every))
    betas = [beta_of(t, niter) for t in times]
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, betas, jobid=5)
    dev = pkg.Lattice(L, R, nc)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(len(times), dtype=np.int32))
    t = 0
    while t < niter:
        n = min(5, niter - t)
        dev.run(t, n, every)
        for tt in range(t, t + n):
            if tt % every == 0:
                k = tt // every
                orc.iteration(o, ctr[k], keys, luts[k])
            else:
                orc.iteration(o)
        t += n
        assert_same_state(dev, o, "t=%d" % t)
        assert_same_measures(dev, orc, o)
    dev.close()


def test_power_of_two_cadence(pkg, orc):
    L, R, nc, niter = 16, 2, 1, 20
    o = random_state(L, R, nc, 5)
    times = [0] + [1 << j for j in range(5)]          # 0,1,2,4,8,16
    betas = [0.9] * len(times)
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, betas)
    dev = pkg.Lattice(L, R, nc)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(len(times), dtype=np.int32))
    dev.run(0, niter, -1)
    for t in range(niter):
        if t in times:
            k = times.index(t)
            orc.iteration(o, ctr[k], keys, luts[k])
        else:
            orc.iteration(o)
    assert_same_state(dev, o)
    dev.close()


def test_philox_construction(pkg, orc):
    L, R, nc = 16, 2, 1
    dev = pkg.Lattice(L, R, nc)
    dev.init_philox(0x1234567890ABCDEF)
    o = OLattice(L, R, nc)
    orc.init_philox(o, 0x1234567890ABCDEF)
    S, D0, D1 = dev.download()
    assert np.array_equal(dev.couplings(), o.J)
    assert np.array_equal(S, o.S)
    assert not D0.any() and not D1.any()
    dev.close()


def test_errors_are_reported(pkg):
    with pytest.raises(pkg.MsaError):
        pkg.Lattice(15, 1, 1)
    dev = pkg.Lattice(8, 1, 1)
    with pytest.raises(pkg.MsaError):
        dev.run(0, 1, 1)          # refresh requested, no tables uploaded
    dev.close()
