"""Asynchronous state download (msa_download_state_async / msa_download_wait).

The arrays must hold the state as the work queued before the call leaves it,
also when more sweeps are queued before the wait, into page-locked arrays
(msa_host_alloc) and ordinary ones; the state is checked bit for bit against
the oracle after the same sweeps.
"""
import numpy as np
import pytest

from oracle_lib import Lattice as OLattice

pytestmark = pytest.mark.gpu


def random_state(L, R, nc, seed):
    rng = np.random.default_rng(seed)
    o = OLattice(L, R, nc)
    for a in (o.J, o.S, o.D0, o.D1):
        a[...] = rng.integers(0, 2 ** 64, size=a.shape, dtype=np.uint64)
    return o


@pytest.mark.parametrize("pinned", [True, False])
@pytest.mark.parametrize("L,R,nc", [(8, 1, 1), (16, 2, 2), (24, 2, 1)])
def test_async_download_is_a_snapshot(pkg, orc, L, R, nc, pinned):
    print("case L=%d R=%d nc=%d pinned=%d" % (L, R, nc, pinned), flush=True)
    o = random_state(L, R, nc, 4000 + L + R + nc)
    dev = pkg.Lattice(L, R, nc)
    dev.upload(o.J, o.S, o.D0, o.D1)
    bufs = []
    if pinned:
        bufs = [pkg.HostArray((R, L ** 3)), pkg.HostArray((R, nc, L ** 3)),
                pkg.HostArray((R, nc, L ** 3))]
        S, D0, D1 = (b.array for b in bufs)
    else:
        S = np.zeros((R, L ** 3), np.uint64)
        D0 = np.zeros((R, nc, L ** 3), np.uint64)
        D1 = np.zeros_like(D0)
    for a in (S, D0, D1):
        a[...] = 0
    dev.run(0, 3, 0)
    for _ in range(3):
        orc.iteration(o)
    dev.download_async(S, D0, D1)
    dev.run(3, 5, 0)                    # queued before the wait: not in the snapshot
    dev.download_wait()
    assert np.array_equal(S, o.S) and np.array_equal(D0, o.D0) and np.array_equal(D1, o.D1)
    for _ in range(5):
        orc.iteration(o)
    # a second snapshot into the same arrays, spins only, then the synchronous form
    D0[...] = 0
    dev.download_async(S, None, None)
    dev.download_wait()
    assert np.array_equal(S, o.S) and not D0.any()
    S2, D02, D12 = dev.download()
    assert np.array_equal(S2, o.S) and np.array_equal(D02, o.D0) and np.array_equal(D12, o.D1)
    # a call while a copy is queued waits for it first
    dev.download_async(S, D0, D1)
    dev.download_async(S, D0, D1)
    dev.download_wait()
    dev.download_wait()                 # nothing queued: returns at once
    assert np.array_equal(D0, o.D0) and np.array_equal(D1, o.D1)
    dev.close()
    for b in bufs:
        b.close()


@pytest.mark.parametrize("L,R,nc", [(16, 2, 1), (24, 3, 2)])
def test_async_measure_is_a_snapshot(pkg, orc, L, R, nc):
    """msa_measure_async measures the state the queued work leaves, also with
    sweeps queued before msa_measure_wait; one measurement at a time."""
    from test_gpu_parity import assert_same_measures
    print("case L=%d R=%d nc=%d" % (L, R, nc), flush=True)
    o = random_state(L, R, nc, 4100 + L + R + nc)
    dev = pkg.Lattice(L, R, nc)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.run(0, 4, 0)
    for _ in range(4):
        orc.iteration(o)
    dev.measure_async()
    with pytest.raises(pkg.MsaError):
        dev.measure_async()                 # one is queued
    dev.run(4, 3, 0)                        # queued before the wait: not in the counts
    between = dev.measure()                 # a synchronous one in between leaves it intact
    m = dev.measure_wait()
    assert np.array_equal(m["energy"], orc.energy(o))
    assert np.array_equal(m["demon"], orc.demon_energy(o))
    for p in range(R // 2):
        q, ql = orc.overlap(o, 2 * p, 2 * p + 1)
        assert np.array_equal(m["q"][p], q) and np.array_equal(m["qlink"][p], ql)
    with pytest.raises(pkg.MsaError):
        dev.measure_wait()                  # nothing queued
    for _ in range(3):
        orc.iteration(o)
    assert np.array_equal(between["energy"], orc.energy(o))
    assert_same_measures(dev, orc, o)       # the synchronous form after it
    dev.close()
