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


@pytest.mark.parametrize("L,R,nc", [(4, 1, 1), (6, 3, 1), (16, 4, 2), (40, 5, 6), (160, 2, 1)])
def test_measures_every_shape(pkg, orc, L, R, nc):
    """The one-pass observables on random words: the smallest lattice, odd
    replica counts (a replica without a pair), two pairs, six demon layers and
    the L = 160 headline lattice (blocks of the pass cover a ragged last
    group of rows)."""
    print("case L=%d R=%d nc=%d" % (L, R, nc), flush=True)
    o = random_state(L, R, nc, 7000 + L + R + nc)
    dev = pkg.Lattice(L, R, nc)
    dev.upload(o.J, o.S, o.D0, o.D1)
    assert_same_measures(dev, orc, o)
    m = dev.measure()
    assert m["unsat"].shape == (R, 64) and m["demon"].shape == (R, 64)
    assert ("q" in m) == (R >= 2) and (R < 2 or m["q"].shape == (R // 2, 64))
    # any output may be NULL: the demon counts alone
    import ctypes
    demon = np.zeros((R, 64), np.int32)
    assert dev.lib.msa_measure(dev.ctx, None, demon.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                               None, None) == 0
    assert np.array_equal(demon, m["demon"])
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


@pytest.mark.parametrize("L,R,nc,ty,tz", [
    (16, 1, 1, 16, 16),     # whole planes, whole lattice: halos wrap onto the block itself
    (16, 2, 1, 4, 4),       # row and plane halos
    (24, 2, 3, 6, 8),       # three layers: walk along x, y, z
    (32, 4, 1, 8, 16),
    (20, 2, 6, 5, 10),      # all six walk directions
    (40, 1, 2, 40, 8),
])
@pytest.mark.parametrize("mode", [1, 2])
def test_one_pass_sweep_matches_oracle(pkg, orc, L, R, nc, ty, tz, mode):
    """Both one-pass kernels; the 16- and 8-tile cases also take the XCD
    placement of mode 2 (tile count a multiple of 8)."""
    print("case L=%d R=%d nc=%d tile %dx%d mode %d" % (L, R, nc, ty, tz, mode), flush=True)
    o = random_state(L, R, nc, 4242 + L)
    every, niter = 3, 7
    times = list(range(0, niter, every))
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, [0.8] * len(times), jobid=1)
    dev = pkg.Lattice(L, R, nc)
    dev.set_tiling(mode, ty, tz)
    assert dev.tiling() == (mode, ty, tz)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(len(times), dtype=np.int32))
    for t in range(niter):
        dev.run(t, 1, every)
        if t % every == 0:
            orc.iteration(o, ctr[t // every], keys, luts[t // every])
        else:
            orc.iteration(o)
        assert_same_state(dev, o, "t=%d" % t)
    assert_same_measures(dev, orc, o)
    dev.close()


def test_default_tiling_matches_two_pass(pkg):
    L, R = 32, 2
    a = pkg.Lattice(L, R, 1)
    b = pkg.Lattice(L, R, 1)
    assert a.tiling()[0] == 2
    b.set_tiling(0)
    for d in (a, b):
        d.init_philox(99)
        d.upload(d.couplings(), d.download()[0],
                 *[np.random.default_rng(3).integers(0, 2 ** 64, size=(R, 1, L ** 3), dtype=np.uint64)
                   for _ in range(2)])
        d.run(0, 9, 0)
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)
    a.close()
    b.close()


@pytest.mark.parametrize("L,R,ty,tz", [(96, 2, 16, 8), (64, 1, 8, 16), (48, 4, 6, 12)])
def test_one_pass_kernels_agree_at_larger_sizes(pkg, L, R, ty, tz):
    """Modes 1 and 2 against the two-pass form on bigger lattices (many
    tiles, XCD placement active), 6 sweeps from a Philox state."""
    print("case L=%d R=%d tile %dx%d" % (L, R, ty, tz), flush=True)
    devs = [pkg.Lattice(L, R, 1) for _ in range(3)]
    rng = np.random.default_rng(L)
    d0, d1 = (rng.integers(0, 2 ** 64, size=(R, 1, L ** 3), dtype=np.uint64) for _ in range(2))
    for m, d in enumerate(devs):
        d.set_tiling(m, ty if m else 0, tz if m else 0)
        d.init_philox(1234)
        d.upload(d.couplings(), d.download()[0], d0, d1)
        d.run(0, 6, 0)
    ref = devs[0].download()
    for d in devs[1:]:
        for x, y in zip(d.download(), ref):
            assert np.array_equal(x, y)
    for d in devs:
        d.close()


@pytest.mark.parametrize("L,R,ty,tz", [
    (16, 1, 16, 16),     # whole planes: rows and planes alias onto the block itself
    (16, 2, 4, 4),
    (8, 1, 2, 2),        # the pipeline (tz + 6 planes) wraps the whole lattice
    (24, 2, 3, 8),
    (32, 4, 8, 16),
    (40, 3, 5, 10),      # odd R, XCD placement of (tile, replica) pairs
])
def test_sweep_pairs_match_oracle(pkg, orc, L, R, ty, tz):
    """Mode 3: refresh-free iterations run two per pass; runs of 1..4
    iterations with refreshes every 4 give pairs, single sweeps and refresh
    sweeps in every order."""
    print("case L=%d R=%d pair tile %dx%d" % (L, R, ty, tz), flush=True)
    o = random_state(L, R, 1, 777 + L + R)
    every = 4
    chunks = [1, 2, 3, 4, 2, 1, 4, 3]
    niter = sum(chunks)
    times = list(range(0, niter, every))
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, [0.9] * len(times), jobid=2)
    dev = pkg.Lattice(L, R, 1)
    dev.set_tiling(3, ty, tz)
    assert dev.tiling() == (3, ty, tz)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(len(times), dtype=np.int32))
    t = 0
    for n in chunks:
        dev.run(t, n, every)
        for u in range(t, t + n):
            if u % every == 0:
                orc.iteration(o, ctr[u // every], keys, luts[u // every])
            else:
                orc.iteration(o)
        t += n
        assert_same_state(dev, o, "after t=%d" % t)
    assert_same_measures(dev, orc, o)
    dev.close()


@pytest.mark.parametrize("L,R,ty,tz", [(96, 2, 12, 8), (64, 1, 16, 16), (160, 2, 5, 20)])
def test_sweep_pairs_agree_at_larger_sizes(pkg, L, R, ty, tz):
    """Mode 3 against the two-pass form on bigger lattices, 7 refresh-free
    sweeps (three pairs and one single sweep) from a Philox state."""
    print("case L=%d R=%d pair tile %dx%d" % (L, R, ty, tz), flush=True)
    devs = [pkg.Lattice(L, R, 1) for _ in range(2)]
    rng = np.random.default_rng(L + 1)
    d0, d1 = (rng.integers(0, 2 ** 64, size=(R, 1, L ** 3), dtype=np.uint64) for _ in range(2))
    devs[0].set_tiling(0)
    devs[1].set_tiling(3, ty, tz)
    for d in devs:
        d.init_philox(4321)
        d.upload(d.couplings(), d.download()[0], d0, d1)
        d.run(0, 7, 0)
    for x, y in zip(devs[1].download(), devs[0].download()):
        assert np.array_equal(x, y)
    for d in devs:
        d.close()


def test_L16_anneal_configuration(pkg, orc):
    """The L = 16 parity configuration end to end: 64 +-J samples, R = 1,
    random J and spins, demons drawn at T = 2.0 by the first refresh, linear
    anneal T 2.0 -> 0.7 over 4096 sweeps with a refresh every sweep, per-lane
    integer energy every 64 sweeps; every measurement and the final state
    bit-exact against the oracle."""
    L, R, nc, nsweeps, every = 16, 1, 1, 4096, 64
    o = random_state(L, R, nc, 16016)
    o.D0[...] = 0
    o.D1[...] = 0
    times = list(range(nsweeps))
    betas = [beta_of(t, nsweeps) for t in times]
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, betas, jobid=1)
    dev = pkg.Lattice(L, R, nc)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(nsweeps, dtype=np.int32))
    for t0 in range(0, nsweeps, every):
        dev.run(t0, every, 1)
        for t in range(t0, t0 + every):
            orc.iteration(o, ctr[t], keys, luts[t], nthreads=8)
        m = dev.measure()
        assert np.array_equal(m["energy"], orc.energy(o)), "energy differs after %d sweeps" % (t0 + every)
        assert np.array_equal(m["demon"], orc.demon_energy(o))
    assert_same_state(dev, o, "after %d sweeps" % nsweeps)
    dev.close()


def test_L64_R2_configuration_start(pkg, orc):
    """The start of the L = 64 parity configuration: R = 2 replicas sharing
    J, anneal T 2.0 -> 0.7 over 65536 sweeps with a refresh every sweep; the
    first 256 sweeps, with the integer energies, demon energies, q and q_link
    of the replica pair compared bit-exactly every 128 sweeps."""
    L, R, nc, nsweeps, every = 64, 2, 1, 256, 128
    o = random_state(L, R, nc, 64064)
    o.D0[...] = 0
    o.D1[...] = 0
    times = list(range(nsweeps))
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, [beta_of(t, 65536) for t in times], jobid=2)
    dev = pkg.Lattice(L, R, nc)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(nsweeps, dtype=np.int32))
    for t0 in range(0, nsweeps, every):
        dev.run(t0, every, 1)
        for t in range(t0, t0 + every):
            orc.iteration(o, ctr[t], keys, luts[t], nthreads=16)
        assert_same_measures(dev, orc, o)
    assert_same_state(dev, o, "after %d sweeps" % nsweeps)
    dev.close()


@pytest.mark.parametrize("tile", ["5x40", "autotuned"])
def test_L160_roofline_lattice_matches_oracle(pkg, orc, tile):
    """The L = 160, R = 2, nc = 1 lattice of the roofline configuration against
    the oracle, on the sweep-pair kernel with the 5 x 40 tile and on the kernel
    and tile msa_autotune picks for this geometry (the choice bench.py times).
    Iterations 1..4 are refresh-free (two sweep-pair launches); iteration 5
    refreshes (the two-pass refresh kernel, Philox + LUT demons); 6 and 7 are
    one more pair.  Spins, demons and every per-lane count (unsatisfied bonds,
    demon energy, q, q_link) are compared bit-exactly after each run."""
    L, R, nc, every = 160, 2, 1, 5
    o = random_state(L, R, nc, 160160)
    times = [0, 5]
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, [1 / 1.8, 1 / 1.79], jobid=7)
    dev = pkg.Lattice(L, R, nc)
    if tile == "5x40":
        dev.set_tiling(3, 5, 40)
    else:
        dev.autotune(16)
    print("case L=160 R=2 tiling %s -> %s" % (tile, dev.tiling()), flush=True)
    if tile == "5x40":
        assert dev.tiling() == (3, 5, 40)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(len(times), dtype=np.int32))
    t = 1
    for n in (4, 3):
        dev.run(t, n, every)
        for u in range(t, t + n):
            if u % every == 0:
                orc.iteration(o, ctr[u // every], keys, luts[u // every], nthreads=16)
            else:
                orc.iteration(o, nthreads=16)
        t += n
        assert_same_state(dev, o, "after t=%d" % t)
        assert_same_measures(dev, orc, o)
    dev.close()
