"""The canonical Metropolis thermalization on the device (msa_run_thermal)
against the oracle's (itself pinned against the MSC-layout transcription of
main.cu:202-252 and 847-958), bit-exact."""
import numpy as np
import pytest

from oracle_lib import Lattice as OLattice
from sched_util import oracle_schedule, oracle_thermal

pytestmark = pytest.mark.gpu


def random_state(L, R, seed):
    rng = np.random.default_rng(seed)
    o = OLattice(L, R, 1)
    for a in (o.J, o.S, o.D0, o.D1):
        a[...] = rng.integers(0, 2 ** 64, size=a.shape, dtype=np.uint64)
    return o


@pytest.mark.parametrize("L,R,nterm,beta", [(8, 1, 3, 0.5), (16, 2, 4, 0.9), (24, 1, 2, 0.3),
                                            (32, 3, 2, 1.2),
                                            (16, 2, 3, 2.0)])   # T = 0.5: wide search brackets
def test_thermal_matches_oracle(pkg, orc, L, R, nterm, beta):
    print("case L=%d R=%d nterm=%d beta=%g" % (L, R, nterm, beta), flush=True)
    o = random_state(L, R, 100 + L)
    gens, ctr = oracle_thermal(orc, R, nterm, seed_mc=5005, jobid=7)
    keys = np.array([g[2] for g in gens], np.uint32)
    lut = orc.lut_beta(beta)
    dev = pkg.Lattice(L, R, 1)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.run_thermal(bytes(lut), ctr, keys)
    for t in range(nterm):
        orc.thermal_iteration(o, lut, ctr[t, 0], ctr[t, 1], keys)
    S, D0, D1 = dev.download()
    assert np.array_equal(S, o.S)
    assert np.array_equal(D0, o.D0) and np.array_equal(D1, o.D1)    # demons untouched
    assert np.array_equal(dev.measure()["energy"], orc.energy(o))
    dev.close()


def test_thermal_then_micro_matches_oracle(pkg, orc):
    """Thermal phase, then microcanonical iterations with refreshes whose
    counters continue from the generators the thermal phase advanced."""
    L, R, nterm, niter, every = 16, 2, 3, 9, 4
    o = random_state(L, R, 7)
    gens, tctr = oracle_thermal(orc, R, nterm, seed_mc=3003, jobid=2)
    keys = np.array([g[2] for g in gens], np.uint32)
    times = list(range(0, niter, every))
    ctr, _, raw, luts = oracle_schedule(orc, R, times, [0.8] * len(times), jobid=2, gens=gens)
    lut = orc.lut_beta(1.0)
    dev = pkg.Lattice(L, R, 1)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.run_thermal(bytes(lut), tctr, keys)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(len(times), dtype=np.int32))
    dev.run(0, niter, every)
    for t in range(nterm):
        orc.thermal_iteration(o, lut, tctr[t, 0], tctr[t, 1], keys)
    for t in range(niter):
        if t % every == 0:
            orc.iteration(o, ctr[t // every], keys, luts[t // every])
        else:
            orc.iteration(o)
    S, D0, D1 = dev.download()
    assert np.array_equal(S, o.S)
    assert np.array_equal(D0, o.D0) and np.array_equal(D1, o.D1)
    dev.close()


def test_thermal_rejects_bad_geometry(pkg, orc):
    dev = pkg.Lattice(12, 1, 1)
    with pytest.raises(pkg.MsaError):
        dev.run_thermal(bytes(orc.lut_beta(0.5)), np.zeros((1, 2, 1, 4), np.uint32), [1])
    dev.close()
