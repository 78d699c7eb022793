"""msa_run_thermal_noalloc (the thermalization with its scratch in the
library's fixed device array, no device allocation) against msa_run_thermal
on identical lattices: bit-identical spins, demons and energies on two calls
in a row, the second on a smaller lattice and fewer iterations (the array
keeps the first call's contents), and a call too large for the array is
refused before it runs."""
import numpy as np
import pytest

from oracle_lib import Lattice as OLattice
from sched_util import oracle_thermal

pytestmark = pytest.mark.gpu


def lattice_pair(pkg, L, R, seed):
    rng = np.random.default_rng(seed)
    o = OLattice(L, R, 1)
    for a in (o.J, o.S, o.D0, o.D1):
        a[...] = rng.integers(0, 2 ** 64, size=a.shape, dtype=np.uint64)
    devs = []
    for _ in range(2):
        d = pkg.Lattice(L, R, 1)
        d.upload(o.J, o.S, o.D0, o.D1)
        devs.append(d)
    return o, devs


@pytest.mark.parametrize("first,second", [((32, 2, 3, 0.9), (16, 1, 2, 0.5))])
def test_noalloc_thermal_equals_allocating_entry_point(pkg, orc, first, second):
    for call, (L, R, nterm, beta) in enumerate((first, second)):
        print("call %d: L=%d R=%d nterm=%d" % (call, L, R, nterm), flush=True)
        o, (ref, dev) = lattice_pair(pkg, L, R, 900 + call)
        gens, ctr = oracle_thermal(orc, R, nterm, seed_mc=6006 + call, jobid=3)
        keys = np.array([g[2] for g in gens], np.uint32)
        lut = bytes(orc.lut_beta(beta))
        ref.run_thermal(lut, ctr, keys)
        dev.run_thermal_noalloc(lut, ctr, keys)
        a, b = ref.download(), dev.download()
        for x, y in zip(a, b):
            assert np.array_equal(x, y), "call %d differs from msa_run_thermal" % call
        assert np.array_equal(ref.measure()["energy"], dev.measure()["energy"])
        for t in range(nterm):
            orc.thermal_iteration(o, orc.lut_beta(beta), ctr[t, 0], ctr[t, 1], keys)
        assert np.array_equal(b[0], o.S), "call %d differs from the oracle" % call
        ref.close()
        dev.close()


def test_noalloc_thermal_refuses_what_the_array_cannot_hold(pkg, orc):
    dev = pkg.Lattice(8, 1, 1)
    nterm = 600000                       # 32 B of counters per iteration: 19 MB > 16 MiB
    with pytest.raises(pkg.MsaError, match="fixed scratch"):
        dev.run_thermal_noalloc(bytes(orc.lut_beta(0.5)), np.zeros((nterm, 2, 1, 4), np.uint32),
                                np.ones(1, np.uint32))
    dev.close()
