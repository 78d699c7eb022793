"""Resident launches of small lattices (tiling mode 4, k_sweep_resident and,
for one layer with x rows that tile a wavefront, k_sweep_resident_x): up
to 64 iterations per launch with the state in one CU's LDS, refreshes read
from the demons drawn ahead (refresh modes 1 and 2) or drawn inside the
sweep by the two-pass kernel between launches (mode 0).  Bit-exact against
the oracle: refresh cadences 1, 3, 2, powers of two and none, several
layers and replicas, one and two words per thread, runs split so launches
and draw chunks end at arbitrary iterations."""
import numpy as np
import pytest

from sched_util import beta_of, oracle_schedule
from test_gpu_parity import assert_same_measures, assert_same_state, random_state
from test_gpu_predraw import oracle_run, refresh_times

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L,R,nc,every,niter,splits", [
    (16, 1, 1, 1, 150, (70, 80)),       # the L = 16 configuration's cadence; three draw chunks
    (16, 2, 1, 3, 100, (37, 63)),
    (14, 1, 2, 1, 40, (40,)),           # two words per thread, two layers
    (12, 2, 3, -1, 100, (1, 99)),       # t = 0, 1, 2, 4, ...
    (10, 1, 6, 2, 20, (9, 11)),         # six layers, as the reference's workload
    (8, 3, 2, 0, 130, (130,)),          # no refresh: three launches
    (8, 3, 1, 1, 70, (33, 37)),         # demons in registers, one word per thread
    (6, 2, 2, 1, 30, (13, 17)),
    (4, 1, 1, 1, 20, (20,)),            # 32 sites of a colour in a 64-thread block
])
@pytest.mark.parametrize("mode", [1, 2, 0])
def test_resident_matches_oracle(pkg, orc, L, R, nc, every, niter, splits, mode):
    print("case L=%d R=%d nc=%d every=%d niter=%d mode %d" % (L, R, nc, every, niter, mode),
          flush=True)
    o = random_state(L, R, nc, 5151 + L + R + nc)
    dev = pkg.Lattice(L, R, nc)
    dev.set_tiling(4)
    assert dev.tiling()[0] == 4
    dev.set_refresh_mode(mode)
    dev.upload(o.J, o.S, o.D0, o.D1)
    times, ctr, keys, luts = [], None, None, None
    if every != 0:
        times = refresh_times(niter, every)
        betas = [beta_of(t, niter) for t in times]
        ctr, keys, raw, luts = oracle_schedule(orc, R, times, betas, jobid=5)
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


def test_resident_by_default_where_it_fits(pkg):
    """Default tiling: resident for lattices that fit one CU, not beyond."""
    for L, R, nc, res in [(16, 1, 1, True), (16, 2, 1, True), (8, 1, 6, True),
                          (16, 1, 2, False), (24, 1, 1, False)]:
        d = pkg.Lattice(L, R, nc)
        assert (d.tiling()[0] == 4) == res, (L, R, nc)
        if not res:
            with pytest.raises(pkg.MsaError):
                d.set_tiling(4)
        d.close()
