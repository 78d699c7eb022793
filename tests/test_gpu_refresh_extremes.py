"""Refresh sweeps at the ends of the temperature range that the annealing
schedules and the served temperature grid use (T = 0.5 .. 2.5).  At low T
almost every threshold of a heat-bath table lies in the top buckets of the
presearch, so the bracket searches with several halvings run; at high T the
thresholds spread over the whole range.  Bit-exact against the oracle with
the refresh drawn inside the sweep (refresh mode 0) and drawn ahead (mode
1), one and two demon layers, on the default (resident) launches and on the
one-pass kernel."""
import numpy as np
import pytest

from sched_util import oracle_schedule
from test_gpu_parity import assert_same_measures, assert_same_state, random_state
from test_gpu_predraw import oracle_run

pytestmark = pytest.mark.gpu

TEMPS = [0.5, 0.6, 2.5, 1.0, 0.5, 2.5, 0.55, 0.6, 0.5, 2.0]


@pytest.mark.parametrize("tiling", [None, (2, 0, 0)])
@pytest.mark.parametrize("nc", [1, 2])
@pytest.mark.parametrize("mode", [0, 1])
def test_refresh_at_extreme_temperatures(pkg, orc, tiling, nc, mode):
    L, R = 8, 2
    print("case L=%d R=%d nc=%d tiling=%s mode %d" % (L, R, nc, tiling, mode), flush=True)
    o = random_state(L, R, nc, 6060 + nc)
    times = list(range(len(TEMPS)))
    ctr, keys, raw, luts = oracle_schedule(orc, R, times, [1.0 / T for T in TEMPS], jobid=11)
    dev = pkg.Lattice(L, R, nc)
    if tiling:
        dev.set_tiling(*tiling)
    dev.set_refresh_mode(mode)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_luts(raw)
    dev.set_refresh(ctr, keys, np.arange(len(times), dtype=np.int32))
    t = 0
    for n in (3, 7):
        dev.run(t, n, 1)
        oracle_run(orc, o, t, n, 1, times, ctr, keys, luts)
        t += n
        assert_same_state(dev, o, "t=%d" % t)
    assert_same_measures(dev, orc, o)
    dev.close()
