"""Served form: requests admitted into a running batch as memory frees up.

Every served result must equal the same request run alone on the plain path
(one lattice through msa_run: the single-lattice kernels, not the batch
kernel, with the same counters, keys and grid LUT pairs), bit for bit, and
the small lattices must also equal the oracle run of the request."""
import math

import numpy as np
import pytest

from oracle_lib import Lattice as OLattice
from sched_util import oracle_schedule

pytestmark = pytest.mark.gpu


def modules():
    import importlib

    from conftest import load_package
    load_package()
    return (importlib.import_module("micsa_ea_mc_amd.serve"),
            importlib.import_module("micsa_ea_mc_amd.msa"))


def grid_betas(serve, req):
    """Inverse temperature of every refresh of a served request: the linear
    anneal T0 -> T1 over niter sweeps, refresh every sweep, T rounded to the
    serving grid (msa_sched_tables_grid restated)."""
    nlo = int(round(serve.T_LO / serve.TGRID))
    nlut = int(round(serve.T_HI / serve.TGRID)) - nlo + 1
    out = []
    for t in range(req.niter):
        beta = 1.0 / (req.T0 + (req.T1 - req.T0) * t / req.niter)
        n = min(max(int(math.floor((1.0 / beta) / serve.TGRID + 0.5)) - nlo, 0), nlut - 1)
        out.append(1.0 / (serve.TGRID * (nlo + n)))
    return out


def oracle_run(orc, serve, req):
    """The request on the oracle: state of msa_init_philox(seed), counters
    from the generators of seed_MC[r] = seed + r with jobid = rid mod 2^14."""
    o = OLattice(req.L, req.R, 1)
    orc.init_philox(o, req.seed)
    betas = grid_betas(serve, req)
    times = list(range(req.niter))
    ctr, keys, _, luts = oracle_schedule(orc, req.R, times, betas, seed_mc=req.seed,
                                         jobid=req.rid % 16384)
    for t in times:
        orc.iteration(o, ctr[t], keys, luts[t], nthreads=16)
    return o


def check_against_alone(serve, tables, reqs, done):
    for r in reqs:
        rec = done[r.rid]
        assert rec["finish"] >= rec["admit"] >= r.arrival - 1e-9
        (S, D0, D1), m = serve.run_alone(r, tables, r.rid)
        assert np.array_equal(S, rec["state"][0]), "request %d: spins differ alone vs mixed" % r.rid
        assert np.array_equal(D0, rec["state"][1]) and np.array_equal(D1, rec["state"][2]), \
            "request %d: demons differ alone vs mixed" % r.rid
        assert np.array_equal(m["energy"], rec["energy"]), "request %d: energies differ" % r.rid


def test_serving_mix_equals_alone():
    serve, msa = modules()
    reqs = serve.make_requests(10, rate=400.0, seed=5151, sizes=(8, 16, 24), log2_len=(5, 8))
    tables = serve.Tables()
    # room for about three of the largest lattices at once: admission must queue
    probe = msa.Batch(tables.luts)
    sizes = [probe.job_bytes(tables.job(r, r.rid)) for r in reqs]
    probe.close()
    limit = 3 * max(sizes)
    assert sum(sizes) > limit
    done, summary = serve.serve(reqs, max_bytes=limit, epoch=16, block=128, tables=tables,
                                keep_state=True, solo_min_L=0)
    assert summary["requests"] == len(reqs)
    assert summary["max_bytes_in_use"] <= limit          # admission by footprint held
    assert summary["max_concurrent"] >= 2                 # requests overlapped on the device
    assert any(done[r.rid]["admit"] > r.arrival + 1e-4 for r in reqs)   # someone waited
    check_against_alone(serve, tables, reqs, done)
    assert summary["latency_p99_s"] >= summary["latency_p50_s"] > 0


@pytest.mark.parametrize("share", [0.0, 2.0])
def test_served_configuration_subset(orc, share):
    """The served configuration's mix on 100 requests at short lengths: L
    uniform over {16, 32, 48, 64, 96, 128}, 64 +-J samples, 2^5..2^8 sweeps
    from T = 2.0 to a final T uniform in [0.6, 0.9], Poisson arrivals (seed
    5151), admission by footprint under a limit of three L = 128 lattices, so
    requests queue and overlap; every L = 128 request on the solo lane
    (share 0), or routed by the balance of pending work (share 2).  Every
    request's spins, demons and energies equal the request run alone on the
    plain path; the L <= 32 requests also equal the oracle."""
    serve, msa = modules()
    reqs = serve.make_requests(100, rate=2000.0, seed=5151, log2_len=(5, 8))
    assert {r.L for r in reqs} == {16, 32, 48, 64, 96, 128}
    tables = serve.Tables()
    probe = msa.Batch(tables.luts)
    limit = 3 * max(probe.job_bytes(tables.job(r, r.rid)) for r in reqs)
    probe.close()
    done, summary = serve.serve(reqs, max_bytes=limit, epoch=32, block=1024, tables=tables,
                                keep_state=True, solo_min_iter=0, solo_share=share)
    print("served %d requests: %.3f s, max %d resident, %d solo, %.2f spin updates/ps" % (
        summary["requests"], summary["makespan_s"], summary["max_concurrent"],
        summary["solo_requests"], summary["spin_updates_per_ps"]), flush=True)
    assert summary["requests"] == len(reqs)
    assert summary["max_bytes_in_use"] <= limit
    assert summary["max_concurrent"] >= 3
    n128 = sum(r.L >= 128 for r in reqs)
    if share == 0:
        # every L = 128 request ran one at a time as a plain lattice beside the batch
        assert summary["solo_requests"] == n128 > 0
    else:
        # some went solo, the others to the batch when the solo lane had more
        # than its share of the pending work
        assert 0 < summary["solo_requests"] <= n128
    check_against_alone(serve, tables, reqs, done)
    small = [r for r in reqs if r.L <= 32]
    assert len(small) >= 20
    for r in small:
        o = oracle_run(orc, serve, r)
        rec = done[r.rid]
        assert np.array_equal(rec["state"][0], o.S), "request %d: spins differ from the oracle" % r.rid
        assert np.array_equal(rec["state"][1], o.D0) and np.array_equal(rec["state"][2], o.D1), \
            "request %d: demons differ from the oracle" % r.rid
        assert np.array_equal(rec["energy"], orc.energy(o)), "request %d: energies" % r.rid


def test_solo_lane_in_chunks():
    """Every request on the solo lane (plain lattices one at a time, queued in
    chunks of 2048 iterations): lengths across chunk boundaries, two lattice
    geometries reused by later requests; each result equals the request run
    alone."""
    serve, msa = modules()
    reqs = [serve.Request(i, L, n, 2.0, 0.7, 31 + i, 0.002 * i)
            for i, (L, n) in enumerate([(16, 4100), (24, 2048), (16, 2049), (24, 700), (16, 5)])]
    tables = serve.Tables()
    done, summary = serve.serve(reqs, max_bytes=1 << 30, epoch=16, block=256, tables=tables,
                                keep_state=True, solo_min_L=16, solo_min_iter=0, solo_share=0)
    assert summary["requests"] == len(reqs) and summary["solo_requests"] == len(reqs)
    assert summary["max_concurrent"] == 0                 # nothing went to the batch
    check_against_alone(serve, tables, reqs, done)
