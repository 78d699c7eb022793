"""Served form: requests admitted into a running batch as memory frees up;
every result must equal the same request run alone (bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_serving_mix_equals_alone(pkg):
    import importlib
    serve = importlib.import_module("micsa_ea_mc_amd.serve")
    msa = importlib.import_module("micsa_ea_mc_amd.msa")
    reqs = serve.make_requests(10, rate=400.0, seed=5151, sizes=(8, 16, 24), log2_len=(5, 8))
    tables = serve.Tables()
    # room for about three of the largest lattices at once: admission must queue
    probe = msa.Batch(tables.luts)
    sizes = [probe.job_bytes(tables.job(r, r.rid)) for r in reqs]
    probe.close()
    limit = 3 * max(sizes)
    assert sum(sizes) > limit
    done, summary = serve.serve(reqs, max_bytes=limit, epoch=16, block=128, tables=tables,
                                keep_state=True)
    assert summary["requests"] == len(reqs)
    assert summary["max_bytes_in_use"] <= limit          # admission by footprint held
    assert summary["max_concurrent"] >= 2                 # requests overlapped on the device
    assert any(done[r.rid]["admit"] > r.arrival + 1e-4 for r in reqs)   # someone waited
    for r in reqs:
        rec = done[r.rid]
        assert rec["finish"] >= rec["admit"] >= r.arrival - 1e-9
        (S, D0, D1), m = serve.run_alone(r, tables, r.rid)
        assert np.array_equal(S, rec["state"][0]), "request %d: spins differ alone vs mixed" % r.rid
        assert np.array_equal(D0, rec["state"][1]) and np.array_equal(D1, rec["state"][2])
        assert np.array_equal(m["energy"], rec["energy"])
    assert summary["latency_p99_s"] >= summary["latency_p50_s"] > 0
