"""The reference's demon streams (InitWalkers, ini.c:712-781): the host's
conversion between generators and stream states (CPU), and the device's
jump-ahead refresh against the oracle's sequential restatement (GPU)."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import Lattice as OLattice


def test_stream_state_round_trip(pkg, orc):
    """stream_of / stream_to rebuild generators that continue both streams,
    with the same ring, pointers and last outputs."""
    H = pkg.host()
    streams, gens = pkg.walker_streams(2002, 2, 2)
    H.msa_xo256_next.argtypes = [C.POINTER(pkg.Xo256)]
    H.msa_xo256_next.restype = C.c_uint64
    for st, (pr, xo) in zip(streams, gens):
        for _ in range(1000):                    # move away from the seeding
            H.msa_prng_next(C.byref(pr))
            H.msa_xo256_next(C.byref(xo))
        H.msa_walker_stream_of(C.byref(pr), C.byref(xo), C.byref(st))
        pr2, xo2 = pkg.Prng(), pkg.Xo256()
        H.msa_walker_stream_to(C.byref(st), pr.p, C.byref(pr2), C.byref(xo2))
        assert list(pr2.ring) == list(pr.ring)
        assert (pr2.p, pr2.p24, pr2.p55, pr2.p61) == (pr.p, pr.p24, pr.p55, pr.p61)
        assert pr2.lcg == pr.lcg and pr2.out == pr.out
        assert list(xo2.s) == list(xo.s) and xo2.out == xo.out
        for _ in range(300):
            assert H.msa_prng_next(C.byref(pr2)) == H.msa_prng_next(C.byref(pr))
            assert H.msa_xo256_next(C.byref(xo2)) == H.msa_xo256_next(C.byref(xo))


def test_walker_generators_match_oracle(pkg, orc):
    """Stream (i, r) is seeded from seeds_C[i*R + r] (ini.c:53-58)."""
    streams, _ = pkg.walker_streams(2002, 3, 2)
    seeds = orc.seeds_from_one(2002, 6)
    for k, st in enumerate(streams):
        pr, xo, _ = orc.cpu_generators(seeds[k])
        assert st.lcg == pr.zseed
        assert list(st.xo) == list(xo.s)
        assert [st.ring[m] for m in range(61)] == [pr.ira[(pr.ip - 61 + m) % 256] for m in range(61)]


def oracle_walker_demons(pkg, orc, L, R, nc, lanes, oracle_gens, thr):
    """One refresh from the oracle: InitWalkers per (stream, replica), placed
    in the lanes by the walkers-file converter; other lanes zero."""
    nw = (L ** 3 // 64) * nc
    V = L ** 3
    files = np.zeros((len(lanes), 2 * nw * R), np.uint64)
    for i in range(len(lanes)):
        for r in range(R):
            pr, xo = oracle_gens[i * R + r]
            w0, w1 = orc.init_walkers(pr, xo, thr, nw)
            files[i, r:nw * R:R] = w0                 # word iword at iword * R + r
            files[i, nw * R + r::R] = w1
    d0 = np.zeros((R, nc, V), np.uint64)
    d1 = np.zeros((R, nc, V), np.uint64)
    pkg.msc_to_lanes(L, R, nc, files, lanes, d0, d1)
    return d0, d1


@pytest.mark.gpu
@pytest.mark.parametrize("L,R,nc,lanes", [(8, 2, 1, [0, 1]), (16, 2, 2, [5, 0, 63]),
                                          (24, 3, 1, [7]), (32, 1, 3, list(range(0, 64, 5)))])
def test_walker_refresh_matches_oracle(pkg, orc, L, R, nc, lanes):
    print("case L=%d R=%d nc=%d lanes=%s" % (L, R, nc, lanes), flush=True)
    rng = np.random.default_rng(L * 7 + nc)
    o = OLattice(L, R, nc)
    for arr in (o.J, o.S, o.D0, o.D1):
        arr[...] = rng.integers(0, 2 ** 64, size=arr.shape, dtype=np.uint64)
    n = len(lanes)
    streams, _ = pkg.walker_streams(2002, n, R)
    seeds = orc.seeds_from_one(2002, n * R)
    ogens = [orc.cpu_generators(seeds[k])[:2] for k in range(n * R)]
    betas = [0.9, 0.7, 1.1, 0.8]
    thr = np.array([orc.thresholds(b) for b in betas], np.uint64)
    dev = pkg.Lattice(L, R, nc)
    dev.upload(o.J, o.S, o.D0, o.D1)
    dev.set_walker_streams(lanes, streams, thr)
    k = 0
    for t in range(6):                       # refreshes at t = 0, 1, 2, 4 (every = -1)
        dev.run(t, 1, -1)
        if t == 0 or (t & (t - 1)) == 0:
            o.D0[...], o.D1[...] = oracle_walker_demons(pkg, orc, L, R, nc, lanes, ogens, thr[k])
            k += 1
        orc.iteration(o)
        S, D0, D1 = dev.download()
        assert np.array_equal(D0, o.D0) and np.array_equal(D1, o.D1), "demons t=%d" % t
        assert np.array_equal(S, o.S), "spins t=%d" % t
    # the device's stream states continue the oracle's generators
    for st, (pr, xo) in zip(dev.walker_streams(), ogens):
        assert st.lcg == pr.zseed and list(st.xo) == list(xo.s)
        assert [st.ring[m] for m in range(61)] == [pr.ira[(pr.ip - 61 + m) % 256] for m in range(61)]
    dev.close()
