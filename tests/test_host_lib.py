"""The product's host C library (libmsa_host.so) against the golden fixtures
and the oracle, and the C ABI surface of both libraries (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from oracle_lib import REPO, Lattice as OLattice, read_stream

GOLD = os.path.join(REPO, "tests", "golden")


@pytest.fixture(scope="module")
def H(pkg):
    return pkg.host()


def test_lut_beta_matches_reference_files(pkg):
    with open(os.path.join(GOLD, "lut_betas.txt")) as f:
        names = [ln.split()[0] for ln in f] + ["LUT_for_PRNG_nbits10_NB1.bin"]
    for name in names:
        raw = open(os.path.join(GOLD, name), "rb").read()
        beta = float(np.frombuffer(raw[:8], np.float64)[0])
        assert pkg.lut_beta(beta) == raw[8:], name


@pytest.mark.parametrize("beta", [0.3, 0.5, 1.0 / 1.3, 1.1, 1.0 / 0.6, 2.5])
def test_lut_demon_matches_oracle(pkg, orc, beta):
    a, b = pkg.lut_demon(beta)
    oa, ob = orc.lut_demon(beta)
    assert a == bytes(oa) and b == bytes(ob)


def test_thresholds(H, orc):
    for beta in (0.4, 0.90909090909090, 1.7):
        t = np.zeros(3, np.uint64)
        H.msa_thresholds(beta, t.ctypes.data_as(C.POINTER(C.c_uint64)))
        assert [int(x) for x in t] == orc.thresholds(beta)


def test_seeds_from_one(H):
    want = [int(r[0]) for r in read_stream("rng_seeds_1001.txt")]
    out = np.zeros(len(want), np.uint64)
    H.msa_seeds_from_one(1001, out.ctypes.data_as(C.POINTER(C.c_uint64)), len(want))
    assert [int(x) for x in out] == want


def test_generators_and_counters(pkg, H):
    rows = read_stream("rng_cpu_3003.txt")
    pr, xo, key = pkg.Prng(), pkg.Xo256(), C.c_uint32()
    H.msa_cpu_generators(C.byref(pr), C.byref(xo), C.byref(key), 3003)
    assert key.value == int(rows[0][1])
    xo2 = pkg.Xo256.from_buffer_copy(xo)
    for r in rows[1:50]:
        assert H.msa_prng_next(C.byref(pr)) == int(r[0])
    lib = pkg.host()
    lib.msa_xo256_next.argtypes = [C.POINTER(pkg.Xo256)]
    lib.msa_xo256_next.restype = C.c_uint64
    for r in rows[1:50]:
        assert lib.msa_xo256_next(C.byref(xo2)) == int(r[1])
    for name in ("rng_fill_small.txt", "rng_fill_odd.txt", "rng_fill_big.txt"):
        with open(os.path.join(GOLD, name)) as f:
            hdr = f.readline().split()
        seed0, nrep, it0, parity, job, ncalls = (int(x) for x in hdr[3:9])
        ctrs = [r for r in read_stream(name) if r[0] == "ctr"]
        gens = []
        for ir in range(nrep):
            p, x, k = pkg.Prng(), pkg.Xo256(), C.c_uint32()
            H.msa_cpu_generators(C.byref(p), C.byref(x), C.byref(k), seed0 + ir)
            gens.append((p, x))
        i = 0
        for call in range(ncalls):
            for ir in range(nrep):
                c = np.zeros(4, np.uint32)
                H.msa_fill_counter(C.byref(gens[ir][0]), C.byref(gens[ir][1]), it0 + call, parity,
                                   job, c.ctypes.data_as(C.POINTER(C.c_uint32)))
                assert [int(v) for v in c] == [int(v) for v in ctrs[i][3:7]]
                i += 1


def test_janus_and_hq_streams(pkg, H):
    g = pkg.Prng()
    H.msa_janus_seed(C.byref(g), 15548435256649372992)
    want = [int(r[0]) for r in read_stream("rng_janus.txt")]
    assert [H.msa_janus_next(C.byref(g)) for _ in want] == want
    H.msa_prng_seed_hq(C.byref(g), 2002)
    want = [int(r[0]) for r in read_stream("rng_hq_2002.txt")]
    assert [H.msa_prng_next(C.byref(g)) for _ in want] == want


@pytest.mark.parametrize("seeding", ["hq", "janus"])
def test_prng_skip_equals_stepping(pkg, seeding):
    """msa_prng_skip (jump-ahead of the lagged sum and the LCG) against n
    plain draws, on both sides of its switch to jump-ahead (n = 64) and of
    the ring's 256-word wrap, and after earlier draws."""
    H = pkg.host()
    seed = H.msa_prng_seed_hq if seeding == "hq" else H.msa_janus_seed
    for pre in (0, 5, 300):
        for n in (0, 1, 7, 62, 63, 64, 65, 200, 255, 256, 257, 1000, 12345, 1 << 20):
            a, b = pkg.Prng(), pkg.Prng()
            seed(C.byref(a), 2002 + pre)
            for _ in range(pre):
                H.msa_prng_next(C.byref(a))
            C.pointer(b)[0] = a
            for _ in range(n):
                H.msa_prng_next(C.byref(a))
            H.msa_prng_skip(C.byref(b), n)
            assert (a.lcg, a.out) == (b.lcg, b.out), (pre, n)
            assert [H.msa_prng_next(C.byref(a)) for _ in range(300)] == \
                   [H.msa_prng_next(C.byref(b)) for _ in range(300)], (pre, n)


@pytest.mark.parametrize("L,R,nslot", [(16, 2, 2), (8, 2, 70), (4, 1, 3)])
def test_single_seed_construction_equals_one_stream(pkg, L, R, nslot):
    """The reference workload's construction (msa_build_couplings_one,
    msa_build_spins_slots), built in parallel from jumped-ahead generators,
    against the same streams drawn one by one."""
    H = pkg.host()
    V = L ** 3
    H.msa_build_couplings_one.argtypes = [C.c_int, C.c_uint64, C.c_void_p]
    H.msa_build_spins_slots.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_void_p]
    J = np.zeros((3, V), np.uint64)
    S = np.zeros((R, V), np.uint64)
    H.msa_build_couplings_one(L, 9164166177945067407, J.ctypes.data)
    H.msa_build_spins_slots(L, R, 2002, nslot, S.ctypes.data)
    g = pkg.Prng()
    H.msa_janus_seed(C.byref(g), 9164166177945067407)
    want = np.zeros_like(J)
    for s in range(V):
        for d in range(3):
            want[d, s] = 2 ** 64 - 1 if H.msa_prng_next(C.byref(g)) >> 63 else 0
    assert np.array_equal(J, want)
    H.msa_prng_seed_hq(C.byref(g), 2002)
    want = np.zeros_like(S)
    for r in range(R):
        for s in range(V):
            w = 0
            for b in range(nslot):
                bit = H.msa_prng_next(C.byref(g)) >> 63
                if b < 64:
                    w |= bit << b
            want[r, s] = w
    assert np.array_equal(S, want)


@pytest.mark.parametrize("L,R", [(4, 1), (8, 2), (6, 3), (16, 2), (24, 1)])
def test_state_construction_matches_oracle(pkg, orc, L, R):
    J, S = pkg.build_state(L, R, 1001, 2002)
    o = OLattice(L, R, 1)
    orc.init_ref(o, 1001, 2002)
    assert np.array_equal(J, o.J)
    assert np.array_equal(S, o.S)


@pytest.mark.parametrize("every,niter", [(1, 20), (7, 50), (-1, 40)])
def test_schedule_tables_match_oracle(pkg, orc, every, niter):
    from sched_util import oracle_schedule
    R = 2
    sch = pkg.schedule(2.0, 0.7, niter, every, beta_fixed=1.1)
    if every == -1:
        sch.T0 = 0.0          # fixed-beta, powers-of-two cadence
    ctr, keys, luts, idx = pkg.schedule_tables(sch, niter, R, [3003, 3004], jobid=9)
    H = pkg.host()
    nref = ctr.shape[0]
    times = [H.msa_sched_time(C.byref(sch), k) for k in range(nref)]
    betas = [H.msa_sched_beta(C.byref(sch), k) for k in range(nref)]
    if every > 0:
        assert times == list(range(0, niter, every))
        for t, b in zip(times, betas):
            assert b == 1.0 / (2.0 + (0.7 - 2.0) * t / niter)
    else:
        assert times == [0, 1, 2, 4, 8, 16, 32]
        assert all(b == 1.1 for b in betas)
    octr, okeys, oraw, _ = oracle_schedule(orc, R, times, betas, seed_mc=3003, jobid=9)
    assert np.array_equal(ctr, octr)
    assert np.array_equal(keys, okeys)
    # one LUT pair per distinct beta; refresh k uses pair idx[k]
    pair = 2 * 4096
    assert luts.size // pair == len(set(betas)) and list(idx) == sorted(idx)
    for k in range(nref):
        assert luts[idx[k] * pair:(idx[k] + 1) * pair].tobytes() == oraw[k * pair:(k + 1) * pair]
    if every > 0:
        assert list(idx) == list(range(nref))
    else:
        assert list(idx) == [0] * nref        # fixed beta: a single pair


def test_schedule_holds_T1_past_its_end(pkg):
    """Refreshes after the last sweep of the anneal (t > nsweeps) stay at T1
    instead of extrapolating the line to lower and then negative T; they
    share the LUT pair of T1."""
    H = pkg.host()
    sch = pkg.schedule(2.0, 0.7, 256, 4)
    beta_end = H.msa_sched_beta(C.byref(sch), 64)          # t = 256
    assert abs(1.0 / beta_end - 0.7) < 1e-12
    for k in (65, 100, 120, 1000):                          # t = 260 .. 4000
        assert H.msa_sched_beta(C.byref(sch), k) == beta_end
    ctr, keys, luts, idx = pkg.schedule_tables(sch, 480, 1, [3003], jobid=0)
    assert ctr.shape[0] == 120 and luts.size == 2 * 4096 * 65
    assert list(idx[:65]) == list(range(65)) and all(i == 64 for i in idx[64:])


def test_thermal_counters_and_continuation_match_oracle(pkg, orc):
    """Thermal counters (iteration 2t+p, parity p), then the refresh counters
    continuing from the same generators, against the oracle's _fill_bits."""
    from sched_util import oracle_schedule, oracle_thermal
    R, nterm, jobid = 2, 4, 11
    gens = pkg.generators([3003, 3004])
    tctr = pkg.thermal_counters(gens, nterm, jobid)
    ogens, octr = oracle_thermal(orc, R, nterm, seed_mc=3003, jobid=jobid)
    assert np.array_equal(tctr, octr)
    sch = pkg.schedule(2.0, 0.7, 20, 3)
    ctr, keys, luts, idx = pkg.schedule_tables_from(sch, 20, gens, jobid)
    H = pkg.host()
    nref = ctr.shape[0]
    times = [H.msa_sched_time(C.byref(sch), k) for k in range(nref)]
    betas = [H.msa_sched_beta(C.byref(sch), k) for k in range(nref)]
    rctr, rkeys, raw, _ = oracle_schedule(orc, R, times, betas, jobid=jobid, gens=ogens)
    assert np.array_equal(ctr, rctr) and np.array_equal(keys, rkeys)
    assert luts.tobytes() == raw and list(idx) == list(range(nref))


def _declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msa_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("lib,header", [("libmsa_hip.so", "msa.h"), ("libmsa_host.so", "msa_host.h")])
def test_library_exports_every_declared_symbol(pkg, lib, header):
    path = os.path.join(REPO, "micsa-ea-mc_amd", "lib", lib)
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    exported = set(ln.split()[-1] for ln in out.splitlines())
    missing = [s for s in _declared(header) if s not in exported]
    assert not missing, missing
    C.CDLL(path)   # loads without a GPU


@pytest.mark.parametrize("L,R,nc", [(4, 1, 1), (16, 3, 2), (160, 2, 1), (64, 16, 6)])
def test_footprint_is_the_state_allocation(pkg, L, R, nc):
    """msa_footprint_bytes (what serving admits by) is the one device
    allocation of a lattice: spins and demons twice (ping-pong), J, and the
    measurement counts (R unsat, R demon, R/2 q, R/2 q_link rows of 64 int32),
    each rounded up to 256 bytes.  Host arithmetic only, no GPU."""
    V = L ** 3
    up = lambda n: (n + 255) // 256 * 256  # noqa: E731
    want = (2 * up(8 * V * R) + up(8 * V * 3) + 2 * up(16 * V * R * nc)
            + up(4 * 64 * (2 * R + 2 * (R // 2))))
    assert pkg.footprint_bytes(L, R, nc) == want
    assert pkg.footprint_bytes(L + 1, R, nc) == -1          # odd L is rejected
    assert pkg.footprint_bytes(L, 17, nc) == -1             # at most 16 replicas
