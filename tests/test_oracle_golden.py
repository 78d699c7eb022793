"""Pin the oracle's generators and tables against the reference's own code.

Fixtures in tests/golden/ were produced by the reference's random_generator.cu
and preparation.c compiled in place (tests/golden/gen_golden.py); Philox is
pinned by the Random123 known-answer vectors and xoshiro256++ by values of
Vigna's generator computed by hand from its definition.
"""
import os

import numpy as np
import pytest

from oracle_lib import REPO, Lut, Xoshiro, read_stream

GOLD = os.path.join(REPO, "tests", "golden")


def test_luxury_stream(orc):
    want = [int(r[0]) for r in read_stream("rng_lux.txt")]
    assert orc.lux_stream(12345, len(want)) == want


def test_luescher_u64_stream(orc):
    want = [int(r[0]) for r in read_stream("rng_lux64.txt")]
    assert orc.lux64_stream(777, len(want)) == want


@pytest.mark.parametrize("seed,name", [(3003, "rng_cpu_3003.txt"), (3004, "rng_cpu_3004.txt")])
def test_cpu_generators(orc, seed, name):
    rows = read_stream(name)
    assert rows[0][0] == "key"
    pr, xo, key = orc.cpu_generators(seed)
    assert key == int(rows[0][1])
    for r in rows[1:]:
        orc.lib.orc_prlcg_step(pr)
        orc.lib.orc_xoshiro_step(xo)
        assert (pr.out, xo.out) == (int(r[0]), int(r[1]))


def test_hq_stream(orc):
    want = [int(r[0]) for r in read_stream("rng_hq_2002.txt")]
    assert orc.hq_stream(2002, len(want)) == want


def test_seeds_from_one(orc):
    want = [int(r[0]) for r in read_stream("rng_seeds_1001.txt")]
    assert orc.seeds_from_one(1001, len(want)) == want
    # the first seeds do not depend on how many are asked for
    assert orc.seeds_from_one(1001, 2) == want[:2]


def test_janus_stream(orc):
    want = [int(r[0]) for r in read_stream("rng_janus.txt")]
    assert orc.janus_stream(15548435256649372992, len(want)) == want


def test_xoshiro_published_values(orc):
    g = Xoshiro()
    for i, v in enumerate((1, 2, 3, 4)):
        g.s[i] = v
    orc.lib.orc_xoshiro_step(g)
    assert g.out == 41943041          # rotl(1+4, 23) + 1
    orc.lib.orc_xoshiro_step(g)
    assert g.out == 58720359          # from the updated state (7, 0, 262146, rotl(6, 45))


@pytest.mark.parametrize("name", ["rng_fill_small.txt", "rng_fill_odd.txt", "rng_fill_big.txt"])
def test_fill_counters(orc, name):
    with open(os.path.join(GOLD, name)) as f:
        header = f.readline().split()
    seed0, nrep, it0, parity, job, ncalls = (int(x) for x in header[3:9])
    rows = read_stream(name)
    keys = {int(r[1]): int(r[2]) for r in rows if r[0] == "key"}
    ctrs = [r for r in rows if r[0] == "ctr"]
    gens = [orc.cpu_generators(seed0 + ir) for ir in range(nrep)]
    for ir in range(nrep):
        assert gens[ir][2] == keys[ir]
    k = 0
    for call in range(ncalls):
        for ir in range(nrep):
            got = orc.fill_counter(gens[ir][0], gens[ir][1], it0 + call, parity, job)
            row = ctrs[k]
            assert (int(row[1]), int(row[2])) == (call, ir)
            assert got == [int(x) for x in row[3:7]]
            k += 1


# Random123 kat_vectors for philox4x32 with 10 rounds
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_kat(orc, ctr, key, want):
    assert orc.philox(ctr, key) == want


def _lut_files():
    out = []
    with open(os.path.join(GOLD, "lut_betas.txt")) as f:
        for ln in f:
            name, beta = ln.split()
            out.append(name)
    out.append("LUT_for_PRNG_nbits10_NB1.bin")
    return out


@pytest.mark.parametrize("name", _lut_files())
def test_lut_matches_reference(orc, name):
    raw = open(os.path.join(GOLD, name), "rb").read()
    beta = float(np.frombuffer(raw[:8], np.float64)[0])
    ref = Lut.from_file_bytes(raw)
    got = orc.lut_beta(beta)
    assert bytes(got) == bytes(ref), name


def _umbral(lut, n):
    return (lut.umbrales[2 * n] << 32) | lut.umbrales[2 * n + 1]


@pytest.mark.parametrize("beta", [0.3, 0.5, 0.90909090909090, 1.25, 2.0])
def test_bisection_exhaustive_endpoints(orc, beta):
    """Both ends of every non-empty CDF interval map to its byte
    (the self-check preparation.c:87-217 runs)."""
    for lut in (orc.lut_beta(beta),) + orc.lut_demon(beta):
        prev = 0
        for n in range(256):
            hi = _umbral(lut, n) - 1 if n < 255 else (1 << 64) - 1
            if hi >= prev:
                assert orc.bisection(lut, prev) == n
                assert orc.bisection(lut, hi) == n
            prev = _umbral(lut, n)


@pytest.mark.parametrize("beta", [0.5, 1.1])
def test_lut_closed_form(orc, beta):
    """Interval widths are floor(2^64 q^k (1-q)^(8-k)), k = popcount(byte)."""
    q = np.exp(-4 * beta)
    lut = orc.lut_beta(beta)
    prev = 0
    for n in range(255):
        w = _umbral(lut, n) - prev
        k = bin(n).count("1")
        expect = (2.0 ** 64) * q ** k * (1 - q) ** (8 - k)
        assert abs(w - expect) <= max(1e-9 * expect, 2.0), (n, w, expect)
        prev = _umbral(lut, n)


def test_demon_lut_parameters(orc):
    """Plane-0 and plane-1 tables carry q0 = p/(1+p), q1 = p^2/(1+p^2)."""
    beta = 0.7
    p = np.exp(-4 * beta)
    for lut, q in zip(orc.lut_demon(beta), (p / (1 + p), p * p / (1 + p * p))):
        w0 = _umbral(lut, 0)      # byte 0: all eight bits clear
        assert abs(w0 / 2.0 ** 64 - (1 - q) ** 8) < 1e-12


def test_thresholds_closed_form(orc):
    for beta in (0.5, 0.90909090909090, 1.4):
        p = np.exp(-4 * beta)
        z = 1 + p + p * p + p ** 3
        acc = [p ** 3 / z, (p * p + p ** 3) / z, (p + p * p + p ** 3) / z]
        thr = orc.thresholds(beta)
        for t, a in zip(thr, acc):
            assert abs(t / 2.0 ** 64 - a) < 1e-14


def test_isi_closed_form_exhaustive(orc):
    """flip iff 0 <= D + n - 3 <= 3, D' = D + n - 3 on flip (n = #unsatisfied)."""
    for npos in range(4):
        for nneg in range(4):
            for d in range(4):
                j1, j2, j3, j4 = npos & 1, npos >> 1, nneg & 1, nneg >> 1
                d0, d1, f = orc.isi(j1, j2, j3, j4, d & 1, d >> 1)
                new = d + npos + nneg - 3
                want_flip = 0 <= new <= 3
                assert (f & 1) == int(want_flip), (npos, nneg, d)
                got_d = (d0 & 1) + 2 * (d1 & 1)
                assert got_d == (new if want_flip else d), (npos, nneg, d)


def test_rnd_word_density(orc):
    """Each bit of a word is Bernoulli(q), counter .x advances once per call."""
    beta = 0.6
    q = np.exp(-4 * beta)
    lut = orc.lut_beta(beta)
    ones = 0
    n = 4000
    for t in range(n):
        w, c = orc.rnd_word(lut, [5, 6, 7 + t, 8], [t, 99])
        assert c[0] == 6
        ones += bin(w).count("1")
    frac = ones / (32.0 * n)
    sigma = np.sqrt(q * (1 - q) / (32.0 * n))
    assert abs(frac - q) < 5 * sigma


def test_device_update_circuit_equals_reference_algebra(orc):
    """The reduced circuit of msa::site_update (csrc/msa_device.hpp) against the
    oracle's literal ISI algebra (main.cu:1037-1074), all 64 inputs."""
    for npos in range(4):
        for nneg in range(4):
            for d in range(4):
                j1, j2, j3, j4 = npos & 1, npos >> 1, nneg & 1, nneg >> 1
                D0, D1 = d & 1, d >> 1
                c1 = int((D0 + j1 + j3) in (1, 2))
                c2 = D0 & j1 & j3
                v0 = D1 ^ j2 ^ j4
                v1 = int(D1 + j2 + j4 >= 2)
                flip = v1 ^ c2 ^ (c1 & v0)
                nd0 = (1 - (D0 ^ j1 ^ j3)) if flip else D0
                nd1 = (v0 ^ c1) if flip else D1
                r0, r1, rf = orc.isi(j1, j2, j3, j4, D0, D1)
                assert (nd0, nd1, flip) == (r0 & 1, r1 & 1, rf & 1), (npos, nneg, d)
