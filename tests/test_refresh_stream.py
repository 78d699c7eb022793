"""The Philox demon-refresh stream (the oracle's orc_refresh_layer, which the
device reproduces bit for bit in every refresh parity test) beyond its
single-site law P(D) ~ p^D: every (site, layer, replica, plane, refresh) draw
has a Philox input of its own, and the demons of neighbouring sites, of
different layers, of different replicas, of consecutive refreshes and of
different sample lanes are independent.

A shared counter shows as repeated words: two draws with the same Philox
input give identical words, while two independent words of 64 Bernoulli(q)
bits agree with probability (q^2 + (1-q)^2)^64, which at these temperatures
(q of both planes in 0.31 .. 0.45) is below 2^-51, so a chance repeat among
the few ten thousand words drawn here has probability below 1e-6.
Independence is tested by chi-square on the joint law of (D at a, D at b),
which must factorise; the seeds are fixed, so every test is deterministic."""
import numpy as np
import pytest
from scipy import stats

from oracle_lib import Lattice
from sched_util import oracle_schedule

L, NC, R = 8, 6, 4
BETAS = (0.05, 0.08, 0.1)        # three refreshes, p = e^{-4 beta} from 0.82 to 0.67


@pytest.fixture(scope="module")
def draws(orc):
    """Demons of three consecutive refreshes (iterations 0, 4, 8) of every
    layer and replica: D0, D1 arrays [refresh][R][nc][V]."""
    times = [0, 4, 8]
    ctr, keys, _, luts = oracle_schedule(orc, R, times, list(BETAS), seed_mc=3003, jobid=5)
    lat = Lattice(L, R, NC)
    d0, d1 = [], []
    for k in range(len(times)):
        for layer in range(NC):
            orc.refresh_layer(lat, layer, ctr[k], keys, luts[k])
        d0.append(lat.D0.copy())
        d1.append(lat.D1.copy())
    return np.array(d0), np.array(d1), keys


def bits(w):
    """uint64 array -> uint8 array [..., 64] of its lanes."""
    return ((w[..., None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(np.uint8)


def demon_values(d0, d1):
    return bits(d0) + 2 * bits(d1)


def independent(a, b, alpha=1e-4):
    """Chi-square test that the 4x4 joint law of (a, b) factorises."""
    table = np.zeros((4, 4))
    np.add.at(table, (a.ravel(), b.ravel()), 1)
    _, pval, _, _ = stats.chi2_contingency(table)
    return pval > alpha, pval


def test_no_two_draws_share_a_philox_input(draws):
    d0, d1, _ = draws
    # chance repeats expected among independent words of each plane
    for beta in BETAS:
        p = np.exp(-4 * beta)
        for q in (p / (1 + p), p * p / (1 + p * p)):
            n = d0.size
            assert n * n / 2 * (q * q + (1 - q) ** 2) ** 64 < 1e-6
    for plane in (d0, d1):
        w = plane.ravel()
        assert np.unique(w).size == w.size, "repeated demon words: two draws share a counter"
        # the two 32-bit halves of a plane word are consecutive calls (.x and
        # .x + 1): equal halves would mean the counter did not advance
        assert not np.any((w >> np.uint64(32)) == (w & np.uint64(0xFFFFFFFF)))
    # D0 and D1 of one site come from different calls
    assert not np.any(d0.ravel() == d1.ravel())


def test_neighbouring_sites_are_independent(draws):
    d0, d1, _ = draws
    D = demon_values(d0[0, 0, 0], d1[0, 0, 0]).reshape(L, L, L, 64)   # [z][y][x][lane]
    for axis in (2, 1, 0):                                            # x, y, z neighbours
        ok, p = independent(D, np.roll(D, -1, axis=axis))
        assert ok, "sites and their neighbours along axis %d correlate (p=%g)" % (axis, p)


def test_layers_replicas_and_refreshes_are_independent(draws):
    d0, d1, _ = draws
    D = demon_values(d0, d1)                      # [refresh][R][nc][V][64]
    for l in range(1, NC):
        ok, p = independent(D[0, 0, 0], D[0, 0, l])
        assert ok, "layer 0 and layer %d correlate (p=%g)" % (l, p)
    for r in range(1, R):
        ok, p = independent(D[0, 0, 0], D[0, r, 0])
        assert ok, "replica 0 and replica %d correlate (p=%g)" % (r, p)
    ok, p = independent(D[0, 0, 0], D[1, 0, 0])
    assert ok, "consecutive refreshes correlate (p=%g)" % p


def test_sample_lanes_are_independent(draws):
    """Bits of one word (different samples, one byte of eight Bernoulli
    draws per 8 lanes, rotated) are independent across lanes and bytes."""
    d0, d1, _ = draws
    D = demon_values(d0[0].reshape(-1), d1[0].reshape(-1))          # [words][64]
    for a, b in ((0, 1), (7, 8), (0, 32), (31, 32), (5, 63)):
        ok, p = independent(D[:, a], D[:, b])
        assert ok, "lanes %d and %d correlate (p=%g)" % (a, b, p)


def test_the_law_holds_on_every_layer_and_replica(draws):
    """P(D = d) ~ p^d at each refresh's beta, per layer and per replica."""
    d0, d1, _ = draws
    D = demon_values(d0, d1)
    for k, beta in enumerate(BETAS):
        p = np.exp(-4 * beta)
        probs = np.array([p ** d for d in range(4)])
        probs /= probs.sum()
        for r in range(R):
            for l in range(NC):
                counts = np.bincount(D[k, r, l].ravel(), minlength=4)
                _, pval = stats.chisquare(counts, probs * counts.sum())
                assert pval > 1e-4, (k, r, l, pval)


def test_a_planted_shared_counter_is_caught(orc):
    """The collision test's power: drawing two layers with the same tid (the
    mistake of indexing the counter by site alone) repeats every word."""
    ctr, keys, _, luts = oracle_schedule(orc, 1, [0], [0.05], seed_mc=3003, jobid=5)
    lat = Lattice(L, 1, 2)
    orc.refresh_layer(lat, 0, ctr[0], keys, luts[0])
    orc.refresh_layer(lat, 1, ctr[0], keys, luts[0])
    one = Lattice(L, 1, 1)
    orc.refresh_layer(one, 0, ctr[0], keys, luts[0])
    # layer 1 of a 2-layer lattice differs from layer 0 (tid = layer V + s) ...
    assert not np.array_equal(lat.D0[0, 0], lat.D0[0, 1])
    # ... and a lattice whose counters ignored the layer would draw layer 0's
    # words again: the collision test sees every one of them repeated
    w = np.concatenate([one.D0.ravel(), lat.D0[0, 0]])
    assert np.unique(w).size == w.size // 2
