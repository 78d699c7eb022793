"""The device bisection starts from the presearch bracket without the
reference's umbral[254] pre-test; on the CPU, check that this search returns
the reference bisection's byte on every interval endpoint and bucket edge."""
import random

import pytest


def _umb(lut, n):
    return (lut.umbrales[2 * n] << 32) | lut.umbrales[2 * n + 1]


def _pre(lut, m):
    w = lut.prebusqueda[m >> 1]
    return (w >> 16) if m & 1 else (w & 0xFFFF)


def bracket_search(lut, R):
    e = _pre(lut, R >> 54)
    lo, hi = e & 255, e >> 8
    while hi - lo > 1:
        mid = (hi + lo) >> 1
        if R < _umb(lut, mid):
            hi = mid
        else:
            lo = mid
    return lo if R < _umb(lut, lo) else hi


@pytest.mark.parametrize("beta", [0.2, 0.5, 0.90909090909090, 2.0, 3.0])
def test_bracket_search_equals_reference_bisection(orc, beta):
    rnd = random.Random(7)
    for lut in (orc.lut_beta(beta),) + orc.lut_demon(beta):
        pts = [0, (1 << 64) - 1]
        prev = 0
        for k in range(256):
            u = _umb(lut, k)
            pts += [prev, max(prev, u - 1), u, min(u + 1, (1 << 64) - 1)]
            prev = u
        pts += [m << 54 for m in range(1024)] + [((m + 1) << 54) - 1 for m in range(1024)]
        pts += [rnd.getrandbits(64) for _ in range(2000)]
        for R in pts:
            assert bracket_search(lut, R) == orc.bisection(lut, R), hex(R)


def fine_presearch(lut, bits=15):
    """The device's presearch (k_fine_luts) restated: bucket b of the top
    `bits` bits of R holds the byte of the bucket's first R, b << (64 - bits),
    i.e. the smallest n with that R < umbral[n]; then the sentinel 255."""
    umb = [_umb(lut, n) for n in range(256)]
    pre = []
    for b in range(1 << bits):
        R = b << (64 - bits)
        lo, hi = 0, 255
        while lo < hi:
            m = (lo + hi) >> 1
            if R < umb[m]:
                hi = m
            else:
                lo = m + 1
        pre.append(lo)
    return umb, pre + [255]


def fine_bisect(umb, pre, R, bits=15):
    """msa::bisect: the bracket [pre[b], pre[b + 1]] of R's bucket, at most
    nine halvings, then the final comparison."""
    b = R >> (64 - bits)
    lo, hi = pre[b], pre[b + 1]
    if lo == hi:
        return lo
    for _ in range(9):
        if hi - lo <= 1:
            break
        mid = (hi + lo) >> 1
        if R < umb[mid]:
            hi = mid
        else:
            lo = mid
    return lo if R < umb[lo] else hi


@pytest.mark.parametrize("beta", [0.2, 1.0 / 1.8, 0.90909090909090, 1.0 / 0.6, 3.0])
def test_fine_presearch_equals_reference_bisection(orc, beta):
    """The 2^15-bucket search the device builds from the thresholds gives the
    reference bisection's byte (10-bit presearch, main.cu:615-630) on every
    interval endpoint, on both edges of every bucket and on random R, for the
    reference's heat-bath table and both demon-plane tables."""
    rnd = random.Random(11)
    for lut in (orc.lut_beta(beta),) + orc.lut_demon(beta):
        umb, pre = fine_presearch(lut)
        assert all(pre[b] <= pre[b + 1] for b in range(len(pre) - 1))
        pts = [0, (1 << 64) - 1]
        for u in umb:
            pts += [max(u - 1, 0), u, min(u + 1, (1 << 64) - 1)]
        step = 1 << 49
        for b in range(0, 1 << 15, 3):
            pts += [b * step, (b + 1) * step - 1]
        pts += [rnd.getrandbits(64) for _ in range(3000)]
        for R in pts:
            assert fine_bisect(umb, pre, R) == orc.bisection(lut, R), hex(R)
