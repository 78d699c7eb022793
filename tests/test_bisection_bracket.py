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
