"""Pin the oracle's canonical Metropolis thermalization (Monte_Carlo with
flag = 0, main.cu:202-252): the flip algebra against its closed form, the
random bits against their law, and the lane-layout update against the
literal transcription in the reference's MSC layout (oracle/msc.c)."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import Lattice, unpack_lane, u8p


def thermal_counters(orc, R, seeds, nterm, jobid):
    """Counters of the reference's thermal phase: per iteration t, _fill_bits
    with iter = 2t, parity 0 then iter = 2t+1, parity 1 (main.cu:208-229)."""
    gens = [orc.cpu_generators(s) for s in seeds]
    keys = np.array([g[2] for g in gens], np.uint32)
    ctr = np.zeros((nterm, 2, R, 4), np.uint32)
    for t in range(nterm):
        for p in (0, 1):
            for r in range(R):
                ctr[t, p, r] = orc.fill_counter(gens[r][0], gens[r][1], 2 * t + p, p, jobid)
    return ctr, keys


def test_metropolis_closed_form_exhaustive(orc):
    """flip iff n_u + R >= 3, R = leading ones of (b0, b1, b2): the accept
    probability min(1, exp(-beta dE)) = p^max(0, 3 - n_u) with P(R >= k) = p^k."""
    for j in range(16):
        j1, j2, j3, j4 = j & 1, (j >> 1) & 1, (j >> 2) & 1, j >> 3
        n_u = j1 + 2 * j2 + j3 + 2 * j4
        for b in range(8):
            b0, b1, b2 = b & 1, (b >> 1) & 1, b >> 2
            R = 0 if not b0 else (1 if not b1 else (2 if not b2 else 3))
            f = orc.metropolis_flip(j1, j2, j3, j4, b0, b1, b2) & 1
            assert f == int(n_u + R >= 3), (j1, j2, j3, j4, b0, b1, b2)


@pytest.mark.parametrize("beta", [0.4, 0.9])
def test_rndbits_law(orc, beta):
    """Each bit of b0, b1, b2 is Bernoulli(p), p = exp(-4 beta), independent."""
    L, R = 16, 2
    lut = orc.lut_beta(beta)
    p = np.exp(-4 * beta)
    ctr = np.array([[11, 22, 33, 44], [55, 66, 77, 88]], np.uint32)
    out = orc.gen_rndbits(L, R, lut, ctr, [5, 6])
    bits = np.unpackbits(out.view(np.uint8).reshape(3, -1), axis=1)
    n = bits.shape[1]
    sigma = np.sqrt(p * (1 - p) / n)
    for k in range(3):
        assert abs(bits[k].mean() - p) < 5 * sigma
    both = (bits[0] & bits[1]).mean()
    assert abs(both - p * p) < 5 * np.sqrt(p * p / n)


@pytest.mark.parametrize("L,R,lanes", [(8, 1, (0, 63)), (16, 2, (3, 40)), (24, 1, (17,))])
def test_thermal_lane_equals_msc(orc, L, R, lanes):
    """Three thermal iterations on a random state: every sample of the lane
    oracle equals the MSC transcription of the reference kernels run on that
    sample (the random bits are shared by all samples in both)."""
    rng = np.random.default_rng(L * 7 + R)
    lat = Lattice(L, R, 1)
    for a in (lat.J, lat.S, lat.D0, lat.D1):
        a[...] = rng.integers(0, 2 ** 64, size=a.shape, dtype=np.uint64)
    lut = orc.lut_beta(0.6)
    ctr, keys = thermal_counters(orc, R, [3003 + r for r in range(R)], 3, 9)
    start = lat.copy()
    for t in range(3):
        orc.thermal_iteration(lat, lut, ctr[t, 0], ctr[t, 1], keys)
    V = L ** 3
    p = lambda a: a.ctypes.data_as(u8p)
    for lane in lanes:
        m = orc.lib.orc_msc_create(L, R, 1)
        spin = np.ascontiguousarray(unpack_lane(start.S, lane), np.uint8)
        jx, jy, jz = (np.ascontiguousarray(unpack_lane(start.J[d], lane), np.uint8) for d in range(3))
        d0 = np.ascontiguousarray(unpack_lane(start.D0, lane), np.uint8)
        d1 = np.ascontiguousarray(unpack_lane(start.D1, lane), np.uint8)
        orc.lib.orc_msc_pack(m, p(spin), p(jx), p(jy), p(jz), p(d0), p(d1))
        for t in range(3):
            c0 = np.ascontiguousarray(ctr[t, 0], np.uint32)
            c1 = np.ascontiguousarray(ctr[t, 1], np.uint32)
            k = np.ascontiguousarray(keys, np.uint32)
            orc.lib.orc_msc_thermal(m, C.byref(lut), c0.ctypes.data_as(C.POINTER(C.c_uint32)),
                                    c1.ctypes.data_as(C.POINTER(C.c_uint32)),
                                    k.ctypes.data_as(C.POINTER(C.c_uint32)))
        os_, od0, od1 = np.zeros_like(spin), np.zeros_like(d0), np.zeros_like(d1)
        orc.lib.orc_msc_unpack(m, p(os_), p(od0), p(od1))
        orc.lib.orc_msc_destroy(m)
        assert np.array_equal(os_.reshape(R, V), unpack_lane(lat.S, lane)), "lane %d" % lane
        # demons are not touched by the thermal branch
        assert np.array_equal(od0.reshape(R, 1, V), unpack_lane(start.D0, lane))
    assert np.array_equal(lat.D0, start.D0) and np.array_equal(lat.D1, start.D1)
    assert not np.array_equal(lat.S, start.S)


def test_thermalization_lowers_energy(orc):
    """Physics check: from random spins (e ~ 0), Metropolis at beta = 1
    brings the energy per site well below zero."""
    L, R = 16, 1
    lat = Lattice(L, R, 1)
    orc.init_philox(lat, 77)
    lut = orc.lut_beta(1.0)
    ctr, keys = thermal_counters(orc, R, [3003], 40, 1)
    e0 = orc.energy(lat).astype(np.float64).mean() / L ** 3
    for t in range(40):
        orc.thermal_iteration(lat, lut, ctr[t, 0], ctr[t, 1], keys)
    e1 = orc.energy(lat).astype(np.float64).mean() / L ** 3
    assert abs(e0) < 0.05 and e1 > 1.2     # energy here is 3V - 2 unsat = -H
