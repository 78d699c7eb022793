"""Pin the lane-layout oracle's sweep, walk and energy against the literal
transcription of the reference kernels in the reference's own MSC layout
(oracle/msc.c), plus invariants and brute force.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import Lattice, read_stream, unpack_lane, u8p


def random_lattice(L, R, nc, seed):
    rng = np.random.default_rng(seed)
    lat = Lattice(L, R, nc)
    for a in (lat.J, lat.S, lat.D0, lat.D1):
        a[...] = rng.integers(0, 2 ** 64, size=a.shape, dtype=np.uint64)
    return lat


def _u8(a):
    return np.ascontiguousarray(a, np.uint8)


def msc_iteration(orc, lat, lane, niter=1):
    """Run the MSC transcription on sample `lane` and return real-space bytes."""
    m = orc.lib.orc_msc_create(lat.L, lat.R, lat.nc)
    assert m
    V = lat.V
    spin = _u8(unpack_lane(lat.S, lane))
    jx, jy, jz = (_u8(unpack_lane(lat.J[d], lane)) for d in range(3))
    d0 = _u8(unpack_lane(lat.D0, lane))
    d1 = _u8(unpack_lane(lat.D1, lane))
    p = lambda a: a.ctypes.data_as(u8p)
    orc.lib.orc_msc_pack(m, p(spin), p(jx), p(jy), p(jz), p(d0), p(d1))
    e_before = np.zeros(lat.R, np.int32)
    orc.lib.orc_msc_energy(m, e_before.ctypes.data_as(C.POINTER(C.c_int32)))
    for _ in range(niter):
        orc.lib.orc_msc_monte_carlo(m)
    os_, od0, od1 = np.zeros_like(spin), np.zeros_like(d0), np.zeros_like(d1)
    orc.lib.orc_msc_unpack(m, p(os_), p(od0), p(od1))
    e_after = np.zeros(lat.R, np.int32)
    orc.lib.orc_msc_energy(m, e_after.ctypes.data_as(C.POINTER(C.c_int32)))
    orc.lib.orc_msc_destroy(m)
    return os_.reshape(lat.R, V), od0.reshape(lat.R, lat.nc, V), od1.reshape(lat.R, lat.nc, V), \
        e_before, e_after


@pytest.mark.parametrize("L", [8, 16, 24, 32, 40])
def test_msc_tables(orc, L):
    m = orc.lib.orc_msc_create(L, 2, 1)
    assert m
    assert orc.lib.orc_msc_check_tables(m) == 0
    orc.lib.orc_msc_destroy(m)


def test_msc_pack_roundtrip(orc):
    lat = random_lattice(16, 2, 3, 1)
    lane = 17
    s, d0, d1, _, _ = msc_iteration(orc, lat, lane, niter=0)
    assert np.array_equal(s, unpack_lane(lat.S, lane))
    assert np.array_equal(d0, unpack_lane(lat.D0, lane))
    assert np.array_equal(d1, unpack_lane(lat.D1, lane))


@pytest.mark.parametrize("L,R,nc,lanes", [
    (8, 1, 1, [0, 63]),
    (16, 2, 1, [0, 5, 31, 63]),
    (16, 2, 6, [3, 40]),
    (24, 1, 3, [7]),
    (32, 2, 2, [11]),
])
def test_lane_oracle_equals_msc_kernels(orc, L, R, nc, lanes):
    lat = random_lattice(L, R, nc, 100 + L + nc)
    ref = lat.copy()
    niter = 2
    for _ in range(niter):
        orc.iteration(lat)
    for lane in lanes:
        s, d0, d1, eb, ea = msc_iteration(orc, ref, lane, niter=niter)
        assert np.array_equal(unpack_lane(lat.S, lane), s), ("spins", lane)
        assert np.array_equal(unpack_lane(lat.D0, lane), d0), ("D0", lane)
        assert np.array_equal(unpack_lane(lat.D1, lane), d1), ("D1", lane)
        e = orc.energy(lat)
        assert np.array_equal(e[:, lane], ea), ("energy", lane)
        e0 = orc.energy(ref)
        assert np.array_equal(e0[:, lane], eb)


def test_msc_energy_matches_scalar_energy(orc):
    lat = random_lattice(16, 1, 1, 7)
    for lane in (0, 9, 63):
        _, _, _, eb, _ = msc_iteration(orc, lat, lane, niter=0)
        p = lambda a: np.ascontiguousarray(a, np.uint8).ctypes.data_as(u8p)
        spin = _u8(unpack_lane(lat.S[0], lane))
        js = [_u8(unpack_lane(lat.J[d], lane)) for d in range(3)]
        es = orc.lib.orc_msc_scalar_energy(16, p(spin), p(js[0]), p(js[1]), p(js[2]))
        assert es == eb[0]


def test_walk_closed_form(orc):
    """Colour-0 demons move +2e, colour-1 demons -2e, e_3..5 = -x,-y,-z."""
    L = 8
    lat = random_lattice(L, 1, 6, 3)
    before = lat.copy()
    dirs = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1)]
    for ic in range(6):
        orc.walk(lat, ic)
    x, y, z = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    for ic, (dx, dy, dz) in enumerate(dirs):
        sgn = np.where((x + y + z) % 2 == 0, 1, -1)
        src = ((x - 2 * sgn * dx) % L) + L * (((y - 2 * sgn * dy) % L) + L * ((z - 2 * sgn * dz) % L))
        dst = x + L * (y + L * z)
        assert np.array_equal(lat.D0[0, ic, dst.ravel()], before.D0[0, ic, src.ravel()])


def test_ferromagnet_energy(orc):
    lat = Lattice(8, 2, 1)
    lat.S[1, :] = np.uint64(2 ** 64 - 1)
    e = orc.energy(lat)
    assert (e == 3 * lat.V).all()


def test_microcanonical_invariant(orc):
    """Between refreshes E_spin - 4 * sum(D) is conserved per lane (E = -h_Ener)."""
    lat = random_lattice(16, 2, 2, 11)
    lat.D1[...] = 0
    inv0 = -orc.energy(lat) + 4 * orc.demon_energy(lat)
    changed = False
    for _ in range(3):
        s0 = lat.S.copy()
        orc.iteration(lat)
        changed |= not np.array_equal(s0, lat.S)
        assert np.array_equal(-orc.energy(lat) + 4 * orc.demon_energy(lat), inv0)
    assert changed


def test_overlap_brute_force(orc):
    L = 6
    lat = random_lattice(L, 2, 1, 5)
    q, ql = orc.overlap(lat, 0, 1)
    A = [unpack_lane(lat.S[0], b).reshape(L, L, L) for b in range(64)]   # [z][y][x]
    B = [unpack_lane(lat.S[1], b).reshape(L, L, L) for b in range(64)]
    for b in range(64):
        a, bb = A[b].astype(int), B[b].astype(int)
        assert q[b] == int((a != bb).sum())
        n = 0
        for ax in range(3):
            la = a ^ np.roll(a, -1, axis=ax)
            lb = bb ^ np.roll(bb, -1, axis=ax)
            n += int((la != lb).sum())
        assert ql[b] == n


def test_init_J_uses_reference_stream(orc):
    """Lane 0 of the reference-mode J is HIRANDOM>>63 of ini_random(seeds_J[0])
    (golden stream from the reference's own generator)."""
    L = 4
    lat = Lattice(L, 1, 1)
    orc.init_ref(lat, 1001, 2002)
    draws = [int(r[0]) for r in read_stream("rng_janus.txt")]
    seeds = [int(r[0]) for r in read_stream("rng_seeds_1001.txt")]
    assert seeds[0] == 15548435256649372992
    bits = [d >> 63 for d in draws[:3 * lat.V]]
    for s in range(lat.V):
        for d in range(3):
            assert int(lat.J[d, s]) & 1 == bits[3 * s + d]


def test_init_spins_uses_reference_stream(orc):
    lat = Lattice(4, 1, 1)
    orc.init_ref(lat, 1001, 2002)
    draws = [int(r[0]) for r in read_stream("rng_hq_2002.txt")]
    nsite = len(draws) // 64
    for s in range(nsite):
        w = sum((draws[64 * s + b] >> 63) << b for b in range(64))
        assert int(lat.S[0, s]) == w


def test_refresh_law(orc):
    """Refreshed demons follow P(D) ~ p^D, D in 0..3."""
    beta = 0.5
    lat = Lattice(16, 1, 1)
    luts = orc.lut_demon(beta)
    ctr = np.array([[11, 22, 33, 44]], np.uint32)
    orc.refresh_layer(lat, 0, ctr, np.array([1234], np.uint32), luts)
    d0 = np.unpackbits(lat.D0.view(np.uint8)).astype(int)
    d1 = np.unpackbits(lat.D1.view(np.uint8)).astype(int)
    D = d0 + 2 * d1
    p = np.exp(-4 * beta)
    probs = np.array([p ** k for k in range(4)])
    probs /= probs.sum()
    n = D.size
    hist = np.bincount(D, minlength=4) / n
    for k in range(4):
        assert abs(hist[k] - probs[k]) < 5 * np.sqrt(probs[k] * (1 - probs[k]) / n)
