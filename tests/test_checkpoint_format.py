"""Checkpoint files of the host library (micsa-ea-mc_amd/host/msa_ckpt.c)
against the reference's byte layouts (io.c:17-393, header.h:69-111,
random_generator.h:29-44), the oracle's transcription of the reference's
multispin packing (oracle/msc.c) and the oracle's scalar energy (med.c:85-113).
No GPU needed."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle_lib import u8p, unpack_lane

MASK = (1 << 64) - 1


def rand_words(rng, shape):
    return rng.integers(0, 2 ** 64, size=shape, dtype=np.uint64)


def lane_bits(words, b):
    return ((words >> np.uint64(b)) & np.uint64(1)).astype(np.uint8)


# ------------------------------------------------------------ conversions --
@pytest.mark.parametrize("L,R", [(8, 1), (16, 2)])
def test_janus_payload(pkg, L, R):
    """Janus_packing_for_write (io.c:212-231): site s at bit s&7 of byte s>>3."""
    rng = np.random.default_rng(L + R)
    S = rand_words(rng, (R, L ** 3))
    lanes = [0, 5, 63, 17]
    jan = pkg.janus_from_lanes(L, S, lanes)
    for i, b in enumerate(lanes):
        for r in range(R):
            want = np.packbits(lane_bits(S[r], b), bitorder="little")
            assert np.array_equal(jan[i, r], want)
    # reading back overwrites only the listed lanes
    T = rand_words(rng, (R, L ** 3))
    pkg.janus_to_lanes(L, jan, lanes, T)
    for b in range(64):
        src = S if b in lanes else None
        if src is not None:
            assert np.array_equal(lane_bits(T, b), lane_bits(S, b))
    other = [b for b in range(64) if b not in lanes]
    T2 = T.copy()
    pkg.janus_to_lanes(L, jan, lanes, T2)
    for b in other:
        assert np.array_equal(lane_bits(T2, b), lane_bits(T, b))


@pytest.mark.parametrize("L,R,nc", [(8, 1, 1), (16, 2, 3)])
def test_walker_words_match_reference_packing(pkg, orc, L, R, nc):
    """Demon words of a walkers file equal the oracle's packing_u placement
    (ini.c:301-345) of the same sample's demons, in the file's order."""
    rng = np.random.default_rng(7 * L + nc)
    V = L ** 3
    S = rand_words(rng, (R, V))
    J = rand_words(rng, (3, V))
    D0 = rand_words(rng, (R, nc, V))
    D1 = rand_words(rng, (R, nc, V))
    lanes = [3, 40]
    w = pkg.msc_from_lanes(L, R, nc, D0, D1, lanes)
    nw = pkg.host().msa_msc_words(L, R, nc)
    assert w.shape == (2, nw) and nw == 2 * (V // 64) * R * nc
    for i, b in enumerate(lanes):
        m = orc.lib.orc_msc_create(L, R, nc)
        assert m
        u8 = lambda a: np.ascontiguousarray(unpack_lane(a, b), np.uint8)
        spin, d0, d1 = u8(S), u8(D0), u8(D1)
        jx, jy, jz = (u8(J[d]) for d in range(3))
        p = lambda a: a.ctypes.data_as(u8p)
        orc.lib.orc_msc_pack(m, p(spin), p(jx), p(jy), p(jz), p(d0), p(d1))
        want = np.zeros(nw, np.uint64)
        orc.lib.orc_msc_walker_words(m, want.ctypes.data_as(C.POINTER(C.c_uint64)))
        orc.lib.orc_msc_destroy(m)
        assert np.array_equal(w[i], want), b
    # inverse: only the listed lanes change
    E0 = rand_words(rng, (R, nc, V))
    E1 = rand_words(rng, (R, nc, V))
    F0, F1 = E0.copy(), E1.copy()
    pkg.msc_to_lanes(L, R, nc, w, lanes, F0, F1)
    for b in range(64):
        src0, src1 = (D0, D1) if b in lanes else (E0, E1)
        assert np.array_equal(lane_bits(F0, b), lane_bits(src0, b))
        assert np.array_equal(lane_bits(F1, b), lane_bits(src1, b))


def test_lane_energy_and_countJ(pkg, orc):
    L, R = 8, 2
    rng = np.random.default_rng(3)
    V = L ** 3
    S = rand_words(rng, (R, V))
    J = rand_words(rng, (3, V))
    e = pkg.lane_energy(L, J, S)
    for b in (0, 9, 63):
        js = [np.ascontiguousarray(lane_bits(J[d], b)) for d in range(3)]
        for r in range(R):
            sp = np.ascontiguousarray(lane_bits(S[r], b))
            p = lambda a: a.ctypes.data_as(u8p)
            assert e[r, b] == orc.lib.orc_msc_scalar_energy(L, p(sp), *(p(x) for x in js))
    # closed forms: aligned spins, J = 0 satisfies every bond; J = 1 frustrates every bond
    S0 = np.zeros((R, V), np.uint64)
    assert (pkg.lane_energy(L, np.zeros((3, V), np.uint64), S0) == 3 * V).all()
    assert (pkg.lane_energy(L, np.full((3, V), MASK, np.uint64), S0) == -3 * V).all()
    cnt = pkg.lane_countJ(L, J)
    for b in range(64):
        assert cnt[b] == int(lane_bits(J, b).sum())


# ------------------------------------------------------------ conf / walkers --
def _gens(pkg, seed, skip):
    H = pkg.host()
    pr, xo, key = pkg.Prng(), pkg.Xo256(), C.c_uint32()
    H.msa_cpu_generators(C.byref(pr), C.byref(xo), C.byref(key), seed)
    H.msa_gen_skip(C.byref(pr), C.byref(xo), 1, skip)
    return pr, xo


def test_generator_images(pkg):
    """s_aleatorio_HQ_64bits: ira[256], PR, zseed, final, ip..ip3, pad; s_xoshiro256pp:
    s[4], t, final.  Reading an image back continues the same stream."""
    H = pkg.host()
    pr, xo = _gens(pkg, 3003, 5)
    img = np.zeros(2080, np.uint8)
    H.msa_prng_to_ref(C.byref(pr), img.ctypes.data_as(u8p))
    w = img[:259 * 8].view(np.uint64)
    assert list(w[:256]) == list(pr.ring)
    assert int(w[257]) == pr.lcg and int(w[258]) == pr.out
    assert (int(w[256]) + pr.lcg) & MASK == pr.out          # final = zseed + PR
    assert list(img[2072:2076]) == [pr.p, pr.p24, pr.p55, pr.p61]
    # ip1 = ip - 24, ip2 = ip - 55, ip3 = ip - 61 (mod 256), random_generator.cu:120-122
    assert [(pr.p - x) % 256 for x in (pr.p24, pr.p55, pr.p61)] == [24, 55, 61]
    xi = np.zeros(48, np.uint8)
    H.msa_xo256_to_ref(C.byref(xo), xi.ctypes.data_as(u8p))
    xw = xi.view(np.uint64)
    assert list(xw[:4]) == list(xo.s) and int(xw[5]) == xo.out
    pr2, xo2 = pkg.Prng(), pkg.Xo256()
    H.msa_prng_from_ref(img.ctypes.data_as(u8p), C.byref(pr2))
    H.msa_xo256_from_ref(xi.ctypes.data_as(u8p), C.byref(xo2))
    for _ in range(300):
        assert H.msa_prng_next(C.byref(pr2)) == H.msa_prng_next(C.byref(pr))
    lib = H
    lib.msa_xo256_next.argtypes = [C.POINTER(pkg.Xo256)]
    lib.msa_xo256_next.restype = C.c_uint64
    for _ in range(50):
        assert lib.msa_xo256_next(C.byref(xo2)) == lib.msa_xo256_next(C.byref(xo))


def _sdata(pkg, L, now):
    d = pkg.SData()
    d.maxtime, d.now, d.ntimes, d.nterm, d.nc, d.nr, d.l, d.flag = 1 << 20, now, 7, 0, 2, 2, L, 1
    d.seed_J, d.seed_u = 0x1234567890ABCDEF, 99
    d.seed_MC[0], d.seed_MC[1] = 5, 6
    return d


def test_conf_file_layout(pkg, tmp_path):
    """write_conf (io.c:17-96): byte layout, size 2325 + V/8, hard link 'conf'."""
    L = 8
    V = L ** 3
    rng = np.random.default_rng(11)
    pr, xo = _gens(pkg, 77, 3)
    jan = rng.integers(0, 256, V // 8, dtype=np.uint8)
    d = _sdata(pkg, L, 64)
    rep = str(tmp_path)
    pkg.conf_write(rep, 64, d, 0.875, pr, xo, L, jan, -1234)
    name = os.path.join(rep, "conf.%016d" % 64)
    link = os.path.join(rep, "conf")
    assert os.path.exists(name) and os.stat(link).st_nlink == 2
    assert os.stat(link).st_ino == os.stat(name).st_ino
    raw = open(link, "rb").read()
    assert len(raw) == 184 + 8 + 2080 + 48 + 1 + V // 8 + 4 == pkg.host().msa_conf_bytes(L)
    u64 = lambda off: int(np.frombuffer(raw[off:off + 8], np.uint64)[0])
    i32 = lambda off: int(np.frombuffer(raw[off:off + 4], np.int32)[0])
    assert (u64(0), u64(8)) == (1 << 20, 64)
    assert [i32(16 + 4 * k) for k in range(6)] == [7, 0, 2, 2, L, 1]
    assert (u64(40), u64(48), u64(56), u64(64), u64(72)) == (0x1234567890ABCDEF, 99, 5, 6, 0)
    assert np.frombuffer(raw[184:192], np.float64)[0] == 0.875
    assert u64(192 + 258 * 8) == pr.out                 # PR+LCG final
    assert u64(192 + 2080 + 40) == xo.out               # xoshiro final
    assert raw[2320] == 0                               # clone map
    assert raw[2321:2321 + V // 8] == jan.tobytes()
    assert i32(2321 + V // 8) == -1234
    d2, beta, pr2, xo2, jan2, en = pkg.conf_read(link, L)
    assert bytes(d2) == bytes(d) and beta == 0.875 and en == -1234
    assert bytes(pr2) == bytes(pr) and list(xo2.s) == list(xo.s) and xo2.out == xo.out
    assert np.array_equal(jan2, jan)
    # a later conf replaces the link; the earlier file stays (io.c:64-70)
    pkg.conf_write(rep, 128, _sdata(pkg, L, 128), 0.875, pr, xo, L, jan, 0)
    assert os.path.exists(name) and os.stat(link).st_ino == os.stat(
        os.path.join(rep, "conf.%016d" % 128)).st_ino


def test_conf_size_check(pkg, tmp_path):
    L = 8
    pr, xo = _gens(pkg, 1, 0)
    jan = np.zeros(L ** 3 // 8, np.uint8)
    pkg.conf_write(str(tmp_path), 2, _sdata(pkg, L, 2), 1.0, pr, xo, L, jan, 0)
    link = os.path.join(str(tmp_path), "conf")
    raw = open(link, "rb").read()
    os.remove(link)
    with open(link, "wb") as f:
        f.write(raw[:-1])
    with pytest.raises(pkg.MsaError, match="bytes in"):
        pkg.conf_read(link, L)
    with pytest.raises(pkg.MsaError, match="bytes in"):
        pkg.conf_read(link, 16)


def test_walkers_file_layout(pkg, tmp_path):
    """write_walkers (io.c:254-323): icds, NR generator pairs, demon words;
    the previous measurement's file is removed."""
    L, R, nc = 8, 2, 2
    rng = np.random.default_rng(5)
    nw = pkg.host().msa_msc_words(L, R, nc)
    w = rand_words(rng, nw)
    ic = pkg.Icds(0, 3, R, nc, L, 0, 0xFEDCBA9876543210, 0.9)
    gens = [_gens(pkg, 10 + r, r) for r in range(R)]
    sd = str(tmp_path)
    pkg.walkers_write(sd, 32, 0, ic, L, R, nc, [g[0] for g in gens], [g[1] for g in gens], w)
    pkg.walkers_write(sd, 64, 32, ic, L, R, nc, [g[0] for g in gens], [g[1] for g in gens], w)
    assert not os.path.exists(os.path.join(sd, "walkers.%016d" % 32))
    link = os.path.join(sd, "walkers")
    assert os.stat(link).st_nlink == 2
    raw = open(link, "rb").read()
    assert len(raw) == 40 + R * (2080 + 48) + 8 * nw == pkg.host().msa_walkers_bytes(L, R, nc)
    assert list(np.frombuffer(raw[:24], np.uint32)) == [0, 3, R, nc, L, 0]
    assert int(np.frombuffer(raw[24:32], np.uint64)[0]) == 0xFEDCBA9876543210
    assert np.frombuffer(raw[32:40], np.float64)[0] == 0.9
    assert np.array_equal(np.frombuffer(raw[40 + R * 2128:], np.uint64), w)
    ic2, prs, xos, w2 = pkg.walkers_read(link, L, R, nc)
    assert bytes(ic2) == bytes(ic) and np.array_equal(w2, w)
    for r in range(R):
        assert bytes(prs[r]) == bytes(gens[r][0]) and list(xos[r].s) == list(gens[r][1].s)
