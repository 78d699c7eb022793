"""Offline observables from conf files (host/msa_corr.c, bin/msa_xi): the
replica-overlap correlation function C4(r) and the integral correlation
lengths xi_12, xi_23, against numpy brute force and closed forms."""
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import REPO

XI = os.path.join(REPO, "micsa-ea-mc_amd", "bin", "msa_xi")


def spins(rng, L):
    return rng.integers(0, 2, size=(L, L, L), dtype=np.uint8)     # [z][y][x]


def janus(s):
    return np.packbits(s.reshape(-1), bitorder="little")


def brute_counts(sa, sb):
    q = sa ^ sb
    L = q.shape[0]
    out = np.zeros((3, L // 2 + 1), np.int64)
    for r in range(L // 2 + 1):
        out[0, r] = np.sum(q ^ np.roll(q, -r, axis=2))   # x
        out[1, r] = np.sum(q ^ np.roll(q, -r, axis=1))   # y
        out[2, r] = np.sum(q ^ np.roll(q, -r, axis=0))   # z
    return out


@pytest.mark.parametrize("L", [8, 16])
def test_counts_match_brute_force(pkg, L):
    rng = np.random.default_rng(L)
    a, b = spins(rng, L), spins(rng, L)
    # correlated replicas: b copies a on a slab, so C4 has structure
    b[: L // 2] = a[: L // 2]
    d = pkg.overlap_counts(L, janus(a), janus(b))
    assert np.array_equal(d, brute_counts(a, b))
    c4 = pkg.c4_from_counts(L, 1, d)
    assert c4[0] == 1.0
    assert np.allclose(c4, 1 - 2 * d.sum(axis=0) / (3.0 * L ** 3))


def test_identical_replicas(pkg):
    L = 16
    a = spins(np.random.default_rng(1), L)
    d = pkg.overlap_counts(L, janus(a), janus(a))
    assert not d.any()
    assert np.all(pkg.c4_from_counts(L, 1, d) == 1.0)


def test_xi_integrals(pkg):
    """Trapezoid integrals against numpy, with the cut before C4 <= 0."""
    r = np.arange(41)
    c4 = np.exp(-r / 5.0)
    x12, x23, rc = pkg.xi_integrals(c4)
    I = [np.trapezoid(r ** k * c4, r) for k in range(4)]
    assert rc == 40 and abs(x12 - I[2] / I[1]) < 1e-12 and abs(x23 - I[3] / I[2]) < 1e-12
    c4[25] = -0.01
    x12, x23, rc = pkg.xi_integrals(c4)
    I = [np.trapezoid(r[:25] ** k * c4[:25], r[:25]) for k in range(4)]
    assert rc == 24 and abs(x12 - I[2] / I[1]) < 1e-12


def test_msa_xi_on_conf_files(pkg, tmp_path):
    """The tool reads R%02d/conf of every BIT directory, pairs (0,1), (2,3)."""
    L, R = 8, 4
    rng = np.random.default_rng(7)
    total = np.zeros((3, L // 2 + 1), np.int64)
    dirs = []
    for sample in range(2):
        bd = tmp_path / ("BIT%03d" % sample)
        sp = [spins(rng, L) for _ in range(R)]
        sp[1][: L // 2] = sp[0][: L // 2]
        for r in range(R):
            rd = bd / ("R%02d" % r)
            rd.mkdir(parents=True)
            d = pkg.SData()
            d.now, d.nr, d.l, d.nc = 64, R, L, 1
            pr, xo = pkg.Prng(), pkg.Xo256()
            pkg.conf_write(str(rd), 64, d, 0.9, pr, xo, L, janus(sp[r]), 0)
        for p in range(R // 2):
            total += brute_counts(sp[2 * p], sp[2 * p + 1])
        dirs.append(str(bd))
    out = subprocess.run([XI, str(L), str(R)] + dirs, capture_output=True, text=True, check=True)
    lines = out.stdout.strip().split("\n")
    c4 = np.array([float(ln.split()[1]) for ln in lines[:-1]])
    want = 1 - 2 * total.sum(axis=0) / (3.0 * L ** 3 * 4)
    assert np.allclose(c4, want, atol=1e-9)
    tail = lines[-1].split()
    assert tail[:6] == ["#", "now", "64", "terms", "4", "rc"]
    x12, x23, rc = pkg.xi_integrals(want)
    assert abs(float(tail[tail.index("xi12") + 1]) - x12) < 1e-8
