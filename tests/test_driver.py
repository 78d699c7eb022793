"""The C driver bin/msa_anneal: the reference's command line and files in,
evol_*.txt out.  Error paths run on the CPU (they fail before any device
call); the end-to-end run is a GPU test checked against the oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle_lib import REPO, Lattice as OLattice

DRIVER = os.path.join(REPO, "micsa-ea-mc_amd", "bin", "msa_anneal")
GOLD_LUT = os.path.join(REPO, "tests", "golden", "LUT_for_PRNG_nbits10_NB1.bin")


def write_inputs(d, L=8, nr=2, nc=1, maxtime=16, times=(1, 2, 4, 8, 16), nterm=0, flag=1,
                 seeds=(1001, 2002, 3003, 3004), beta="0.90909090909090"):
    os.makedirs(d, exist_ok=True)
    out = os.path.join(d, "out") + "/"
    os.makedirs(out, exist_ok=True)
    lines = [out, str(maxtime), "0", str(len(times)), str(nterm), str(nc), str(nr), str(L),
             str(flag), str(seeds[0]), str(seeds[1])] + [str(seeds[2] + r) for r in range(nr)]
    with open(os.path.join(d, "input.in"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(d, "beta.in"), "w") as f:
        f.write(beta + "\n")
    with open(os.path.join(d, "times.in"), "w") as f:
        f.write("\n".join(str(t) for t in times) + "\n")
    shutil.copy(GOLD_LUT, os.path.join(d, "lut.bin"))
    return out


def run_driver(d, extra=(), sample=3, nbits=2):
    args = [DRIVER, str(sample), str(nbits), os.path.join(d, "beta.in"), os.path.join(d, "input.in"),
            os.path.join(d, "lut.bin"), os.path.join(d, "times.in"), "0"] + list(extra)
    return subprocess.run(args, capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("change,msg", [
    (dict(nterm=5, L=12), "multiple of 8"),
    (dict(nterm=-1), "Bad data input"),
    (dict(flag=0), "impossible to do a backup"),
    (dict(flag=-1), "not recognize"),
    (dict(flag=0, L=12), "multiple of 8"),
    (dict(beta="0.5"), "betas[0]"),
    (dict(nc=7), "nc=7"),
])
def test_driver_rejects_before_touching_the_gpu(tmp_path, change, msg):
    d = str(tmp_path)
    write_inputs(d, **change)
    p = run_driver(d)
    assert p.returncode != 0
    assert msg in p.stderr


def test_driver_usage(tmp_path):
    p = subprocess.run([DRIVER], capture_output=True, text=True)
    assert p.returncode != 0 and "Usage" in p.stderr
    p = subprocess.run([DRIVER, "-h"], capture_output=True, text=True)
    assert p.returncode == 0 and "Usage" in p.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("use_sched,nterm", [(False, 0), (True, 0), (False, 3)])
def test_driver_end_to_end_matches_oracle(tmp_path, orc, pkg, use_sched, nterm):
    """The Philox-refresh runs (a schedule file): an annealing schedule, or
    fixed beta (T0 = T1 = 0) with and without thermalization."""
    from sched_util import oracle_schedule, oracle_thermal
    d = str(tmp_path)
    L, nr, maxtime = 8, 2, 16
    times = (1, 2, 4, 8, 16)
    out = write_inputs(d, L=L, nr=nr, maxtime=maxtime, times=times, nterm=nterm)
    with open(os.path.join(d, "sched.in"), "w") as f:
        f.write(("2.0\n0.7\n%d\n2\n" if use_sched else "0\n0\n%d\n-1\n") % maxtime)
    extra = ["0", "", os.path.join(d, "sched.in")]
    p = run_driver(d, extra, sample=3, nbits=2)
    assert p.returncode == 0, p.stderr
    # oracle side: same seeds, same counters, same tables
    o = OLattice(L, nr, 1)
    orc.init_ref(o, 1001, 2002)
    o.D0[...] = 0
    o.D1[...] = 0
    gens, tctr = oracle_thermal(orc, nr, nterm, seed_mc=3003, jobid=3)
    if nterm:
        # thermalization at the LUT file's beta (main.cu:124-128), then the micro phase
        lut = orc.lut_beta(0.90909090909090)
        for t in range(nterm):
            orc.thermal_iteration(o, lut, tctr[t, 0], tctr[t, 1], [g[2] for g in gens])
    H = pkg.host()
    sch = pkg.schedule(2.0, 0.7, maxtime, 2) if use_sched else pkg.schedule(0, 0, maxtime, -1, 0.90909090909090)
    import ctypes as C
    nref = H.msa_sched_count(C.byref(sch), maxtime)
    rt = [H.msa_sched_time(C.byref(sch), k) for k in range(nref)]
    rb = [H.msa_sched_beta(C.byref(sch), k) for k in range(nref)]
    ctr, keys, _, luts = oracle_schedule(orc, nr, rt, rb, seed_mc=3003, jobid=3, gens=gens)
    want = {}
    for t in range(maxtime):
        if t in rt:
            k = rt.index(t)
            orc.iteration(o, ctr[k], keys, luts[k])
        else:
            orc.iteration(o)
        if t + 1 in times:
            want[t + 1] = (orc.energy(o), orc.demon_energy(o))
    for lane in (0, 1):
        rows = open(os.path.join(out, "I003", "BIT%03d" % lane, "evol_E_int.txt")).read().split("\n")
        rows = [r.split() for r in rows if r]
        assert [int(r[0]) for r in rows] == list(times)
        for r in rows:
            e, dem = want[int(r[0])]
            got = [int(x) for x in r[1:]]
            assert got == [v for rr in range(nr) for v in (int(e[rr, lane]), int(dem[rr, lane]))]
        es = open(os.path.join(out, "I003", "BIT%03d" % lane, "evol_E_spins.txt")).read().split("\n")
        first = es[0].split()
        e1 = want[1][0][:, lane] / float(L ** 3)
        assert abs(float(first[1]) - float(-e1.mean())) < 1e-8
    assert "finished at iteration 16" in p.stdout


def reference_workload_oracle(orc, pkg, L, R, nterm, maxtime, times, lanes=(0, 1), seed_u=2002,
                              seed_mc=3003, jobid=3, beta=0.90909090909090):
    """The reference program's own run on the oracle: J of the JANUS seed for
    every sample (io.c:797-799), spins from Init_u with two slots
    (ini.c:275-299), thermalization, demons from InitWalkers at iterations
    0, 1, 2, 4, ... (main.cu:140-142), micro sweeps.  Returns the energies at
    the measurement times, the final lattice, the thermal and walker
    generators."""
    from sched_util import oracle_thermal
    from test_walkers import oracle_walker_demons
    V = L ** 3
    o = OLattice(L, R, 1)
    js = np.array(orc.janus_stream(9164166177945067407, 3 * V), np.uint64) >> np.uint64(63)
    js = js.reshape(V, 3)
    for d in range(3):
        o.J[d] = np.where(js[:, d] == 1, np.uint64(2 ** 64 - 1), np.uint64(0))
    hq = (np.array(orc.hq_stream(seed_u, R * V * 2), np.uint64) >> np.uint64(63)).reshape(R, V, 2)
    o.S[...] = hq[:, :, 0] | (hq[:, :, 1] << np.uint64(1))
    o.D0[...] = 0
    o.D1[...] = 0
    gens, tctr = oracle_thermal(orc, R, nterm, seed_mc=seed_mc, jobid=jobid)
    if nterm:
        lut = orc.lut_beta(beta)
        for t in range(nterm):
            orc.thermal_iteration(o, lut, tctr[t, 0], tctr[t, 1], [g[2] for g in gens])
    seeds = orc.seeds_from_one(seed_u, len(lanes) * R)
    wgens = [orc.cpu_generators(seeds[k])[:2] for k in range(len(lanes) * R)]
    thr = np.array(orc.thresholds(beta), np.uint64)
    want = {}
    for t in range(maxtime):
        if t == 0 or (t & (t - 1)) == 0:
            o.D0[...], o.D1[...] = oracle_walker_demons(pkg, orc, L, R, 1, list(lanes), wgens, thr)
        orc.iteration(o)
        if t + 1 in times:
            want[t + 1] = (orc.energy(o), orc.demon_energy(o))
    return want, o, gens, wgens


@pytest.mark.gpu
@pytest.mark.parametrize("nterm", [0, 3])
def test_reference_workload_matches_oracle(tmp_path, orc, pkg, nterm):
    """No schedule file: the reference program's run bit for bit -- energies,
    and the conf and walkers files' spins, demons and generator states."""
    d = str(tmp_path)
    L, R, maxtime = 8, 2, 16
    times = (1, 2, 4, 8, 16)
    out = write_inputs(d, L=L, nr=R, maxtime=maxtime, times=times, nterm=nterm)
    p = run_driver(d, sample=3, nbits=2)
    assert p.returncode == 0, p.stderr
    want, o, gens, wgens = reference_workload_oracle(orc, pkg, L, R, nterm, maxtime, times)
    sd = os.path.join(out, "I003")
    for i, lane in enumerate((0, 1)):
        bd = os.path.join(sd, "BIT%03d" % lane)
        rows = [r.split() for r in open(os.path.join(bd, "evol_E_int.txt")).read().split("\n") if r]
        assert [int(r[0]) for r in rows] == list(times)
        for r in rows:
            e, dem = want[int(r[0])]
            assert [int(x) for x in r[1:]] == [v for rr in range(R)
                                               for v in (int(e[rr, lane]), int(dem[rr, lane]))]
        # walkers file: the reference's layout, the streams after 5 refreshes
        ic, prs, xos, w = pkg.walkers_read(os.path.join(bd, "walkers"), L, R, 1)
        for r in range(R):
            opr, oxo = wgens[i * R + r]
            assert list(prs[r].ring) == list(opr.ira)
            assert (prs[r].p, prs[r].p24, prs[r].p55, prs[r].p61) == (opr.ip, opr.ip1, opr.ip2, opr.ip3)
            assert (prs[r].lcg, prs[r].out) == (opr.zseed, opr.out)
            assert list(xos[r].s) == list(oxo.s) and xos[r].out == oxo.out
        D0 = np.zeros((R, 1, L ** 3), np.uint64)
        D1 = np.zeros((R, 1, L ** 3), np.uint64)
        pkg.msc_to_lanes(L, R, 1, w[None, :], [lane], D0, D1)
        bit = np.uint64(1 << lane)
        assert np.array_equal(D0 & bit, o.D0 & bit) and np.array_equal(D1 & bit, o.D1 & bit)
        # conf files: thermal generators after the thermalization, the spins
        for r in range(R):
            sdata, beta, pr, xo, jan, en = pkg.conf_read(os.path.join(bd, "R%02d" % r, "conf"), L)
            assert sdata.now == 16 and sdata.seed_J == 9164166177945067407
            assert list(pr.ring) == list(gens[r][0].ira) and pr.lcg == gens[r][0].zseed
            assert list(xo.s) == list(gens[r][1].s)
            bits = np.unpackbits(jan, bitorder="little")[:L ** 3]
            assert np.array_equal(bits, ((o.S[r] >> np.uint64(lane)) & np.uint64(1)).astype(np.uint8))
    assert "finished at iteration 16" in p.stdout


@pytest.mark.parametrize("nbits,samples,msg", [(3, None, "nbits out of range"),
                                               (2, "0 2", "samples must be 0 and/or 1")])
def test_reference_workload_limits(tmp_path, nbits, samples, msg):
    d = str(tmp_path)
    write_inputs(d)
    extra = []
    if samples:
        with open(os.path.join(d, "samples.in"), "w") as f:
            f.write(samples.replace(" ", "\n") + "\n")
        extra = ["0", os.path.join(d, "samples.in")]
    p = run_driver(d, extra, nbits=nbits)
    assert p.returncode == 1 and msg in p.stderr
