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
    (dict(flag=0, L=12), "Checkpoints need L to be a multiple of 8"),
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
    from sched_util import oracle_schedule, oracle_thermal
    d = str(tmp_path)
    L, nr, maxtime = 8, 2, 16
    times = (1, 2, 4, 8, 16)
    out = write_inputs(d, L=L, nr=nr, maxtime=maxtime, times=times, nterm=nterm)
    extra = []
    if use_sched:
        with open(os.path.join(d, "sched.in"), "w") as f:
            f.write("2.0\n0.7\n%d\n2\n" % maxtime)
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
