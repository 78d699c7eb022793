"""changentimes.in: the run-time change of a run's end and measurement list
(stop, PROGRAM/src/io.c:930-968; read before the run, main.cu:107, and
after every measurement, main.cu:169).  The reference's reader dereferences
an uninitialised pointer; the driver defines the file as
"maxtime [t_1 t_2 ...]" (msa_read_changentimes).  CPU: the reader and the
driver's error paths before any device call.  GPU: a run shortened by the
file from its start, and a run extended by a file written while it runs,
equal the uninterrupted runs of the same length (every evol_*.txt, the last
conf and walkers files)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from test_driver import DRIVER, run_driver, write_inputs
from test_driver_restart import assert_same_outputs, sample_dir


def read_file(pkg, path):
    H = pkg.host()
    H.msa_read_changentimes.argtypes = [C.c_char_p, C.POINTER(C.c_ulonglong),
                                        C.POINTER(C.POINTER(C.c_ulonglong)), C.POINTER(C.c_int),
                                        C.c_char_p, C.c_int]
    mt, tp, n = C.c_ulonglong(7), C.POINTER(C.c_ulonglong)(), C.c_int()
    err = C.create_string_buffer(512)
    rc = H.msa_read_changentimes(path.encode(), C.byref(mt), C.byref(tp), C.byref(n), err, 512)
    times = [tp[i] for i in range(n.value)] if n.value else []
    return rc, mt.value, times, err.value.decode()


def test_reader(pkg, tmp_path):
    p = str(tmp_path / "changentimes.in")
    assert read_file(pkg, p)[0] == 0                                  # no file
    open(p, "w").write("64\n1\n2\n4\n8 16 32\n64\n")
    assert read_file(pkg, p)[:3] == (1, 64, [1, 2, 4, 8, 16, 32, 64])
    open(p, "w").write("128\n")
    assert read_file(pkg, p)[:3] == (1, 128, [])                     # the end alone
    open(p, "w").write("64\n1\n4\n2\n")
    rc, _, _, e = read_file(pkg, p)
    assert rc == -1 and "not sorted" in e
    open(p, "w").write("64\n1\nx\n")
    assert read_file(pkg, p)[0] == -1
    open(p, "w").write("\n")
    assert read_file(pkg, p)[0] == 2 and not os.path.exists(p)       # removed


def _sample_dir_file(d, text):
    sd = sample_dir(os.path.join(d, "out") + "/")
    os.makedirs(sd, exist_ok=True)
    with open(os.path.join(sd, "changentimes.in"), "w") as f:
        f.write(text)
    return sd


def test_unsorted_list_stops_the_run_before_the_device(tmp_path):
    d = str(tmp_path)
    write_inputs(d)
    sd = _sample_dir_file(d, "16\n1\n4\n2\n")
    p = run_driver(d)
    assert p.returncode == 1 and "not sorted" in p.stderr
    assert "not sorted" in open(os.path.join(sd, "error")).read()
    assert not os.path.exists(os.path.join(sd, "running"))


def test_a_file_without_a_number_is_removed(tmp_path):
    d = str(tmp_path)
    write_inputs(d)
    sd = _sample_dir_file(d, "\n")
    p = run_driver(d)
    assert "holds no number: removed" in p.stdout
    assert not os.path.exists(os.path.join(sd, "changentimes.in"))


POW = lambda k: tuple(2 ** j for j in range(k + 1))  # noqa: E731


def sched_args(d, mode):
    if mode == "reference":
        return []
    with open(os.path.join(d, "sched.in"), "w") as f:
        f.write("0\n0\n%d\n-1\n" % (1 << 17))
    return ["0", "", os.path.join(d, "sched.in")]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["reference", "philox"])
def test_shortened_from_the_start(tmp_path, mode):
    L, R = 16, 2
    dA, dB = str(tmp_path / "A"), str(tmp_path / "B")
    outA = write_inputs(dA, L=L, nr=R, maxtime=8, times=POW(3))
    p = run_driver(dA, sched_args(dA, mode))
    assert p.returncode == 0, p.stderr
    outB = write_inputs(dB, L=L, nr=R, maxtime=64, times=POW(6))
    _sample_dir_file(dB, "8\n1\n2\n4\n8\n")
    p = run_driver(dB, sched_args(dB, mode))
    assert p.returncode == 0 and "finished at iteration 8" in p.stdout, p.stderr
    assert_same_outputs(outA, outB, (0, 1), R, restarted=False)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["reference", "philox"])
def test_extended_while_running(tmp_path, mode):
    """Input maxtime 2^16; once block 16 is measured the file extends the run
    to 2^17 with times up to 2^17, so the refresh tables (counters and LUTs,
    or the walker streams) are rebuilt mid-run."""
    L, R = 32, 2
    dA, dB = str(tmp_path / "A"), str(tmp_path / "B")
    outA = write_inputs(dA, L=L, nr=R, maxtime=1 << 17, times=POW(17))
    p = run_driver(dA, sched_args(dA, mode))
    assert p.returncode == 0, p.stderr
    outB = write_inputs(dB, L=L, nr=R, maxtime=1 << 16, times=POW(16))
    args = [DRIVER, "3", "2", os.path.join(dB, "beta.in"), os.path.join(dB, "input.in"),
            os.path.join(dB, "lut.bin"), os.path.join(dB, "times.in"), "0"] + sched_args(dB, mode)
    proc = subprocess.Popen(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    lines = []
    for ln in proc.stdout:
        lines.append(ln)
        if ln.rstrip().endswith("Block = 16"):
            _sample_dir_file(dB, "%d\n%s\n" % (1 << 17, "\n".join(map(str, POW(17)))))
            break
    so, se = proc.communicate(timeout=120)
    out = "".join(lines) + so
    assert proc.returncode == 0, se
    assert "finished at iteration %d" % (1 << 17) in out, out[-2000:]
    assert "maxtime %d" % (1 << 17) in out
    assert_same_outputs(outA, outB, (0, 1), R, restarted=False)
