"""bin/msa_serve, the native request loop: argument and request-file errors
are reported before any device call (CPU).  Its results are checked on the
GPU in tests/test_gpu_serve.py."""
import os
import subprocess

from oracle_lib import REPO

SERVE = os.path.join(REPO, "micsa-ea-mc_amd", "bin", "msa_serve")


def run(args):
    return subprocess.run([SERVE] + args, capture_output=True, text=True, timeout=60)


def test_usage():
    p = run([])
    assert p.returncode == 1 and "usage" in p.stderr


def test_bad_request_lines(tmp_path):
    rq = str(tmp_path / "r.txt")
    for text, msg in (("0 16 1 64 2.0 0.7 7\n", "bad request line 1"),
                      ("# only a comment\n", "no requests"),
                      ("0 15 1 64 2.0 0.7 7 0.0\n", "L=15"),
                      ("0 16 17 64 2.0 0.7 7 0.0\n", "R=17"),
                      ("0 16 1 0 2.0 0.7 7 0.0\n", "niter=0"),
                      ("3 16 1 64 2.0 0.7 7 0.0\n3 16 1 64 2.0 0.7 8 0.1\n", "appears twice")):
        open(rq, "w").write(text)
        p = run([rq, str(tmp_path / "out.txt")])
        assert p.returncode == 1 and msg in p.stderr, (text, p.stderr)


def test_bad_options(tmp_path):
    rq = str(tmp_path / "r.txt")
    open(rq, "w").write("0 16 1 64 2.0 0.7 7 0.0\n")
    p = run([rq, str(tmp_path / "out.txt"), "--epoch"])
    assert p.returncode == 1 and "needs a value" in p.stderr
    p = run([rq, str(tmp_path / "out.txt"), "--speed", "1"])
    assert p.returncode == 1 and "unknown option" in p.stderr
    p = run([rq, str(tmp_path / "out.txt"), "--epoch", "0"])
    assert p.returncode == 1 and "bad option" in p.stderr
