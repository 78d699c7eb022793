#!/usr/bin/env python3
"""Regenerate the golden fixtures in tests/golden/ from the reference's own code.

Runs oracle/_ref/ref_rng_dump (the reference's PROGRAM/src/random_generator.cu
and its macros, compiled in place by `make -C oracle ref`) and
oracle/_ref/ref_prepare_lut (PROGRAM/input/LUT/preparation.c compiled in place).
Only this script touches those binaries; it needs the reference checkout and
is run by hand in the build container.  The fixtures it writes are plain data
(decimal streams and 4104-byte LUT files) and are committed.

usage: python tests/golden/gen_golden.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(REPO, "oracle", "_ref")
REF_LUT = "/root/reference/PROGRAM/input/LUT/LUT_for_PRNG_nbits10_NB1.bin"

STREAMS = [
    # name, args
    ("rng_lux.txt", ["lux", "12345", "400"]),
    ("rng_lux64.txt", ["lux64", "777", "64"]),
    ("rng_cpu_3003.txt", ["cpu", "3003", "400"]),
    ("rng_cpu_3004.txt", ["cpu", "3004", "50"]),
    ("rng_hq_2002.txt", ["hq", "2002", "400"]),
    ("rng_seeds_1001.txt", ["seeds", "1001", "64"]),
    ("rng_janus.txt", ["janus", "15548435256649372992", "400"]),
    ("rng_fill_small.txt", ["fill", "3003", "2", "0", "0", "7", "4"]),
    ("rng_fill_odd.txt", ["fill", "3003", "2", "5", "1", "16383", "3"]),
    ("rng_fill_big.txt", ["fill", "3003", "2", "123456789012", "0", "4095", "2"]),
]

BETAS = ["0.5", "0.90909090909090", "1.25", "1.4285714285714286", "0.3"]


def main():
    dump = os.path.join(REF_BIN, "ref_rng_dump")
    prep = os.path.join(REF_BIN, "ref_prepare_lut")
    if not (os.path.exists(dump) and os.path.exists(prep)):
        sys.exit("build the reference first: make -C oracle ref")
    for name, args in STREAMS:
        out = subprocess.run([dump] + args, check=True, capture_output=True, text=True).stdout
        with open(os.path.join(HERE, name), "w") as f:
            f.write("# " + " ".join(["ref_rng_dump"] + args) + "\n")
            f.write(out)
    for i, beta in enumerate(BETAS):
        with tempfile.TemporaryDirectory() as td:
            with open(os.path.join(td, "beta.in"), "w") as f:
                f.write(beta + "\n")
            subprocess.run([prep, "beta.in"], check=True, cwd=td, capture_output=True)
            shutil.copy(os.path.join(td, "LUT_for_PRNG_nbits10_NB1.bin"),
                        os.path.join(HERE, "lut_beta_%d.bin" % i))
    with open(os.path.join(HERE, "lut_betas.txt"), "w") as f:
        for i, beta in enumerate(BETAS):
            f.write("lut_beta_%d.bin %s\n" % (i, beta))
    # the reference's committed table (beta = 0.90909090909090, PROGRAM/input/BETAS/beta.in)
    shutil.copy(REF_LUT, os.path.join(HERE, "LUT_for_PRNG_nbits10_NB1.bin"))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
