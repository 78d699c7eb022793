#!/usr/bin/env python3
"""Energy relaxation and coherence length of a micSA run at fixed temperature.

The reference's own workload (PROGRAM/input/input.in: L = 160, nc = 6 demons
per site, beta = 0.90909 from input/BETAS/beta.in, demons refreshed at
iterations 0, 1, 2, 4, ... (main.cu:140-142), measurements at the times of
input/measures_times.in: powers of two to 1024, then 8 per octave), run on
the device with the 64-lane layout: 64 disorder samples x R replicas, the
Philox demon refresh at the same cadence.  At each measurement time the
per-lane integer energies give e(t) = -(3V - 2 n_unsat) / V averaged over
samples and replicas (error: over samples); at the octave times the spins of
replica pairs give C4(r) and the integral coherence length xi_12 (the
quantities of the reference's figures 2 and 3, FIGURES/fig2/fig2.gpt:129-130,
FIGURES/fig3/fig3.gpt:50-58).

Fits: e(t) = A t^-z + e_inf over t >= --fit-from (the form of fig3.gpt:52-54,
where the canonical data at T = 1.1 give e_inf = -1.70227 with A ~ 0.15,
z ~ 0.38), and xi_12(t) = a t^b.  A micSA iteration of time t matches a
canonical time k* t (FIGURES/fig3/fig3.gpt:292, k* = 5.876 at T = 1.1), so
with the canonical amplitude A_can the micSA amplitude gives
k* = (A_can / A)^(1/z).  This is a statistical check of the dynamics, not a
parity test.  Prints one JSON object.
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_package():
    if "micsa_ea_mc_amd" in sys.modules:
        return sys.modules["micsa_ea_mc_amd"]
    spec = importlib.util.spec_from_file_location(
        "micsa_ea_mc_amd", os.path.join(HERE, "__init__.py"), submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["micsa_ea_mc_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def measure_times(tmax):
    """Powers of two up to 1024, then 8 logarithmically spaced times per octave
    (the pattern of PROGRAM/input/measures_times.in), up to tmax."""
    out = [1 << k for k in range(11)]
    t = 1024.0
    while True:
        t *= 2 ** (1 / 8)
        n = int(round(t))
        if n > tmax:
            break
        if n > out[-1]:
            out.append(n)
    return [t for t in out if t <= tmax]


def fit_power_offset(t, y, err, p0):
    from scipy.optimize import curve_fit

    def f(x, A, z, c):
        return A * x ** (-z) + c
    p, cov = curve_fit(f, t, y, p0=p0, sigma=err, absolute_sigma=True, maxfev=20000)
    return p, np.sqrt(np.diag(cov))


def fit_power(t, y, err):
    from scipy.optimize import curve_fit

    def f(x, a, b):
        return a * x ** b
    p, cov = curve_fit(f, t, y, p0=(1.0, 0.2), sigma=err, absolute_sigma=True, maxfev=20000)
    return p, np.sqrt(np.diag(cov))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=160)
    ap.add_argument("--R", type=int, default=2)
    ap.add_argument("--nc", type=int, default=6)
    ap.add_argument("--beta", type=float, default=0.90909090909090)
    ap.add_argument("--iters", type=int, default=1 << 19)
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--xi-samples", type=int, default=16,
                    help="samples whose replica overlaps give C4 and xi_12")
    ap.add_argument("--fit-from", type=float, default=1e3)
    ap.add_argument("--A-can", type=float, default=0.15, help="canonical amplitude at this T")
    args = ap.parse_args()
    pkg = load_package()
    msa = pkg.msa
    L, R, nc = args.L, args.R, args.nc
    V = L ** 3
    lat = pkg.Lattice(L, R, nc)
    lat.init_philox(args.seed)
    tuned = lat.autotune(8)
    sched = msa.schedule(0.0, 0.0, args.iters, -1, beta_fixed=args.beta)
    ctr, keys, luts, idx = msa.schedule_tables(sched, args.iters + 1, R,
                                               [3003 + r for r in range(R)], 0)
    lat.set_luts(luts)
    lat.set_refresh(ctr, keys, idx)
    times = measure_times(args.iters)
    octaves = {1 << k for k in range(40)}
    rows, xis = [], []
    t, w0 = 0, time.perf_counter()
    for tm in times:
        lat.run(t, tm - t, -1)
        t = tm
        m = lat.measure()
        e = -(m["energy"].astype(np.float64)) / V          # [R][64]
        per_sample = e.mean(axis=0)
        ec = m["demon"].astype(np.float64).mean(axis=0) / (V * nc)
        rows.append((tm, float(per_sample.mean()), float(per_sample.std(ddof=1) / 8.0),
                     float(ec.mean())))
        if tm in octaves and tm >= 8 and R >= 2:
            S = lat.download()[0]
            ns = args.xi_samples
            jan = msa.janus_from_lanes(L, S, list(range(ns)))
            diff = np.zeros((3, L // 2 + 1), np.int64)
            for i in range(ns):
                diff += msa.overlap_counts(L, jan[i, 0], jan[i, 1])
            c4 = msa.c4_from_counts(L, ns, diff)
            xi12, xi23, rc = msa.xi_integrals(c4)
            xis.append((tm, float(xi12), float(xi23)))
        print("t=%d e=%.6f +- %.6f (%.0f s)" % (tm, rows[-1][1], rows[-1][2],
                                               time.perf_counter() - w0), file=sys.stderr, flush=True)
    lat.close()
    wall = time.perf_counter() - w0
    tt = np.array([r[0] for r in rows], float)
    ee = np.array([r[1] for r in rows])
    er = np.array([r[2] for r in rows])
    sel = tt >= args.fit_from
    (A, z, einf), (dA, dz, de) = fit_power_offset(tt[sel], ee[sel], er[sel], (0.1, 0.38, -1.70))
    res = {"workload": "micSA L=%d R=%d nc=%d at beta=%.14g, 64 samples, demon refresh at "
                       "t = 0 and powers of two, %d iterations" % (L, R, nc, args.beta, args.iters),
           "tiling": list(tuned[:3]), "wall_s": round(wall, 1),
           "energy": [[int(r[0]), round(r[1], 7), round(r[2], 7), round(r[3], 7)] for r in rows],
           "fit_energy": {"range": [args.fit_from, args.iters], "A": A, "A_err": dA, "z": z,
                          "z_err": dz, "e_inf": einf, "e_inf_err": de,
                          "reference_e_inf_T1.1": -1.70227, "reference_z_T1.1": 0.38,
                          "k_star": (args.A_can / A) ** (1.0 / z) if A > 0 else None,
                          "reference_k_star_T1.1": 5.876}}
    if len(xis) >= 3:
        xt = np.array([x[0] for x in xis], float)
        xv = np.array([x[1] for x in xis])
        s2 = xt >= args.fit_from
        if s2.sum() >= 3:
            (a, b), (da, db) = fit_power(xt[s2], xv[s2], 0.02 * xv[s2])
            res["fit_xi12"] = {"a": a, "a_err": da, "b": b, "b_err": db,
                               "note": "xi_12 = a t^b, 2% weights"}
        res["xi"] = [[int(x[0]), round(x[1], 4), round(x[2], 4)] for x in xis]
    print(json.dumps(res))


if __name__ == "__main__":
    main()
