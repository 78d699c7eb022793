"""ctypes bindings of the micSA MI355X libraries.

libmsa_hip.so  (include/msa.h)       device layer: HIP kernels behind a C ABI
libmsa_host.so (include/msa_host.h)  host C library: generators, tables, schedules

Nothing here falls back to a CPU path: if a library is missing, loading raises.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)


class MsaError(RuntimeError):
    pass


class Geom(C.Structure):
    _fields_ = [("L", C.c_int32), ("R", C.c_int32), ("nc", C.c_int32)]


LUT_BYTES = 4096        # one msa_lut / s_lut_heat_bath


class Job(C.Structure):
    _fields_ = [("geom", Geom), ("seed", C.c_uint64), ("niter", C.c_int64),
                ("refresh_every", C.c_int64), ("nref", C.c_int32), ("ctr", u32p),
                ("key", u32p), ("lut_idx", i32p)]


def _p(a, t=u64p):
    return None if a is None else a.ctypes.data_as(t)


def _load(name):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise MsaError("%s is not built (run __graft_entry__.build() or make -C micsa-ea-mc_amd)"
                       % path)
    return C.CDLL(path)


_hip = None


def hip():
    """The device library; raises if it is not built."""
    global _hip
    if _hip is None:
        L = _load("libmsa_hip.so")
        vp = C.c_void_p
        L.msa_create.argtypes = [C.c_int, C.POINTER(Geom), C.POINTER(vp)]
        L.msa_destroy.argtypes = [vp]
        L.msa_upload_state.argtypes = [vp, u64p, u64p, u64p, u64p]
        L.msa_download_state.argtypes = [vp, u64p, u64p, u64p]
        L.msa_download_couplings.argtypes = [vp, u64p]
        L.msa_init_philox.argtypes = [vp, C.c_uint64]
        L.msa_set_luts.argtypes = [vp, C.c_int32, C.c_void_p]
        L.msa_set_refresh.argtypes = [vp, C.c_int32, u32p, u32p, i32p]
        L.msa_run.argtypes = [vp, C.c_uint64, C.c_int64, C.c_int64]
        L.msa_run_timed.argtypes = [vp, C.c_uint64, C.c_int64, C.c_int64, C.POINTER(C.c_float)]
        L.msa_measure.argtypes = [vp, i32p, i32p, i32p, i32p]
        L.msa_sync.argtypes = [vp]
        L.msa_set_tiling.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
        L.msa_get_tiling.argtypes = [vp, i32p, i32p, i32p]
        L.msa_run_thermal.argtypes = [vp, C.c_int64, C.c_void_p, u32p, u32p]
        bp = C.c_void_p
        L.msa_batch_create.argtypes = [C.c_int, C.c_int64, C.POINTER(bp)]
        L.msa_batch_destroy.argtypes = [bp]
        L.msa_batch_set_luts.argtypes = [bp, C.c_int32, C.c_void_p]
        L.msa_batch_add.argtypes = [bp, C.POINTER(Job), i32p]
        L.msa_batch_job_bytes.argtypes = [C.POINTER(Job)]
        L.msa_batch_job_bytes.restype = C.c_int64
        L.msa_batch_bytes_in_use.argtypes = [bp]
        L.msa_batch_bytes_in_use.restype = C.c_int64
        L.msa_batch_upload_state.argtypes = [bp, C.c_int32, u64p, u64p, u64p, u64p]
        L.msa_batch_set_shape.argtypes = [bp, C.c_int32, C.c_int32]
        L.msa_batch_run.argtypes = [bp, C.c_int64]
        L.msa_batch_sync.argtypes = [bp]
        L.msa_batch_progress.argtypes = [bp, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.msa_batch_measure.argtypes = [bp, C.c_int32, i32p, i32p, i32p, i32p]
        L.msa_batch_download_state.argtypes = [bp, C.c_int32, u64p, u64p, u64p]
        L.msa_batch_remove.argtypes = [bp, C.c_int32]
        L.msa_copy_probe.argtypes = [C.c_int, C.c_int64, C.c_int32, C.POINTER(C.c_float)]
        L.msa_autotune.argtypes = [vp, C.c_int32, i32p, i32p, i32p, C.POINTER(C.c_float)]
        L.msa_footprint_bytes.argtypes = [C.POINTER(Geom)]
        L.msa_footprint_bytes.restype = C.c_int64
        L.msa_strerror.argtypes = [C.c_int]
        L.msa_strerror.restype = C.c_char_p
        _hip = L
    return _hip


def _check(rc):
    if rc != 0:
        raise MsaError("msa error %d: %s" % (rc, hip().msa_strerror(rc).decode()))


class Lattice:
    """One device lattice: L^3 sites x 64 sample lanes x R replicas, nc demon layers."""

    def __init__(self, L, R=1, nc=1, device=0):
        self.L, self.R, self.nc = L, R, nc
        self.V = L ** 3
        self.lib = hip()
        self.ctx = C.c_void_p()
        g = Geom(L, R, nc)
        _check(self.lib.msa_create(device, C.byref(g), C.byref(self.ctx)))

    def close(self):
        if self.ctx:
            self.lib.msa_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, J, S, D0=None, D1=None):
        J = np.ascontiguousarray(J, np.uint64)
        S = np.ascontiguousarray(S, np.uint64)
        assert J.size == 3 * self.V and S.size == self.R * self.V
        if D0 is not None:
            D0 = np.ascontiguousarray(D0, np.uint64)
            D1 = np.ascontiguousarray(D1, np.uint64)
            assert D0.size == D1.size == self.R * self.nc * self.V
        _check(self.lib.msa_upload_state(self.ctx, _p(J), _p(S), _p(D0), _p(D1)))

    def init_philox(self, seed):
        _check(self.lib.msa_init_philox(self.ctx, seed))

    def download(self):
        S = np.zeros((self.R, self.V), np.uint64)
        D0 = np.zeros((self.R, self.nc, self.V), np.uint64)
        D1 = np.zeros_like(D0)
        _check(self.lib.msa_download_state(self.ctx, _p(S), _p(D0), _p(D1)))
        return S, D0, D1

    def couplings(self):
        J = np.zeros((3, self.V), np.uint64)
        _check(self.lib.msa_download_couplings(self.ctx, _p(J)))
        return J

    def set_luts(self, luts):
        """luts: bytes / uint8 array of 2*n tables of 4096 bytes (plane 0, plane 1)."""
        raw = np.frombuffer(bytes(luts), np.uint8) if not isinstance(luts, np.ndarray) else luts
        raw = np.ascontiguousarray(raw, np.uint8)
        assert raw.size % (2 * LUT_BYTES) == 0
        self._luts = raw
        _check(self.lib.msa_set_luts(self.ctx, raw.size // (2 * LUT_BYTES),
                                     raw.ctypes.data_as(C.c_void_p)))

    def set_refresh(self, ctr, key, lut_idx):
        ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, self.R, 4)
        key = np.ascontiguousarray(key, np.uint32)
        lut_idx = np.ascontiguousarray(lut_idx, np.int32)
        assert key.size == self.R and lut_idx.size == ctr.shape[0]
        _check(self.lib.msa_set_refresh(self.ctx, ctr.shape[0], _p(ctr, u32p), _p(key, u32p),
                                        _p(lut_idx, i32p)))

    def run(self, t0, niter, refresh_every):
        _check(self.lib.msa_run(self.ctx, t0, niter, refresh_every))

    def run_thermal(self, lut, ctr, key):
        """Canonical Metropolis thermalization (msa_run_thermal): `lut` is the
        4096 bytes of the heat-bath table, ctr (nterm, 2, R, 4), key (R,)."""
        raw = np.ascontiguousarray(np.frombuffer(bytes(lut), np.uint8))
        assert raw.size == LUT_BYTES
        ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 2, self.R, 4)
        key = np.ascontiguousarray(key, np.uint32)
        _check(self.lib.msa_run_thermal(self.ctx, ctr.shape[0], raw.ctypes.data_as(C.c_void_p),
                                        _p(ctr, u32p), _p(key, u32p)))

    def run_timed(self, t0, niter, refresh_every):
        ms = C.c_float()
        _check(self.lib.msa_run_timed(self.ctx, t0, niter, refresh_every, C.byref(ms)))
        return ms.value

    def set_tiling(self, mode, ty=0, tz=0):
        _check(self.lib.msa_set_tiling(self.ctx, mode, ty, tz))

    def autotune(self, sweeps=8):
        m, ty, tz, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_float()
        _check(self.lib.msa_autotune(self.ctx, sweeps, C.byref(m), C.byref(ty), C.byref(tz),
                                     C.byref(ms)))
        return m.value, ty.value, tz.value, ms.value

    def tiling(self):
        m, ty, tz = C.c_int32(), C.c_int32(), C.c_int32()
        _check(self.lib.msa_get_tiling(self.ctx, C.byref(m), C.byref(ty), C.byref(tz)))
        return m.value, ty.value, tz.value

    def sync(self):
        _check(self.lib.msa_sync(self.ctx))

    def measure(self):
        """Per-lane integer observables: dict of int32 arrays."""
        npair = self.R // 2
        unsat = np.zeros((self.R, 64), np.int32)
        demon = np.zeros((self.R, 64), np.int32)
        q = np.zeros((max(npair, 1), 64), np.int32)
        ql = np.zeros((max(npair, 1), 64), np.int32)
        _check(self.lib.msa_measure(self.ctx, _p(unsat, i32p), _p(demon, i32p),
                                    _p(q, i32p) if npair else None,
                                    _p(ql, i32p) if npair else None))
        out = {"unsat": unsat, "energy": 3 * self.V - 2 * unsat, "demon": demon}
        if npair:
            out["q"] = q[:npair]
            out["qlink"] = ql[:npair]
        return out


def copy_probe(nbytes, iters=20, device=0):
    """ms per streaming copy of nbytes (reads nbytes, writes nbytes)."""
    ms = C.c_float()
    _check(hip().msa_copy_probe(device, nbytes, iters, C.byref(ms)))
    return ms.value


def footprint_bytes(L, R, nc):
    return hip().msa_footprint_bytes(C.byref(Geom(L, R, nc)))


# ---------------------------------------------------------------- host lib --
class Prng(C.Structure):
    _fields_ = [("ring", C.c_uint64 * 256), ("lcg", C.c_uint64), ("out", C.c_uint64),
                ("p", C.c_uint8), ("p24", C.c_uint8), ("p55", C.c_uint8), ("p61", C.c_uint8)]


class Xo256(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4), ("out", C.c_uint64)]


class Sched(C.Structure):
    _fields_ = [("T0", C.c_double), ("T1", C.c_double), ("nsweeps", C.c_int64),
                ("refresh_every", C.c_int64), ("beta_fixed", C.c_double)]


_host = None


def host():
    """The host C library (generators, tables, schedules); raises if not built."""
    global _host
    if _host is None:
        L = _load("libmsa_host.so")
        L.msa_lut_beta.argtypes = [C.c_double, C.c_void_p]
        L.msa_lut_demon.argtypes = [C.c_double, C.c_void_p, C.c_void_p]
        L.msa_thresholds.argtypes = [C.c_double, u64p]
        L.msa_seeds_from_one.argtypes = [C.c_uint64, u64p, C.c_int]
        L.msa_cpu_generators.argtypes = [C.POINTER(Prng), C.POINTER(Xo256), u32p, C.c_uint64]
        L.msa_prng_seed_hq.argtypes = [C.POINTER(Prng), C.c_uint64]
        L.msa_prng_next.argtypes = [C.POINTER(Prng)]
        L.msa_prng_next.restype = C.c_uint64
        L.msa_janus_seed.argtypes = [C.POINTER(Prng), C.c_uint64]
        L.msa_janus_next.argtypes = [C.POINTER(Prng)]
        L.msa_janus_next.restype = C.c_uint64
        L.msa_fill_counter.argtypes = [C.POINTER(Prng), C.POINTER(Xo256), C.c_uint64, C.c_uint32,
                                       C.c_uint32, u32p]
        L.msa_build_couplings.argtypes = [C.c_int, C.c_uint64, u64p]
        L.msa_build_spins.argtypes = [C.c_int, C.c_int, C.c_uint64, u64p]
        L.msa_sched_count.argtypes = [C.POINTER(Sched), C.c_int64]
        L.msa_sched_count.restype = C.c_int64
        L.msa_sched_time.argtypes = [C.POINTER(Sched), C.c_int64]
        L.msa_sched_time.restype = C.c_int64
        L.msa_sched_beta.argtypes = [C.POINTER(Sched), C.c_int64]
        L.msa_sched_beta.restype = C.c_double
        L.msa_sched_tables.argtypes = [C.POINTER(Sched), C.c_int64, C.c_int, u64p, C.c_uint32,
                                       u32p, u32p, C.c_void_p, i32p]
        L.msa_lut_grid.argtypes = [C.c_double, C.c_int32, C.c_int32, C.c_void_p]
        L.msa_thermal_counters.argtypes = [C.POINTER(Prng), C.POINTER(Xo256), C.c_int, C.c_int64,
                                           C.c_uint32, u32p]
        L.msa_sched_tables_from.argtypes = [C.POINTER(Sched), C.c_int64, C.c_int, C.POINTER(Prng),
                                            C.POINTER(Xo256), C.c_uint32, u32p, C.c_void_p, i32p]
        L.msa_sched_tables_grid.argtypes = [C.POINTER(Sched), C.c_int64, C.c_int, u64p, C.c_uint32,
                                            C.c_double, C.c_int32, C.c_int32, u32p, u32p, i32p]
        _host = L
    return _host


def lut_grid(tgrid, nlo, nlut):
    """Shared LUT pairs for temperatures tgrid*(nlo .. nlo+nlut-1), as bytes."""
    out = np.zeros(2 * nlut * LUT_BYTES, np.uint8)
    host().msa_lut_grid(tgrid, nlo, nlut, out.ctypes.data_as(C.c_void_p))
    return out


def schedule_tables_grid(sched, niter, R, seed_mc, jobid, tgrid, nlo, nlut):
    """Counters (nref,R,4), keys (R,), indices into the shared grid table (nref,)."""
    H = host()
    nref = H.msa_sched_count(C.byref(sched), niter)
    ctr = np.zeros((max(nref, 1), R, 4), np.uint32)
    keys = np.zeros(R, np.uint32)
    idx = np.zeros(max(nref, 1), np.int32)
    seeds = np.ascontiguousarray(seed_mc, np.uint64)
    rc = H.msa_sched_tables_grid(C.byref(sched), nref, R, _p(seeds), jobid, tgrid, nlo, nlut,
                                 _p(ctr, u32p), _p(keys, u32p), _p(idx, i32p))
    if rc:
        raise MsaError("grid schedule tables failed")
    return ctr, keys, idx


def lut_demon(beta):
    """bytes of the plane-0 and plane-1 heat-bath tables for the demon refresh."""
    a = np.zeros(LUT_BYTES, np.uint8)
    b = np.zeros(LUT_BYTES, np.uint8)
    host().msa_lut_demon(beta, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p))
    return a.tobytes(), b.tobytes()


def lut_beta(beta):
    a = np.zeros(LUT_BYTES, np.uint8)
    host().msa_lut_beta(beta, a.ctypes.data_as(C.c_void_p))
    return a.tobytes()


def build_state(L, R, seed_J, seed_u):
    """Reference-mode couplings [3][V] and spins [R][V]."""
    V = L ** 3
    J = np.zeros((3, V), np.uint64)
    S = np.zeros((R, V), np.uint64)
    host().msa_build_couplings(L, seed_J, _p(J))
    host().msa_build_spins(L, R, seed_u, _p(S))
    return J, S


def schedule(T0, T1, nsweeps, refresh_every, beta_fixed=0.0):
    return Sched(T0, T1, nsweeps, refresh_every, beta_fixed)


def generators(seed_mc):
    """Per-replica (Prng, Xo256) states and keys (Inicia_generadores_CPU)."""
    R = len(seed_mc)
    pr, xo = (Prng * R)(), (Xo256 * R)()
    keys = np.zeros(R, np.uint32)
    for r, sd in enumerate(seed_mc):
        host().msa_cpu_generators(C.byref(pr[r]), C.byref(xo[r]), _p(keys[r:r + 1], u32p), sd)
    return pr, xo, keys


def thermal_counters(gens, nterm, jobid=0):
    """Counters (nterm, 2, R, 4) of the thermal phase; advances `gens`."""
    pr, xo, keys = gens
    R = len(keys)
    ctr = np.zeros((max(nterm, 1), 2, R, 4), np.uint32)
    host().msa_thermal_counters(pr, xo, R, nterm, jobid, _p(ctr, u32p))
    return ctr[:nterm]


def schedule_tables_from(sched, niter, gens, jobid=0):
    """schedule_tables continuing from generator states (after a thermal phase)."""
    pr, xo, keys = gens
    R = len(keys)
    H = host()
    nref = H.msa_sched_count(C.byref(sched), niter)
    ctr = np.zeros((nref, R, 4), np.uint32)
    luts = np.zeros(2 * nref * LUT_BYTES, np.uint8)
    idx = np.zeros(nref, np.int32)
    rc = H.msa_sched_tables_from(C.byref(sched), nref, R, pr, xo, jobid, _p(ctr, u32p),
                                 luts.ctypes.data_as(C.c_void_p), _p(idx, i32p))
    if rc:
        raise MsaError("schedule tables failed")
    return ctr, keys.copy(), luts, idx


def schedule_tables(sched, niter, R, seed_mc, jobid=0):
    """Counters (nref,R,4) u32, keys (R,) u32, LUT bytes (2*nref tables), lut_idx (nref,)."""
    H = host()
    nref = H.msa_sched_count(C.byref(sched), niter)
    ctr = np.zeros((nref, R, 4), np.uint32)
    keys = np.zeros(R, np.uint32)
    luts = np.zeros(2 * nref * LUT_BYTES, np.uint8)
    idx = np.zeros(nref, np.int32)
    seeds = np.ascontiguousarray(seed_mc, np.uint64)
    rc = H.msa_sched_tables(C.byref(sched), nref, R, _p(seeds), jobid, _p(ctr, u32p),
                            _p(keys, u32p), luts.ctypes.data_as(C.c_void_p), _p(idx, i32p))
    if rc:
        raise MsaError("schedule tables failed")
    return ctr, keys, luts, idx


class Batch:
    """Many lattices advanced together (msa_batch_* of include/msa.h)."""

    def __init__(self, luts, max_bytes=0, device=0):
        self.lib = hip()
        self.b = C.c_void_p()
        _check(self.lib.msa_batch_create(device, max_bytes, C.byref(self.b)))
        raw = np.frombuffer(bytes(luts), np.uint8) if not isinstance(luts, np.ndarray) else luts
        self._luts = np.ascontiguousarray(raw, np.uint8)
        _check(self.lib.msa_batch_set_luts(self.b, self._luts.size // (2 * LUT_BYTES),
                                           self._luts.ctypes.data_as(C.c_void_p)))
        self.geom = {}

    def close(self):
        if self.b:
            self.lib.msa_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def make_job(L, R, seed, niter, refresh_every, ctr, key, lut_idx):
        ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, R, 4)
        key = np.ascontiguousarray(key, np.uint32)
        lut_idx = np.ascontiguousarray(lut_idx, np.int32)
        j = Job(Geom(L, R, 1), seed, niter, refresh_every, ctr.shape[0], _p(ctr, u32p),
                _p(key, u32p), _p(lut_idx, i32p))
        j._keep = (ctr, key, lut_idx)
        return j

    def job_bytes(self, job):
        return self.lib.msa_batch_job_bytes(C.byref(job))

    def add(self, job):
        """Slot of the admitted lattice, or None when it does not fit (admission by footprint)."""
        slot = C.c_int32()
        rc = self.lib.msa_batch_add(self.b, C.byref(job), C.byref(slot))
        if rc == -3:
            return None
        _check(rc)
        self.geom[slot.value] = (job.geom.L, job.geom.R)
        return slot.value

    def bytes_in_use(self):
        return self.lib.msa_batch_bytes_in_use(self.b)

    def upload(self, slot, J, S, D0=None, D1=None):
        J = np.ascontiguousarray(J, np.uint64)
        S = np.ascontiguousarray(S, np.uint64)
        if D0 is not None:
            D0 = np.ascontiguousarray(D0, np.uint64)
            D1 = np.ascontiguousarray(D1, np.uint64)
        _check(self.lib.msa_batch_upload_state(self.b, slot, _p(J), _p(S), _p(D0), _p(D1)))

    def set_shape(self, block, chunk):
        _check(self.lib.msa_batch_set_shape(self.b, block, chunk))

    def run(self, niter):
        _check(self.lib.msa_batch_run(self.b, niter))

    def sync(self):
        _check(self.lib.msa_batch_sync(self.b))

    def progress(self, nslots):
        t = np.zeros(nslots, np.int64)
        te = np.zeros(nslots, np.int64)
        _check(self.lib.msa_batch_progress(self.b, nslots, t.ctypes.data_as(C.POINTER(C.c_int64)),
                                           te.ctypes.data_as(C.POINTER(C.c_int64))))
        return t, te

    def download(self, slot):
        L, R = self.geom[slot]
        V = L ** 3
        S = np.zeros((R, V), np.uint64)
        D0 = np.zeros((R, 1, V), np.uint64)
        D1 = np.zeros_like(D0)
        _check(self.lib.msa_batch_download_state(self.b, slot, _p(S), _p(D0), _p(D1)))
        return S, D0, D1

    def measure(self, slot):
        L, R = self.geom[slot]
        unsat = np.zeros((R, 64), np.int32)
        demon = np.zeros((R, 64), np.int32)
        _check(self.lib.msa_batch_measure(self.b, slot, _p(unsat, i32p), _p(demon, i32p), None, None))
        return {"unsat": unsat, "energy": 3 * L ** 3 - 2 * unsat, "demon": demon}

    def remove(self, slot):
        _check(self.lib.msa_batch_remove(self.b, slot))
        self.geom.pop(slot, None)
