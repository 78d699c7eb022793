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
        L.msa_download_state_async.argtypes = [vp, u64p, u64p, u64p]
        L.msa_download_wait.argtypes = [vp]
        L.msa_host_alloc.argtypes = [C.c_int64]
        L.msa_host_alloc.restype = C.c_void_p
        L.msa_host_free.argtypes = [C.c_void_p]
        L.msa_init_philox.argtypes = [vp, C.c_uint64]
        L.msa_set_luts.argtypes = [vp, C.c_int32, C.c_void_p]
        L.msa_set_refresh.argtypes = [vp, C.c_int32, u32p, u32p, i32p]
        L.msa_run.argtypes = [vp, C.c_uint64, C.c_int64, C.c_int64]
        L.msa_run_timed.argtypes = [vp, C.c_uint64, C.c_int64, C.c_int64, C.POINTER(C.c_float)]
        L.msa_measure.argtypes = [vp, i32p, i32p, i32p, i32p]
        L.msa_measure_async.argtypes = [vp]
        L.msa_measure_wait.argtypes = [vp, i32p, i32p, i32p, i32p]
        L.msa_sync.argtypes = [vp]
        L.msa_set_tiling.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
        L.msa_get_tiling.argtypes = [vp, i32p, i32p, i32p]
        L.msa_set_refresh_mode.argtypes = [vp, C.c_int32]
        L.msa_get_refresh_mode.argtypes = [vp, i32p]
        L.msa_run_thermal.argtypes = [vp, C.c_int64, C.c_void_p, u32p, u32p]
        L.msa_run_thermal_noalloc.argtypes = [vp, C.c_int64, C.c_void_p, u32p, u32p]
        L.msa_set_walker_streams.argtypes = [vp, C.c_int32, i32p, C.c_void_p, C.c_int32, u64p]
        L.msa_get_walker_streams.argtypes = [vp, C.c_void_p]
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
        L.msa_batch_planned.argtypes = [bp, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.msa_batch_wait_previous.argtypes = [bp]
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


class HostArray:
    """Page-locked host memory (msa_host_alloc) seen as a uint64 numpy array."""

    def __init__(self, shape):
        n = int(np.prod(shape))
        self.lib = hip()
        self.ptr = self.lib.msa_host_alloc(8 * n)
        if not self.ptr:
            raise MsaError("msa_host_alloc of %d bytes failed" % (8 * n))
        self.array = np.ctypeslib.as_array((C.c_uint64 * n).from_address(self.ptr)).reshape(shape)

    def close(self):
        if self.ptr:
            self.array = None
            self.lib.msa_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()


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

    def download_async(self, S, D0, D1):
        """Queue a copy of the state into S [R][V], D0, D1 [R][nc][V] (uint64,
        any may be None) without waiting; download_wait() completes it."""
        for a, n in ((S, self.R * self.V), (D0, self.R * self.nc * self.V),
                     (D1, self.R * self.nc * self.V)):
            assert a is None or (a.dtype == np.uint64 and a.size == n and a.flags.c_contiguous)
        self._xfer = (S, D0, D1)          # kept alive until the copy is waited for
        _check(self.lib.msa_download_state_async(self.ctx, _p(S), _p(D0), _p(D1)))

    def download_wait(self):
        _check(self.lib.msa_download_wait(self.ctx))
        self._xfer = None

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

    def set_walker_streams(self, lanes, streams, thr):
        """Demon refresh from the reference's walker streams (msa_set_walker_streams):
        lanes (n,), streams: n*R WalkerStream (stream i, replica r at i*R + r),
        thr (nref, 3) uint64 thresholds.  Empty lanes return to the Philox refresh."""
        lanes = np.ascontiguousarray(lanes, np.int32)
        thr = np.ascontiguousarray(thr, np.uint64).reshape(-1, 3)
        arr = (WalkerStream * max(1, len(streams)))(*streams)
        _check(self.lib.msa_set_walker_streams(self.ctx, lanes.size, _p(lanes, i32p),
                                               C.cast(arr, C.c_void_p), thr.shape[0], _p(thr)))
        self._walker_n = lanes.size

    def walker_streams(self):
        arr = (WalkerStream * (self._walker_n * self.R))()
        _check(self.lib.msa_get_walker_streams(self.ctx, C.cast(arr, C.c_void_p)))
        return list(arr)

    def run_thermal(self, lut, ctr, key):
        """Canonical Metropolis thermalization (msa_run_thermal): `lut` is the
        4096 bytes of the heat-bath table, ctr (nterm, 2, R, 4), key (R,)."""
        raw = np.ascontiguousarray(np.frombuffer(bytes(lut), np.uint8))
        assert raw.size == LUT_BYTES
        ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 2, self.R, 4)
        key = np.ascontiguousarray(key, np.uint32)
        _check(self.lib.msa_run_thermal(self.ctx, ctr.shape[0], raw.ctypes.data_as(C.c_void_p),
                                        _p(ctr, u32p), _p(key, u32p)))

    def run_thermal_noalloc(self, lut, ctr, key):
        """msa_run_thermal_noalloc: the thermalization without device
        allocation (scratch in the library's fixed device array)."""
        raw = np.ascontiguousarray(np.frombuffer(bytes(lut), np.uint8))
        assert raw.size == LUT_BYTES
        ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 2, self.R, 4)
        key = np.ascontiguousarray(key, np.uint32)
        _check(self.lib.msa_run_thermal_noalloc(self.ctx, ctr.shape[0],
                                                raw.ctypes.data_as(C.c_void_p),
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

    def set_refresh_mode(self, mode):
        """0: refresh drawn inside the sweep; 1: drawn ahead on a second stream;
        2: as 1, two refresh sweeps in a row as one sweep pair; -1: the
        library's choice for the geometry."""
        _check(self.lib.msa_set_refresh_mode(self.ctx, mode))

    def refresh_mode(self):
        m = C.c_int32()
        _check(self.lib.msa_get_refresh_mode(self.ctx, C.byref(m)))
        return m.value

    def tiling(self):
        m, ty, tz = C.c_int32(), C.c_int32(), C.c_int32()
        _check(self.lib.msa_get_tiling(self.ctx, C.byref(m), C.byref(ty), C.byref(tz)))
        return m.value, ty.value, tz.value

    def sync(self):
        _check(self.lib.msa_sync(self.ctx))

    def measure(self, fn="msa_measure"):
        """Per-lane integer observables: dict of int32 arrays."""
        npair = self.R // 2
        unsat = np.zeros((self.R, 64), np.int32)
        demon = np.zeros((self.R, 64), np.int32)
        q = np.zeros((max(npair, 1), 64), np.int32)
        ql = np.zeros((max(npair, 1), 64), np.int32)
        _check(getattr(self.lib, fn)(self.ctx, _p(unsat, i32p), _p(demon, i32p),
                                     _p(q, i32p) if npair else None,
                                     _p(ql, i32p) if npair else None))
        out = {"unsat": unsat, "energy": 3 * self.V - 2 * unsat, "demon": demon}
        if npair:
            out["q"] = q[:npair]
            out["qlink"] = ql[:npair]
        return out

    def measure_async(self):
        """Queue a measurement of the state the queued work leaves (msa_measure_async)."""
        _check(self.lib.msa_measure_async(self.ctx))

    def measure_wait(self):
        """The queued measurement's observables, as measure() returns them."""
        return self.measure("msa_measure_wait")


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


class WalkerStream(C.Structure):
    """msa_walker_stream: last 61 lagged-sum values, LCG state, xoshiro256++ state."""
    _fields_ = [("ring", C.c_uint64 * 61), ("lcg", C.c_uint64), ("xo", C.c_uint64 * 4)]


class SData(C.Structure):
    """s_data as stored in conf files (184 bytes)."""
    _fields_ = [("maxtime", C.c_uint64), ("now", C.c_uint64), ("ntimes", C.c_int32),
                ("nterm", C.c_int32), ("nc", C.c_int32), ("nr", C.c_int32), ("l", C.c_int32),
                ("flag", C.c_int32), ("seed_J", C.c_uint64), ("seed_u", C.c_uint64),
                ("seed_MC", C.c_uint64 * 16)]


class Icds(C.Structure):
    """icds, the head of walkers files (40 bytes)."""
    _fields_ = [("first", C.c_uint32), ("isample", C.c_uint32), ("nrep", C.c_uint32),
                ("walkers_per_site", C.c_uint32), ("lv", C.c_uint32), ("t", C.c_uint32),
                ("seed", C.c_uint64), ("beta", C.c_double)]


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
        L.msa_prng_skip.argtypes = [C.POINTER(Prng), C.c_uint64]
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
        L.msa_sched_tables_resume.argtypes = [C.POINTER(Sched), C.c_int64, C.c_int64, C.c_int,
                                              C.POINTER(Prng), C.POINTER(Xo256), C.c_uint32, u32p,
                                              C.c_void_p, i32p]
        L.msa_gen_skip.argtypes = [C.POINTER(Prng), C.POINTER(Xo256), C.c_int, C.c_int64]
        L.msa_walker_generators.argtypes = [C.c_uint64, C.c_int, C.c_int, C.POINTER(Prng),
                                            C.POINTER(Xo256)]
        L.msa_walker_stream_of.argtypes = [C.POINTER(Prng), C.POINTER(Xo256), C.POINTER(WalkerStream)]
        L.msa_walker_stream_to.argtypes = [C.POINTER(WalkerStream), C.c_uint8, C.POINTER(Prng),
                                           C.POINTER(Xo256)]
        u8p = C.POINTER(C.c_uint8)
        L.msa_conf_bytes.argtypes = [C.c_int]
        L.msa_conf_bytes.restype = C.c_uint64
        L.msa_walkers_bytes.argtypes = [C.c_int, C.c_int, C.c_int]
        L.msa_walkers_bytes.restype = C.c_uint64
        L.msa_msc_words.argtypes = [C.c_int, C.c_int, C.c_int]
        L.msa_msc_words.restype = C.c_uint64
        L.msa_prng_to_ref.argtypes = [C.POINTER(Prng), u8p]
        L.msa_prng_from_ref.argtypes = [u8p, C.POINTER(Prng)]
        L.msa_xo256_to_ref.argtypes = [C.POINTER(Xo256), u8p]
        L.msa_xo256_from_ref.argtypes = [u8p, C.POINTER(Xo256)]
        L.msa_conf_write.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(SData), C.c_double,
                                     C.POINTER(Prng), C.POINTER(Xo256), C.c_int, u8p, C.c_int32,
                                     C.c_char_p, C.c_int]
        L.msa_conf_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(SData), C.POINTER(C.c_double),
                                    C.POINTER(Prng), C.POINTER(Xo256), u8p, i32p, C.c_char_p,
                                    C.c_int]
        L.msa_walkers_write.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64, C.POINTER(Icds), C.c_int,
                                        C.c_int, C.c_int, C.POINTER(Prng), C.POINTER(Xo256), u64p,
                                        C.c_char_p, C.c_int]
        L.msa_walkers_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(Icds),
                                       C.POINTER(Prng), C.POINTER(Xo256), u64p, C.c_char_p, C.c_int]
        L.msa_janus_from_lanes.argtypes = [C.c_int, C.c_int, u64p, C.c_int, i32p, u8p]
        L.msa_janus_to_lanes.argtypes = [C.c_int, C.c_int, u8p, C.c_int, i32p, u64p]
        L.msa_msc_from_lanes.argtypes = [C.c_int, C.c_int, C.c_int, u64p, u64p, C.c_int, i32p, u64p]
        L.msa_msc_to_lanes.argtypes = [C.c_int, C.c_int, C.c_int, u64p, C.c_int, i32p, u64p, u64p]
        L.msa_lane_energy.argtypes = [C.c_int, C.c_int, u64p, u64p, i32p]
        L.msa_lane_countJ.argtypes = [C.c_int, u64p, i32p]
        i64p = C.POINTER(C.c_int64)
        L.msa_overlap_counts.argtypes = [C.c_int, u8p, u8p, i64p]
        L.msa_c4_from_counts.argtypes = [C.c_int, C.c_int64, i64p, C.POINTER(C.c_double)]
        L.msa_xi_integrals.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
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
    npair = H.msa_sched_tables_from(C.byref(sched), nref, R, pr, xo, jobid, _p(ctr, u32p),
                                    luts.ctypes.data_as(C.c_void_p), _p(idx, i32p))
    if npair < 0:
        raise MsaError("schedule tables failed")
    return ctr, keys.copy(), luts[:2 * npair * LUT_BYTES], idx


def schedule_tables(sched, niter, R, seed_mc, jobid=0):
    """Counters (nref,R,4) u32, keys (R,) u32, LUT bytes (one pair of tables per
    distinct beta), lut_idx (nref,) = the pair of each refresh."""
    H = host()
    nref = H.msa_sched_count(C.byref(sched), niter)
    ctr = np.zeros((nref, R, 4), np.uint32)
    keys = np.zeros(R, np.uint32)
    luts = np.zeros(2 * nref * LUT_BYTES, np.uint8)
    idx = np.zeros(nref, np.int32)
    seeds = np.ascontiguousarray(seed_mc, np.uint64)
    npair = H.msa_sched_tables(C.byref(sched), nref, R, _p(seeds), jobid, _p(ctr, u32p),
                               _p(keys, u32p), luts.ctypes.data_as(C.c_void_p), _p(idx, i32p))
    if npair < 0:
        raise MsaError("schedule tables failed")
    return ctr, keys, luts[:2 * npair * LUT_BYTES], idx


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
        """Iterations reached and lengths, from the device (syncs)."""
        t = np.zeros(nslots, np.int64)
        te = np.zeros(nslots, np.int64)
        _check(self.lib.msa_batch_progress(self.b, nslots, t.ctypes.data_as(C.POINTER(C.c_int64)),
                                           te.ctypes.data_as(C.POINTER(C.c_int64))))
        return t, te

    def planned(self, nslots):
        """Iterations enqueued so far and lengths (no sync)."""
        t = np.zeros(nslots, np.int64)
        te = np.zeros(nslots, np.int64)
        _check(self.lib.msa_batch_planned(self.b, nslots, t.ctypes.data_as(C.POINTER(C.c_int64)),
                                          te.ctypes.data_as(C.POINTER(C.c_int64))))
        return t, te

    def wait_previous(self):
        """Wait for the work enqueued before the latest run."""
        _check(self.lib.msa_batch_wait_previous(self.b))

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


# ------------------------------------------------------------- checkpoints --
_u8p = C.POINTER(C.c_uint8)


def _lanes(lanes):
    return np.ascontiguousarray(lanes, np.int32)


def janus_from_lanes(L, spin, lanes):
    """Conf spin payload (V/8 bytes per lane and replica): uint8 [n][R][V/8]."""
    R, V = spin.shape
    ln = _lanes(lanes)
    out = np.zeros((len(ln), R, V // 8), np.uint8)
    host().msa_janus_from_lanes(L, R, _p(np.ascontiguousarray(spin, np.uint64)), len(ln),
                                _p(ln, i32p), _p(out, _u8p))
    return out


def janus_to_lanes(L, janus, lanes, spin):
    """Write the conf payloads janus [n][R][V/8] into the listed lanes of spin [R][V] (in place)."""
    R = spin.shape[0]
    ln = _lanes(lanes)
    host().msa_janus_to_lanes(L, R, _p(np.ascontiguousarray(janus, np.uint8), _u8p), len(ln),
                              _p(ln, i32p), _p(spin))
    return spin


def msc_from_lanes(L, R, nc, d0, d1, lanes):
    """Walkers-file demon words of the listed lanes: uint64 [n][msa_msc_words]."""
    ln = _lanes(lanes)
    nw = host().msa_msc_words(L, R, nc)
    out = np.zeros((len(ln), nw), np.uint64)
    host().msa_msc_from_lanes(L, R, nc, _p(np.ascontiguousarray(d0, np.uint64)),
                              _p(np.ascontiguousarray(d1, np.uint64)), len(ln), _p(ln, i32p), _p(out))
    return out


def msc_to_lanes(L, R, nc, w, lanes, d0, d1):
    ln = _lanes(lanes)
    host().msa_msc_to_lanes(L, R, nc, _p(np.ascontiguousarray(w, np.uint64)), len(ln), _p(ln, i32p),
                            _p(d0), _p(d1))
    return d0, d1


def lane_energy(L, J, spin):
    """calculate_scalar_energy of every lane: int32 [R][64]."""
    R = spin.shape[0]
    e = np.zeros((R, 64), np.int32)
    host().msa_lane_energy(L, R, _p(np.ascontiguousarray(J, np.uint64)),
                           _p(np.ascontiguousarray(spin, np.uint64)), _p(e, i32p))
    return e


def lane_countJ(L, J):
    c = np.zeros(64, np.int32)
    host().msa_lane_countJ(L, _p(np.ascontiguousarray(J, np.uint64)), _p(c, i32p))
    return c


def _errbuf():
    return C.create_string_buffer(512)


def conf_write(rep_dir, now, sdata, beta, pr, xo, L, janus, energy):
    e = _errbuf()
    j = np.ascontiguousarray(janus, np.uint8)
    if host().msa_conf_write(rep_dir.encode(), now, C.byref(sdata), beta, C.byref(pr), C.byref(xo),
                             L, _p(j, _u8p), int(energy), e, len(e)):
        raise MsaError(e.value.decode())


def conf_read(path, L):
    """(SData, beta, Prng, Xo256, janus uint8 [V/8], energy)."""
    e = _errbuf()
    d, beta, pr, xo, en = SData(), C.c_double(), Prng(), Xo256(), C.c_int32()
    janus = np.zeros(L ** 3 // 8, np.uint8)
    if host().msa_conf_read(path.encode(), L, C.byref(d), C.byref(beta), C.byref(pr), C.byref(xo),
                            _p(janus, _u8p), C.byref(en), e, len(e)):
        raise MsaError(e.value.decode())
    return d, beta.value, pr, xo, janus, en.value


def walkers_write(sample_dir, now, prev, icds, L, R, nc, prs, xos, w):
    e = _errbuf()
    PA, XA = Prng * R, Xo256 * R
    pa, xa = PA(*prs), XA(*xos)
    ww = np.ascontiguousarray(w, np.uint64)
    if host().msa_walkers_write(sample_dir.encode(), now, prev, C.byref(icds), L, R, nc, pa, xa,
                                _p(ww), e, len(e)):
        raise MsaError(e.value.decode())


def walkers_read(path, L, R, nc):
    """(Icds, [Prng]*R, [Xo256]*R, uint64 words)."""
    e = _errbuf()
    ic = Icds()
    pa, xa = (Prng * R)(), (Xo256 * R)()
    w = np.zeros(host().msa_msc_words(L, R, nc), np.uint64)
    if host().msa_walkers_read(path.encode(), L, R, nc, C.byref(ic), pa, xa, _p(w), e, len(e)):
        raise MsaError(e.value.decode())
    return ic, list(pa), list(xa), w


# ---------------------------------------------------------- walker streams --
def walker_streams(seed_u, n, R):
    """The reference's walker streams of n samples x R replicas (stream i*R + r):
    (list of WalkerStream, list of (Prng, Xo256) they started from)."""
    H = host()
    pr, xo = (Prng * (n * R))(), (Xo256 * (n * R))()
    H.msa_walker_generators(seed_u, n, R, pr, xo)
    out = []
    for k in range(n * R):
        st = WalkerStream()
        H.msa_walker_stream_of(C.byref(pr[k]), C.byref(xo[k]), C.byref(st))
        out.append(st)
    return out, [(pr[k], xo[k]) for k in range(n * R)]


def thresholds(beta):
    t = np.zeros(3, np.uint64)
    host().msa_thresholds(beta, _p(t))
    return t


# ------------------------------------------------------ offline observables --
def overlap_counts(L, janus_a, janus_b):
    """int64 [3][L/2+1]: sites where q_x != q_{x + r e_mu}."""
    d = np.zeros((3, L // 2 + 1), np.int64)
    a = np.ascontiguousarray(janus_a, np.uint8)
    b = np.ascontiguousarray(janus_b, np.uint8)
    if host().msa_overlap_counts(L, _p(a, _u8p), _p(b, _u8p), _p(d, C.POINTER(C.c_int64))):
        raise MsaError("overlap counts: bad L=%d" % L)
    return d


def c4_from_counts(L, nterms, diff):
    c4 = np.zeros(L // 2 + 1, np.float64)
    d = np.ascontiguousarray(diff, np.int64)
    host().msa_c4_from_counts(L, nterms, _p(d, C.POINTER(C.c_int64)), _p(c4, C.POINTER(C.c_double)))
    return c4


def xi_integrals(c4):
    """(xi12, xi23, rc) of a correlation function c4[r], r = 0..nr-1."""
    c = np.ascontiguousarray(c4, np.float64)
    x12, x23 = C.c_double(), C.c_double()
    rc = host().msa_xi_integrals(c.size, _p(c, C.POINTER(C.c_double)), C.byref(x12), C.byref(x23))
    return x12.value, x23.value, rc
