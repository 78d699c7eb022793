"""ctypes view of oracle/liboracle.so (the CPU restatement used as the checker).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)


class Lux(C.Structure):
    _fields_ = [("x", C.c_uint32 * 256), ("borrow", C.c_int),
                ("ip", C.c_uint8), ("ip1", C.c_uint8), ("ip2", C.c_uint8)]


class PrLcg(C.Structure):
    _fields_ = [("ira", C.c_uint64 * 256), ("pr", C.c_uint64), ("zseed", C.c_uint64),
                ("out", C.c_uint64), ("ip", C.c_uint8), ("ip1", C.c_uint8),
                ("ip2", C.c_uint8), ("ip3", C.c_uint8)]


class Xoshiro(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4), ("out", C.c_uint64)]


class Janus(C.Structure):
    _fields_ = [("ira", C.c_uint64 * 256), ("zseed", C.c_uint64), ("ip", C.c_uint8),
                ("ip1", C.c_uint8), ("ip2", C.c_uint8), ("ip3", C.c_uint8)]


class Lut(C.Structure):
    _fields_ = [("umbrales", C.c_uint32 * 512), ("prebusqueda", C.c_uint32 * 512)]

    def to_bytes(self):
        return bytes(self)

    @classmethod
    def from_file_bytes(cls, raw):
        """A reference LUT file is `double beta` followed by s_lut_heat_bath."""
        assert len(raw) == 8 + 4096
        return cls.from_buffer_copy(raw[8:])


class LatticeS(C.Structure):
    _fields_ = [("L", C.c_int), ("R", C.c_int), ("nc", C.c_int), ("V", C.c_uint64),
                ("J", u64p), ("S", u64p), ("D0", u64p), ("D1", u64p)]


def _p(a, t=u64p):
    return a.ctypes.data_as(t)


class Lattice:
    """Sample-per-lane lattice held in numpy arrays (canonical site order)."""

    def __init__(self, L, R, nc):
        self.L, self.R, self.nc = L, R, nc
        self.V = L ** 3
        self.J = np.zeros((3, self.V), np.uint64)
        self.S = np.zeros((R, self.V), np.uint64)
        self.D0 = np.zeros((R, nc, self.V), np.uint64)
        self.D1 = np.zeros((R, nc, self.V), np.uint64)

    def struct(self):
        for a in (self.J, self.S, self.D0, self.D1):
            assert a.flags.c_contiguous and a.dtype == np.uint64
        self._s = LatticeS(self.L, self.R, self.nc, self.V, _p(self.J), _p(self.S),
                           _p(self.D0), _p(self.D1))
        return C.byref(self._s)

    def copy(self):
        o = Lattice(self.L, self.R, self.nc)
        o.J[...] = self.J
        o.S[...] = self.S
        o.D0[...] = self.D0
        o.D1[...] = self.D1
        return o


class Oracle:
    def __init__(self, lib):
        self.lib = lib
        L = lib
        L.orc_lux_init.argtypes = [C.POINTER(Lux), C.c_uint64]
        L.orc_lux_next.argtypes = [C.POINTER(Lux)]
        L.orc_lux_next.restype = C.c_uint32
        L.orc_lux_u64.argtypes = [C.POINTER(Lux)]
        L.orc_lux_u64.restype = C.c_uint64
        L.orc_prlcg_step.argtypes = [C.POINTER(PrLcg)]
        L.orc_xoshiro_step.argtypes = [C.POINTER(Xoshiro)]
        L.orc_prlcg_init_hq.argtypes = [C.POINTER(PrLcg), C.c_uint64]
        L.orc_init_cpu_generators.argtypes = [C.POINTER(PrLcg), C.POINTER(Xoshiro), u32p, C.c_uint64]
        L.orc_seeds_from_one.argtypes = [C.c_uint64, u64p, C.c_int]
        L.orc_init_walkers.argtypes = [C.POINTER(PrLcg), C.POINTER(Xoshiro), u64p, C.c_uint64,
                                       u64p, u64p]
        L.orc_janus_init.argtypes = [C.POINTER(Janus), C.c_uint64]
        L.orc_janus_next.argtypes = [C.POINTER(Janus)]
        L.orc_janus_next.restype = C.c_uint64
        L.orc_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.orc_lut_build_beta.argtypes = [C.c_double, C.POINTER(Lut)]
        L.orc_lut_build_demon.argtypes = [C.c_double, C.POINTER(Lut), C.POINTER(Lut)]
        L.orc_thresholds.argtypes = [C.c_double, u64p]
        L.orc_bisection.argtypes = [C.POINTER(Lut), C.c_uint32, C.c_uint32]
        L.orc_bisection.restype = C.c_uint32
        L.orc_rnd_word.argtypes = [C.POINTER(Lut), u32p, u32p]
        L.orc_rnd_word.restype = C.c_uint32
        L.orc_fill_counter.argtypes = [C.POINTER(PrLcg), C.POINTER(Xoshiro), C.c_uint64,
                                       C.c_uint64, C.c_uint32, u32p]
        L.orc_isi_update.argtypes = [C.c_uint64] * 4 + [u64p, u64p, u64p]
        lat = C.c_void_p
        L.orc_halfsweep.argtypes = [lat, C.c_int, C.c_int]
        L.orc_walk.argtypes = [lat, C.c_int]
        L.orc_refresh_layer.argtypes = [lat, C.c_int, u32p, u32p, C.POINTER(Lut), C.POINTER(Lut)]
        L.orc_iteration.argtypes = [lat, C.c_int, u32p, u32p, C.POINTER(Lut), C.POINTER(Lut)]
        L.orc_iteration_omp.argtypes = [lat, C.c_int, u32p, u32p, C.POINTER(Lut),
                                        C.POINTER(Lut), C.c_int]
        L.orc_energy.argtypes = [lat, i32p]
        L.orc_demon_energy.argtypes = [lat, i32p]
        L.orc_overlap.argtypes = [lat, C.c_int, C.c_int, i32p, i32p]
        L.orc_init_spins_ref.argtypes = [lat, C.c_uint64]
        L.orc_init_J_ref.argtypes = [lat, C.c_uint64]
        L.orc_init_philox.argtypes = [lat, C.c_uint64]
        L.orc_gen_rndbits.argtypes = [C.c_int, C.c_int, C.POINTER(Lut), u32p, u32p, u64p]
        L.orc_metropolis_flip.argtypes = [C.c_uint64] * 7
        L.orc_metropolis_flip.restype = C.c_uint64
        L.orc_thermal_iteration.argtypes = [lat, C.POINTER(Lut), u32p, u32p, u32p]
        L.orc_msc_thermal.argtypes = [C.c_void_p, C.POINTER(Lut), u32p, u32p, u32p]
        L.orc_msc_create.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_msc_create.restype = C.c_void_p
        L.orc_msc_destroy.argtypes = [C.c_void_p]
        L.orc_msc_pack.argtypes = [C.c_void_p] + [u8p] * 6
        L.orc_msc_unpack.argtypes = [C.c_void_p, u8p, u8p, u8p]
        L.orc_msc_walker_words.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.orc_msc_monte_carlo.argtypes = [C.c_void_p]
        L.orc_msc_energy.argtypes = [C.c_void_p, i32p]
        L.orc_msc_scalar_energy.argtypes = [C.c_int] + [u8p] * 4
        L.orc_msc_scalar_energy.restype = C.c_int
        L.orc_msc_check_tables.argtypes = [C.c_void_p]
        L.orc_msc_check_tables.restype = C.c_int

    # ---- generators
    def lux_stream(self, seed, n):
        g = Lux()
        self.lib.orc_lux_init(C.byref(g), seed)
        return [self.lib.orc_lux_next(C.byref(g)) for _ in range(n)]

    def lux64_stream(self, seed, n):
        g = Lux()
        self.lib.orc_lux_init(C.byref(g), seed)
        return [self.lib.orc_lux_u64(C.byref(g)) for _ in range(n)]

    def cpu_generators(self, seed):
        pr, xo, key = PrLcg(), Xoshiro(), C.c_uint32()
        self.lib.orc_init_cpu_generators(C.byref(pr), C.byref(xo), C.byref(key), seed)
        return pr, xo, key.value

    def hq_stream(self, seed, n):
        g = PrLcg()
        self.lib.orc_prlcg_init_hq(C.byref(g), seed)
        out = []
        for _ in range(n):
            self.lib.orc_prlcg_step(C.byref(g))
            out.append(g.out)
        return out

    def seeds_from_one(self, seed, n):
        a = np.zeros(n, np.uint64)
        self.lib.orc_seeds_from_one(seed, _p(a), n)
        return [int(x) for x in a]

    def janus_stream(self, seed, n):
        g = Janus()
        self.lib.orc_janus_init(C.byref(g), seed)
        return [self.lib.orc_janus_next(C.byref(g)) for _ in range(n)]

    def philox(self, ctr, key):
        c = np.array(ctr, np.uint32)
        k = np.array(key, np.uint32)
        o = np.zeros(4, np.uint32)
        self.lib.orc_philox4x32_10(_p(c, u32p), _p(k, u32p), _p(o, u32p))
        return [int(x) for x in o]

    def lut_beta(self, beta):
        t = Lut()
        self.lib.orc_lut_build_beta(beta, C.byref(t))
        return t

    def lut_demon(self, beta):
        a, b = Lut(), Lut()
        self.lib.orc_lut_build_demon(beta, C.byref(a), C.byref(b))
        return a, b

    def thresholds(self, beta):
        t = np.zeros(3, np.uint64)
        self.lib.orc_thresholds(beta, _p(t))
        return [int(x) for x in t]

    def bisection(self, lut, r64):
        return self.lib.orc_bisection(C.byref(lut), (r64 >> 32) & 0xFFFFFFFF, r64 & 0xFFFFFFFF)

    def rnd_word(self, lut, ctr, key):
        c = np.array(ctr, np.uint32)
        k = np.array(key, np.uint32)
        w = self.lib.orc_rnd_word(C.byref(lut), _p(c, u32p), _p(k, u32p))
        return w, [int(x) for x in c]

    def init_walkers(self, pr, xo, thr, nwords):
        """InitWalkers for one replica: (w0, w1) uint64 [nwords]; advances pr, xo."""
        w0 = np.zeros(nwords, np.uint64)
        w1 = np.zeros(nwords, np.uint64)
        t = np.ascontiguousarray(thr, np.uint64)
        self.lib.orc_init_walkers(C.byref(pr), C.byref(xo), _p(t), nwords, _p(w0), _p(w1))
        return w0, w1

    def fill_counter(self, pr, xo, it, parity, jobid):
        c = np.zeros(4, np.uint32)
        self.lib.orc_fill_counter(C.byref(pr), C.byref(xo), it, parity, jobid, _p(c, u32p))
        return [int(x) for x in c]

    def isi(self, j1, j2, j3, j4, d0, d1):
        a, b, f = C.c_uint64(d0), C.c_uint64(d1), C.c_uint64(0)
        self.lib.orc_isi_update(j1, j2, j3, j4, C.byref(a), C.byref(b), C.byref(f))
        return a.value, b.value, f.value

    # ---- lattice
    def iteration(self, lat, ctr=None, key=None, luts=None, nthreads=0):
        if ctr is None:
            args = (None, None, None, None)
            do = 0
        else:
            ctr = np.ascontiguousarray(ctr, np.uint32)
            key = np.ascontiguousarray(key, np.uint32)
            args = (_p(ctr, u32p), _p(key, u32p), C.byref(luts[0]), C.byref(luts[1]))
            do = 1
        if nthreads:
            self.lib.orc_iteration_omp(lat.struct(), do, *args, nthreads)
        else:
            self.lib.orc_iteration(lat.struct(), do, *args)

    def refresh_layer(self, lat, layer, ctr, key, luts):
        ctr = np.ascontiguousarray(ctr, np.uint32)
        key = np.ascontiguousarray(key, np.uint32)
        self.lib.orc_refresh_layer(lat.struct(), layer, _p(ctr, u32p), _p(key, u32p),
                                   C.byref(luts[0]), C.byref(luts[1]))

    def halfsweep(self, lat, colour, layer):
        self.lib.orc_halfsweep(lat.struct(), colour, layer)

    def walk(self, lat, layer):
        self.lib.orc_walk(lat.struct(), layer)

    def energy(self, lat):
        e = np.zeros((lat.R, 64), np.int32)
        self.lib.orc_energy(lat.struct(), _p(e, i32p))
        return e

    def demon_energy(self, lat):
        e = np.zeros((lat.R, 64), np.int32)
        self.lib.orc_demon_energy(lat.struct(), _p(e, i32p))
        return e

    def overlap(self, lat, ra, rb):
        q = np.zeros(64, np.int32)
        ql = np.zeros(64, np.int32)
        self.lib.orc_overlap(lat.struct(), ra, rb, _p(q, i32p), _p(ql, i32p))
        return q, ql

    def init_ref(self, lat, seed_J, seed_u):
        self.lib.orc_init_J_ref(lat.struct(), seed_J)
        self.lib.orc_init_spins_ref(lat.struct(), seed_u)

    def init_philox(self, lat, seed):
        self.lib.orc_init_philox(lat.struct(), seed)

    # ---- canonical Metropolis thermalization
    def gen_rndbits(self, L, R, lut, ctr, key):
        out = np.zeros((3, L ** 3 // 128 * R), np.uint64)
        ctr = np.ascontiguousarray(ctr, np.uint32)
        key = np.ascontiguousarray(key, np.uint32)
        self.lib.orc_gen_rndbits(L, R, C.byref(lut), _p(ctr, u32p), _p(key, u32p), _p(out))
        return out

    def metropolis_flip(self, j1, j2, j3, j4, b0, b1, b2):
        return self.lib.orc_metropolis_flip(j1, j2, j3, j4, b0, b1, b2)

    def thermal_iteration(self, lat, lut, ctr0, ctr1, key):
        ctr0 = np.ascontiguousarray(ctr0, np.uint32)
        ctr1 = np.ascontiguousarray(ctr1, np.uint32)
        key = np.ascontiguousarray(key, np.uint32)
        self.lib.orc_thermal_iteration(lat.struct(), C.byref(lut), _p(ctr0, u32p), _p(ctr1, u32p),
                                       _p(key, u32p))


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load():
    if not os.path.exists(LIB):
        build()
    return Oracle(C.CDLL(LIB))


def unpack_lane(words, lane):
    """uint64 array -> uint8 array of bit `lane`."""
    return ((words >> np.uint64(lane)) & np.uint64(1)).astype(np.uint8)


def read_stream(name):
    path = os.path.join(REPO, "tests", "golden", name)
    with open(path) as f:
        return [ln.split() for ln in f if not ln.startswith("#")]
