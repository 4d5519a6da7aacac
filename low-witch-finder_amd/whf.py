"""Python binding of the MI355X Seed Check library (libwhf_hip.so, C ABI in include/whf.h).

Thin ctypes layer over the C ABI, used by the benchmark and the tests. It loads the in-tree library and
raises if it is missing: there is no fallback implementation.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwhf_hip.so")
PHASE1_SCAN = os.path.join(HERE, "bin", "phase1_scan")
DEEPDIVE_SCAN = os.path.join(HERE, "bin", "deepdive_scan")
BIOME_MAP = os.path.join(HERE, "bin", "biome_map")

OK, EINVAL, ENOMEM, EDEVICE, EOVERFLOW, ENOTSUP = 0, -1, -2, -3, -4, -5
SEEDS_LIST, SEEDS_SPLITMIX, SEEDS_COUNTER = 0, 1, 2
V_118, V_262, V_LARGE, V_SINGLE = 1, 2, 4, 8
PREC_FP64, PREC_FP32_GATE = 0, 1
KIND_N = 11
CFG_PROFILE = 1
CFG_ONE_STREAM = 2   # every stage of a scan on one stream (no batch pipeline)

# Exported symbols of include/whf.h, checked by the tests.
SYMBOLS = ("whf_open", "whf_close", "whf_scan", "whf_probe", "whf_version_flags", "whf_strerror",
           "whf_device_name", "whf_noise_error", "whf_fp32_bound", "whf_deepdive", "whf_deepdive_noalloc", "whf_dd_terrain",
           "whf_room_floors",
           "whf_server_open", "whf_server_close", "whf_submit", "whf_submit_deepdive", "whf_poll",
           "whf_biome_tree", "whf_biome", "whf_surface_heights", "whf_post_carver_heights", "whf_hut_eval",
           "whf_get_tuning", "whf_set_tuning", "whf_get_tuning_for", "whf_set_tuning_for", "whf_load_tuning",
           "whf_build_id", "whf_solve", "whf_solve_chunks")
HIT_DEEPDIVE, HIT_PHASE2 = 1, 2
P2_NONE, P2_OR_LOWER, P2_EXACTLY = 0, 1, 2
BIOME_SWAMP, BIOME_QY0, BIOME_QY_N = 6, -17, 97
BIO_OCC_ONLY = 0x100
OCC_SURFACE = -(1 << 31)
BLOCK_SOLID, BLOCK_AIR, BLOCK_WATER, BLOCK_LAVA = 0, 1, 2, 3


class Tuning(ctypes.Structure):
    _fields_ = [("ero_wg", ctypes.c_int), ("ero_ilp", ctypes.c_int), ("ero_wg_per_cu", ctypes.c_int),
                ("gate_wg_per_cu", ctypes.c_int), ("chunk_checks", ctypes.c_uint64), ("reserved", ctypes.c_int * 8)]


TUNING_FIELDS = ("ero_wg", "ero_ilp", "ero_wg_per_cu", "gate_wg_per_cu", "chunk_checks")


class Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("flags", ctypes.c_int), ("hbm_budget", ctypes.c_uint64)]


class Request(ctypes.Structure):
    _fields_ = [("seed_mode", ctypes.c_int), ("precision", ctypes.c_int),
                ("seeds", ctypes.POINTER(ctypes.c_int64)), ("n_seeds", ctypes.c_uint64),
                ("stream", ctypes.c_uint64), ("offset", ctypes.c_uint64),
                ("rx0", ctypes.c_int), ("rx1", ctypes.c_int), ("rz0", ctypes.c_int), ("rz1", ctypes.c_int),
                ("max_height", ctypes.c_int), ("vflags", ctypes.c_int), ("salt", ctypes.c_int),
                ("phase2", ctypes.c_int), ("phase2_max_y", ctypes.c_int), ("biome_tree", ctypes.c_int)]


class Hit(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int64), ("rx", ctypes.c_int32), ("rz", ctypes.c_int32),
                ("cx", ctypes.c_int32), ("cz", ctypes.c_int32), ("floor_y", ctypes.c_int32),
                ("flags", ctypes.c_int32)]


class Sink(ctypes.Structure):
    _fields_ = [("hits", ctypes.POINTER(Hit)), ("capacity", ctypes.c_uint64), ("count", ctypes.c_uint64)]


class Stats(ctypes.Structure):
    _fields_ = [("seeds", ctypes.c_uint64), ("checks", ctypes.c_uint64), ("accepted", ctypes.c_uint64),
                ("reached_ladder", ctypes.c_uint64), ("funnel", ctypes.c_uint64 * KIND_N),
                ("tiers", ctypes.c_uint64 * 8), ("samplers_built", ctypes.c_uint64 * 3), ("device_ms", ctypes.c_double),
                ("kernel_ms", ctypes.c_double * 8), ("fp64_rechecks", ctypes.c_uint64),
                ("phase2", ctypes.c_uint64 * 5), ("phase2_ms", ctypes.c_double)]


class DdRequest(ctypes.Structure):
    _fields_ = [("s_first", ctypes.c_uint64), ("count", ctypes.c_uint64), ("rx0", ctypes.c_int), ("rx1", ctypes.c_int),
                ("rz0", ctypes.c_int), ("rz1", ctypes.c_int), ("floor_max", ctypes.c_int), ("expand", ctypes.c_int),
                ("max_height", ctypes.c_int), ("vflags", ctypes.c_int), ("salt", ctypes.c_int),
                ("biome_tree", ctypes.c_int)]


class DdSpot(ctypes.Structure):
    _fields_ = [("s48", ctypes.c_uint64), ("rx", ctypes.c_int32), ("rz", ctypes.c_int32), ("cx", ctypes.c_int32),
                ("cz", ctypes.c_int32), ("carver", ctypes.c_uint64), ("room_floor", ctypes.c_int32),
                ("pieces", ctypes.c_int32)]


class DdSink(ctypes.Structure):
    _fields_ = [("spots", ctypes.POINTER(DdSpot)), ("capacity", ctypes.c_uint64), ("count", ctypes.c_uint64)]


class DdStats(ctypes.Structure):
    _fields_ = [("structure_seeds", ctypes.c_uint64), ("tests", ctypes.c_uint64), ("starts", ctypes.c_uint64),
                ("walks", ctypes.c_uint64 * 3), ("spots", ctypes.c_uint64), ("draws", ctypes.c_uint64),
                ("pieces", ctypes.c_uint64), ("expanded_seeds", ctypes.c_uint64), ("accepted", ctypes.c_uint64),
                ("funnel", ctypes.c_uint64 * KIND_N), ("tiers", ctypes.c_uint64 * 8),
                ("samplers_built", ctypes.c_uint64 * 3), ("device_ms", ctypes.c_double),
                ("stage_ms", ctypes.c_double * 3), ("biome_checked", ctypes.c_uint64),
                ("biome_rejected", ctypes.c_uint64)]


class BiomeQuery(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int64), ("hut_x", ctypes.c_int32), ("hut_z", ctypes.c_int32),
                ("occ_y", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class BiomeResult(ctypes.Structure):
    _fields_ = [("target", ctypes.c_int64 * 6)] + \
               [(n, ctypes.c_int32) for n in ("qx", "qy", "qz", "occ_y", "winner", "ok")] + \
               [("swamp_qy", ctypes.c_uint64 * 2)]


class Column(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int64), ("x", ctypes.c_int32), ("z", ctypes.c_int32)]


class HutQuery(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int64), ("cx", ctypes.c_int32), ("cz", ctypes.c_int32)]


class HutResult(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("dir", "size_x", "size_z", "biome_ok", "winner", "occ_y", "wet", "flooded",
                                               "carve_skipped", "avg_y")] + \
               [("sum_h", ctypes.c_int64), ("ok", ctypes.c_int32), ("reject", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("slots", ctypes.c_int), ("chunk_checks", ctypes.c_uint64),
                ("hbm_budget", ctypes.c_uint64), ("flags", ctypes.c_int), ("reserved", ctypes.c_int)]


class Event(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int), ("reserved", ctypes.c_int), ("stats", Stats),
                ("submit_s", ctypes.c_double), ("start_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class SolveBundle(ctypes.Structure):
    _fields_ = [("layout", ctypes.c_uint64), ("regime_x", ctypes.c_int), ("regime_z", ctypes.c_int),
                ("tier", ctypes.c_int), ("reserved", ctypes.c_int)]


class SolveHit(ctypes.Structure):
    _fields_ = [("s48", ctypes.c_uint64), ("layout", ctypes.c_uint64), ("cx", ctypes.c_int32), ("cz", ctypes.c_int32),
                ("bundle", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class SolveSink(ctypes.Structure):
    _fields_ = [("hits", ctypes.POINTER(SolveHit)), ("capacity", ctypes.c_uint64), ("count", ctypes.c_uint64)]


class SolveStats(ctypes.Structure):
    _fields_ = [("groups", ctypes.c_uint64), ("regions", ctypes.c_uint64), ("tasks", ctypes.c_uint64),
                ("candidates", ctypes.c_uint64), ("leaves", ctypes.c_uint64), ("hits", ctypes.c_uint64),
                ("device_ms", ctypes.c_double)]


TIERS = {"odd": 1, "u2": 2, "eq": 3}


def parse_plan(text):
    """Plan lines "X XZ class ..." (hut_lookup_v4_source.cu:1297-1305): [(X, regime_x, regime_z, tier)]. Lines
    starting with '#' and blank lines are skipped; the class defaults to odd."""
    out = []
    for ln in text.splitlines():
        if not ln.strip() or ln.startswith("#"):
            continue
        f = ln.split()
        rg = f[1]
        if len(rg) != 2 or rg[0] not in "012" or rg[1] not in "012":
            raise ValueError("bad plan line: %r" % ln)
        cls = f[2] if len(f) > 2 else "odd"
        if cls not in TIERS:
            raise ValueError("plan class %r unknown (odd | u2 | eq)" % cls)
        out.append((int(f[0]) & ((1 << 48) - 1), int(rg[0]), int(rg[1]), TIERS[cls]))
    return out


KERNEL_FAMILIES = ("seeds", "build_erosion", "erosion_gate", "compact", "build_climate", "gate2", "build_ladder",
                   "gate3")

_lib = None


def load():
    """The loaded C library; raises if the in-tree build is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libwhf_hip.so is not built (run __graft_entry__.build()): %s" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.whf_open.restype = ctypes.c_int
        L.whf_open.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_void_p)]
        L.whf_close.restype = None
        L.whf_close.argtypes = [ctypes.c_void_p]
        L.whf_scan.restype = ctypes.c_int
        L.whf_scan.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(Sink), ctypes.POINTER(Stats)]
        L.whf_probe.restype = ctypes.c_int
        L.whf_probe.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32),
                                ctypes.POINTER(ctypes.c_int32), ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_double)]
        L.whf_version_flags.restype = ctypes.c_int
        L.whf_version_flags.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        L.whf_strerror.restype = ctypes.c_char_p
        L.whf_strerror.argtypes = [ctypes.c_int]
        L.whf_device_name.restype = ctypes.c_char_p
        L.whf_device_name.argtypes = [ctypes.c_void_p]
        L.whf_noise_error.restype = ctypes.c_int
        L.whf_noise_error.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_uint64,
                                      ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.whf_fp32_bound.restype = ctypes.c_double
        L.whf_fp32_bound.argtypes = []
        L.whf_server_open.restype = ctypes.c_int
        L.whf_server_open.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(ctypes.c_void_p)]
        L.whf_server_close.restype = None
        L.whf_server_close.argtypes = [ctypes.c_void_p]
        L.whf_submit.restype = ctypes.c_int
        L.whf_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(Sink),
                                 ctypes.POINTER(ctypes.c_uint64)]
        L.whf_submit_deepdive.restype = ctypes.c_int
        L.whf_submit_deepdive.argtypes = [ctypes.c_void_p, ctypes.POINTER(DdRequest), ctypes.POINTER(DdSink),
                                          ctypes.POINTER(Sink), ctypes.POINTER(DdStats), ctypes.POINTER(ctypes.c_uint64)]
        L.whf_poll.restype = ctypes.c_int
        L.whf_poll.argtypes = [ctypes.c_void_p, ctypes.POINTER(Event), ctypes.c_int, ctypes.c_int]
        L.whf_deepdive.restype = ctypes.c_int
        L.whf_deepdive.argtypes = [ctypes.c_void_p, ctypes.POINTER(DdRequest), ctypes.POINTER(DdSink),
                                   ctypes.POINTER(Sink), ctypes.POINTER(DdStats)]
        L.whf_deepdive_noalloc.restype = ctypes.c_int
        L.whf_deepdive_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(DdRequest), ctypes.POINTER(DdSink),
                                           ctypes.POINTER(DdStats)]
        L.whf_dd_terrain.restype = ctypes.c_int
        L.whf_dd_terrain.argtypes = [ctypes.c_void_p, ctypes.POINTER(DdRequest), ctypes.POINTER(DdSpot), ctypes.c_uint64,
                                     ctypes.POINTER(Sink), ctypes.POINTER(DdStats)]
        L.whf_room_floors.restype = ctypes.c_int
        L.whf_room_floors.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32)]
        L.whf_surface_heights.restype = ctypes.c_int
        L.whf_surface_heights.argtypes = [ctypes.c_void_p, ctypes.POINTER(Column), ctypes.c_uint64, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        L.whf_post_carver_heights.restype = ctypes.c_int
        L.whf_post_carver_heights.argtypes = [ctypes.c_void_p, ctypes.POINTER(Column), ctypes.c_uint64, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        L.whf_hut_eval.restype = ctypes.c_int
        L.whf_hut_eval.argtypes = [ctypes.c_void_p, ctypes.POINTER(HutQuery), ctypes.c_uint64, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.POINTER(HutResult)]
        L.whf_get_tuning.restype = ctypes.c_int
        L.whf_get_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(Tuning)]
        L.whf_set_tuning.restype = ctypes.c_int
        L.whf_set_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(Tuning)]
        L.whf_get_tuning_for.restype = ctypes.c_int
        L.whf_get_tuning_for.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Tuning)]
        L.whf_set_tuning_for.restype = ctypes.c_int
        L.whf_set_tuning_for.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Tuning)]
        L.whf_load_tuning.restype = ctypes.c_int
        L.whf_load_tuning.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(Tuning)]
        L.whf_build_id.restype = ctypes.c_char_p
        L.whf_build_id.argtypes = []
        L.whf_solve.restype = ctypes.c_int
        L.whf_solve.argtypes = [ctypes.c_void_p, ctypes.POINTER(SolveBundle), ctypes.c_uint64, ctypes.c_uint64,
                                ctypes.c_uint64, ctypes.POINTER(SolveSink), ctypes.POINTER(SolveStats)]
        L.whf_solve_chunks.restype = ctypes.c_int
        L.whf_solve_chunks.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int32), ctypes.c_uint64, ctypes.POINTER(SolveSink),
                                       ctypes.POINTER(SolveStats)]
        L.whf_biome_tree.restype = ctypes.c_int
        L.whf_biome_tree.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        L.whf_biome.restype = ctypes.c_int
        L.whf_biome.argtypes = [ctypes.c_void_p, ctypes.POINTER(BiomeQuery), ctypes.c_uint64, ctypes.c_int,
                                ctypes.c_int, ctypes.POINTER(BiomeResult), ctypes.POINTER(ctypes.c_uint8)]
        _lib = L
    return _lib


class WhfError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s: %s (%d)" % (what, load().whf_strerror(code).decode(), code))
        self.code = code


def version_flags(name):
    v = ctypes.c_int(0)
    rc = load().whf_version_flags(name.encode(), ctypes.byref(v))
    if rc:
        raise WhfError(rc, "version %r" % name)
    return v.value


def _solve_stats(st):
    return {f: getattr(st, f) for f, _ in SolveStats._fields_}


def _dd_stats(st, rc, fsink):
    return dict(structure_seeds=st.structure_seeds, tests=st.tests, starts=st.starts, walks=list(st.walks),
                spots=st.spots, draws=st.draws, pieces=st.pieces, expanded_seeds=st.expanded_seeds,
                accepted=st.accepted, funnel=list(st.funnel), tiers=list(st.tiers)[:7],
                samplers_built=list(st.samplers_built), device_ms=st.device_ms, stage_ms=list(st.stage_ms),
                biome_checked=st.biome_checked, biome_rejected=st.biome_rejected,
                overflow=rc == EOVERFLOW, find_count=fsink.count)


class Context:
    """One device context (whf_open / whf_close)."""

    def __init__(self, device=0, profile=False, one_stream=False):
        self.L = load()
        self.h = ctypes.c_void_p()
        cfg = Config(device, (CFG_PROFILE if profile else 0) | (CFG_ONE_STREAM if one_stream else 0), 0)
        rc = self.L.whf_open(ctypes.byref(cfg), ctypes.byref(self.h))
        if rc:
            raise WhfError(rc, "whf_open")

    def close(self):
        if self.h:
            self.L.whf_close(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def device_name(self):
        return self.L.whf_device_name(self.h).decode()

    def tuning(self, vflags=0):
        """The context's launch shapes for a rule layout (whf_get_tuning_for) as a dict."""
        t = Tuning()
        rc = self.L.whf_get_tuning_for(self.h, vflags, ctypes.byref(t))
        if rc:
            raise WhfError(rc, "whf_get_tuning_for")
        return {f: getattr(t, f) for f in TUNING_FIELDS}

    def set_tuning(self, vflags=None, **kw):
        """Sets launch shapes: for one rule layout (whf_set_tuning_for), or every layout when vflags is None
        (whf_set_tuning); fields not given keep their current value."""
        cur = self.tuning(vflags or 0)
        cur.update(kw)
        t = Tuning(*[cur[f] for f in TUNING_FIELDS])
        rc = (self.L.whf_set_tuning(self.h, ctypes.byref(t)) if vflags is None else
              self.L.whf_set_tuning_for(self.h, vflags, ctypes.byref(t)))
        if rc:
            raise WhfError(rc, "whf_set_tuning %r" % kw)

    def scan(self, rect, seeds=None, splitmix=None, counter=None, max_height=-50, vflags=0, capacity=1 << 16,
             salt=0, precision=PREC_FP64, phase2=None):
        """Seed Check. rect = (rx0, rx1, rz0, rz1); seeds = int64 array, or splitmix/counter = (stream, count[,
        offset]); phase2 = (P2_OR_LOWER or P2_EXACTLY, max_y[, tree]) adds lysh's hut verdict. Returns (hits, stats):
        hits are tuples (seed, rx, rz, cx, cz), with the hut's floor appended under phase 2."""
        req, keep = make_request(rect, seeds=seeds, splitmix=splitmix, counter=counter, max_height=max_height,
                                 vflags=vflags, salt=salt, precision=precision, phase2=phase2)
        buf = (Hit * capacity)()
        sink = Sink(buf, capacity, 0)
        st = Stats()
        rc = self.L.whf_scan(self.h, ctypes.byref(req), ctypes.byref(sink), ctypes.byref(st))
        if rc not in (OK, EOVERFLOW):
            raise WhfError(rc, "whf_scan")
        n = min(sink.count, capacity)
        if phase2 is not None:
            hits = sorted((buf[i].seed, buf[i].rx, buf[i].rz, buf[i].cx, buf[i].cz, buf[i].floor_y) for i in range(n))
        else:
            hits = sorted((buf[i].seed, buf[i].rx, buf[i].rz, buf[i].cx, buf[i].cz) for i in range(n))
        stats = dict(seeds=st.seeds, checks=st.checks, accepted=st.accepted, reached_ladder=st.reached_ladder,
                     funnel=list(st.funnel), tiers=list(st.tiers)[:7], samplers_built=list(st.samplers_built),
                     device_ms=st.device_ms, kernel_ms=dict(zip(KERNEL_FAMILIES, list(st.kernel_ms))),
                     hit_count=sink.count, overflow=rc == EOVERFLOW, fp64_rechecks=st.fp64_rechecks,
                     phase2=list(st.phase2), phase2_ms=st.phase2_ms)
        return hits, stats

    def probe(self, seeds, xs, zs, max_height=-50, vflags=0):
        """The 38 raw phase-1 values per (seed, hut_x, hut_z), as uint64 bit patterns (n x 38)."""
        s = np.ascontiguousarray(seeds, dtype=np.int64)
        x = np.ascontiguousarray(xs, dtype=np.int32)
        z = np.ascontiguousarray(zs, dtype=np.int32)
        out = np.zeros((len(s), 38), dtype=np.float64)
        rc = self.L.whf_probe(self.h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                              x.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                              z.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(s), max_height, vflags,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc:
            raise WhfError(rc, "whf_probe")
        return out.view(np.uint64)

    def noise_error(self, seeds, points_per_seed, vflags=0):
        """Max |FP32-gate - FP64| per erosion sampler (8 values) over points spread over the world."""
        s = np.ascontiguousarray(seeds, dtype=np.int64)
        out = (ctypes.c_double * 8)()
        rc = self.L.whf_noise_error(self.h, s.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(s), points_per_seed,
                                    vflags, out)
        if rc:
            raise WhfError(rc, "whf_noise_error")
        return list(out)


    def deepdive(self, s_first, count, rect, floor_max=-55, expand=False, max_height=-50, vflags=0, salt=0,
                 spot_capacity=1 << 16, find_capacity=1 << 16, biome=True):
        """Deep Dive over structure seeds S = s_first + i (mod 2^48), i < count, hut regions rect. biome: True for the
        rule flags' default tree, an int for tree `biome`, False for no swamp check of the expansion's survivors.
        Returns (spots as tuples (s48, rx, rz, cx, cz, carver, room_floor, pieces), finds as tuples
        (seed, rx, rz, cx, cz, room_floor), stats dict)."""
        req = DdRequest(s_first & 0xFFFFFFFFFFFFFFFF, count, rect[0], rect[1], rect[2], rect[3], floor_max,
                        1 if expand else 0, max_height, vflags, salt, _dd_biome(biome))
        sbuf = (DdSpot * spot_capacity)()
        ssink = DdSink(sbuf, spot_capacity, 0)
        fbuf = (Hit * find_capacity)()
        fsink = Sink(fbuf, find_capacity, 0)
        st = DdStats()
        rc = self.L.whf_deepdive(self.h, ctypes.byref(req), ctypes.byref(ssink), ctypes.byref(fsink), ctypes.byref(st))
        if rc not in (OK, EOVERFLOW):
            raise WhfError(rc, "whf_deepdive")
        spots = sorted((p.s48, p.rx, p.rz, p.cx, p.cz, p.carver, p.room_floor, p.pieces)
                       for p in sbuf[:min(ssink.count, spot_capacity)])
        finds = sorted((h.seed, h.rx, h.rz, h.cx, h.cz, h.floor_y) for h in fbuf[:min(fsink.count, find_capacity)])
        return spots, finds, _dd_stats(st, rc, fsink)

    def deepdive_noalloc(self, s_first, count, rect, floor_max=-55, salt=0, spot_capacity=1 << 16):
        """The Deep Dive sweep through whf_deepdive_noalloc (no allocation in the call). Returns (spots as in
        deepdive, stats dict)."""
        req = DdRequest(s_first & 0xFFFFFFFFFFFFFFFF, count, rect[0], rect[1], rect[2], rect[3], floor_max, 0, -50, 0,
                        salt, 0)
        sbuf = (DdSpot * max(1, spot_capacity))()
        ssink = DdSink(sbuf, spot_capacity, 0)
        st = DdStats()
        rc = self.L.whf_deepdive_noalloc(self.h, ctypes.byref(req), ctypes.byref(ssink), ctypes.byref(st))
        if rc not in (OK, EOVERFLOW):
            raise WhfError(rc, "whf_deepdive_noalloc")
        spots = sorted((p.s48, p.rx, p.rz, p.cx, p.cz, p.carver, p.room_floor, p.pieces)
                       for p in sbuf[:min(ssink.count, spot_capacity)])
        return spots, _dd_stats(st, rc, Sink(None, 0, 0))

    def dd_terrain(self, spots, max_height=-50, vflags=0, find_capacity=1 << 16):
        """The --spots terrain stage: spots are (s48, cx, cz). Returns (the spots filled in, as tuples
        (s48, rx, rz, cx, cz, carver, room_floor, pieces), in input order, pieces = 0 where the spot is not a
        mineshaft start; finds as tuples (seed, rx, rz, cx, cz, room_floor), sorted; stats dict)."""
        req = DdRequest(0, 0, 0, 1, 0, 1, 0, 1, max_height, vflags, 0, 0)
        n = len(spots)
        sbuf = (DdSpot * max(1, n))()
        for i, (s48, cx, cz) in enumerate(spots):
            sbuf[i].s48 = s48 & 0xFFFFFFFFFFFFFFFF
            sbuf[i].cx = cx
            sbuf[i].cz = cz
        fbuf = (Hit * find_capacity)()
        fsink = Sink(fbuf, find_capacity, 0)
        st = DdStats()
        rc = self.L.whf_dd_terrain(self.h, ctypes.byref(req), sbuf, n, ctypes.byref(fsink), ctypes.byref(st))
        if rc not in (OK, EOVERFLOW):
            raise WhfError(rc, "whf_dd_terrain")
        out = [(p.s48, p.rx, p.rz, p.cx, p.cz, p.carver, p.room_floor, p.pieces) for p in sbuf[:n]]
        finds = sorted((h.seed, h.rx, h.rz, h.cx, h.cz, h.floor_y) for h in fbuf[:min(fsink.count, find_capacity)])
        return out, finds, _dd_stats(st, rc, fsink)

    def solve(self, bundles, first_group, n_groups, capacity=1 << 22):
        """The Deep Dive solver (whf_solve) over groups [first_group, first_group + n_groups) of the plan bundles
        [(X, regime_x, regime_z, tier)]: (hits as sorted tuples (bundle, S, cx, cz, X), stats dict)."""
        arr = (SolveBundle * len(bundles))()
        for i, (x, rgx, rgz, tier) in enumerate(bundles):
            arr[i] = SolveBundle(x, rgx, rgz, tier, 0)
        buf = (SolveHit * capacity)()
        sink = SolveSink(buf, capacity, 0)
        st = SolveStats()
        rc = self.L.whf_solve(self.h, arr, len(bundles), first_group, n_groups, ctypes.byref(sink), ctypes.byref(st))
        if rc:
            raise WhfError(rc, "whf_solve")
        return [(h.bundle, h.s48, h.cx, h.cz, h.layout) for h in buf[:sink.count]], _solve_stats(st)

    def solve_chunks(self, layout, chunks, capacity=1 << 20):
        """The solver's search at given chunks [(cx, cz)] of one layout (whf_solve_chunks): (hits as tuples
        (chunk index, S, cx, cz) sorted, stats dict)."""
        n = len(chunks)
        cx = np.ascontiguousarray([c[0] for c in chunks], dtype=np.int32)
        cz = np.ascontiguousarray([c[1] for c in chunks], dtype=np.int32)
        buf = (SolveHit * capacity)()
        sink = SolveSink(buf, capacity, 0)
        st = SolveStats()
        p32 = ctypes.POINTER(ctypes.c_int32)
        rc = self.L.whf_solve_chunks(self.h, layout, cx.ctypes.data_as(p32), cz.ctypes.data_as(p32), n,
                                     ctypes.byref(sink), ctypes.byref(st))
        if rc:
            raise WhfError(rc, "whf_solve_chunks")
        return [(h.bundle, h.s48, h.cx, h.cz) for h in buf[:sink.count]], _solve_stats(st)

    def room_floors(self, carvers):
        """Full layout walks of the given carver seeds: arrays (status, room_floor, pieces)."""
        c = np.ascontiguousarray(carvers, dtype=np.uint64)
        n = len(c)
        st = np.zeros(n, np.int32)
        fl = np.zeros(n, np.int32)
        pc = np.zeros(n, np.int32)
        p32 = ctypes.POINTER(ctypes.c_int32)
        rc = self.L.whf_room_floors(self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n,
                                    st.ctypes.data_as(p32), fl.ctypes.data_as(p32), pc.ctypes.data_as(p32))
        if rc:
            raise WhfError(rc, "whf_room_floors")
        return st, fl, pc


    def biome(self, queries, tree, vflags=0, winners=False):
        """Swamp-hut biome checks of (seed, hut_x, hut_z, occ_y) queries against biome tree `tree`: a list of result
        dicts (targets, quart position, winner, ok, swamp quart heights as a 97-bit int), and with winners=True
        also the winner at every quart height, an (n, 97) uint8 array."""
        n = len(queries)
        q = (BiomeQuery * max(n, 1))()
        for i, (seed, hx, hz, occ) in enumerate(queries):
            q[i].seed = _s64(seed)
            q[i].hut_x, q[i].hut_z, q[i].occ_y = hx, hz, occ
        res = (BiomeResult * max(n, 1))()
        win = np.zeros((max(n, 1), BIOME_QY_N), np.uint8) if winners else None
        rc = self.L.whf_biome(self.h, q, n, tree, vflags, res,
                              win.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)) if winners else None)
        if rc:
            raise WhfError(rc, "whf_biome")
        out = [dict(target=list(r.target), qx=r.qx, qy=r.qy, qz=r.qz, occ_y=r.occ_y, winner=r.winner, ok=r.ok,
                    swamp_qy=r.swamp_qy[0] | r.swamp_qy[1] << 64) for r in res[:n]]
        return (out, win[:n]) if winners else out


    def surface_heights(self, cols, vflags=0):
        """Pre-carver surface heights of (seed, x, z) columns: arrays (height, top block state)."""
        n = len(cols)
        q = (Column * max(n, 1))()
        for i, (seed, x, z) in enumerate(cols):
            q[i].seed = _s64(seed)
            q[i].x, q[i].z = x, z
        h = np.zeros(max(n, 1), np.int32)
        st = np.zeros(max(n, 1), np.int32)
        p32 = ctypes.POINTER(ctypes.c_int32)
        rc = self.L.whf_surface_heights(self.h, q, n, vflags, h.ctypes.data_as(p32), st.ctypes.data_as(p32))
        if rc:
            raise WhfError(rc, "whf_surface_heights")
        return h[:n], st[:n]


    def post_carver_heights(self, cols, vflags=0):
        """Post- and pre-carver surface heights of (seed, x, z) columns: arrays (post, pre)."""
        n = len(cols)
        q = (Column * max(n, 1))()
        for i, (seed, x, z) in enumerate(cols):
            q[i].seed = _s64(seed)
            q[i].x, q[i].z = x, z
        post = np.zeros(max(n, 1), np.int32)
        pre = np.zeros(max(n, 1), np.int32)
        p32 = ctypes.POINTER(ctypes.c_int32)
        rc = self.L.whf_post_carver_heights(self.h, q, n, vflags, post.ctypes.data_as(p32), pre.ctypes.data_as(p32))
        if rc:
            raise WhfError(rc, "whf_post_carver_heights")
        return post[:n], pre[:n]

    def hut_eval(self, huts, tree, max_y=-50, vflags=0):
        """lysh's hut verdicts for (seed, cx, cz) huts: a list of result dicts."""
        n = len(huts)
        q = (HutQuery * max(n, 1))()
        for i, (seed, cx, cz) in enumerate(huts):
            q[i].seed = _s64(seed)
            q[i].cx, q[i].cz = cx, cz
        res = (HutResult * max(n, 1))()
        rc = self.L.whf_hut_eval(self.h, q, n, tree, vflags, max_y, res)
        if rc:
            raise WhfError(rc, "whf_hut_eval")
        return [{f: getattr(r, f) for f, _ in HutResult._fields_} for r in res[:n]]


def biome_tree(version):
    t = ctypes.c_int(0)
    rc = load().whf_biome_tree(version.encode(), ctypes.byref(t))
    if rc:
        raise WhfError(rc, "whf_biome_tree")
    return t.value


def _dd_biome(biome):
    """whf_dd_request.biome_tree of a `biome` argument: True = the rules' default tree, False = none, N = tree N."""
    if biome is True:
        return 0
    if biome is False or biome is None:
        return -1
    return int(biome) + 1


def _s64(v):
    v &= 0xFFFFFFFFFFFFFFFF
    return v - (1 << 64) if v >= 1 << 63 else v


def make_request(rect, seeds=None, splitmix=None, counter=None, max_height=-50, vflags=0, salt=0,
                 precision=PREC_FP64, phase2=None):
    """A whf_request and the array it points to (keep both alive while the request runs)."""
    req = Request()
    req.precision = precision
    keep = None
    if seeds is not None:
        keep = np.ascontiguousarray(seeds, dtype=np.int64)
        req.seed_mode = SEEDS_LIST
        req.seeds = keep.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        req.n_seeds = len(keep)
    else:
        gen = splitmix if splitmix is not None else counter
        req.seed_mode = SEEDS_SPLITMIX if splitmix is not None else SEEDS_COUNTER
        req.stream = gen[0] & 0xFFFFFFFFFFFFFFFF
        req.n_seeds = gen[1]
        req.offset = gen[2] if len(gen) > 2 else 0
    req.rx0, req.rx1, req.rz0, req.rz1 = rect
    req.max_height = max_height
    req.vflags = vflags
    req.salt = salt
    if phase2 is not None:   # (mode, max_y, tree or None)
        req.phase2, req.phase2_max_y = phase2[0], phase2[1]
        req.biome_tree = 0 if len(phase2) < 3 or phase2[2] is None else phase2[2] + 1
    return req, keep


def stats_dict(st):
    return dict(seeds=st.seeds, checks=st.checks, accepted=st.accepted, reached_ladder=st.reached_ladder,
                funnel=list(st.funnel), tiers=list(st.tiers)[:7], samplers_built=list(st.samplers_built),
                device_ms=st.device_ms, fp64_rechecks=st.fp64_rechecks, phase2=list(st.phase2))


SERVE_SLOTS = 1   # whf_server_config.flags: the host-driven slot scheduler instead of the device-resident loop


class Server:
    """Continuous batching (whf_server_open / whf_submit / whf_poll / whf_server_close). By default the
    device-resident loop; slots_mode=True selects the host-driven slot scheduler (slots contexts)."""

    def __init__(self, device=0, slots=4, chunk_checks=0, hbm_budget=0, slots_mode=False):
        self.L = load()
        self.h = ctypes.c_void_p()
        cfg = ServerConfig(device, slots, chunk_checks, hbm_budget, SERVE_SLOTS if slots_mode else 0, 0)
        rc = self.L.whf_server_open(ctypes.byref(cfg), ctypes.byref(self.h))
        if rc:
            raise WhfError(rc, "whf_server_open")
        self.pending = {}

    def submit(self, rect, capacity=4096, **kw):
        """Queues a Seed Check request (arguments as Context.scan); returns its id."""
        req, keep = make_request(rect, **kw)
        buf = (Hit * capacity)()
        sink = Sink(buf, capacity, 0)
        rid = ctypes.c_uint64(0)
        rc = self.L.whf_submit(self.h, ctypes.byref(req), ctypes.byref(sink), ctypes.byref(rid))
        if rc:
            raise WhfError(rc, "whf_submit")
        self.pending[rid.value] = (req, keep, buf, sink, kw.get("phase2") is not None)
        return rid.value

    def submit_deepdive(self, s_first, count, rect, floor_max=-55, expand=True, max_height=-50, vflags=0, salt=0,
                        spot_capacity=1 << 20, find_capacity=1 << 16, biome=True):
        """Queues a background Deep Dive; its event (from poll) carries spots, finds and dd_stats."""
        req = DdRequest(s_first & 0xFFFFFFFFFFFFFFFF, count, rect[0], rect[1], rect[2], rect[3], floor_max,
                        1 if expand else 0, max_height, vflags, salt, _dd_biome(biome))
        sbuf = (DdSpot * spot_capacity)()
        ssink = DdSink(sbuf, spot_capacity, 0)
        fbuf = (Hit * find_capacity)()
        fsink = Sink(fbuf, find_capacity, 0)
        st = DdStats()
        rid = ctypes.c_uint64(0)
        rc = self.L.whf_submit_deepdive(self.h, ctypes.byref(req), ctypes.byref(ssink), ctypes.byref(fsink),
                                        ctypes.byref(st), ctypes.byref(rid))
        if rc:
            raise WhfError(rc, "whf_submit_deepdive")
        self.pending[rid.value] = ("dd", req, sbuf, ssink, fbuf, fsink, st)
        return rid.value

    def poll(self, timeout_ms=-1, max_events=256):
        """Completed requests: list of dicts (id, status, hits, stats, submit_s, start_s, done_s); a Deep Dive's
        has spots, finds and dd_stats instead of hits and stats."""
        ev = (Event * max_events)()
        n = self.L.whf_poll(self.h, ev, max_events, timeout_ms)
        if n < 0:
            raise WhfError(n, "whf_poll")
        out = []
        for e in ev[:n]:
            p = self.pending.pop(e.id)
            if p[0] == "dd":
                _, req, sbuf, ssink, fbuf, fsink, st = p
                spots = sorted((q.s48, q.rx, q.rz, q.cx, q.cz, q.carver, q.room_floor, q.pieces)
                               for q in sbuf[:min(ssink.count, ssink.capacity)])
                finds = sorted((h.seed, h.rx, h.rz, h.cx, h.cz, h.floor_y) for h in fbuf[:min(fsink.count, fsink.capacity)])
                out.append(dict(id=e.id, status=e.status, deepdive=True, spots=spots, finds=finds,
                                dd_stats=dict(structure_seeds=st.structure_seeds, starts=st.starts, walks=list(st.walks),
                                              spots=st.spots, expanded_seeds=st.expanded_seeds, accepted=st.accepted,
                                              funnel=list(st.funnel), tiers=list(st.tiers)[:7],
                                              biome_checked=st.biome_checked, biome_rejected=st.biome_rejected),
                                submit_s=e.submit_s, start_s=e.start_s, done_s=e.done_s))
                continue
            req, keep, buf, sink, p2 = p
            m = min(sink.count, sink.capacity)
            if p2:
                hits = sorted((buf[i].seed, buf[i].rx, buf[i].rz, buf[i].cx, buf[i].cz, buf[i].floor_y)
                              for i in range(m))
            else:
                hits = sorted((buf[i].seed, buf[i].rx, buf[i].rz, buf[i].cx, buf[i].cz) for i in range(m))
            out.append(dict(id=e.id, status=e.status, hits=hits, stats=stats_dict(e.stats), submit_s=e.submit_s,
                            start_s=e.start_s, done_s=e.done_s))
        return out

    def close(self):
        if self.h:
            self.L.whf_server_close(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def fp32_bound():
    return load().whf_fp32_bound()


def build_id():
    return load().whf_build_id().decode()


def load_tuning(path, device, vflags=0):
    """The launch shapes whf_open would apply on `device` from the file at `path`: (applied, dict)."""
    t = Tuning()
    rc = load().whf_load_tuning(path.encode() if path else None, device.encode(), vflags, ctypes.byref(t))
    return rc == OK, {f: getattr(t, f) for f in TUNING_FIELDS}
