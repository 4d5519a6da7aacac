"""ctypes binding of the CPU oracle (oracle/_build/libwhf_oracle.so) for the tests.

The oracle is the checker: tests, the smoke check and the benchmark's CPU-baseline leg use it to decide
whether the HIP path is right. Nothing in the product path imports this module.
"""
import ctypes
import os
import struct

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "_build", "libwhf_oracle.so")

KIND_N = 11
V_118, V_262, V_LARGE, V_SINGLE = 1, 2, 4, 8

(N_TEMPERATURE, N_VEGETATION, N_CONTINENTALNESS, N_EROSION, N_TEMPERATURE_LARGE, N_VEGETATION_LARGE,
 N_CONTINENTALNESS_LARGE, N_EROSION_LARGE, N_RIDGE, N_FLOODEDNESS, N_CAVE_ENTRANCE, N_SPAG3D_RARITY,
 N_SPAG3D_THICKNESS, N_SPAG3D_1, N_SPAG3D_2, N_SPAG_ROUGH_MOD, N_SPAG_ROUGH, N_CAVE_LAYER, N_CAVE_CHEESE,
 N_OFFSET) = range(20)


class Funnel(ctypes.Structure):
    _fields_ = [("checks", ctypes.c_uint64), ("accepted", ctypes.c_uint64), ("reached_ladder", ctypes.c_uint64),
                ("funnel", ctypes.c_uint64 * KIND_N), ("tiers", ctypes.c_uint64 * 8)]


class Hit(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int64), ("rx", ctypes.c_int32), ("rz", ctypes.c_int32),
                ("cx", ctypes.c_int32), ("cz", ctypes.c_int32)]


class MsPiece(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("x0", "y0", "z0", "x1", "y1", "z1")] + \
               [(n, ctypes.c_int8) for n in ("type", "facing", "depth", "rails")] + [("parent", ctypes.c_int16)]


class MsResult(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("status", "pieces", "span_y", "min_y", "max_y", "room_floor")] + \
               [("draws", ctypes.c_uint64), ("overlap_tests", ctypes.c_uint64)]


class DdStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("structure_seeds", "tests", "starts", "survivors")] + \
               [("status", ctypes.c_uint64 * 3), ("draws", ctypes.c_uint64), ("pieces", ctypes.c_uint64)]


class DdHit(ctypes.Structure):
    _fields_ = [("s48", ctypes.c_uint64)] + [(n, ctypes.c_int32) for n in ("rx", "rz", "cx", "cz")] + \
               [("carver", ctypes.c_uint64), ("room_floor", ctypes.c_int32), ("pieces", ctypes.c_int32)]


class BiomeResult(ctypes.Structure):
    _fields_ = [("t", ctypes.c_int64 * 6)] + [(n, ctypes.c_int) for n in ("qx", "qy", "qz", "occ_y", "winner", "ok")]


class HutResult(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("dir", "size_x", "size_z", "biome_ok", "winner", "occ_y", "wet", "flooded",
                                             "carve_skipped")] + \
               [("sum_h", ctypes.c_longlong)] + [(n, ctypes.c_int) for n in ("avg_y", "ok", "reject")]


class XR(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(ORACLE_SO)
        u64, i64, i32, dbl = ctypes.c_uint64, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        L.wo_splitmix64_at.restype = u64
        L.wo_splitmix64_at.argtypes = [u64, u64]
        L.wo_scramble_index.restype = u64
        L.wo_scramble_index.argtypes = [u64]
        L.wo_md5.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p]
        L.wo_xr_seed.argtypes = [ctypes.POINTER(XR), u64]
        L.wo_xr_next.restype = u64
        L.wo_xr_next.argtypes = [ctypes.POINTER(XR)]
        L.wo_xr_next_int.restype = i32
        L.wo_xr_next_int.argtypes = [ctypes.POINTER(XR), i32]
        L.wo_xr_next_double.restype = dbl
        L.wo_xr_next_double.argtypes = [ctypes.POINTER(XR)]
        L.wo_lcg_next_int_seeded.restype = i32
        L.wo_lcg_next_int_seeded.argtypes = [ctypes.POINTER(u64), i32]
        L.wo_lcg_next_long.restype = u64
        L.wo_lcg_next_long.argtypes = [ctypes.POINTER(u64)]
        L.wo_hut_chunk.argtypes = [u64, i32, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.wo_version_flags.restype = i32
        L.wo_version_flags.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32)]
        L.wo_scan.restype = i32
        L.wo_scan.argtypes = [ctypes.POINTER(i64), i64, i32, i32, i32, i32, i32, i32, i32, ctypes.POINTER(Funnel),
                              ctypes.POINTER(Hit), i64, ctypes.POINTER(i64)]
        L.wo_scan_regions.restype = i32
        L.wo_scan_regions.argtypes = [i64, i32, i32, i32, i32, i32, i32, ctypes.c_void_p, ctypes.c_void_p]
        L.wo_sampler_fingerprint.restype = u64
        L.wo_sampler_fingerprint.argtypes = [u64, i32, i32, i32]
        L.wo_probe_alloc.restype = ctypes.c_void_p
        L.wo_probe_alloc.argtypes = [u64, i32]
        L.wo_probe_free.argtypes = [ctypes.c_void_p]
        L.wo_phase1_probe.argtypes = [ctypes.c_void_p, i32, i32, i32, ctypes.POINTER(dbl)]
        L.wo_noise_alloc.restype = ctypes.c_void_p
        L.wo_noise_alloc.argtypes = [u64, i32]
        L.wo_noise_amplitude.restype = dbl
        L.wo_noise_amplitude.argtypes = [ctypes.c_void_p]
        L.wo_noise_sample.restype = dbl
        L.wo_noise_sample.argtypes = [ctypes.c_void_p, dbl, dbl, dbl]
        L.wo_noise_sample_y0.restype = dbl
        L.wo_noise_sample_y0.argtypes = [ctypes.c_void_p, dbl, dbl]
        L.wo_erosion_tiered.restype = i32
        L.wo_erosion_tiered.argtypes = [ctypes.c_void_p, dbl, dbl, ctypes.POINTER(dbl), ctypes.POINTER(i32)]
        L.wo_noise_perm.argtypes = [ctypes.c_void_p, i32, i32, ctypes.c_char_p]
        L.wo_ms_carver_seed.restype = u64
        L.wo_ms_carver_seed.argtypes = [u64, ctypes.c_int32, ctypes.c_int32]
        L.wo_ms_start_double.restype = dbl
        L.wo_ms_start_double.argtypes = [u64]
        L.wo_ms_is_start.restype = i32
        L.wo_ms_is_start.argtypes = [u64]
        L.wo_ms_layout.restype = i32
        L.wo_ms_layout.argtypes = [u64, ctypes.c_int32, ctypes.c_int32, i32, ctypes.POINTER(MsPiece), i32,
                                   ctypes.POINTER(MsResult)]
        L.wo_ms_room_floor.restype = i32
        L.wo_ms_room_floor.argtypes = [u64, i32, ctypes.POINTER(MsResult)]
        L.wo_deepdive_sweep.restype = i32
        L.wo_deepdive_sweep.argtypes = [u64, u64, i32, i32, i32, i32, i32, i32, i32, ctypes.POINTER(DdStats),
                                        ctypes.POINTER(DdHit), i64, ctypes.POINTER(i64)]
        L.wo_biome_alloc.restype = ctypes.c_void_p
        L.wo_biome_alloc.argtypes = [u64, i32]
        L.wo_biome_point.argtypes = [ctypes.c_void_p, i32, i32, ctypes.POINTER(dbl)]
        L.wo_biome_check.restype = i32
        L.wo_biome_check.argtypes = [ctypes.c_void_p, i32, i32, i32, i32, ctypes.POINTER(BiomeResult)]
        L.wo_biome_heights.argtypes = [ctypes.c_void_p, i32, i32, i32, ctypes.c_char_p]
        L.wo_biome_tree.restype = i32
        L.wo_biome_tree.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32)]
        L.wo_terrain_spline.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]
        L.wo_depth_gradient.restype = dbl
        L.wo_depth_gradient.argtypes = [dbl]
        i64a = ctypes.POINTER(i64)
        L.wo_tree_dist.restype = i64
        L.wo_tree_dist.argtypes = [i32, i32, i64a]
        L.wo_tree_leaf.restype = i32
        L.wo_tree_leaf.argtypes = [i32, i64a]
        L.wo_tree_biome.restype = i32
        L.wo_tree_biome.argtypes = [i32, i64a]
        L.wo_tree_min_dist.restype = i64
        L.wo_tree_min_dist.argtypes = [i32, i64a, ctypes.POINTER(i32)]
        L.wo_p2_alloc.restype = ctypes.c_void_p
        L.wo_p2_alloc.argtypes = [u64, i32]
        L.wo_p2_terrain_point.argtypes = [ctypes.c_void_p, i32, i32, ctypes.POINTER(ctypes.c_float)]
        L.wo_p2_blended.restype = dbl
        L.wo_p2_blended.argtypes = [ctypes.c_void_p, i32, i32, i32]
        L.wo_p2_noodle_child.restype = dbl
        L.wo_p2_noodle_child.argtypes = [ctypes.c_void_p, i32, i32, i32, i32]
        L.wo_p2_corner_column.argtypes = [ctypes.c_void_p, i32, i32, ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
        L.wo_p2_density.restype = dbl
        L.wo_p2_density.argtypes = [ctypes.c_void_p, i32, i32, i32]
        L.wo_p2_prelim_surface.restype = i32
        L.wo_p2_prelim_surface.argtypes = [ctypes.c_void_p, i32, i32]
        L.wo_p2_skip_above.restype = i32
        L.wo_p2_skip_above.argtypes = [ctypes.c_void_p, i32, i32]
        L.wo_p2_fluid.argtypes = [ctypes.c_void_p, i32, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.wo_p2_substance.restype = i32
        L.wo_p2_substance.argtypes = [ctypes.c_void_p, i32, i32, i32, dbl]
        L.wo_p2_height.restype = i32
        L.wo_p2_height.argtypes = [ctypes.c_void_p, i32, i32, ctypes.POINTER(i32)]
        L.wo_set_cube_mode.restype = None
        L.wo_set_cube_mode.argtypes = [ctypes.c_int]
        L.wo_cube_stats.restype = None
        L.wo_cube_stats.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)]
        L.wo_post_carver_height.restype = i32
        L.wo_post_carver_height.argtypes = [ctypes.c_void_p, u64, i32, i32]
        L.wo_hut_orientation.restype = i32
        L.wo_hut_orientation.argtypes = [u64, i32, i32]
        L.wo_hut_eval.restype = i32
        L.wo_hut_eval.argtypes = [ctypes.c_void_p, i32, u64, i32, i32, i32, ctypes.POINTER(HutResult)]
        L.wo_solve_perm_group.restype = ctypes.c_uint32
        L.wo_solve_perm_group.argtypes = [ctypes.c_uint32]
        L.wo_solve_creg.restype = u64
        L.wo_solve_creg.argtypes = [ctypes.c_int32, ctypes.c_int32]
        L.wo_solve_regions.restype = i32
        L.wo_solve_regions.argtypes = [ctypes.c_uint32, i32, i32, ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int32), i32]
        L.wo_solve_offsets.restype = i32
        L.wo_solve_offsets.argtypes = [i32, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint8)]
        L.wo_solve_task.restype = i32
        L.wo_solve_task.argtypes = [u64, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(u64), i32]
        L.wo_solve_prefixes.restype = i32
        L.wo_solve_prefixes.argtypes = [u64, ctypes.c_int32, ctypes.c_int32, i32, ctypes.POINTER(u64), i32]
        L.wo_solve_tasks.restype = i64
        L.wo_solve_tasks.argtypes = [u64, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), i32, i32, i32,
                                     ctypes.POINTER(u64), ctypes.POINTER(i32)]
        _lib = L
    return _lib


def u64(v):
    return v & 0xFFFFFFFFFFFFFFFF


def s64(v):
    v &= 0xFFFFFFFFFFFFFFFF
    return v - (1 << 64) if v >= 1 << 63 else v


def splitmix(stream, index):
    return lib().wo_splitmix64_at(u64(stream), index)


def seeds_splitmix(stream, count, offset=0):
    return np.array([s64(splitmix(stream, offset + i)) for i in range(count)], dtype=np.int64)


def version_flags(name):
    v = ctypes.c_int(0)
    if lib().wo_version_flags(name.encode(), ctypes.byref(v)):
        raise ValueError("unknown version %r" % name)
    return v.value


def scan(seeds, rect, max_y=-50, vflags=0, threads=8, cap=1 << 16):
    seeds = np.ascontiguousarray(seeds, dtype=np.int64)
    f = Funnel()
    hits = (Hit * cap)()
    nh = ctypes.c_int64(0)
    rc = lib().wo_scan(seeds.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(seeds), rect[0], rect[1], rect[2],
                       rect[3], max_y, vflags, threads, ctypes.byref(f), hits, cap, ctypes.byref(nh))
    assert rc == 0
    out = [(hits[i].seed, hits[i].rx, hits[i].rz, hits[i].cx, hits[i].cz) for i in range(min(nh.value, cap))]
    return dict(checks=f.checks, accepted=f.accepted, reached_ladder=f.reached_ladder, funnel=list(f.funnel),
                tiers=list(f.tiers)[:7], hits=sorted(out))


def scan_regions(seed, rect, max_y=-50, vflags=0):
    n = (rect[1] - rect[0]) * (rect[3] - rect[2])
    kind = np.zeros(n, np.int8)
    tier = np.zeros(n, np.int8)
    assert lib().wo_scan_regions(s64(seed), rect[0], rect[1], rect[2], rect[3], max_y, vflags,
                                 kind.ctypes.data, tier.ctypes.data) == 0
    return kind, tier


def hut_chunk(seed, rx, rz, salt=14357620):
    cx, cz = ctypes.c_int(0), ctypes.c_int(0)
    lib().wo_hut_chunk(u64(seed), rx, rz, salt, ctypes.byref(cx), ctypes.byref(cz))
    return cx.value, cz.value


def probe(seed, x, z, max_y, vflags=0):
    L = lib()
    h = L.wo_probe_alloc(u64(seed), vflags)
    out = (ctypes.c_double * 38)()
    L.wo_phase1_probe(h, x, z, max_y, out)
    L.wo_probe_free(h)
    return [struct.unpack("<Q", struct.pack("<d", v))[0] for v in out]


class Noise:
    def __init__(self, seed, noise_id):
        self.h = lib().wo_noise_alloc(u64(seed), noise_id)

    def __del__(self):
        if getattr(self, "h", None):
            lib().wo_probe_free(self.h)

    @property
    def amplitude(self):
        return lib().wo_noise_amplitude(self.h)

    def sample(self, x, y, z):
        return lib().wo_noise_sample(self.h, x, y, z)

    def sample_y0(self, x, z):
        return lib().wo_noise_sample_y0(self.h, x, z)

    def tiered(self, x, z):
        v, t = ctypes.c_double(0), ctypes.c_int(0)
        fired = lib().wo_erosion_tiered(self.h, x, z, ctypes.byref(v), ctypes.byref(t))
        return fired, t.value, v.value

    def perm(self, sub, slot):
        buf = ctypes.create_string_buffer(256)
        lib().wo_noise_perm(self.h, sub, slot, buf)
        return bytes(buf.raw)


def md5(data):
    out = ctypes.create_string_buffer(16)
    lib().wo_md5(data, len(data), out)
    return out.raw.hex()


# ---- Deep Dive (whf_oracle_ms.c) ----
NO_CUT = -(1 << 31)
MS_KIND = {0: "R", 1: "C", 3: "S"}


def carver_seed(s48, cx, cz):
    return lib().wo_ms_carver_seed(u64(s48), cx, cz)


def is_start(carver):
    return bool(lib().wo_ms_is_start(u64(carver)))


def start_double(carver):
    return lib().wo_ms_start_double(u64(carver))


def ms_layout(carver, cx, cz, cut=NO_CUT):
    """(status, room_floor, pieces as mstree lines, result struct) for the mineshaft started by `carver` in
    chunk (cx, cz); the lines follow mstree's emit format (kind x0 y0 z0 x1 y1 z1 facing parent rails)."""
    buf = (MsPiece * 512)()
    r = MsResult()
    st = lib().wo_ms_layout(u64(carver), cx, cz, cut, buf, 512, ctypes.byref(r))
    lines = []
    for i in range(r.pieces):
        p = buf[i]
        kind = ("Y" if p.y1 - p.y0 + 1 > 3 else "X") if p.type == 2 else MS_KIND[p.type]
        lines.append("%s %d %d %d %d %d %d %d %d %d" % (kind, p.x0, p.y0, p.z0, p.x1, p.y1, p.z1, p.facing,
                                                        p.parent, p.rails))
    return st, r.room_floor, lines, r


def room_floor(carver, cut=NO_CUT):
    r = MsResult()
    st = lib().wo_ms_room_floor(u64(carver), cut, ctypes.byref(r))
    return st, r.room_floor, r.pieces, r


def deepdive_sweep(s_first, count, rect, floor_max=-55, cut=NO_CUT, threads=8, cap=1 << 16):
    st = DdStats()
    hits = (DdHit * cap)()
    n = ctypes.c_int64(0)
    rc = lib().wo_deepdive_sweep(u64(s_first), count, rect[0], rect[1], rect[2], rect[3], floor_max, cut, threads,
                                 ctypes.byref(st), hits, cap, ctypes.byref(n))
    assert rc == 0
    spots = sorted((h.s48, h.rx, h.rz, h.cx, h.cz, h.carver, h.room_floor, h.pieces)
                   for h in hits[:min(n.value, cap)])
    return dict(structure_seeds=st.structure_seeds, tests=st.tests, starts=st.starts, survivors=st.survivors,
                status=list(st.status), draws=st.draws, pieces=st.pieces, spots=spots)


# ---- biome stage (whf_oracle_biome.c) ----
BIOME_QY0, BIOME_QY_N = -17, 97
SWAMP = 6


def _np6(t):
    return (ctypes.c_int64 * 6)(*[int(v) for v in t])


class Biome:
    """The six climate noises of a world seed, for the biome check."""

    def __init__(self, seed, vflags=0):
        self.h = lib().wo_biome_alloc(u64(seed), vflags)

    def __del__(self):
        if getattr(self, "h", None):
            lib().wo_probe_free(self.h)

    def point(self, qx, qz):
        out = (ctypes.c_double * 8)()
        lib().wo_biome_point(self.h, qx, qz, out)
        return list(out)

    def check(self, tree, hut_x, hut_z, occ_y):
        r = BiomeResult()
        lib().wo_biome_check(self.h, tree, hut_x, hut_z, occ_y, ctypes.byref(r))
        return dict(t=list(r.t), qx=r.qx, qy=r.qy, qz=r.qz, occ_y=r.occ_y, winner=r.winner, ok=r.ok)

    def heights(self, tree, hut_x, hut_z):
        buf = ctypes.create_string_buffer(BIOME_QY_N)
        lib().wo_biome_heights(self.h, tree, hut_x, hut_z, buf)
        return list(buf.raw)


def biome_tree(version):
    t = ctypes.c_int(0)
    if lib().wo_biome_tree(version.encode(), ctypes.byref(t)):
        raise ValueError(version)
    return t.value


def terrain_spline(c, e, w):
    out = (ctypes.c_float * 3)()
    lib().wo_terrain_spline(c, e, w, out)
    return list(out)


def depth_gradient(y):
    return lib().wo_depth_gradient(y)


def tree_leaf(tree, t):
    return lib().wo_tree_leaf(tree, _np6(t))


def tree_biome(tree, t):
    return lib().wo_tree_biome(tree, _np6(t))


def tree_dist(tree, node, t):
    return lib().wo_tree_dist(tree, node, _np6(t))


def tree_min_dist(tree, t):
    n = ctypes.c_int(0)
    d = lib().wo_tree_min_dist(tree, _np6(t), ctypes.byref(n))
    return d, n.value


# ---- terrain stage (whf_oracle_p2.c) ----
SOLID, AIR, WATER, LAVA = 0, 1, 2, 3


class Terrain:
    """The phase-2 noises of a world seed: final density, preliminary surface, aquifer, surface height."""

    def __init__(self, seed, vflags=0):
        self.seed = seed
        self.h = lib().wo_p2_alloc(u64(seed), vflags)

    def __del__(self):
        if getattr(self, "h", None):
            lib().wo_probe_free(self.h)

    def point(self, qx, qz):
        out = (ctypes.c_float * 3)()
        lib().wo_p2_terrain_point(self.h, qx, qz, out)
        return list(out)

    def blended(self, x, y, z):
        return lib().wo_p2_blended(self.h, x, y, z)

    def noodle(self, which, x, y, z):
        return lib().wo_p2_noodle_child(self.h, which, x, y, z)

    def corner_column(self, cx, cz):
        main = (ctypes.c_double * 49)()
        nood = (ctypes.c_double * 147)()
        lib().wo_p2_corner_column(self.h, cx, cz, main, nood)
        return list(main), [list(nood[49 * c:49 * c + 49]) for c in range(3)]

    def density(self, x, y, z):
        return lib().wo_p2_density(self.h, x, y, z)

    def prelim_surface(self, x, z):
        return lib().wo_p2_prelim_surface(self.h, x, z)

    def skip_above(self, cx, cz):
        return lib().wo_p2_skip_above(self.h, cx, cz)

    def fluid(self, x, y, z):
        lv, ty = ctypes.c_int(0), ctypes.c_int(0)
        lib().wo_p2_fluid(self.h, x, y, z, ctypes.byref(lv), ctypes.byref(ty))
        return lv.value, ty.value

    def substance(self, x, y, z, density):
        return lib().wo_p2_substance(self.h, x, y, z, density)

    def height(self, x, z):
        st = ctypes.c_int(0)
        h = lib().wo_p2_height(self.h, x, z, ctypes.byref(st))
        return h, st.value

    def post_carver_height(self, x, z):
        return lib().wo_post_carver_height(self.h, u64(self.seed), x, z)

    def hut_eval(self, tree, hut_x, hut_z, max_y=-50):
        r = HutResult()
        lib().wo_hut_eval(self.h, tree, u64(self.seed), hut_x, hut_z, max_y, ctypes.byref(r))
        return {n: getattr(r, n) for n, _ in HutResult._fields_}


WO_HUT_OK, WO_HUT_REJECT_Y, WO_HUT_REJECT_FLOOD, WO_HUT_REJECT_BIOME = 0, 1, 2, 3


def hut_orientation(seed, hut_x, hut_z):
    return lib().wo_hut_orientation(u64(seed), hut_x, hut_z)


def set_cube_mode(mode):
    """The oracle's pillar cube: 0 = the C library's pow as the reference, 1 = the device's compensated cube."""
    lib().wo_set_cube_mode(mode)


def cube_stats():
    """(cubes evaluated in mode 1, cubes differing from pow) since the last call."""
    out = (ctypes.c_ulonglong * 2)()
    lib().wo_cube_stats(out)
    return out[0], out[1]


# ---- Deep Dive solver (whf_oracle_solve.c) ----

def perm_group(p):
    """The group index's 20-bit region-constant residue (taskspace.h:24)."""
    return lib().wo_solve_perm_group(p)


def creg(rx, rz):
    return lib().wo_solve_creg(rx, rz)


def solve_regions(c20, regime_x, regime_z):
    """The hut regions [(rx, rz)] whose region constant is c20 modulo 2^20, in the regimes (taskspace.h:27-37)."""
    cap = 60000
    rx = (ctypes.c_int32 * cap)()
    rz = (ctypes.c_int32 * cap)()
    n = lib().wo_solve_regions(c20, regime_x, regime_z, rx, rz, cap)
    return [(rx[i], rz[i]) for i in range(min(n, cap))]


def solve_offsets(tier):
    """The chunk offsets [(a, b)] of a tier, a-major (solver_fused.cuh:25-33)."""
    ox = (ctypes.c_uint8 * 576)()
    oz = (ctypes.c_uint8 * 576)()
    n = lib().wo_solve_offsets(tier, ox, oz)
    return [(ox[i], oz[i]) for i in range(n)]


def solve_task(layout, cx, cz, cap=256):
    """Every structure seed of one chunk task, ascending; None for an offset no tier holds."""
    out = (ctypes.c_uint64 * cap)()
    n = lib().wo_solve_task(layout, cx, cz, out, cap)
    if n < 0:
        return None
    assert n <= cap
    return [out[i] for i in range(n)]


def solve_tasks(layout, chunks, threads=16, per_task_cap=64):
    """wo_solve_task over chunks [(cx, cz)] on host threads: [sorted seed list per chunk]."""
    n = len(chunks)
    cx = (ctypes.c_int32 * n)(*[c[0] for c in chunks])
    cz = (ctypes.c_int32 * n)(*[c[1] for c in chunks])
    seeds = (ctypes.c_uint64 * (n * per_task_cap))()
    counts = (ctypes.c_int * n)()
    lib().wo_solve_tasks(layout, cx, cz, n, per_task_cap, threads, seeds, counts)
    out = []
    for i in range(n):
        assert counts[i] <= per_task_cap
        out.append(None if counts[i] < 0 else [seeds[i * per_task_cap + k] for k in range(counts[i])])
    return out


def solve_prefixes(layout, cx, cz, width, cap=1 << 16):
    """The solver's search stopped at `width` bits: the kept states t mod 2^width, ascending."""
    out = (ctypes.c_uint64 * cap)()
    n = lib().wo_solve_prefixes(layout, cx, cz, width, out, cap)
    assert 0 <= n <= cap
    return [out[i] for i in range(n)]
