"""MI355X linking-number engine: Python binding of the C ABI in include/flk.h.

The product path is libflk.so (hand-written HIP kernels for gfx950) plus the
C++ drop-in host library; this module only loads them with ctypes and converts
numpy arrays.  Loading fails loudly when the built library is missing: there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
LIBFLK_PATH = os.path.join(LIB_DIR, "libflk.so")
LIBVC_PATH = os.path.join(LIB_DIR, "libverifycurves.so")
VERIFYCURVES_BIN = os.path.join(LIB_DIR, "verifycurves")
OBJ2BCC_BIN = os.path.join(LIB_DIR, "obj2bcc")

FLK_DIRECTSUM = 0
FLK_BARNESHUT = 1
FLK_CRIT_BETA = 0
FLK_CRIT_MLP = 1
FLK_PAIR_FALLBACK = 1
FLK_PAIR_NAN = 2
FLK_PAIR_RETRY = 4
FLK_PAIR_DIRECT = 8
MLP_NPARAM = 5057
MLP_NFEAT = 12


class FlkParams(C.Structure):
    _fields_ = [
        ("method", C.c_int32),
        ("criterion", C.c_int32),
        ("beta_init", C.c_double),
        ("beta_limit", C.c_double),
        ("eps_sq", C.c_double),
        ("target_error", C.c_double),
        ("err_scale", C.c_double),
        ("fallback_tol", C.c_double),
        ("bh_direct_threshold", C.c_int64),
        ("leaf_size", C.c_int32),
        ("fp64", C.c_int32),
        ("mlp_log_tol", C.c_double),
        ("bh_batch_ksegs", C.c_int32),
        ("timing", C.c_int32),
        ("graph", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]


class FlkStats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double),
        ("ms_kernel", C.c_double),
        ("ms_ds_kernel", C.c_double),
        ("ms_bh_traverse", C.c_double),
        ("ms_bh_build", C.c_double),
        ("ms_fallback", C.c_double),
        ("seg_pairs", C.c_int64),
        ("ds_seg_pairs", C.c_int64),
        ("bh_visits", C.c_int64),
        ("bh_far", C.c_int64),
        ("bh_leaf_pairs", C.c_int64),
        ("fallbacks", C.c_int64),
        ("retries", C.c_int64),
        ("ds_launches", C.c_int32),
        ("ds_waves", C.c_int32),
        ("ms_bh_level", C.c_double),
        ("ms_mlp", C.c_double),
        ("mlp_evals", C.c_int64),
        ("bh_host_passes", C.c_int64),
        ("graph_launches", C.c_int64),
        ("ms_graph_launch", C.c_double),
        ("reserved", C.c_int64 * 2),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


# Every function declared in include/flk.h: (name, restype, argtypes).
_P = C.POINTER
class FlkBccInfo(C.Structure):
    """Result of flk_load_bcc (include/flk.h flk_bcc_info)."""
    _fields_ = [
        ("ncurves", C.c_int32),
        ("segment_type", C.c_int32),
        ("points", C.c_int64),
        ("offsets", C.POINTER(C.c_int64)),
        ("looped", C.POINTER(C.c_uint8)),
        ("reserved", C.c_int64 * 4),
    ]


class FlkTune(C.Structure):
    """Kernel launch shapes (include/flk.h flk_tune)."""
    _fields_ = [
        ("ds_rows_per_lane", C.c_int32),
        ("ds_blocks_per_cu", C.c_int32),
        ("ds_tasks_per_wave", C.c_int32),
        ("bh_items_per_thread", C.c_int32),
        ("ingest_threads", C.c_int32),
        ("bh_level_capacity", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class FlkDiscInfo(C.Structure):
    """Counts of a device discretization (include/flk.h flk_disc_info)."""
    _fields_ = [
        ("segments", C.c_int64),
        ("vertices", C.c_int64),
        ("rounds", C.c_int32),
        ("small_curve", C.c_int32),
        ("small_segment", C.c_int32),
        ("pad", C.c_int32),
        ("small_size", C.c_double),
        ("overlaps", C.c_int64 * 55),
        ("pieces", C.c_int64),
        ("reserved", C.c_int64 * 3),
    ]


class FlkTreeNode(C.Structure):
    """One Barnes–Hut tree node in input coordinates (include/flk.h flk_tree_node)."""
    _fields_ = [
        ("center", C.c_double * 3),
        ("radius", C.c_double),
        ("rsq", C.c_double),
        ("c", C.c_double * 3),
        ("D", C.c_double * 9),
        ("Q", C.c_double * 27),
        ("cn", C.c_double),
        ("dn", C.c_double),
        ("qn", C.c_double),
        ("child", C.c_int32 * 2),
        ("seg0", C.c_int32),
        ("nseg", C.c_int32),
    ]


SEG_BSPLINE, SEG_CATMULLROM, SEG_BEZIER, SEG_POLYLINE = 0, 1, 2, 3

FLK_EXPORTS = [
    ("flk_default_params", C.c_int32, [_P(FlkParams)]),
    ("flk_create", C.c_int32, [C.c_int32, _P(C.c_void_p)]),
    ("flk_destroy", None, [C.c_void_p]),
    ("flk_last_error", C.c_char_p, [C.c_void_p]),
    ("flk_device_name", C.c_int32, [C.c_void_p, C.c_char_p, C.c_int32]),
    ("flk_set_curves", C.c_int32, [C.c_void_p, _P(C.c_double), _P(C.c_int64), C.c_int32]),
    ("flk_set_curves_f32", C.c_int32, [C.c_void_p, _P(C.c_float), _P(C.c_int64), C.c_int32]),
    ("flk_load_bcc", C.c_int32, [C.c_void_p, C.c_char_p, _P(FlkBccInfo)]),
    ("flk_curve_boxes", C.c_int32, [C.c_void_p, _P(C.c_double)]),
    ("flk_potential_links", C.c_int32,
     [C.c_void_p, _P(C.c_double), C.c_int32, _P(_P(C.c_int32)), _P(C.c_int64)]),
    ("flk_linking_numbers", C.c_int32,
     [C.c_void_p, _P(C.c_int32), C.c_int64, _P(FlkParams), _P(C.c_double), _P(C.c_double),
      _P(C.c_uint32)]),
    ("flk_linking_numbers_noalloc", C.c_int32,
     [C.c_void_p, _P(C.c_int32), C.c_int64, _P(FlkParams), _P(C.c_double), _P(C.c_double),
      _P(C.c_uint32)]),
    ("flk_get_stats", C.c_int32, [C.c_void_p, _P(FlkStats)]),
    ("flk_tree_moments", C.c_int32,
     [C.c_void_p, C.c_int32, C.c_int32, _P(FlkTreeNode), C.c_int32, _P(C.c_int32), _P(C.c_int32)]),
    ("flk_discretize", C.c_int32,
     [C.c_void_p, _P(C.c_double), _P(C.c_int64), _P(C.c_int32), _P(C.c_uint8), C.c_int32, _P(C.c_int32),
      C.c_int64, _P(_P(C.c_double)), _P(_P(C.c_int64)), _P(FlkDiscInfo)]),
    ("flk_get_tune", C.c_int32, [C.c_void_p, _P(FlkTune)]),
    ("flk_set_tune", C.c_int32, [C.c_void_p, _P(FlkTune)]),
    ("flk_autotune", C.c_int32, [C.c_void_p, _P(C.c_int32), C.c_int64, _P(FlkParams), _P(FlkTune)]),
    ("flk_mlp_set_weights", C.c_int32, [C.c_void_p, _P(C.c_float), _P(C.c_float)]),
    ("flk_mlp_get_weights", C.c_int32, [C.c_void_p, _P(C.c_float), _P(C.c_float)]),
    ("flk_mlp_forward", C.c_int32, [C.c_void_p, _P(C.c_float), C.c_int64, _P(C.c_float)]),
    ("flk_mlp_gradient", C.c_int32,
     [C.c_void_p, _P(C.c_float), _P(C.c_float), C.c_int64, _P(C.c_float), _P(C.c_float)]),
    ("flk_mlp_train", C.c_int32,
     [C.c_void_p, _P(C.c_float), _P(C.c_float), C.c_int64, C.c_int32, C.c_int32, C.c_float,
      C.c_uint64, _P(C.c_float)]),
    ("flk_mlp_samples", C.c_int32,
     [C.c_void_p, C.c_int32, C.c_int32, C.c_int64, C.c_uint64, _P(C.c_float), _P(C.c_float),
      _P(C.c_int64)]),
]

_lib = None


def load_library(path: str = LIBFLK_PATH) -> C.CDLL:
    """Loads libflk.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP engine first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, res, args in FLK_EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class FlkError(RuntimeError):
    pass


def _ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def default_params(**kw) -> FlkParams:
    lib = load_library()
    p = FlkParams()
    lib.flk_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "method" and isinstance(v, str):
            v = {"directsum": FLK_DIRECTSUM, "barneshut": FLK_BARNESHUT}[v]
        if k == "criterion" and isinstance(v, str):
            v = {"beta": FLK_CRIT_BETA, "mlp": FLK_CRIT_MLP}[v]
        setattr(p, k, v)
    return p


def pack_curves(curves: Sequence[np.ndarray]) -> tuple[np.ndarray, np.ndarray]:
    """List of (n_i, 3) arrays -> (xyz AoS float64, offsets int64)."""
    offs = np.zeros(len(curves) + 1, dtype=np.int64)
    for i, c in enumerate(curves):
        offs[i + 1] = offs[i] + len(c)
    xyz = np.zeros((int(offs[-1]), 3), dtype=np.float64)
    for i, c in enumerate(curves):
        xyz[offs[i]:offs[i + 1]] = np.asarray(c, dtype=np.float64).reshape(-1, 3)
    return np.ascontiguousarray(xyz.reshape(-1)), offs


@dataclass
class LinkResult:
    lk: np.ndarray
    err: np.ndarray
    flags: np.ndarray
    stats: dict = field(default_factory=dict)


class Engine:
    """One engine context = one HIP stream on one device (include/flk.h)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.flk_create(device, C.byref(h))
        if rc != 0:
            raise FlkError(f"flk_create failed ({rc}): {self.lib.flk_last_error(None).decode()}")
        self.h = h
        self.ncurves = 0
        self.sizes = np.zeros(0, dtype=np.int64)

    def close(self):
        if getattr(self, "h", None):
            self.lib.flk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise FlkError(f"{what} failed ({rc}): {self.lib.flk_last_error(self.h).decode()}")

    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        self._check(self.lib.flk_device_name(self.h, buf, 256), "flk_device_name")
        return buf.value.decode()

    def set_curves(self, curves: Sequence[np.ndarray]):
        xyz, offs = pack_curves(curves)
        self.set_curves_packed(xyz, offs)

    def set_curves_packed(self, xyz: np.ndarray, offs: np.ndarray):
        """Uploads packed curves: xyz (total, 3) and offsets (ncurves + 1).
        float32 input goes through flk_set_curves_f32, anything else as
        float64 through flk_set_curves."""
        offs = np.ascontiguousarray(offs, dtype=np.int64)
        n = len(offs) - 1
        if np.asarray(xyz).dtype == np.float32:
            xyz = np.ascontiguousarray(xyz, dtype=np.float32)
            self._check(self.lib.flk_set_curves_f32(self.h, _ptr(xyz, C.c_float), _ptr(offs, C.c_int64), n),
                        "flk_set_curves_f32")
        else:
            xyz = np.ascontiguousarray(xyz, dtype=np.float64)
            self._check(self.lib.flk_set_curves(self.h, _ptr(xyz, C.c_double), _ptr(offs, C.c_int64), n),
                        "flk_set_curves")
        self.ncurves = n
        self.sizes = np.diff(offs)

    def load_bcc(self, path: str) -> dict:
        """Reads a BCC file straight into the engine's curve set (control
        points); returns the segment type, the offsets and the loop flags."""
        info = FlkBccInfo()
        self._check(self.lib.flk_load_bcc(self.h, os.fsencode(path), C.byref(info)), "flk_load_bcc")
        n = info.ncurves
        offs = np.ctypeslib.as_array(info.offsets, shape=(n + 1,)).copy() if n else np.zeros(1, np.int64)
        looped = np.ctypeslib.as_array(info.looped, shape=(n,)).copy().astype(bool) if n else np.zeros(0, bool)
        self.ncurves = n
        self.sizes = np.diff(offs)
        return {"segment_type": info.segment_type, "points": info.points, "offsets": offs, "looped": looped}

    def discretize_current(self, types, looped, pairs):
        """Device discretization of the engine's current curve set (control
        points, e.g. from load_bcc); the polylines replace it."""
        n = self.ncurves
        types = np.broadcast_to(np.asarray(types, dtype=np.int32), (n,)).copy()
        looped = np.broadcast_to(np.asarray(looped, dtype=np.uint8), (n,)).copy()
        offs = np.zeros(n + 1, np.int64)
        offs[1:] = np.cumsum(self.sizes)
        pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        ooff = C.POINTER(C.c_int64)()
        info = FlkDiscInfo()
        rc = self.lib.flk_discretize(self.h, None, _ptr(offs, C.c_int64), _ptr(types, C.c_int32),
                                     _ptr(looped, C.c_uint8), n, _ptr(pairs, C.c_int32), len(pairs),
                                     None, C.byref(ooff), C.byref(info))
        self._check(rc, "flk_discretize")
        off = np.ctypeslib.as_array(ooff, shape=(n + 1,)).copy()
        self.sizes = np.diff(off)
        return {"rounds": info.rounds, "overlaps": list(info.overlaps[: info.rounds]), "segments": info.segments,
                "pieces": info.pieces, "vertices": info.vertices}

    def curve_boxes(self) -> np.ndarray:
        out = np.zeros((self.ncurves, 6), dtype=np.float64)
        self._check(self.lib.flk_curve_boxes(self.h, _ptr(out, C.c_double)), "flk_curve_boxes")
        return out

    def potential_links(self, boxes: np.ndarray | None = None) -> np.ndarray:
        p = C.POINTER(C.c_int32)()
        n = C.c_int64()
        if boxes is None:
            rc = self.lib.flk_potential_links(self.h, None, self.ncurves, C.byref(p), C.byref(n))
        else:
            b = np.ascontiguousarray(boxes, dtype=np.float64)
            rc = self.lib.flk_potential_links(self.h, _ptr(b, C.c_double), len(b), C.byref(p),
                                              C.byref(n))
        self._check(rc, "flk_potential_links")
        if n.value == 0:
            return np.zeros((0, 2), dtype=np.int32)
        return np.ctypeslib.as_array(p, shape=(n.value * 2,)).reshape(-1, 2).copy()

    def linking_numbers(self, pairs: np.ndarray, params: FlkParams | None = None, **kw) -> LinkResult:
        pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        m = len(pairs)
        if params is None:
            params = default_params(**kw)
        lk = np.zeros(m, dtype=np.float64)
        err = np.zeros(m, dtype=np.float64)
        flags = np.zeros(m, dtype=np.uint32)
        rc = self.lib.flk_linking_numbers(self.h, _ptr(pairs, C.c_int32), m, C.byref(params),
                                          _ptr(lk, C.c_double), _ptr(err, C.c_double),
                                          _ptr(flags, C.c_uint32))
        self._check(rc, "flk_linking_numbers")
        return LinkResult(lk, err, flags, self.stats())

    def linking_numbers_noalloc(self, pairs: np.ndarray, params: FlkParams | None = None, **kw) -> LinkResult:
        """Direct sum through flk_linking_numbers_noalloc: no device
        allocation, the batch's arrays in the library's fixed per-device
        array (at most 65,536 pairs and 1,048,576 tasks per call)."""
        pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        m = len(pairs)
        if params is None:
            kw.setdefault("method", "directsum")
            params = default_params(**kw)
        lk = np.zeros(m, dtype=np.float64)
        err = np.zeros(m, dtype=np.float64)
        flags = np.zeros(m, dtype=np.uint32)
        rc = self.lib.flk_linking_numbers_noalloc(self.h, _ptr(pairs, C.c_int32), m, C.byref(params),
                                                  _ptr(lk, C.c_double), _ptr(err, C.c_double),
                                                  _ptr(flags, C.c_uint32))
        self._check(rc, "flk_linking_numbers_noalloc")
        return LinkResult(lk, err, flags, self.stats())

    def tree_moments(self, curve: int, leaf_size: int = 1) -> tuple[dict, np.ndarray]:
        """The Barnes–Hut tree of one uploaded curve, as flk_linking_numbers
        builds it: per-node arrays (center, radius, rsq, c, D (n,3,3),
        Q (n,3,3,3), cn, dn, qn, child (n,2), seg0, nseg), root first, and the
        curve segment at each position of the tree's order."""
        nn = C.c_int32()
        self._check(self.lib.flk_tree_moments(self.h, curve, leaf_size, None, 0, C.byref(nn), None),
                    "flk_tree_moments")
        n = nn.value
        nodes = (FlkTreeNode * max(n, 1))()
        order = np.zeros(int(self.sizes[curve]) if n else 0, dtype=np.int32)
        self._check(self.lib.flk_tree_moments(self.h, curve, leaf_size, nodes, n, C.byref(nn),
                                              _ptr(order, C.c_int32) if n else None), "flk_tree_moments")
        raw = np.ctypeslib.as_array(C.cast(nodes, C.POINTER(C.c_uint8)), shape=(C.sizeof(FlkTreeNode) * n,))
        rec = np.frombuffer(raw.tobytes(), dtype=np.dtype(
            [("center", "<f8", 3), ("radius", "<f8"), ("rsq", "<f8"), ("c", "<f8", 3), ("D", "<f8", (3, 3)),
             ("Q", "<f8", (3, 3, 3)), ("cn", "<f8"), ("dn", "<f8"), ("qn", "<f8"), ("child", "<i4", 2),
             ("seg0", "<i4"), ("nseg", "<i4")]), count=n)
        return {k: rec[k].copy() for k in rec.dtype.names}, order

    def stats(self) -> dict:
        s = FlkStats()
        self._check(self.lib.flk_get_stats(self.h, C.byref(s)), "flk_get_stats")
        return s.as_dict()

    # ---- discretization
    def discretize(self, curves, types, looped, pairs):
        """Device discretization of control polygons `curves` (segment types
        SEG_*, looped flags) against potentially linked `pairs`; returns the
        polylines and the counts (rounds, overlaps per round, pieces).  The
        polylines also become the engine's curve set."""
        n = len(curves)
        types = np.broadcast_to(np.asarray(types, dtype=np.int32), (n,)).copy()
        looped = np.broadcast_to(np.asarray(looped, dtype=np.uint8), (n,)).copy()
        xyz, offs = pack_curves(curves)
        pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        out = C.POINTER(C.c_double)()
        ooff = C.POINTER(C.c_int64)()
        info = FlkDiscInfo()
        rc = self.lib.flk_discretize(self.h, _ptr(xyz, C.c_double), _ptr(offs, C.c_int64), _ptr(types, C.c_int32),
                                     _ptr(looped, C.c_uint8), n, _ptr(pairs, C.c_int32), len(pairs),
                                     C.byref(out), C.byref(ooff), C.byref(info))
        d = {"segments": info.segments, "vertices": info.vertices, "rounds": info.rounds,
             "pieces": info.pieces, "overlaps": list(info.overlaps[: info.rounds]),
             "small_curve": info.small_curve, "small_segment": info.small_segment,
             "small_size": info.small_size}
        self._check(rc, "flk_discretize")
        off = np.ctypeslib.as_array(ooff, shape=(n + 1,)).copy()
        v = np.ctypeslib.as_array(out, shape=(3 * int(off[-1]),)).copy().reshape(-1, 3) if off[-1] else np.zeros((0, 3))
        self.ncurves = n
        self.sizes = np.diff(off)
        return [v[off[c]:off[c + 1]] for c in range(n)], d

    # ---- launch shapes
    def get_tune(self) -> dict:
        t = FlkTune()
        self._check(self.lib.flk_get_tune(self.h, C.byref(t)), "flk_get_tune")
        return t.as_dict()

    def set_tune(self, **kw) -> dict:
        t = FlkTune()
        self._check(self.lib.flk_get_tune(self.h, C.byref(t)), "flk_get_tune")
        for k, v in kw.items():
            setattr(t, k, v)
        self._check(self.lib.flk_set_tune(self.h, C.byref(t)), "flk_set_tune")
        return t.as_dict()

    def autotune(self, pairs: np.ndarray, params: FlkParams | None = None, **kw) -> dict:
        """Times candidate launch shapes on these pairs and keeps the fastest."""
        pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        if params is None:
            params = default_params(**kw)
        t = FlkTune()
        self._check(self.lib.flk_autotune(self.h, _ptr(pairs, C.c_int32), len(pairs), C.byref(params),
                                          C.byref(t)), "flk_autotune")
        return t.as_dict()

    # ---- learned criterion
    def mlp_set_weights(self, params: np.ndarray, norm: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float32)
        nm = np.ascontiguousarray(norm, dtype=np.float32)
        assert p.size == MLP_NPARAM and nm.size == 2 * MLP_NFEAT
        self._check(self.lib.flk_mlp_set_weights(self.h, _ptr(p, C.c_float), _ptr(nm, C.c_float)),
                    "flk_mlp_set_weights")

    def mlp_get_weights(self) -> tuple[np.ndarray, np.ndarray]:
        p = np.zeros(MLP_NPARAM, dtype=np.float32)
        nm = np.zeros(2 * MLP_NFEAT, dtype=np.float32)
        self._check(self.lib.flk_mlp_get_weights(self.h, _ptr(p, C.c_float), _ptr(nm, C.c_float)),
                    "flk_mlp_get_weights")
        return p, nm

    def mlp_forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_NFEAT)
        y = np.zeros(len(x), dtype=np.float32)
        self._check(self.lib.flk_mlp_forward(self.h, _ptr(x, C.c_float), len(x), _ptr(y, C.c_float)),
                    "flk_mlp_forward")
        return y

    def mlp_gradient(self, x: np.ndarray, y: np.ndarray) -> tuple[np.ndarray, float]:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_NFEAT)
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
        g = np.zeros(MLP_NPARAM, dtype=np.float32)
        loss = C.c_float()
        self._check(self.lib.flk_mlp_gradient(self.h, _ptr(x, C.c_float), _ptr(y, C.c_float), len(x),
                                              _ptr(g, C.c_float), C.byref(loss)), "flk_mlp_gradient")
        return g, loss.value

    def mlp_train(self, x: np.ndarray, y: np.ndarray, epochs: int = 10, batch: int = 4096,
                  lr: float = 1e-3, seed: int = 5) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_NFEAT)
        y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
        loss = C.c_float()
        self._check(self.lib.flk_mlp_train(self.h, _ptr(x, C.c_float), _ptr(y, C.c_float), len(x),
                                           epochs, batch, lr, seed, C.byref(loss)), "flk_mlp_train")
        return loss.value

    def mlp_samples(self, a: int, b: int, cap: int, seed: int = 5) -> tuple[np.ndarray, np.ndarray]:
        x = np.zeros((cap, MLP_NFEAT), dtype=np.float32)
        y = np.zeros(cap, dtype=np.float32)
        n = C.c_int64()
        self._check(self.lib.flk_mlp_samples(self.h, a, b, cap, seed, _ptr(x, C.c_float),
                                             _ptr(y, C.c_float), C.byref(n)), "flk_mlp_samples")
        return x[: n.value].copy(), y[: n.value].copy()


def certificate_from_lk(pairs: np.ndarray, lk: np.ndarray) -> list[tuple[int, int, int]]:
    """Rounds pair sums into certificate triplets as the reference does
    (src/linking_number_certificate.cpp:185-213, :294-322): nearest integer
    (half away from zero), kept when |round| >= 0.25, sorted by (i, j); a
    non-finite sum is kept with the value -2**31."""
    p = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    v = np.asarray(lk, dtype=np.float64).reshape(-1)
    fin = np.isfinite(v)
    r = np.where(fin, np.sign(v) * np.floor(np.abs(np.where(fin, v, 0.0)) + 0.5), 0.0)
    keep = ~fin | (np.abs(r) >= 0.25)
    i = np.minimum(p[keep, 0], p[keep, 1])
    j = np.maximum(p[keep, 0], p[keep, 1])
    val = np.where(fin[keep], r[keep], 0.0).astype(np.int64)
    val[~fin[keep]] = -2 ** 31
    order = np.lexsort((val, j, i))
    return list(zip(i[order].tolist(), j[order].tolist(), val[order].tolist()))
