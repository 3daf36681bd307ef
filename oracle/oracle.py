"""CPU ORACLE — ctypes binding of oracle/_build/liboracle.so (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module; the product package never does.  The C++ restatement it loads is
documented function by function (with reference file:line) in oracle/oracle.cpp.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "liboracle.so")
CLI = os.path.join(BUILD, "verifycurves_oracle")

BSPLINE, CATMULLROM, BEZIER, POLYLINE = 0, 1, 2, 3
_P = C.POINTER
_D = _P(C.c_double)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_model_new": (C.c_void_p, []),
            "orc_model_free": (None, [C.c_void_p]),
            "orc_model_add_curve": (C.c_int, [C.c_void_p, _D, C.c_int64, C.c_int, C.c_int]),
            "orc_model_load_bcc": (C.c_int, [C.c_void_p, C.c_char_p]),
            "orc_model_save_bcc": (C.c_int, [C.c_void_p, C.c_char_p]),
            "orc_model_load_obj": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
            "orc_model_ncurves": (C.c_int32, [C.c_void_p]),
            "orc_model_curve_size": (C.c_int64, [C.c_void_p, C.c_int32]),
            "orc_model_curve_info": (C.c_int, [C.c_void_p, C.c_int32, _P(C.c_int), _P(C.c_int)]),
            "orc_model_get_curve": (C.c_int, [C.c_void_p, C.c_int32, _D]),
            "orc_point": (C.c_int, [_D, C.c_int64, C.c_int, C.c_int, C.c_int64, C.c_double, _D]),
            "orc_segment_box": (C.c_int, [_D, C.c_int64, C.c_int, C.c_int, C.c_int64, C.c_double,
                                          C.c_double, _D]),
            "orc_curve_boxes": (C.c_int, [C.c_void_p, _D]),
            "orc_pls": (C.c_int64, [C.c_void_p, C.c_int, _P(C.c_int32), C.c_int64]),
            "orc_discretize": (C.c_int, [C.c_void_p, C.c_void_p]),
            "orc_ds_one_arctan": (C.c_double, [_D, C.c_int64, _D, C.c_int64, C.c_int]),
            "orc_ds_fast": (C.c_double, [_D, C.c_int64, _D, C.c_int64]),
            "orc_ds_rows": (C.c_double, [_D, C.c_int64, _D, C.c_int64, C.c_int64, C.c_int64, C.c_int]),
            "orc_segment_pair_term": (C.c_double, [_D, _D, _D, _D]),
            "orc_ds_pairs": (C.c_int, [_D, _P(C.c_int64), C.c_int32, _P(C.c_int32), C.c_int64, _D]),
            "orc_bh_leaf_term": (C.c_double, [_D, _D, _D, _D]),
            "orc_bh_pair": (C.c_double, [_D, C.c_int64, _D, C.c_int64, C.c_double, C.c_int, _D,
                                         _P(C.c_int64)]),
            "orc_tree_moments": (C.c_int, [_D, C.c_int64, _D, _D, _P(C.c_int32), _P(C.c_int32)]),
            "orc_far_field": (C.c_double, [_D, _D, C.c_double, C.c_double, C.c_double, _D]),
            "orc_certificate": (C.c_int64, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                            _P(C.c_int32), C.c_int64, _D, _D, _D, _P(C.c_int32),
                                            C.c_int64, _P(C.c_int64)]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _pts(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, 3))


class Model:
    """Oracle curve collection (reference include/model.h)."""

    def __init__(self):
        self.h = lib().orc_model_new()

    def __del__(self):
        try:
            lib().orc_model_free(self.h)
        except Exception:
            pass

    def _ck(self, rc):
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    @classmethod
    def from_curves(cls, curves, seg_type=POLYLINE, looped=True):
        m = cls()
        for c in curves:
            p = _pts(c)
            m._ck(lib().orc_model_add_curve(m.h, _dp(p), len(p), seg_type, int(looped)))
        return m

    @classmethod
    def load_bcc(cls, path):
        m = cls()
        m._ck(lib().orc_model_load_bcc(m.h, path.encode()))
        return m

    @classmethod
    def load_obj(cls, path, seg_type=POLYLINE):
        m = cls()
        m._ck(lib().orc_model_load_obj(m.h, path.encode(), seg_type))
        return m

    def save_bcc(self, path):
        self._ck(lib().orc_model_save_bcc(self.h, path.encode()))

    def __len__(self):
        return lib().orc_model_ncurves(self.h)

    def curve(self, c):
        n = lib().orc_model_curve_size(self.h, c)
        out = np.zeros((n, 3))
        lib().orc_model_get_curve(self.h, c, _dp(out))
        return out

    def curves(self):
        return [self.curve(c) for c in range(len(self))]

    def info(self, c):
        t, l = C.c_int(), C.c_int()
        lib().orc_model_curve_info(self.h, c, C.byref(t), C.byref(l))
        return t.value, bool(l.value)

    def boxes(self):
        out = np.zeros((len(self), 6))
        self._ck(lib().orc_curve_boxes(self.h, _dp(out)))
        return out

    def pls(self, use_tree=False):
        n = len(self)
        cap = max(16, n * 64)
        out = np.zeros(2 * cap, dtype=np.int32)
        k = lib().orc_pls(self.h, int(use_tree), out.ctypes.data_as(_P(C.c_int32)), cap)
        if k < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        if k > cap:
            out = np.zeros(2 * k, dtype=np.int32)
            k = lib().orc_pls(self.h, int(use_tree), out.ctypes.data_as(_P(C.c_int32)), k)
        return out[: 2 * k].reshape(-1, 2)

    def discretize(self):
        out = Model()
        self._ck(lib().orc_discretize(self.h, out.h))
        return out

    def certificate(self, method="barneshut", discretize=True, beta_init=2.0, beta_limit=10.0):
        """Returns (triplets, mean_err, max_err, candidate_pairs, candidate_lk)."""
        cap = 1 << 20
        trip = np.zeros(3 * cap, dtype=np.int32)
        me, mx = C.c_double(), C.c_double()
        ncand = C.c_int64()
        ccap = 1 << 20
        lk = np.zeros(ccap)
        cand = np.zeros(2 * ccap, dtype=np.int32)
        k = lib().orc_certificate(self.h, int(method == "directsum"), int(discretize), beta_init,
                                  beta_limit, trip.ctypes.data_as(_P(C.c_int32)), cap, C.byref(me),
                                  C.byref(mx), _dp(lk), cand.ctypes.data_as(_P(C.c_int32)), ccap,
                                  C.byref(ncand))
        if k < 0:
            raise RuntimeError(lib().orc_last_error().decode())
        nc = ncand.value
        return ([tuple(int(v) for v in trip[3 * i:3 * i + 3]) for i in range(k)], me.value, mx.value,
                cand[: 2 * nc].reshape(-1, 2), lk[:nc])


def ds(p, q, parallel=True):
    """FP64 one-arctan direct sum (reference compute_link.cpp:126-178)."""
    p, q = _pts(p), _pts(q)
    return lib().orc_ds_one_arctan(_dp(p), len(p), _dp(q), len(q), int(parallel))


def ds_fast(p, q):
    """FP64 running-product direct sum (reference compute_link.cpp:200-280)."""
    p, q = _pts(p), _pts(q)
    return lib().orc_ds_fast(_dp(p), len(p), _dp(q), len(q))


def ds_rows(p, q, r0, r1, parallel=True):
    p, q = _pts(p), _pts(q)
    return lib().orc_ds_rows(_dp(p), len(p), _dp(q), len(q), r0, r1, int(parallel))


def ds_pairs(xyz, offsets, pairs):
    """The reference's direct-sum pair loop over a packed curve set."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
    out = np.zeros(len(pairs) // 2)
    rc = lib().orc_ds_pairs(_dp(xyz), offsets.ctypes.data_as(_P(C.c_int64)), len(offsets) - 1,
                            pairs.ctypes.data_as(_P(C.c_int32)), len(pairs) // 2, _dp(out))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return out


def segment_pair_term(p0, p1, q0, q1):
    a = [np.ascontiguousarray(np.asarray(v, dtype=np.float64)) for v in (p0, p1, q0, q1)]
    return lib().orc_segment_pair_term(*[_dp(v) for v in a])


def bh_leaf_term(m1, t1, m2, t2):
    a = [np.ascontiguousarray(np.asarray(v, dtype=np.float64)) for v in (m1, t1, m2, t2)]
    return lib().orc_bh_leaf_term(*[_dp(v) for v in a])


def bh_pair(p, q, beta=2.0, parallel=True):
    """Returns (lk, err_estimate, visits) of the reference Barnes–Hut."""
    p, q = _pts(p), _pts(q)
    err = C.c_double()
    vis = C.c_int64()
    v = lib().orc_bh_pair(_dp(p), len(p), _dp(q), len(q), beta, int(parallel), C.byref(err),
                          C.byref(vis))
    return v, err.value, vis.value


def tree_moments(p):
    p = _pts(p)
    n = len(p)
    mom = np.zeros((2 * n - 1, 45))
    boxes = np.zeros((n - 1, 6))
    ch = np.zeros((n - 1, 2), dtype=np.int32)
    seg = np.zeros(n, dtype=np.int32)
    rc = lib().orc_tree_moments(_dp(p), n, _dp(mom), _dp(boxes), ch.ctypes.data_as(_P(C.c_int32)),
                                seg.ctypes.data_as(_P(C.c_int32)))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return mom, boxes, ch, seg


def far_field(m1, m2, r1, r2, epssq=1e-24):
    a, b = (np.ascontiguousarray(np.asarray(v, dtype=np.float64)) for v in (m1, m2))
    err = C.c_double()
    v = lib().orc_far_field(_dp(a), _dp(b), r1, r2, epssq, C.byref(err))
    return v, err.value


def run_cli(input_path, output_path, method="barneshut", discretize=True, obj_type=None, extra=()):
    if not os.path.exists(CLI):
        build()
    cmd = [CLI, f"--input={input_path}", f"--output={output_path}", f"--method={method}",
           f"--discretize={'true' if discretize else 'false'}"]
    if obj_type:
        cmd.append(f"--obj_segment_type={obj_type}")
    cmd += list(extra)
    return subprocess.run(cmd, capture_output=True, text=True, check=True)
