"""ctypes wrapper of the CPU oracle (liboracle.so built from vkgs_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product never imports this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VKGS_ORACLE_LIB points the tests at another build of vkgs_oracle.c (used to
# check that the oracle tests fail on a deliberately mutated oracle)
LIB_PATH = os.environ.get("VKGS_ORACLE_LIB", os.path.join(HERE, "liboracle.so"))

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.oracle_f32_to_f16.restype = C.c_uint16
        _lib.oracle_f32_to_f16.argtypes = [C.c_float]
        _lib.oracle_f16_to_f32.restype = C.c_float
        _lib.oracle_f16_to_f32.argtypes = [C.c_uint16]
        _lib.oracle_bin.restype = C.c_uint64
        _lib.oracle_num_threads.restype = C.c_int
    return _lib


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32)


def camera_bytes(cam) -> np.ndarray:
    """192-byte Camera UBO as a uint8 array (accepts the ctypes struct)."""
    return np.frombuffer(bytes(cam), np.uint8).copy()


IDENTITY = np.eye(4, dtype=np.float32).ravel()


def parse_ply(rows: np.ndarray, offsets: np.ndarray):
    rows = _f32(rows)
    n = rows.shape[0]
    off = np.ascontiguousarray(offsets, np.uint32)
    pos = np.empty((n, 3), np.float32)
    cov = np.empty((n, 6), np.float32)
    opa = np.empty(n, np.float32)
    sh = np.empty((n, 48), np.uint16)
    lib().oracle_parse_ply(_p(rows), _p(off), C.c_uint32(n), _p(pos), _p(cov), _p(opa), _p(sh))
    return pos, cov, opa, sh


def rank(cam, pos: np.ndarray, model: np.ndarray = IDENTITY):
    pos = _f32(pos)
    n = pos.shape[0]
    cb = camera_bytes(cam)
    m = _f32(model).ravel()
    count = np.zeros(1, np.uint32)
    key = np.empty(max(n, 1), np.uint32)
    index = np.empty(max(n, 1), np.uint32)
    lib().oracle_rank(_p(cb), _p(m), _p(pos), C.c_uint32(n), _p(count), _p(key), _p(index))
    v = int(count[0])
    return v, key[:v].copy(), index[:v].copy()


def sort_kv(key: np.ndarray, value: np.ndarray):
    k = np.ascontiguousarray(key, np.uint32).copy()
    v = np.ascontiguousarray(value, np.uint32).copy()
    lib().oracle_sort_kv(C.c_uint32(k.size), _p(k), _p(v))
    return k, v


def inverse(index: np.ndarray, n: int) -> np.ndarray:
    idx = np.ascontiguousarray(index, np.uint32)
    inv = np.empty(max(n, 1), np.int32)
    lib().oracle_inverse(C.c_uint32(idx.size), _p(idx), C.c_uint32(n), _p(inv))
    return inv[:n]


def projection(cam, pos, cov, opa, sh, inv, count: int, model=IDENTITY, model_inv=IDENTITY):
    pos, cov, opa = _f32(pos), _f32(cov), _f32(opa)
    sh = np.ascontiguousarray(sh, np.uint16)
    inv = np.ascontiguousarray(inv, np.int32)
    n = pos.shape[0]
    cb = camera_bytes(cam)
    inst = np.zeros((max(count, 1), 12), np.float32)
    di = np.zeros(12, np.uint32)
    lib().oracle_projection(_p(cb), _p(_f32(model).ravel()), _p(_f32(model_inv).ravel()), _p(pos), _p(cov),
                            _p(opa), _p(sh), _p(inv), C.c_uint32(n), C.c_uint32(count), _p(inst), _p(di))
    return inst[:count], di


def draw(inst: np.ndarray, width: int, height: int) -> np.ndarray:
    inst = _f32(inst).reshape(-1, 12)
    img = np.empty((height, width, 4), np.float32)
    lib().oracle_draw(_p(inst), C.c_uint32(inst.shape[0]), C.c_uint32(width), C.c_uint32(height), _p(img))
    return img


def draw_triangles(inst: np.ndarray, width: int, height: int, subpixel_bits: int = 8) -> np.ndarray:
    """The draw as two rasterized triangles per instance (snapped vertices, top-left rule)."""
    inst = _f32(inst).reshape(-1, 12)
    img = np.empty((height, width, 4), np.float32)
    lib().oracle_draw_triangles(_p(inst), C.c_uint32(inst.shape[0]), C.c_uint32(width), C.c_uint32(height),
                                C.c_uint32(subpixel_bits), _p(img))
    return img


def draw_bruteforce(inst: np.ndarray, width: int, height: int) -> np.ndarray:
    inst = _f32(inst).reshape(-1, 12)
    img = np.empty((height, width, 4), np.float32)
    lib().oracle_draw_bruteforce(_p(inst), C.c_uint32(inst.shape[0]), C.c_uint32(width), C.c_uint32(height),
                                 _p(img))
    return img


def bin_tiles(inst: np.ndarray, width: int, height: int, tile_w: int, tile_h: int):
    """(ranges (tiles, 2) uint32, list of instance indices) of the tile binning."""
    inst = _f32(inst).reshape(-1, 12)
    tx = (width + tile_w - 1) // tile_w
    ty = (height + tile_h - 1) // tile_h
    ranges = np.zeros((tx * ty, 2), np.uint32)
    total = lib().oracle_bin(_p(inst), C.c_uint32(inst.shape[0]), C.c_uint32(width), C.c_uint32(height),
                             C.c_uint32(tile_w), C.c_uint32(tile_h), _p(ranges), None)
    lst = np.zeros(max(int(total), 1), np.uint32)
    lib().oracle_bin(_p(inst), C.c_uint32(inst.shape[0]), C.c_uint32(width), C.c_uint32(height),
                     C.c_uint32(tile_w), C.c_uint32(tile_h), _p(ranges), _p(lst))
    return ranges, lst[:int(total)]


def f32_to_f16(x: float) -> int:
    return int(lib().oracle_f32_to_f16(C.c_float(x)))


def f16_to_f32(h: int) -> float:
    return float(lib().oracle_f16_to_f32(C.c_uint16(h)))


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def frame(cam, pos, cov, opa, sh, width: int, height: int, model=IDENTITY, model_inv=IDENTITY):
    """The reference's whole frame: rank -> stable sort -> inverse -> projection -> draw."""
    n = pos.shape[0]
    v, key, index = rank(cam, pos, model)
    skey, sidx = sort_kv(key, index)
    inv = inverse(sidx, n)
    inst, di = projection(cam, pos, cov, opa, sh, inv, v, model, model_inv)
    img = draw(inst, width, height)
    return dict(visible=v, keys=skey, order=sidx, inverse=inv, instances=inst, draw_indirect=di, image=img)
