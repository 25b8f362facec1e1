"""Python mirror of vkgs's renderer stages over the C ABI (device memory from torch).

Each function maps to one reference stage (engine.cc:1137-1290) and calls the
matching HIP entry point in libvkgs_hip.so on torch's current stream.  u32
buffers are held in torch.int32 tensors (same bits).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr

IDENTITY = np.eye(4, dtype=np.float32).ravel()
STAGES = ["count", "project", "depth_sort", "bin", "tile_sort", "raster", "total"]
RASTER_TRIANGLES = 2  # VKGS_RASTER_TRIANGLES
LOAD_STATES = ("idle", "reading", "ready", "done", "canceled", "failed")  # vkgs_load_state


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _dev() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("vkgs_amd needs a HIP device (there is no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def _model_ptr(model):
    if model is None:
        return None, None
    m = np.ascontiguousarray(model, np.float32).ravel()
    return m, m.ctypes.data


@dataclass
class SplatStorage:
    """vkgs splat storage (engine.cc:485-492) on the device."""

    position: torch.Tensor  # (N, 3) f32
    cov3d: torch.Tensor     # (N, 6) f32
    opacity: torch.Tensor   # (N,) f32
    sh: torch.Tensor        # (N, 48) f16 bits in int16

    @property
    def n(self) -> int:
        return int(self.position.shape[0])


def empty_storage(n: int) -> SplatStorage:
    d = _dev()
    return SplatStorage(torch.empty((n, 3), dtype=torch.float32, device=d),
                        torch.empty((n, 6), dtype=torch.float32, device=d),
                        torch.empty((n,), dtype=torch.float32, device=d),
                        torch.empty((n, 48), dtype=torch.int16, device=d))


def parse_ply(ply_buffer: np.ndarray, n: int) -> SplatStorage:
    """Stage 0 (parse_ply.comp) on a host ply buffer (60 offsets + rows)."""
    d_ply = torch.from_numpy(np.ascontiguousarray(ply_buffer).view(np.int32)).to(_dev())
    st = empty_storage(n)
    check(lib().vkgs_parse_ply(ptr(d_ply), n, ptr(st.position), ptr(st.cov3d), ptr(st.opacity), ptr(st.sh),
                               _stream()), "vkgs_parse_ply")
    return st


def rank(cam, st: SplatStorage, model=None):
    """Stage 1 (rank.comp): returns (count (1,), key (N,), index (N,)) device tensors."""
    n = st.n
    d = _dev()
    count = torch.zeros(1, dtype=torch.int32, device=d)
    key = torch.empty(max(n, 1), dtype=torch.int32, device=d)
    index = torch.empty(max(n, 1), dtype=torch.int32, device=d)
    nbytes = lib().vkgs_rank_storage_bytes(n)
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=d)
    m, mp = _model_ptr(model)
    check(lib().vkgs_rank(C.byref(cam), mp, ptr(st.position), n, ptr(count), ptr(key), ptr(index), ptr(scratch),
                          nbytes, _stream()), "vkgs_rank")
    return count, key, index


def sort_kv(count: torch.Tensor, key: torch.Tensor, value: torch.Tensor, key_bits: int = 32) -> None:
    """Stage 2 (vrdxCmdSortKeyValueIndirect): stable in-place sort of the first *count pairs."""
    max_n = int(key.numel())
    nbytes = lib().vkgs_sort_kv_storage_bytes(max_n)
    scratch = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=key.device)
    check(lib().vkgs_sort_kv(max_n, ptr(count), ptr(key), ptr(value), key_bits, ptr(scratch), nbytes, _stream()),
          "vkgs_sort_kv")


def inverse_index(count: torch.Tensor, index: torch.Tensor, n: int) -> torch.Tensor:
    """Stage 3 (inverse_index.comp)."""
    inv = torch.empty(max(n, 1), dtype=torch.int32, device=index.device)
    check(lib().vkgs_inverse_index(ptr(count), ptr(index), n, ptr(inv), _stream()), "vkgs_inverse_index")
    return inv[:n]


def projection(cam, st: SplatStorage, inverse: torch.Tensor, count: torch.Tensor, model=None):
    """Stage 4 (projection.comp): (instances (N, 12) f32, draw_indirect (12,) u32 bits)."""
    n = st.n
    d = _dev()
    inst = torch.zeros((max(n, 1), 12), dtype=torch.float32, device=d)
    di = torch.zeros(12, dtype=torch.int32, device=d)
    m, mp = _model_ptr(model)
    check(lib().vkgs_projection(C.byref(cam), mp, ptr(st.position), ptr(st.cov3d), ptr(st.opacity), ptr(st.sh),
                                ptr(inverse), n, ptr(count), ptr(inst), ptr(di), _stream()), "vkgs_projection")
    return inst, di


class DrawStorage:
    def __init__(self, max_instances: int, width: int, height: int, max_pairs: int, cfg=None):
        self.cfg = cfg if cfg is not None else _lib.default_raster_config()
        self.max_instances, self.width, self.height, self.max_pairs = max_instances, width, height, max_pairs
        self.nbytes = lib().vkgs_draw_storage_bytes(max_instances, width, height, max_pairs, C.byref(self.cfg))
        if self.nbytes == 0:
            raise ValueError("invalid draw configuration")
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=_dev())

    def status(self):
        pairs = C.c_uint64(0)
        st = lib().vkgs_draw_status(ptr(self.buf), self.nbytes, C.byref(pairs), _stream())
        return st, int(pairs.value)

    def lists(self):
        """(ranges (tiles, 2) int64 numpy, list of instance indices numpy) of the last draw."""
        rp, lp = C.c_void_p(), C.c_void_p()
        tx, ty = C.c_uint32(), C.c_uint32()
        check(lib().vkgs_draw_lists(ptr(self.buf), self.nbytes, C.byref(rp), C.byref(lp), C.byref(tx), C.byref(ty)),
              "vkgs_draw_lists")
        st, pairs = self.status()
        tiles = tx.value * ty.value
        base = self.buf.data_ptr()
        r_off = rp.value - base
        l_off = lp.value - base
        ranges = self.buf[r_off:r_off + tiles * 8].view(torch.int32).cpu().numpy().view(np.uint32).reshape(tiles, 2)
        lst = self.buf[l_off:l_off + pairs * 4].view(torch.int32).cpu().numpy().view(np.uint32)
        return ranges.astype(np.int64), lst.copy()


def draw(instances: torch.Tensor, draw_indirect: torch.Tensor, width: int, height: int, cfg=None,
         max_pairs: int | None = None, storage: DrawStorage | None = None):
    """Stage 5 (the splat draw): (image (H, W, 4) f32, DrawStorage)."""
    max_inst = int(instances.shape[0])
    if storage is None:
        if max_pairs is None:
            max_pairs = 4 * max_inst + (1 << 20)
        storage = DrawStorage(max_inst, width, height, max_pairs, cfg)
    img = torch.empty((height, width, 4), dtype=torch.float32, device=instances.device)
    check(lib().vkgs_draw(ptr(instances), ptr(draw_indirect), max_inst, width, height, storage.max_pairs,
                          C.byref(storage.cfg), ptr(img), ptr(storage.buf), storage.nbytes, _stream()),
          "vkgs_draw")
    return img, storage


def reference_frame(cam, st: SplatStorage, width: int, height: int, cfg=None, model=None):
    """rank -> sort -> inverse -> projection -> draw, the reference's stage order (engine.cc:1164-1290)."""
    count, key, index = rank(cam, st, model)
    sort_kv(count, key, index, 30)
    inv = inverse_index(count, index, st.n)
    inst, di = projection(cam, st, inv, count, model)
    img, ds = draw(inst, di, width, height, cfg)
    return dict(count=count, key=key, index=index, inverse=inv, instances=inst, draw_indirect=di, image=img,
                draw_storage=ds)


def u32(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy().view(np.uint32)


class Renderer:
    """The Engine analogue (vkgs_renderer_*): owns splat storage, renders fused frames."""

    def __init__(self, capacity: int):
        _dev()
        self._h = C.c_void_p()
        check(lib().vkgs_renderer_create(capacity, C.byref(self._h)), "vkgs_renderer_create")
        self.capacity = capacity
        self.tuned = None  # the RasterConfig of the last autotune()

    def close(self):
        if self._h:
            lib().vkgs_renderer_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_ply(self, path: str) -> None:
        check(lib().vkgs_renderer_load_ply(self._h, path.encode()), "vkgs_renderer_load_ply")

    def load_ply_async(self, path: str) -> None:
        """Starts loading a PLY file on the renderer's loader thread (a load in progress is canceled);
        the first render after it finishes shows it."""
        check(lib().vkgs_renderer_load_ply_async(self._h, path.encode()), "vkgs_renderer_load_ply_async")

    def cancel_load(self) -> None:
        check(lib().vkgs_renderer_cancel_load(self._h), "vkgs_renderer_cancel_load")

    def load_progress(self):
        """(rows in the file, rows read, state), state one of LOAD_STATES; raises for a failed load."""
        t, l, st = C.c_uint64(), C.c_uint64(), C.c_int()
        rc = lib().vkgs_renderer_load_progress(self._h, C.byref(t), C.byref(l), C.byref(st))
        check(rc, "vkgs_renderer_load_progress")
        return int(t.value), int(l.value), LOAD_STATES[st.value]

    def load_buffer(self, ply_buffer: np.ndarray, n: int) -> None:
        buf = np.ascontiguousarray(ply_buffer)
        check(lib().vkgs_renderer_load_ply_buffer(self._h, buf.ctypes.data, n), "vkgs_renderer_load_ply_buffer")

    @property
    def n(self) -> int:
        return int(lib().vkgs_renderer_splat_count(self._h))

    def storage(self) -> SplatStorage:
        """Views of the renderer's splat storage as torch tensors (copies)."""
        ps = [C.c_void_p() for _ in range(4)]
        check(lib().vkgs_renderer_storage(self._h, *[C.byref(p) for p in ps]), "vkgs_renderer_storage")
        n = self.n
        out = empty_storage(n)
        torch.cuda.synchronize()
        for t, p in zip((out.position, out.cov3d, out.opacity, out.sh), ps):
            _copy_device(t.data_ptr(), p.value, t.numel() * t.element_size())
        return out

    def set_config(self, tile_w=16, tile_h=16, t_min=1.0 / 8192.0, triangles=False, waves_per_cu=0,
                   config=None) -> None:
        """Tile shape, early-exit threshold, triangles=True for the reference's triangle coverage
        (VKGS_RASTER_TRIANGLES) instead of the analytic square, and a resident-wave cap of the
        raster (0: none); or a whole RasterConfig (config=, e.g. the autotuner's)."""
        cfg = _lib.RasterConfig()
        if config is not None:
            C.memmove(C.byref(cfg), C.byref(config), C.sizeof(cfg))
        else:
            cfg.tile_w, cfg.tile_h, cfg.t_min = tile_w, tile_h, t_min
            cfg.flags = RASTER_TRIANGLES if triangles else 0
            cfg.waves_per_cu = waves_per_cu
        check(lib().vkgs_renderer_set_config(self._h, C.byref(cfg)), "vkgs_renderer_set_config")

    def render(self, cam, out: torch.Tensor | None = None, model=None) -> torch.Tensor:
        w, h = cam.screen_size[0], cam.screen_size[1]
        if out is None:
            out = torch.empty((h, w, 4), dtype=torch.float32, device=_dev())
        m, mp = _model_ptr(model)
        check(lib().vkgs_renderer_render(self._h, C.byref(cam), mp, ptr(out), _stream()), "vkgs_renderer_render")
        return out

    def frame_info(self):
        v = C.c_uint32()
        k = C.c_uint64()
        st = lib().vkgs_renderer_frame_info(self._h, C.byref(v), C.byref(k), _stream())
        return st, int(v.value), int(k.value)

    def sorted_ids(self):
        ip, kp = C.c_void_p(), C.c_void_p()
        check(lib().vkgs_renderer_sorted(self._h, C.byref(ip), C.byref(kp)), "vkgs_renderer_sorted")
        st, v, _ = self.frame_info()
        check(st, "vkgs_renderer_frame_info")
        ids = _download_u32(ip.value, v)
        keys = _download_u32(kp.value, v)
        return ids, keys

    def tile_lists(self):
        rp, lp = C.c_void_p(), C.c_void_p()
        tx, ty = C.c_uint32(), C.c_uint32()
        check(lib().vkgs_renderer_tile_lists(self._h, C.byref(rp), C.byref(lp), C.byref(tx), C.byref(ty)),
              "vkgs_renderer_tile_lists")
        # an overflowing frame's pair count exceeds the list storage: its lists are incomplete
        st, _, k = self.frame_info()
        check(st, "vkgs_renderer_frame_info")
        tiles = tx.value * ty.value
        ranges = _download_u32(rp.value, 2 * tiles).reshape(tiles, 2).astype(np.int64)
        return ranges, _download_u32(lp.value, k)

    def autotune(self, cam, out: torch.Tensor | None = None, frames: int = 10, model=None):
        """Picks the fastest tile shape (8x8 one or two tiles per wave, ...) and resident-wave cap for
        this camera and keeps it; returns (tile_w, tile_h, ms_per_frame).  The whole chosen
        RasterConfig is self.tuned afterwards."""
        w, h = cam.screen_size[0], cam.screen_size[1]
        if out is None:
            out = torch.empty((h, w, 4), dtype=torch.float32, device=_dev())
        m, mp = _model_ptr(model)
        cfg = _lib.RasterConfig()
        ms = C.c_float()
        check(lib().vkgs_renderer_autotune(self._h, C.byref(cam), mp, ptr(out), frames, C.byref(cfg), C.byref(ms),
                                           _stream()), "vkgs_renderer_autotune")
        self.tuned = cfg
        return cfg.tile_w, cfg.tile_h, float(ms.value)

    def set_timing(self, on: bool) -> None:
        check(lib().vkgs_renderer_set_timing(self._h, 1 if on else 0), "vkgs_renderer_set_timing")

    def stage_ms(self):
        a = (C.c_float * 8)()
        check(lib().vkgs_renderer_stage_ms(self._h, a), "vkgs_renderer_stage_ms")
        return dict(zip(STAGES, list(a)))

    def raster_work(self) -> int:
        """List entries the raster's waves composited in the last timed frame (64 lanes each)."""
        w = C.c_uint64()
        check(lib().vkgs_renderer_raster_work(self._h, C.byref(w)), "vkgs_renderer_raster_work")
        return int(w.value)


def _copy_device(dst: int, src: int, nbytes: int) -> None:
    """Synchronous device-to-device copy (hipMemcpyDeviceToDevice = 3)."""
    if nbytes == 0:
        return
    rc = _hip().hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")


def _download_u32(addr: int, count: int) -> np.ndarray:
    out = np.empty(max(count, 0), np.uint32)
    if count:
        rc = _hip().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(addr), C.c_size_t(4 * count), 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")
    return out


_hip_lib = None


def _hip():
    global _hip_lib
    if _hip_lib is None:
        _hip_lib = C.CDLL("libamdhip64.so")
        _hip_lib.hipMemcpy.restype = C.c_int
        _hip_lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip_lib
