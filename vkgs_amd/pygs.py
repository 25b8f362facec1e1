"""The reference's pygs surface (binding/python/pygs/__init__.py:14-32) on the HIP renderer.

    show()          starts the render loop on its own thread (pygs_cpp/main.cc:48-68);
                    this build has no window, so frames go to latest_frame()
    load(path)      FileNotFoundError for a missing file (__init__.py:18-25), then an
                    asynchronous load like Engine::LoadSplatsAsync (engine.cc:546-549)
    close()         stops the loop (main.cc:76-80)

Added for programmatic use:
    render(...)     one synchronous frame as a (H, W, 4) cuda tensor
    rasterize(...)  the differentiable renderer (vkgs_amd.diff.render)
"""
from __future__ import annotations

import errno
import math
import os
import threading

from . import diff as _diff
from . import scene as _scene

_lock = threading.Lock()
_state = {"renderer": None, "thread": None, "stop": False, "pending": None, "frame": None, "camera": None,
          "size": (1600, 900), "error": None}

rasterize = _diff.render
rasterize_batch = _diff.render_batch


def _renderer(capacity: int):
    from .pipeline import Renderer
    r = _state["renderer"]
    if r is None or r.capacity < capacity:
        if r is not None:
            r.close()
        r = Renderer(max(capacity, 1))
        _state["renderer"] = r
    return r


def _load_now(path: str) -> None:
    import ctypes as C
    from . import _lib
    info = _lib.PlyInfo()
    _lib.check(_lib.lib().vkgs_ply_read_header(path.encode(), C.byref(info)), "vkgs_ply_read_header")
    r = _renderer(int(info.point_count))
    r.load_ply(path)


def _render_checked(r, cam):
    """One frame; a frame whose tile lists outgrew the renderer's pair storage is rendered again
    (the renderer then sizes the storage for the whole list), so no incomplete frame is returned."""
    for attempt in range(2):
        img = r.render(cam)
        st, _, _ = r.frame_info()  # synchronous
        if st == 0:
            return img
        if st != 6 or attempt == 1:  # VKGS_ERR_OVERFLOW twice, or another error
            from ._lib import check
            check(st, "vkgs_renderer_frame_info")
    return img


def _loop() -> None:
    import torch
    while True:
        with _lock:
            if _state["stop"]:
                break
            pending, _state["pending"] = _state["pending"], None
        if pending is not None:
            _load_now(pending)
        r = _state["renderer"]
        if r is None or r.n == 0:
            threading.Event().wait(0.01)
            continue
        cam = _state["camera"] or _scene.orbit_camera(*_state["size"], 2.0)
        try:
            img = _render_checked(r, cam)
            torch.cuda.synchronize()
        except Exception as e:  # reported through latest_error(); the loop stops
            with _lock:
                _state["error"] = e
                _state["stop"] = True
            break
        with _lock:
            _state["frame"] = img


def show() -> None:
    """Starts the render loop if it is not running (main.cc:48-68)."""
    with _lock:
        t = _state["thread"]
        if t is not None and t.is_alive():
            print("[pygs] viewer is already running")
            return
        _state["stop"] = False
        _state["error"] = None
        t = threading.Thread(target=_loop, name="vkgs-render", daemon=True)
        _state["thread"] = t
    t.start()


def load(ply_filepath: str) -> None:
    ply_filepath = os.path.abspath(ply_filepath)
    if not os.path.exists(ply_filepath):
        raise FileNotFoundError(errno.ENOENT, os.strerror(errno.ENOENT), ply_filepath)
    print(f"load {ply_filepath}")
    with _lock:
        running = _state["thread"] is not None and _state["thread"].is_alive()
        if running:
            _state["pending"] = ply_filepath
            return
    _load_now(ply_filepath)


def close() -> None:
    with _lock:
        _state["stop"] = True
        t = _state["thread"]
    if t is not None:
        t.join()
    with _lock:
        _state["thread"] = None


def set_camera(width: int = 1600, height: int = 900, r: float = 2.0, phi: float = math.radians(45),
               theta: float = math.radians(45), center=(0.0, 0.0, 0.0), fovy: float = math.radians(60)) -> None:
    """Orbit camera of the loop (camera.h:42-57 defaults)."""
    cam = _scene.orbit_camera(width, height, r, phi, theta, center, fovy)
    with _lock:
        _state["camera"] = cam
        _state["size"] = (width, height)


def latest_frame():
    with _lock:
        return _state["frame"]


def latest_error():
    """The exception that stopped the render loop, if any."""
    with _lock:
        return _state["error"]


def render(camera=None, width: int = 1600, height: int = 900, r: float = 2.0):
    """One frame of the loaded splats, synchronously."""
    rnd = _state["renderer"]
    if rnd is None or rnd.n == 0:
        raise RuntimeError("no splats loaded: call pygs.load(path) first")
    cam = camera or _scene.orbit_camera(width, height, r)
    return _render_checked(rnd, cam)


__all__ = ["show", "load", "close", "render", "set_camera", "latest_frame", "latest_error", "rasterize",
           "rasterize_batch"]
