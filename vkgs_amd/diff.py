"""torch.autograd binding of the differentiable, batched splat path (vkgs_diff_*).

    img = render(means, log_scales, quats, opacity_logits, sh, camera)            # one view
    imgs = render_batch([scene0, scene1], [(0, cam_a), (0, cam_b), (1, cam_c)])  # many views

Parameters follow the PLY conventions (quats are rot_0..rot_3 = w, x, y, z; SH
are (N, 48) in the vkgs channel-major order).  Images are (H, W, 4) RGBA fp32
with the reference's clear colour (0, 0, 0, 1).  Forward and backward run
hand-written HIP kernels; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

_P = C.c_void_p


class DiffScene(C.Structure):
    _fields_ = [("means", _P), ("log_scales", _P), ("quats", _P), ("opacity_logits", _P), ("sh", _P),
                ("grad_means", _P), ("grad_log_scales", _P), ("grad_quats", _P), ("grad_opacity_logits", _P),
                ("grad_sh", _P), ("n", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class DiffView(C.Structure):
    _fields_ = [("camera", _lib.Camera), ("model", C.c_float * 16), ("scene", C.c_uint32),
                ("reserved", C.c_uint32 * 3), ("rgba", _P), ("grad_rgba", _P)]


_SIGS = {
    "vkgs_diff_create": (C.c_int, [_P]),
    "vkgs_diff_destroy": (C.c_int, [_P]),
    "vkgs_diff_forward": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, _P]),
    "vkgs_diff_backward": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P]),
    "vkgs_diff_info": (C.c_int, [_P, _P, _P, _P]),
    "vkgs_diff_check": (C.c_int, [_P, _P, _P]),
    "vkgs_diff_set_timing": (C.c_int, [_P, C.c_int]),
    "vkgs_diff_stage_ms": (C.c_int, [_P, _P]),
    "vkgs_diff_raster_work": (C.c_int, [_P, _P, _P]),
    "vkgs_diff_autotune": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, _P, _P, _P]),
}


def _L():
    h = lib()
    for name, (res, args) in _SIGS.items():
        fn = getattr(h, name)
        fn.restype, fn.argtypes = res, args
    return h


class _Ctx:
    """One vkgs_diff_ctx: holds a forward's state until its backward."""

    def __init__(self):
        self.h = C.c_void_p()
        check(_L().vkgs_diff_create(C.byref(self.h)), "vkgs_diff_create")

    def __del__(self):
        try:
            if self.h:
                _L().vkgs_diff_destroy(self.h)
        except Exception:
            pass


class _Pool:
    def __init__(self):
        self.free = []
        self.lock = threading.Lock()

    def get(self) -> _Ctx:
        with self.lock:
            return self.free.pop() if self.free else _Ctx()

    def put(self, c: _Ctx) -> None:
        with self.lock:
            self.free.append(c)


_pool = _Pool()
# visible (view, splat) items and (tile, splat) pairs of the last forward
last_info = {"visible": 0, "pairs": 0, "views": 0}

PARAM_NAMES = ("means", "log_scales", "quats", "opacity_logits", "sh")
IDENTITY = np.eye(4, dtype=np.float32).ravel()


def _aligned(t: torch.Tensor) -> torch.Tensor:
    """t itself if its data is 16-byte aligned (the SH rows are read as float4), else an aligned copy."""
    return t if t.data_ptr() % 16 == 0 else t.clone()


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


SPATIAL_ORDER = 1  # VKGS_DIFF_SPATIAL_ORDER
GLOBAL_TILE_KEYS = 16  # VKGS_DIFF_GLOBAL_TILE_KEYS


def _config(tile=(8, 8), t_min=1.0 / 8192.0, spatial_order=False, config=None):
    """The raster configuration of a call: `config` (a RasterConfig, e.g. from autotune()) as
    given, else one built from tile, t_min and spatial_order."""
    if config is not None:
        cfg = _lib.RasterConfig()
        C.memmove(C.byref(cfg), C.byref(config), C.sizeof(cfg))
        return cfg
    cfg = _lib.RasterConfig()
    cfg.tile_w, cfg.tile_h, cfg.t_min = tile[0], tile[1], t_min
    cfg.flags = SPATIAL_ORDER if spatial_order else 0
    return cfg


def _scene_structs(groups, grads=None):
    arr = (DiffScene * len(groups))()
    for i, g in enumerate(groups):
        s = arr[i]
        s.means, s.log_scales, s.quats, s.opacity_logits, s.sh = (t.data_ptr() for t in g)
        s.n = g[0].shape[0]
        if grads is not None:
            (s.grad_means, s.grad_log_scales, s.grad_quats, s.grad_opacity_logits,
             s.grad_sh) = (t.data_ptr() for t in grads[i])
    return arr


def _view_structs(views, outs, grads=None):
    arr = (DiffView * len(views))()
    for i, (scene, cam, model) in enumerate(views):
        v = arr[i]
        v.camera = cam
        m = IDENTITY if model is None else np.ascontiguousarray(model, np.float32).ravel()
        v.model[:] = [float(x) for x in m]
        v.scene = scene
        v.rgba = outs[i].data_ptr()
        if grads is not None:
            v.grad_rgba = grads[i].data_ptr()
    return arr


def _spans(views):
    """(offset, height, width) of each view's pixels in a flat (P, 4) buffer."""
    spans, off = [], 0
    for _, cam, _ in views:
        w, h = cam.screen_size[0], cam.screen_size[1]
        spans.append((off, h, w))
        off += h * w
    return spans, off


class _RenderBatch(torch.autograd.Function):
    """meta = (views, n_scenes, cfg, flat, own): with flat the views' images are
    slices of one (P, 4) tensor, which is the single output.  own = None: a
    pooled context, and each forward waits for its pair count (an overflowing
    forward is run again with room for it); own = a context kept by the caller,
    which checks for overflow later (vkgs_diff_check), so neither pass waits
    for the device."""

    @staticmethod
    def forward(ctx, meta, *params):
        views, n_scenes, cfg, flat, own = meta
        groups = [tuple(_aligned(p.detach().contiguous().float()) for p in params[5 * i:5 * i + 5])
                  for i in range(n_scenes)]
        dev = groups[0][0].device
        if flat:
            spans, total = _spans(views)
            buf = torch.empty((total, 4), dtype=torch.float32, device=dev)
            outs = [buf[o:o + h * w].view(h, w, 4) for o, h, w in spans]
        else:
            outs = [torch.empty((cam.screen_size[1], cam.screen_size[0], 4), dtype=torch.float32, device=dev)
                    for _, cam, _ in views]
        c = own if own is not None else _pool.get()
        sarr = _scene_structs(groups)
        varr = _view_structs(views, outs)
        if own is not None:
            check(_L().vkgs_diff_forward(c.h, sarr, n_scenes, varr, len(views), C.byref(cfg), _stream()),
                  "vkgs_diff_forward")
        # an overflowing forward reports its whole pair count, and the next forward sizes the pair
        # storage for it, so one retry suffices; a result is never returned from a truncated list
        for attempt in range(2 if own is None else 0):
            check(_L().vkgs_diff_forward(c.h, sarr, n_scenes, varr, len(views), C.byref(cfg), _stream()),
                  "vkgs_diff_forward")
            vis, pairs = C.c_uint32(), C.c_uint64()
            st = _L().vkgs_diff_info(c.h, C.byref(vis), C.byref(pairs), _stream())
            if st == 0:
                last_info.update(visible=int(vis.value), pairs=int(pairs.value), views=len(views))
                break
            if st != 6 or attempt == 1:  # VKGS_ERR_OVERFLOW twice: raise, never a truncated image
                _pool.put(c)
                check(st, "vkgs_diff_info")
        ctx.c = c
        ctx.meta = meta
        ctx.groups = groups
        ctx.save_for_backward(*params)
        return buf if flat else tuple(outs)

    @staticmethod
    def backward(ctx, *grad_outs):
        views, n_scenes, cfg, flat, own = ctx.meta
        groups = ctx.groups
        if flat:
            spans, total = _spans(views)
            g = grad_outs[0]
            g = (torch.zeros((total, 4), dtype=torch.float32, device=groups[0][0].device) if g is None
                 else g.contiguous().float())
            grad_outs = [g[o:o + h * w].view(h, w, 4) for o, h, w in spans]
        gouts = []
        for (scene, cam, _), g in zip(views, grad_outs):
            if g is None:
                g = torch.zeros((cam.screen_size[1], cam.screen_size[0], 4), dtype=torch.float32,
                                device=groups[0][0].device)
            gouts.append(g.contiguous().float())
        # the backward writes every gradient element, so no zero fill
        grads = [tuple(torch.empty_like(t) for t in g) for g in groups]
        outs_dummy = gouts  # rgba pointers are not used by the backward
        sarr = _scene_structs(groups, grads)
        varr = _view_structs(views, outs_dummy, gouts)
        check(_L().vkgs_diff_backward(ctx.c.h, sarr, n_scenes, varr, len(views), _stream()), "vkgs_diff_backward")
        if own is None:
            _pool.put(ctx.c)
        ctx.c = None
        flat = []
        for g, p in zip([t for gg in grads for t in gg], ctx.saved_tensors):
            flat.append(g.to(p.dtype) if p.requires_grad else None)
        return (None, *flat)


def _batch_args(scenes, views):
    vs = [(v[0], v[1], v[2] if len(v) > 2 else None) for v in views]
    params = []
    for s in scenes:
        params += [s[k] for k in PARAM_NAMES]
    return vs, params


def render_batch(scenes, views, tile=(8, 8), t_min=1.0 / 8192.0, spatial_order=False, config=None):
    """scenes: list of dicts with PARAM_NAMES tensors (cuda);
    views: list of (scene_index, camera[, model]).  Returns a list of (H, W, 4) images.
    spatial_order: process the splats of scenes with four or more views in this call in a spatial
    (Morton) order (faster; equal depth keys then blend in that order instead of ascending id).
    config: a RasterConfig (from autotune()) used instead of tile, t_min and spatial_order."""
    vs, params = _batch_args(scenes, views)
    cfg = _config(tile, t_min, spatial_order, config)
    return list(_RenderBatch.apply((vs, len(scenes), cfg, False, None), *params))


def render_batch_flat(scenes, views, tile=(8, 8), t_min=1.0 / 8192.0, spatial_order=False, config=None):
    """Same as render_batch, with every view's pixels in one (P, 4) tensor: returns (flat, spans),
    spans[i] = (offset, H, W) of view i, whose image is flat[offset:offset + H * W].view(H, W, 4)
    (see view_images).  A loss over all views can then be formed with a few whole-batch tensor
    operations instead of a few per view."""
    vs, params = _batch_args(scenes, views)
    flat = _RenderBatch.apply((vs, len(scenes), _config(tile, t_min, spatial_order, config), True, None), *params)
    return flat, _spans(vs)[0]


def autotune(scenes, views, t_min=1.0 / 8192.0, spatial_order=False, steps=3, grad=None):
    """Chooses the tile shape, the backward's LDS batch depth and the resident-wave cap of
    forward + backward steps over this batch (vkgs_diff_autotune): returns (RasterConfig, ms per
    step) for the config= argument of render_batch, render_batch_flat and GraphedStep.  grad: the
    (P, 4) dL/d(flat output) the steps use (default: ones).  Synchronous; the parameters' .grad
    are not touched."""
    vs, params = _batch_args(scenes, views)
    n_scenes = len(scenes)
    groups = [tuple(_aligned(p.detach().contiguous().float()) for p in params[5 * i:5 * i + 5])
              for i in range(n_scenes)]
    dev = groups[0][0].device
    spans, total = _spans(vs)
    buf = torch.empty((total, 4), dtype=torch.float32, device=dev)
    g = torch.ones((total, 4), dtype=torch.float32, device=dev) if grad is None else grad.contiguous().float()
    outs = [buf[o:o + h * w].view(h, w, 4) for o, h, w in spans]
    gouts = [g[o:o + h * w].view(h, w, 4) for o, h, w in spans]
    grads = [tuple(torch.empty_like(t) for t in gr) for gr in groups]
    sarr = _scene_structs(groups, grads)
    varr = _view_structs(vs, outs, gouts)
    base = _config(t_min=t_min, spatial_order=spatial_order)
    chosen, ms = _lib.RasterConfig(), C.c_float()
    c = _Ctx()
    check(_L().vkgs_diff_autotune(c.h, sarr, n_scenes, varr, len(vs), C.byref(base), steps, C.byref(chosen),
                                  C.byref(ms), _stream()), "vkgs_diff_autotune")
    return chosen, float(ms.value)


def describe(cfg) -> str:
    """A RasterConfig as text: tile shape, pair form, LDS batch, wave cap, t_min."""
    pairs = {4: " pairs", 8: " single"}.get(cfg.flags & 12, "")
    return (f"{cfg.tile_w}x{cfg.tile_h}{pairs} lds_batch={cfg.lds_batch or 16} "
            f"waves/CU={cfg.waves_per_cu or 'max'} t_min={cfg.t_min:g}")


class GraphedStep:
    """A forward + backward step over a fixed batch of views, captured once into a hipGraph and
    replayed: every launch of both passes, and the elementwise gradient of the loss between them,
    go to the GPU as one graph launch with no host synchronization.

        step = GraphedStep([params_a, params_b], views, lambda flat: torch.sign(flat - target) * w)
        step.step()     # replays; p.grad of every parameter tensor holds the step's gradients
        step.check()    # raises VkgsError if a replay since the last check overflowed its pair list

    grad_fn(flat) -> dL/d(flat) for render_batch_flat's (P, 4) output, made of elementwise tensor
    operations (no reductions: a reduction's internal memsets are not safe to replay).  The
    construction runs `warmup` eager steps (sizing every buffer of the step's own context, with the
    overflow retry of the eager path), then captures.  Each replay overwrites the .grad tensors
    (they are the graph's outputs, not accumulated).  The views, parameter tensors and the spatial
    processing order are those of the capture; after check() reports an overflow, build a new step
    (its warm-up sizes the pair storage for the longer list)."""

    def __init__(self, scenes, views, grad_fn, tile=(8, 8), t_min=1.0 / 8192.0, spatial_order=False, warmup=2,
                 config=None):
        self.vs, self.params = _batch_args(scenes, views)
        self.n_scenes = len(scenes)
        self.cfg = _config(tile, t_min, spatial_order, config)
        self.grad_fn = grad_fn
        self.ctx = _Ctx()
        self.spans = _spans(self.vs)[0]
        # eager warm-up on a side stream, as torch's graph capture requires (work of the warm-up on
        # the default stream before a capture breaks it)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(max(warmup, 1)):
                self._zero_grad()
                flat = _RenderBatch.apply((self.vs, self.n_scenes, self.cfg, True, None), *self.params)
                flat.backward(grad_fn(flat.detach()))
            # the step's own context, sized by eager steps of its own (an overflowing one sizes the
            # pair storage of the next for its whole list)
            for attempt in range(2):
                self._zero_grad()
                self._run()
                st = _L().vkgs_diff_check(self.ctx.h, None, _stream())
                if st == 0:
                    break
                if st != 6 or attempt == 1:
                    check(st, "vkgs_diff_check")
        torch.cuda.current_stream().wait_stream(side)
        self._zero_grad()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.flat = self._run()

    def _zero_grad(self):
        for p in self.params:
            p.grad = None

    def _run(self):
        flat = _RenderBatch.apply((self.vs, self.n_scenes, self.cfg, True, self.ctx), *self.params)
        flat.backward(self.grad_fn(flat.detach()))
        return flat

    def step(self):
        """Replays the captured step; returns the (P, 4) rendered output (overwritten by the next replay)."""
        self.graph.replay()
        return self.flat.detach()

    def check(self) -> None:
        """Synchronous: raises VkgsError (VKGS_ERR_OVERFLOW) if any replay since the last check
        overflowed its pair list."""
        check(_L().vkgs_diff_check(self.ctx.h, None, _stream()), "vkgs_diff_check")


# backward_reduce is empty: the per-slot sum of the pair partials runs inside the projection backward
DIFF_STAGES = ("forward_prep", "forward_raster", "forward_total", "backward_raster", "backward_reduce",
               "backward_project", "backward_total")


def profile_step(scenes, views, grad_fn, tile=(8, 8), t_min=1.0 / 8192.0, spatial_order=False, config=None):
    """One eager forward + backward of render_batch_flat's form with per-stage HIP events and the
    raster kernels' work counters (vkgs_diff_set_timing): returns {"ms": {stage: ms}, "forward_entries",
    "backward_entries"} (list entries stepped through, summed over waves; 64 x each = executed (pixel,
    entry) evaluations).  grad_fn(flat) -> dL/d(flat).  Synchronous; the parameters' .grad receive
    the step's gradients (accumulated as autograd does)."""
    vs, params = _batch_args(scenes, views)
    cfg = _config(tile, t_min, spatial_order, config)
    c = _Ctx()
    check(_L().vkgs_diff_set_timing(c.h, 1), "vkgs_diff_set_timing")
    for attempt in range(2):
        flat = _RenderBatch.apply((vs, len(scenes), cfg, True, c), *params)
        flat.backward(grad_fn(flat.detach()))
        st = _L().vkgs_diff_check(c.h, None, _stream())
        if st == 0:
            break
        if st != 6 or attempt == 1:  # the retry is sized for the whole list
            check(st, "vkgs_diff_check")
    ms = (C.c_float * 8)()
    check(_L().vkgs_diff_stage_ms(c.h, ms), "vkgs_diff_stage_ms")
    fe, be = C.c_uint64(), C.c_uint64()
    check(_L().vkgs_diff_raster_work(c.h, C.byref(fe), C.byref(be)), "vkgs_diff_raster_work")
    return {"ms": dict(zip(DIFF_STAGES, list(ms))), "forward_entries": int(fe.value),
            "backward_entries": int(be.value)}


def view_images(flat, spans):
    """The (H, W, 4) image of each view as a differentiable view of render_batch_flat's output."""
    return [flat[o:o + h * w].view(h, w, 4) for o, h, w in spans]


def render(means, log_scales, quats, opacity_logits, sh, camera, model=None, tile=(8, 8), t_min=1.0 / 8192.0,
           config=None):
    """One view of one scene; differentiable w.r.t. every parameter tensor."""
    scene = dict(means=means, log_scales=log_scales, quats=quats, opacity_logits=opacity_logits, sh=sh)
    return render_batch([scene], [(0, camera, model)], tile, t_min, config=config)[0]


def scene_tensors(sc, device="cuda", requires_grad=True):
    """Raw parameters of a vkgs_amd.scene.Scene as cuda tensors."""
    arrs = dict(means=sc.means(), log_scales=sc.log_scales(), quats=sc.quats_wxyz(),
                opacity_logits=sc.opacity_logits(), sh=sc.sh_channel_major())
    return {k: torch.tensor(np.ascontiguousarray(v), dtype=torch.float32, device=device,
                            requires_grad=requires_grad) for k, v in arrs.items()}


def smoke_backward() -> None:
    """One forward + backward of a small scene; gradients must be finite and nonzero."""
    from . import scene as S
    sc = S.gaussian_cloud(2000, seed=3, pos_std=1.0, logscale=(-2.5, 0.4))
    cam = S.orbit_camera(64, 64, 4.0)
    p = scene_tensors(sc)
    img = render(**p, camera=cam)
    loss = (img * torch.linspace(0.5, 1.5, img.numel(), device=img.device).reshape(img.shape)).sum()
    loss.backward()
    for k, t in p.items():
        assert t.grad is not None and torch.isfinite(t.grad).all(), k
    assert float(p["means"].grad.abs().sum()) > 0
