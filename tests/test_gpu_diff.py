"""Differentiable, batched path on the GPU vs the float64 torch restatement.

Bars: forward image within fp32-vs-fp64 agreement (all but a handful of
coverage-edge pixels within 1e-4); gradients of L = sum(w * image) with
relative error ||g - g_ref|| / ||g_ref|| <= 1e-3 per parameter group; a batch
of views of several scenes gives the same images as one view at a time.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import oracle  # noqa: E402
from oracle import torch_ref  # noqa: E402
from vkgs_amd import diff as D  # noqa: E402
from vkgs_amd import pipeline as P  # noqa: E402
from vkgs_amd import scene as S  # noqa: E402

from _spatial import draw_order, morton_rank  # noqa: E402
from _stack import held_rgb, stack_closed_form, stacked_scene  # noqa: E402

NAMES = D.PARAM_NAMES
DEFAULT_T_MIN = 1.0 / 8192.0
DIFF_T_MIN_FLOOR = 2.0 ** -20  # the differentiable raster stops at max(t_min, 2^-20)


def small_scene(n, seed, w, h, r=4.0):
    sc = S.gaussian_cloud(n, seed=seed, pos_std=1.0, logscale=(-2.5, 0.4))
    c = S.COL
    sc.rows[:, c["opacity"]] = np.clip(sc.rows[:, c["opacity"]], -6, 4)
    sc.rows[:, c["f_dc_0"]:c["f_rest_0"] + 45] = sc.rows[:, c["f_dc_0"]:c["f_rest_0"] + 45].astype(np.float16)
    return sc, S.orbit_camera(w, h, r)


def ref_inputs(sc, cam):
    pos, cov, opa, sh = oracle.parse_ply(sc.rows, S.standard_offsets())
    v, key, index = oracle.rank(cam, pos)
    _, order = oracle.sort_kv(key, index)
    Pm, Vm, eye = S.camera_arrays(cam)
    cam64 = (torch.tensor(Pm.reshape(4, 4).T, dtype=torch.float64), torch.tensor(Vm.reshape(4, 4).T, dtype=torch.float64),
             torch.tensor(eye, dtype=torch.float64))
    p64 = [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, requires_grad=True)
           for a in (sc.means(), sc.log_scales(), sc.quats_wxyz(), sc.opacity_logits(), sc.sh_channel_major())]
    return order.astype(np.int64), cam64, p64


def test_diff_forward_and_gradients_match_float64():
    W, H = 64, 64
    sc, cam = small_scene(2000, 31, W, H)
    order, cam64, p64 = ref_inputs(sc, cam)
    rng = np.random.default_rng(1)
    w = rng.uniform(0.5, 1.5, (H, W, 4))
    ref_img = torch_ref.render(*p64, *cam64, W, H, order)
    (ref_img * torch.tensor(w)).sum().backward()
    p = D.scene_tensors(sc)
    img = D.render(**p, camera=cam, t_min=0.0)
    (img * torch.tensor(w, dtype=torch.float32, device="cuda")).sum().backward()
    torch.cuda.synchronize()
    d = np.abs(img.detach().cpu().numpy() - ref_img.detach().numpy()).max(-1)
    print("forward: pixels > 1e-4:", int((d > 1e-4).sum()), "max", float(d.max()), flush=True)
    assert (d > 1e-4).mean() < 0.005 and d.max() < 0.02
    for name, t64 in zip(NAMES, p64):
        g = p[name].grad.detach().cpu().double()
        gr = t64.grad.detach()
        rel = float(torch.linalg.norm(g - gr) / torch.linalg.norm(gr))
        print(f"grad {name}: rel err {rel:.2e}", flush=True)
        assert rel <= 1e-3, (name, rel)


def test_diff_forward_matches_plain_renderer():
    W, H = 320, 200
    sc, cam = small_scene(20000, 32, W, H, r=5.0)
    r = P.Renderer(sc.n)
    r.load_buffer(sc.ply_buffer(), sc.n)
    r.set_config(8, 8, 0.0)
    plain = r.render(cam).cpu().numpy()
    p = D.scene_tensors(sc, requires_grad=False)
    img = D.render(**p, camera=cam, t_min=0.0).cpu().numpy()
    d = np.abs(img - plain).max(-1)
    print("diff vs plain: pixels > 1e-5:", int((d > 1e-5).sum()), "max", float(d.max()), flush=True)
    assert (d > 1e-5).mean() < 0.005 and d.max() < 0.02
    r.close()


def test_batch_equals_single_views():
    sa, cam_a = small_scene(3000, 33, 96, 64)
    sb, _ = small_scene(1500, 34, 96, 64)
    cam_b = S.orbit_camera(128, 96, 5.0, theta=1.1)
    cam_c = S.orbit_camera(64, 64, 3.5, phi=1.2, theta=-0.4)
    views = [(0, cam_a), (1, cam_b), (0, cam_c), (1, cam_a)]
    pa, pb = D.scene_tensors(sa), D.scene_tensors(sb)
    rng = np.random.default_rng(2)
    outs = D.render_batch([pa, pb], views, t_min=0.0)
    ws = [torch.tensor(rng.uniform(0.5, 1.5, tuple(o.shape)), dtype=torch.float32, device="cuda") for o in outs]
    sum((o * w).sum() for o, w in zip(outs, ws)).backward()
    gb = {k: (pa[k].grad.clone(), pb[k].grad.clone()) for k in NAMES}
    singles = []
    for k in NAMES:
        pa[k].grad = None
        pb[k].grad = None
    for (si, cam), w in zip(views, ws):
        prm = pa if si == 0 else pb
        img = D.render(**prm, camera=cam, t_min=0.0)
        (img * w).sum().backward()
        singles.append(img.detach())
    for o, s1 in zip(outs, singles):
        np.testing.assert_array_equal(o.detach().cpu().numpy(), s1.cpu().numpy())
    for k in NAMES:
        for gbat, gsin in zip(gb[k], (pa[k].grad, pb[k].grad)):
            rel = float(torch.linalg.norm(gbat - gsin) / torch.linalg.norm(gsin))
            assert rel < 1e-5, (k, rel)


def test_early_termination_gradients_are_finite():
    sc, cam = small_scene(5000, 35, 128, 96)
    p = D.scene_tensors(sc)
    img = D.render(**p, camera=cam)  # default t_min = 1/8192
    img.sum().backward()
    for k in NAMES:
        assert torch.isfinite(p[k].grad).all(), k


def test_backward_writes_zeros_for_untouched_splats_and_viewless_scenes():
    """The backward writes every gradient element: splats behind the camera of
    the only view, and every splat of a scene that no view renders, get exact
    zeros (the gradient tensors are not pre-zeroed)."""
    sa, cam = small_scene(1000, 36, 64, 48)
    sb, _ = small_scene(700, 37, 64, 48)
    pa, pb = D.scene_tensors(sa), D.scene_tensors(sb)
    # a third of scene a far behind the camera: culled, so untouched
    eye = np.asarray(S.camera_arrays(cam)[2][:3], np.float64)
    behind = np.arange(0, sa.n, 3)
    with torch.no_grad():
        pa["means"][behind] = torch.tensor(2.0 * eye, dtype=torch.float32, device="cuda")
    outs = D.render_batch([pa, pb], [(0, cam)], t_min=0.0)  # scene b: no view
    outs[0].sum().backward()
    for k in NAMES:
        assert torch.isfinite(pa[k].grad).all(), k
        assert torch.count_nonzero(pa[k].grad[behind]) == 0, k
        assert torch.count_nonzero(pb[k].grad) == 0, k
    assert torch.count_nonzero(pa["means"].grad) > 0


def test_backward_is_deterministic_with_8x8_tiles():
    """No global atomics and one writer per LDS partial: two backward passes
    over the same forward inputs give bit-identical gradients."""
    sa, cam_a = small_scene(4000, 38, 96, 80)
    cam_b = S.orbit_camera(128, 64, 3.0, theta=0.7)
    grads = []
    for _ in range(2):
        p = D.scene_tensors(sa)
        outs = D.render_batch([p], [(0, cam_a), (0, cam_b)], tile=(8, 8))
        sum(o.square().sum() for o in outs).backward()
        grads.append({k: p[k].grad.detach().cpu().numpy().copy() for k in NAMES})
    for k in NAMES:
        np.testing.assert_array_equal(grads[0][k], grads[1][k])


@pytest.mark.parametrize("tile", [(16, 16), (16, 8), (32, 4)])
def test_multi_wave_tiles_match_8x8(tile):
    """Tiles of several waves take the LDS-atomic path of the backward and one
    list entry per thread per batch; images match the 8x8 path exactly and
    gradients to float rounding."""
    sc, cam = small_scene(3000, 39, 112, 72)
    ref_p, p = D.scene_tensors(sc), D.scene_tensors(sc)
    w = torch.tensor(np.random.default_rng(3).uniform(0.5, 1.5, (72, 112, 4)), dtype=torch.float32, device="cuda")
    img_ref = D.render(**ref_p, camera=cam, tile=(8, 8), t_min=0.0)
    (img_ref * w).sum().backward()
    img = D.render(**p, camera=cam, tile=tile, t_min=0.0)
    (img * w).sum().backward()
    np.testing.assert_array_equal(img.detach().cpu().numpy(), img_ref.detach().cpu().numpy())
    for k in NAMES:
        g, gr = p[k].grad, ref_p[k].grad
        rel = float(torch.linalg.norm(g - gr) / torch.linalg.norm(gr))
        assert rel < 1e-5, (k, rel)


def test_pair_overflow_retries_and_matches_a_roomy_context(monkeypatch):
    """Metre-sized splats on a fresh context need more than 8x its first pair storage: the
    overflowing forward stays in bounds and reports the whole pair count, the one retry sizes
    the storage for it, and the result and its backward match those of a context that already
    has room."""
    W, H = 768, 512
    sc = S.gaussian_cloud(4000, seed=40, pos_std=1.0, logscale=(0.0, 0.0))
    cam = S.orbit_camera(W, H, 4.0)
    w = torch.tensor(np.random.default_rng(4).uniform(0.5, 1.5, (H, W, 4)), dtype=torch.float32, device="cuda")
    monkeypatch.setattr(D, "_pool", D._Pool())  # fresh context: smallest pair storage
    runs = []
    for _ in range(2):
        p = D.scene_tensors(sc)
        img = D.render(**p, camera=cam, tile=(8, 8), t_min=0.0)
        (img * w).sum().backward()
        runs.append((img.detach().cpu().numpy(), {k: p[k].grad.detach().cpu().numpy() for k in NAMES}))
        pairs = D.last_info["pairs"]
        print("pairs", pairs, flush=True)
        assert pairs > 8 * (4 * 4096 + (1 << 20))  # 8x the first storage of a 4000-splat context
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    for k in NAMES:
        assert np.isfinite(runs[0][1][k]).all(), k
        np.testing.assert_array_equal(runs[0][1][k], runs[1][1][k])


def test_flat_batch_matches_list_batch():
    """render_batch_flat: the same images as render_batch (bit-exact) and, for the same loss,
    the same gradients; view_images gives each view's image."""
    sa, cam_a = small_scene(2500, 41, 96, 64)
    sb, _ = small_scene(1200, 42, 96, 64)
    cam_b = S.orbit_camera(80, 48, 4.5, theta=0.9)
    views = [(0, cam_a), (1, cam_b), (0, cam_b)]
    pa, pb = D.scene_tensors(sa), D.scene_tensors(sb)
    rng = np.random.default_rng(5)
    ws = [torch.tensor(rng.uniform(0.5, 1.5, (c.screen_size[1], c.screen_size[0], 4)), dtype=torch.float32,
                       device="cuda") for _, c in views]
    outs = D.render_batch([pa, pb], views, t_min=0.0)
    sum((o * w).sum() for o, w in zip(outs, ws)).backward()
    ref = {k: (pa[k].grad.clone(), pb[k].grad.clone()) for k in NAMES}
    for k in NAMES:
        pa[k].grad = None
        pb[k].grad = None
    flat, spans = D.render_batch_flat([pa, pb], views, t_min=0.0)
    assert flat.shape == (sum(h * w for _, h, w in spans), 4)
    imgs = D.view_images(flat, spans)
    for o, im in zip(outs, imgs):
        np.testing.assert_array_equal(o.detach().cpu().numpy(), im.detach().cpu().numpy())
    w_flat = torch.cat([w.reshape(-1, 4) for w in ws])
    (flat * w_flat).sum().backward()
    for k in NAMES:
        for g, gr in zip((pa[k].grad, pb[k].grad), ref[k]):
            rel = float(torch.linalg.norm(g - gr) / torch.linalg.norm(gr))
            assert rel < 1e-5, (k, rel)


def test_unpacked_rect_path_matches_packed():
    """At 2560x1600 a tile rectangle of the 8x8 grid (320 x 200 tiles) needs more than 32 bits, so
    the forward gathers (rect, count, view) words in depth order; the 16x16 grid (160 x 100) packs
    the rectangle into the depth sort's value.  Same image, gradients to float rounding."""
    W, H = 2560, 1600
    sc = S.gaussian_cloud(6000, seed=43, pos_std=1.0, logscale=(-2.5, 0.2))
    cam = S.orbit_camera(W, H, 4.0)
    ref_p, p = D.scene_tensors(sc), D.scene_tensors(sc)
    w = torch.tensor(np.random.default_rng(6).uniform(0.5, 1.5, (H, W, 4)), dtype=torch.float32, device="cuda")
    img_ref = D.render(**ref_p, camera=cam, tile=(16, 16), t_min=0.0)
    (img_ref * w).sum().backward()
    img = D.render(**p, camera=cam, tile=(8, 8), t_min=0.0)
    (img * w).sum().backward()
    np.testing.assert_array_equal(img.detach().cpu().numpy(), img_ref.detach().cpu().numpy())
    for k in NAMES:
        g, gr = p[k].grad, ref_p[k].grad
        rel = float(torch.linalg.norm(g - gr) / torch.linalg.norm(gr))
        assert rel < 1e-5, (k, rel)


def test_spatial_order_matches_default_order():
    """VKGS_DIFF_SPATIAL_ORDER: a scene with four or more views is processed in Morton order.  Equal
    depth keys then blend in that order, so images and gradients differ from the default order only
    by tie effects (both orders are valid blends of the same splats; the accuracy bar against float64
    applies to the default order); the order is cached, so a second call gives identical results."""
    sc, cam = small_scene(6000, 44, 96, 72)
    cams = [cam] + [S.orbit_camera(96, 72, 4.0 + 0.3 * i, theta=0.4 * i) for i in range(1, 5)]
    views = [(0, c) for c in cams]
    w = [torch.tensor(np.random.default_rng(7 + i).uniform(0.5, 1.5, (72, 96, 4)), dtype=torch.float32,
                      device="cuda") for i in range(len(views))]
    runs = []
    for spatial in (False, True, True):
        p = D.scene_tensors(sc)
        outs = D.render_batch([p], views, t_min=0.0, spatial_order=spatial)
        sum((o * wi).sum() for o, wi in zip(outs, w)).backward()
        runs.append(([o.detach().cpu().numpy() for o in outs], {k: p[k].grad.detach().cpu().numpy() for k in NAMES}))
    (img0, g0), (img1, g1), (img2, g2) = runs
    for a, b in zip(img0, img1):
        d = np.abs(a - b).max(-1)
        print("spatial vs default: pixels > 1e-5:", int((d > 1e-5).sum()), "max", float(d.max()), flush=True)
        assert (d > 1e-5).mean() < 0.02 and d.max() < 0.02
    for k in NAMES:
        rel = float(np.linalg.norm(g1[k] - g0[k]) / np.linalg.norm(g0[k]))
        assert rel < 5e-3, (k, rel)
        np.testing.assert_array_equal(g1[k], g2[k])
    for a, b in zip(img1, img2):
        np.testing.assert_array_equal(a, b)


def test_spatial_order_mixed_scenes():
    """Spatial order applies per scene: a scene with five views is reordered, a one-view scene in the
    same call is not (its image and gradients equal the default order's bit for bit)."""
    sa, cam = small_scene(3000, 45, 80, 64)
    sb, _ = small_scene(800, 46, 80, 64)
    cams = [cam] + [S.orbit_camera(80, 64, 4.0 + 0.25 * i, theta=0.5 * i) for i in range(1, 5)]
    views = [(0, c) for c in cams] + [(1, S.orbit_camera(64, 48, 3.5, phi=1.0))]
    res = []
    for spatial in (False, True):
        pa, pb = D.scene_tensors(sa), D.scene_tensors(sb)
        outs = D.render_batch([pa, pb], views, t_min=0.0, spatial_order=spatial)
        sum(o.square().sum() for o in outs).backward()
        res.append(([o.detach().cpu().numpy() for o in outs],
                    {k: (pa[k].grad.detach().cpu().numpy(), pb[k].grad.detach().cpu().numpy()) for k in NAMES}))
    (i0, g0), (i1, g1) = res
    np.testing.assert_array_equal(i0[-1], i1[-1])  # the one-view scene keeps the default order
    for k in NAMES:
        np.testing.assert_array_equal(g0[k][1], g1[k][1])
        rel = float(np.linalg.norm(g1[k][0] - g0[k][0]) / np.linalg.norm(g0[k][0]))
        assert rel < 5e-3, (k, rel)  # tie-order effects (see test_spatial_order_matches_default_order)
    for a, b in zip(i0[:-1], i1[:-1]):
        assert np.abs(a - b).max() < 0.02


def test_full_batch_workload_equals_single_views():
    """The fwd+bwd batch workload at full size (48 views: 32 of a 3M-splat scene, 2 each of 8 scenes of
    50k-2M splats, 256^2 to 1600x900): the flat batch's images equal one-view-at-a-time renders bit for
    bit, and the gradients of one weighted loss agree within 1e-5 relative per parameter group (the
    batch sums each splat's per-view terms in another order)."""
    import bench

    scenes, views = bench.batch_workload()
    params = [D.scene_tensors(s) for s in scenes]
    rng = np.random.default_rng(7)
    weights = [torch.tensor(rng.uniform(0.5, 1.5, (c.screen_size[1], c.screen_size[0], 4)), dtype=torch.float32,
                            device="cuda") for _, c in views]
    flat, spans = D.render_batch_flat(params, views)
    imgs = D.view_images(flat, spans)
    sum((img * w).sum() for img, w in zip(imgs, weights)).backward()
    print("full batch visible", D.last_info["visible"], "pairs", D.last_info["pairs"], flush=True)
    gb = [{k: p[k].grad.clone() for k in NAMES} for p in params]
    batch_imgs = [img.detach().cpu().numpy() for img in imgs]
    del flat, imgs
    for p in params:
        for k in NAMES:
            p[k].grad = None
    for (si, cam), w, bimg in zip(views, weights, batch_imgs):
        img = D.render(**params[si], camera=cam)
        (img * w).sum().backward()
        np.testing.assert_array_equal(bimg, img.detach().cpu().numpy())
    for si, p in enumerate(params):
        for k in NAMES:
            den = float(torch.linalg.norm(p[k].grad))
            rel = float(torch.linalg.norm(gb[si][k] - p[k].grad)) / (den if den > 0 else 1.0)
            assert rel < 1e-5, (si, k, rel)


def test_view_major_tile_keys_match_global_keys():
    """Five 1600x900 views on 8x8 tiles hold 113,000 tiles (17-bit global tile ids, three radix
    passes) but at most 22,600 per view (15 bits, two passes), so the batch sorts view-major keys:
    the view above a 25-bit integer depth, then the tile within the view alone.  The tile lists are
    the ones global tile keys give (VKGS_DIFF_GLOBAL_TILE_KEYS), so images and gradients are the same
    bit for bit, in the default and in the spatial processing order; and the images equal one-view
    renders bit for bit."""
    sa = S.gaussian_cloud(40000, seed=61, pos_std=1.5, logscale=(-3.0, 0.5))
    sb = S.gaussian_cloud(9000, seed=62, pos_std=1.0, logscale=(-2.5, 0.4))
    cams = [S.orbit_camera(1600, 900, 4.0 + 0.4 * i, theta=0.7 * i) for i in range(4)]
    views = [(0, c) for c in cams] + [(1, S.orbit_camera(1600, 900, 3.5, phi=1.0))]
    rng = np.random.default_rng(8)
    weights = [torch.tensor(rng.uniform(0.5, 1.5, (900, 1600, 4)), dtype=torch.float32, device="cuda")
               for _ in views]
    def run(spatial, flags):
        cfg = D._config(t_min=1.0 / 8192.0, spatial_order=spatial)
        cfg.flags |= flags
        params = [D.scene_tensors(sa), D.scene_tensors(sb)]
        flat, spans = D.render_batch_flat(params, views, config=cfg)
        imgs = D.view_images(flat, spans)
        sum((img * w).sum() for img, w in zip(imgs, weights)).backward()
        return ([img.detach().cpu().numpy() for img in imgs],
                [{k: p[k].grad.detach().cpu().numpy() for k in NAMES} for p in params])

    for spatial in (False, True):
        local_i, local_g = run(spatial, 0)
        glob_i, glob_g = run(spatial, D.GLOBAL_TILE_KEYS)
        print("view-major vs global keys, spatial", spatial, "pairs", D.last_info["pairs"], flush=True)
        for a, b in zip(local_i, glob_i):
            np.testing.assert_array_equal(a, b)
        for ga, gb in zip(local_g, glob_g):
            for k in NAMES:
                np.testing.assert_array_equal(ga[k], gb[k])
        if not spatial:
            params = [D.scene_tensors(sa), D.scene_tensors(sb)]
            for (si, cam), img in zip(views, local_i):
                one = D.render(**params[si], camera=cam).detach().cpu().numpy()
                np.testing.assert_array_equal(img, one)


@pytest.mark.parametrize("n_views", [128, 129])
def test_view_major_keys_at_the_view_limit(n_views):
    """View-major depth keys hold the view index above the 25-bit depth: 128 views fill the 32-bit key
    (the largest view index in bits 25-31), 129 views do not fit and the batch keeps global tile keys.
    64x48 views on 8x8 tiles (48 tiles each, 6 bits, one radix pass; 6,144-6,192 tiles together, 13 bits,
    two passes), so view-major keys are chosen whenever they fit.  Either way the lists are the global
    keys' lists: images and gradients bit-identical to VKGS_DIFF_GLOBAL_TILE_KEYS, and every image equal
    to its one-view render."""
    sc, _ = small_scene(3000, 63, 64, 48)
    cams = [S.orbit_camera(64, 48, 3.5 + 0.01 * i, phi=0.5 + 0.004 * i, theta=0.049 * i) for i in range(n_views)]
    views = [(0, c) for c in cams]
    w = torch.tensor(np.random.default_rng(10).uniform(0.5, 1.5, (n_views * 48 * 64, 4)), dtype=torch.float32,
                     device="cuda")
    runs = []
    for flags in (0, D.GLOBAL_TILE_KEYS):
        cfg = D._config(t_min=1.0 / 8192.0)
        cfg.flags |= flags
        p = D.scene_tensors(sc)
        flat, spans = D.render_batch_flat([p], views, config=cfg)
        (flat * w).sum().backward()
        runs.append((flat.detach().cpu().numpy(), {k: p[k].grad.detach().cpu().numpy() for k in NAMES}, spans))
    print("views", n_views, "pairs", D.last_info["pairs"], flush=True)
    (fa, ga, spans), (fb, gb, _) = runs
    np.testing.assert_array_equal(fa, fb)
    for k in NAMES:
        np.testing.assert_array_equal(ga[k], gb[k])
    p = D.scene_tensors(sc, requires_grad=False)
    for (o, h, wd), cam in list(zip(spans, cams))[::16] + [(spans[-1], cams[-1])]:
        one = D.render(**p, camera=cam).cpu().numpy()
        np.testing.assert_array_equal(fa[o:o + h * wd].reshape(h, wd, 4), one)


def cam_t64(cam):
    Pm, Vm, eye = S.camera_arrays(cam)
    return (torch.tensor(Pm.reshape(4, 4).T, dtype=torch.float64),
            torch.tensor(Vm.reshape(4, 4).T, dtype=torch.float64), torch.tensor(eye, dtype=torch.float64))


def params_t64(sc):
    return [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, requires_grad=True)
            for a in (sc.means(), sc.log_scales(), sc.quats_wxyz(), sc.opacity_logits(), sc.sh_channel_major())]


@pytest.mark.parametrize("t_min,m_want", [(DEFAULT_T_MIN, 14), (0.0, 21)])
def test_diff_stop_rule_on_stacked_splats(t_min, m_want):
    """30 equal splats (alpha 0.5) on the centre pixel: the raster composites exactly the splats
    whose transmittance in front is >= max(t_min, 2^-20) -- 14 at the default t_min = 2^-13, 21 at
    t_min = 0 -- as the closed form says (and the float64 reference's stop rule, pinned to it)."""
    sc = stacked_scene(30, 0.0, (0.8, 0.3, 0.55))
    cam = S.orbit_camera(33, 33, 4.0)
    p = D.scene_tensors(sc, requires_grad=False)
    img = D.render(**p, camera=cam, t_min=t_min).cpu().numpy()
    want, m = stack_closed_form(0.5, held_rgb(sc), 30, max(t_min, DIFF_T_MIN_FLOOR))
    assert m == m_want
    print("centre", img[16, 16], "closed form", want, flush=True)
    np.testing.assert_allclose(img[16, 16], want, rtol=0, atol=2e-6)


def test_diff_gradients_through_saturated_stack():
    """24 near-opaque splats (alpha 0.9975) stacked on the image centre, composited with t_min = 0:
    the transmittance falls below 2^-20 after three of them, the raster stops there, and the
    backward's recovered transmittance stays exact, so the front splats' gradients (and every
    group's) match float64 autograd within 1e-3 (with no stop, the stored T is 0 and the front
    splat's gradients vanish)."""
    rgb = np.random.default_rng(11).uniform(0.1, 0.9, (24, 3))
    sc = stacked_scene(24, 6.0, rgb, log_scale=(-1.5, -1.8, -1.6), jitter=0.02, seed=12)
    W = H = 33
    cam = S.orbit_camera(W, H, 4.0)
    pos, _, _, _ = oracle.parse_ply(sc.rows, S.standard_offsets())
    order = draw_order(cam, pos)
    assert len(order) == 24
    w = np.random.default_rng(13).uniform(0.5, 1.5, (H, W, 4))
    p64 = params_t64(sc)
    ref = torch_ref.render(*p64, *cam_t64(cam), W, H, order, t_min=DIFF_T_MIN_FLOOR)
    (ref * torch.tensor(w)).sum().backward()
    p = D.scene_tensors(sc)
    img = D.render(**p, camera=cam, t_min=0.0)
    (img * torch.tensor(w, dtype=torch.float32, device="cuda")).sum().backward()
    d = np.abs(img.detach().cpu().numpy() - ref.detach().numpy()).max()
    print("stack image max abs", d, flush=True)
    assert d < 1e-4
    front = int(order[-1])
    for name, t64 in zip(NAMES, p64):
        g = p[name].grad.detach().cpu().double()
        gr = t64.grad.detach()
        rel = float(torch.linalg.norm(g - gr) / torch.linalg.norm(gr))
        rel_front = float(torch.linalg.norm(g[front] - gr[front]) / torch.linalg.norm(gr[front]))
        print(f"grad {name}: rel err {rel:.2e}, front splat {rel_front:.2e}", flush=True)
        assert float(torch.linalg.norm(gr[front])) > 0
        assert rel <= 1e-3 and rel_front <= 1e-3, (name, rel, rel_front)


def test_bench_batched_form_matches_float64():
    """The benchmarked batched fwd+bwd form, small: render_batch_flat with spatial_order=True at the
    default t_min, five views of one scene (processed in Morton order) and two of a second scene
    (default order), the weighted L1 loss over the flat output.  Against the float64 reference fed
    each view's draw order (ascending depth key; equal keys in Morton order for the first scene,
    ascending id for the second) and the same stop rule: images within 1/255 away from coverage
    edges and PSNR >= 50 dB, every gradient group within 1e-3."""
    sa, cam0 = small_scene(1500, 51, 48, 48)
    sb, _ = small_scene(800, 52, 64, 40)
    cams_a = [cam0] + [S.orbit_camera(48, 48, 4.0 + 0.3 * i, theta=0.4 * i) for i in range(1, 5)]
    cams_b = [S.orbit_camera(64, 40, 3.8, phi=1.0), S.orbit_camera(64, 40, 4.4, theta=-0.6)]
    views = [(0, c) for c in cams_a] + [(1, c) for c in cams_b]
    scenes = [sa, sb]
    ranks = [morton_rank(sa.means()), None]  # the five-view scene is the spatially ordered one
    rng = np.random.default_rng(14)
    targets = [rng.uniform(0.0, 1.0, (c.screen_size[1], c.screen_size[0], 4)) for _, c in views]
    weights = [1.0 / (4 * c.screen_size[0] * c.screen_size[1]) for _, c in views]

    params = [D.scene_tensors(s) for s in scenes]
    flat, spans = D.render_batch_flat(params, views, spatial_order=True)
    tgt_flat = torch.cat([torch.tensor(t.reshape(-1, 4), dtype=torch.float32, device="cuda") for t in targets])
    w_flat = torch.cat([torch.full((h * w_, 1), wt, device="cuda") for (_, h, w_), wt in zip(spans, weights)])
    ((flat - tgt_flat).abs() * w_flat).sum().backward()
    imgs = [im.detach().cpu().numpy() for im in D.view_images(flat, spans)]

    p64 = [params_t64(s) for s in scenes]
    pos = [oracle.parse_ply(s.rows, S.standard_offsets())[0] for s in scenes]
    loss = 0.0
    for vi, ((si, cam), tgt, wt) in enumerate(zip(views, targets, weights)):
        W, H = cam.screen_size[0], cam.screen_size[1]
        order = draw_order(cam, pos[si], ranks[si])
        ref, edge = torch_ref.render(*p64[si], *cam_t64(cam), W, H, order, t_min=DEFAULT_T_MIN,
                                     return_edge=True)
        loss = loss + ((ref - torch.tensor(tgt)).abs() * wt).sum()
        r = ref.detach().numpy()
        d = np.abs(imgs[vi] - r)
        away = edge.numpy() > 1e-4
        psnr = 10 * np.log10(1.0 / max(float(np.mean(d * d)), 1e-30))
        print(f"view {vi}: max abs {d[away].max():.2e} (edge pixels {int((~away).sum())}), PSNR {psnr:.1f}",
              flush=True)
        assert d[away].max() <= 1.0 / 255.0 and psnr >= 50.0
        assert (~away).mean() < 0.01
    loss.backward()
    for si in range(2):
        for name, t64 in zip(NAMES, p64[si]):
            g = params[si][name].grad.detach().cpu().double()
            gr = t64.grad.detach()
            rel = float(torch.linalg.norm(g - gr) / torch.linalg.norm(gr))
            print(f"scene {si} grad {name}: rel err {rel:.2e}", flush=True)
            assert rel <= 1e-3, (si, name, rel)


def test_c4_full_size_fwd_bwd_step():
    """The differentiation workload at full size: 500k splats, one 1024x1024 view, L1 loss against
    the image of parameters perturbed by N(0, 0.01) (positions, log-scales) and N(0, 0.1) (SH,
    opacity logits).  Gradients are finite and nonzero, a two-view batch of the same scene renders
    each view bit for bit as one-view calls do, and a small gradient step lowers the loss."""
    import bench

    wl = S.WORKLOADS["c4_500k"]
    sc = wl.make()
    cam = S.orbit_camera(wl.width, wl.height, wl.radius)
    cam2 = S.orbit_camera(wl.width, wl.height, wl.radius * 1.2, theta=1.2)
    with torch.no_grad():
        target = D.render(**D.scene_tensors(bench.c4_perturbed(sc), requires_grad=False), camera=cam)
    p = D.scene_tensors(sc)
    img = D.render(**p, camera=cam)
    loss = (img - target).abs().mean()
    loss.backward()
    print("c4 visible", D.last_info["visible"], "pairs", D.last_info["pairs"], "loss", float(loss), flush=True)
    for k in NAMES:
        assert torch.isfinite(p[k].grad).all(), k
        assert float(p[k].grad.abs().sum()) > 0, k
    with torch.no_grad():
        outs = D.render_batch([p], [(0, cam), (0, cam2)])
        single2 = D.render(**p, camera=cam2)
        np.testing.assert_array_equal(outs[0].cpu().numpy(), img.detach().cpu().numpy())
        np.testing.assert_array_equal(outs[1].cpu().numpy(), single2.cpu().numpy())
        # steps along -grad moving each group by eta of its norm: a gradient of the wrong sign or
        # scale raises the loss at every small step
        base = {k: p[k].detach().clone() for k in NAMES}
        losses = []
        for eta in (1e-3, 1e-4, 1e-5):
            q = {k: base[k] - eta * float(torch.linalg.norm(base[k])) / float(torch.linalg.norm(p[k].grad))
                 * p[k].grad for k in NAMES}
            losses.append(float((D.render(**q, camera=cam) - target).abs().mean()))
    print("c4 loss after steps", losses, flush=True)
    assert min(losses) < float(loss)


def test_graphed_step_replays_match_eager_steps():
    """diff.GraphedStep: forward, elementwise loss gradient and backward captured into one hipGraph.
    Four replays, with eager steps of other views and host synchronizations between them, each give
    the gradients of the same step run eagerly, bit for bit (the passes never wait for the device, so
    nothing of them is left out of the graph)."""
    sa, cam = small_scene(3000, 71, 80, 64)
    sb, _ = small_scene(1200, 72, 80, 64)
    cams = [cam] + [S.orbit_camera(80, 64, 4.0 + 0.3 * i, theta=0.5 * i) for i in range(1, 5)]
    views = [(0, c) for c in cams] + [(1, S.orbit_camera(64, 48, 3.6, phi=1.1))]
    pa, pb = D.scene_tensors(sa), D.scene_tensors(sb)
    spans, total = D._spans([(v[0], v[1], None) for v in views])
    tgt = torch.rand((total, 4), device="cuda")
    w = torch.rand((total, 1), device="cuda") + 0.5

    def grad_fn(flat):
        return torch.sign(flat - tgt) * w

    def eager():
        for p in (pa, pb):
            for t in p.values():
                t.grad = None
        flat, _ = D.render_batch_flat([pa, pb], views, spatial_order=True)
        flat.backward(grad_fn(flat.detach()))
        return flat.detach().clone(), {k: (pa[k].grad.clone(), pb[k].grad.clone()) for k in NAMES}

    ref_img, ref_g = eager()
    step = D.GraphedStep([pa, pb], views, grad_fn, spatial_order=True)
    other = [(0, S.orbit_camera(96, 72, 5.0, theta=2.0))]
    for rep in range(4):
        flat = step.step()
        torch.cuda.synchronize()
        print("replay", rep, flush=True)
        np.testing.assert_array_equal(flat.cpu().numpy(), ref_img.cpu().numpy())
        for k in NAMES:
            np.testing.assert_array_equal(pa[k].grad.cpu().numpy(), ref_g[k][0].cpu().numpy())
            np.testing.assert_array_equal(pb[k].grad.cpu().numpy(), ref_g[k][1].cpu().numpy())
        # eager work of the program between replays (another view set through the pooled contexts)
        o, _ = D.render_batch_flat([pa], other)
        o.sum().backward()
        torch.cuda.synchronize()
    step.check()
