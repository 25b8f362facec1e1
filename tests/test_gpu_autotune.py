"""Launch choices the autotuners pick from (tile shape, 8x8 tiles one or two per
wave, the backward's LDS batch depth, a resident-wave cap) change how the work
is laid over waves, not the per-pixel arithmetic.

Bars: the plain renderer's image is bit-identical across wave caps, and
across pair forms within 2 t_min (it stops compositing when every pixel of a
wave is below t_min, and a wave holds one tile or two: a pixel may add a few
splats more in one form, each past T < t_min, so the two images differ by less
than t_min x max(colour, 1)); the differentiable forward's image is
bit-identical across every choice (its pixels stop one by one);
its gradients on 8x8 tiles are bit-identical across LDS batch depths and wave
caps within one pair form (each (tile, splat) partial is summed over the same
half rows in the same order whatever the batch) and within 1e-5 relative
between the two forms (the pair kernel adds a lane's two pixels with FMAs
first) and on multi-wave tiles (LDS atomics).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

from vkgs_amd import _lib  # noqa: E402
from vkgs_amd import diff as D  # noqa: E402
from vkgs_amd import pipeline as P  # noqa: E402
from vkgs_amd import scene as S  # noqa: E402

NAMES = D.PARAM_NAMES
PAIRS_ALWAYS, PAIRS_NEVER = 4, 8


def cfg_of(tile=(8, 8), flags=0, lds_batch=0, waves=0, t_min=1.0 / 8192.0):
    c = _lib.RasterConfig()
    c.tile_w, c.tile_h, c.t_min = tile[0], tile[1], t_min
    c.flags, c.lds_batch, c.waves_per_cu = flags, lds_batch, waves
    return c


def diff_scene(n, seed):
    sc = S.gaussian_cloud(n, seed=seed, pos_std=1.0, logscale=(-2.5, 0.4))
    c = S.COL
    sc.rows[:, c["opacity"]] = np.clip(sc.rows[:, c["opacity"]], -6, 4)
    return sc


def fwd_bwd(sc, views, cfg, weights):
    p = D.scene_tensors(sc)
    imgs = D.render_batch([p], views, config=cfg)
    sum((img * w).sum() for img, w in zip(imgs, weights)).backward()
    return [i.detach().cpu().numpy() for i in imgs], {k: p[k].grad.detach().cpu().numpy() for k in NAMES}


def test_diff_launch_choices_give_the_same_results():
    sc = diff_scene(6000, 51)
    views = [(0, S.orbit_camera(112, 72, 4.0)), (0, S.orbit_camera(128, 96, 3.5, theta=0.6))]
    rng = np.random.default_rng(5)
    weights = [torch.tensor(rng.uniform(0.5, 1.5, (c.screen_size[1], c.screen_size[0], 4)), dtype=torch.float32,
                            device="cuda") for _, c in views]
    base_i, base_g = fwd_bwd(sc, views, cfg_of(flags=PAIRS_NEVER), weights)
    pair_i, pair_g = fwd_bwd(sc, views, cfg_of(flags=PAIRS_ALWAYS), weights)
    for a, b in zip(base_i, pair_i):
        np.testing.assert_array_equal(a, b)
    for k in NAMES:
        rel = np.linalg.norm(pair_g[k] - base_g[k]) / np.linalg.norm(base_g[k])
        assert rel < 1e-5, (k, rel)
    for flags, ref_g in ((PAIRS_NEVER, base_g), (PAIRS_ALWAYS, pair_g)):
        for lds_batch, waves in ((8, 0), (32, 0), (16, 8), (8, 24), (32, 12)):
            print("diff choice", flags, lds_batch, waves, flush=True)
            imgs, g = fwd_bwd(sc, views, cfg_of(flags=flags, lds_batch=lds_batch, waves=waves), weights)
            for a, b in zip(base_i, imgs):
                np.testing.assert_array_equal(a, b)
            for k in NAMES:
                np.testing.assert_array_equal(g[k], ref_g[k])
    # multi-wave tiles under a wave cap: same image; their backward adds the waves' partials with LDS
    # atomics, in no fixed order, so gradients agree to float rounding
    _, g16 = fwd_bwd(sc, views, cfg_of(tile=(16, 16)), weights)
    imgs, g = fwd_bwd(sc, views, cfg_of(tile=(16, 16), waves=8), weights)
    for a, b in zip(base_i, imgs):
        np.testing.assert_array_equal(a, b)
    for k in NAMES:
        rel = np.linalg.norm(g[k] - g16[k]) / np.linalg.norm(g16[k])
        assert rel < 1e-5, (k, rel)


def test_plain_renderer_pair_forms_and_wave_caps_are_bit_identical():
    sc = S.gaussian_cloud(200_000, seed=8)
    cam = S.orbit_camera(1600, 900, 6.0)  # 200 x 113 tiles of 8x8: two per wave by default
    r = P.Renderer(sc.n)
    r.load_buffer(sc.ply_buffer(), sc.n)
    t_min = 1.0 / 8192.0
    ref = {}
    for flags, waves in ((0, 0), (PAIRS_NEVER, 0), (PAIRS_ALWAYS, 0), (0, 8), (PAIRS_NEVER, 16), (0, 24),
                         (PAIRS_NEVER, 8)):
        print("plain choice", flags, waves, flush=True)
        r.set_config(config=cfg_of(flags=flags, waves=waves, t_min=t_min))
        img = r.render(cam).cpu().numpy()
        if r.frame_info()[0] == 6:  # VKGS_ERR_OVERFLOW: the first frame sizes the pair storage
            img = r.render(cam).cpu().numpy()
        assert r.frame_info()[0] == 0
        form = "single" if flags == PAIRS_NEVER else "pairs"  # 22600 tiles: pairs by default
        if form in ref:
            np.testing.assert_array_equal(img, ref[form])
        else:
            ref[form] = img
    d = np.abs(ref["single"] - ref["pairs"]).max()
    print("pair forms differ by", d, flush=True)
    assert d < 2 * t_min
    r.close()


def test_diff_autotune_choice_renders_the_same_image():
    sc = diff_scene(20000, 52)
    views = [(0, S.orbit_camera(320, 240, 4.0, theta=0.3 * i)) for i in range(4)]
    p = D.scene_tensors(sc)
    cfg, ms = D.autotune([p], views, steps=2)
    print("diff autotune", D.describe(cfg), f"{ms:.3f} ms", flush=True)
    assert ms > 0
    assert (cfg.tile_w, cfg.tile_h) in {(8, 8), (16, 8), (8, 16), (16, 4), (16, 16), (32, 4)}
    assert cfg.lds_batch in (0, 8, 16, 32) and cfg.waves_per_cu in (0, 24, 16, 12, 8)
    assert all(p[k].grad is None for k in NAMES)  # the tuner does not touch .grad
    tuned = [i.detach().cpu().numpy() for i in D.render_batch([p], views, config=cfg)]
    plain = [i.detach().cpu().numpy() for i in D.render_batch([p], views)]
    for a, b in zip(tuned, plain):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("flags", [PAIRS_NEVER, PAIRS_ALWAYS])
def test_diff_timing_counts_work_and_keeps_results(flags):
    """vkgs_diff_set_timing: per-stage times are positive and add up, the raster work counters obey
    0 < backward entries <= forward entries <= pairs (the backward walks each tile's list only up to
    its pixels' last entry), and the gradients equal an untimed step's."""
    sc = diff_scene(8000, 53)
    views = [(0, S.orbit_camera(160, 120, 4.0)), (0, S.orbit_camera(200, 100, 3.0, theta=0.9))]
    spans, total = D._spans([(0, c, None) for _, c in views])
    w = torch.tensor(np.random.default_rng(9).uniform(0.5, 1.5, (total, 4)), dtype=torch.float32, device="cuda")
    cfg = cfg_of(flags=flags)
    p = D.scene_tensors(sc)
    prof = D.profile_step([p], views, lambda f: w, config=cfg)
    g_timed = {k: p[k].grad.detach().cpu().numpy() for k in NAMES}
    ms = prof["ms"]
    print("diff timing", flags, ms, prof["forward_entries"], prof["backward_entries"], D.last_info, flush=True)
    assert all(v > 0 for v in ms.values())
    assert ms["forward_total"] >= ms["forward_raster"] and ms["backward_total"] >= ms["backward_raster"]
    p2 = D.scene_tensors(sc)
    flat, _ = D.render_batch_flat([p2], views, config=cfg)
    (flat * w).sum().backward()
    assert 0 < prof["backward_entries"] <= prof["forward_entries"] <= D.last_info["pairs"]
    for k in NAMES:
        np.testing.assert_array_equal(g_timed[k], p2[k].grad.detach().cpu().numpy())
