"""Edge cases of the splat path on the GPU and the host-facing surfaces around it: degenerate splats,
capacity limits, PLY files with permuted properties, the pygs loop and the per-stage telemetry.

Degenerate splats follow the reference's quirks: an isotropic 2D covariance (D = 0) makes the
eigendecomposition NaN (projection.comp:124-134), so the splat is counted, sorted and slotted but draws
nothing; a centre exactly on the far plane passes rank's 0 <= z <= 1 (rank.comp:37) and fails the depth
test LESS against the cleared 1.0 (graphics_pipeline.cc:81); NaN positions are culled.  The NaN
rot_scale words are the canonical quiet NaN on both sides, so instances compare bit for bit.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import oracle  # noqa: E402
from vkgs_amd import _lib  # noqa: E402
from vkgs_amd import pipeline as P  # noqa: E402
from vkgs_amd import scene as S  # noqa: E402

W = H = 64
QNAN = 0x7FC00000


def _rows(pos, log_scales):
    rows = np.zeros((len(pos), S.STRIDE), np.float32)
    c = S.COL
    rows[:, 0:3] = pos
    rows[:, c["scale_0"]:c["scale_0"] + 3] = log_scales
    rows[:, c["rot_0"]] = 1.0
    rows[:, c["opacity"]] = 2.0
    rows[:, c["f_dc_0"]] = 1.0
    return rows


def degenerate_rows(cam):
    """Isotropic splats on the optical axis (D = 0), zero-scale splats (cov3d = 0: the 2D covariance is the
    isotropic low-pass of a square image), centres on the far plane (z_ndc = 1 exactly), a NaN position and
    a splat behind the camera."""
    eye = S.camera_arrays(cam)[2].astype(np.float64)
    d = -eye / np.linalg.norm(eye)
    pos, logs = [], []
    for t in (0.0, 0.5, -0.7, 1.3):
        pos.append(eye + (4.0 + t) * d)
        logs.append((-2.0, -2.0, -2.0))
    for t in (0.0, 0.9):
        pos.append(eye + (4.0 + t) * d)
        logs.append((-np.inf, -np.inf, -np.inf))
    for dd in np.linspace(99.9995, 100.0005, 41):
        pos.append(eye + dd * d)
        logs.append((-2.0, -2.2, -2.1))
    pos.append((np.nan, 0.0, 0.0))
    logs.append((-2.0, -2.2, -2.1))
    pos.append(eye - 2.0 * d)
    logs.append((-2.0, -2.2, -2.1))
    return _rows(np.array(pos), np.array(logs, np.float32))


def _oracle_frame(cam, rows):
    pos, cov, opa, sh = oracle.parse_ply(rows, S.standard_offsets())
    v, key, index = oracle.rank(cam, pos)
    skey, sidx = oracle.sort_kv(key, index)
    inv = oracle.inverse(sidx, rows.shape[0])
    inst, di = oracle.projection(cam, pos, cov, opa, sh, inv, v)
    return dict(visible=v, keys=skey, order=sidx, inst=inst[:v], image=oracle.draw(inst[:v], W, H),
                storage=(pos, cov, opa, sh))


def _stage_frame(cam, rows, storage):
    sc = S.Scene(rows)
    st = P.parse_ply(sc.ply_buffer(), sc.n)
    pos, cov, opa, sh = storage  # the oracle's parsed bits (exp may differ in the last bit on the GPU)
    st.position.copy_(torch.from_numpy(pos))
    st.cov3d.copy_(torch.from_numpy(cov))
    st.opacity.copy_(torch.from_numpy(opa))
    st.sh.copy_(torch.from_numpy(sh.view(np.int16)))
    return P.reference_frame(cam, st, W, H)


def test_degenerate_splats_match_oracle_and_draw_nothing():
    cam = S.orbit_camera(W, H, 4.0)
    rows = degenerate_rows(cam)
    ref = _oracle_frame(cam, rows)
    v = ref["visible"]
    inst_bits = ref["inst"].view(np.uint32)
    n_nan = int((inst_bits[:, 4:8] == QNAN).all(1).sum())
    n_far = int((ref["keys"] == 0).sum())
    print("visible", v, "NaN rot_scale", n_nan, "on the far plane", n_far, flush=True)
    assert n_nan == 6 and n_far >= 3 and v == n_nan + n_far  # the NaN position and the one behind are culled
    out = _stage_frame(cam, rows, ref["storage"])
    assert int(P.u32(out["count"])[0]) == v
    np.testing.assert_array_equal(P.u32(out["key"])[:v], ref["keys"])
    np.testing.assert_array_equal(P.u32(out["index"])[:v], ref["order"])
    got = out["instances"].cpu().numpy()[:v].view(np.uint32)
    np.testing.assert_array_equal(got, inst_bits)  # NaN words included
    clear = np.zeros((H, W, 4), np.float32)
    clear[..., 3] = 1.0
    np.testing.assert_array_equal(out["image"].cpu().numpy(), clear)
    np.testing.assert_array_equal(ref["image"], clear)
    r = P.Renderer(rows.shape[0])
    r.load_buffer(S.Scene(rows).ply_buffer(), rows.shape[0])
    img = r.render(cam).cpu().numpy()
    st, rv, rk = r.frame_info()
    r.close()
    assert st == 0 and rv == v and rk == 0  # slotted and sorted, but no tile pairs
    np.testing.assert_array_equal(img, clear)


def test_degenerate_splats_among_normal_ones():
    cam = S.orbit_camera(W, H, 4.0)
    normal = S.gaussian_cloud(600, seed=61, pos_std=1.0, logscale=(-2.5, 0.3))
    rows = np.concatenate([degenerate_rows(cam), normal.rows])
    ref = _oracle_frame(cam, rows)
    v = ref["visible"]
    out = _stage_frame(cam, rows, ref["storage"])
    np.testing.assert_array_equal(P.u32(out["index"])[:v], ref["order"])
    np.testing.assert_array_equal(out["instances"].cpu().numpy()[:v].view(np.uint32), ref["inst"].view(np.uint32))
    img = out["image"].cpu().numpy()
    assert np.abs(img - ref["image"]).max() <= 1.0 / 255.0


def test_renderer_refuses_more_splats_than_capacity():
    sc = S.gaussian_cloud(1001, seed=62)
    r = P.Renderer(1000)
    with pytest.raises(_lib.VkgsError) as e:
        r.load_buffer(sc.ply_buffer(), sc.n)
    assert e.value.status == 2  # VKGS_ERR_CAPACITY
    r.close()


def test_load_ply_with_permuted_properties_matches_load_buffer(tmp_path):
    """A PLY whose properties come in another order (and with extra ones) loads through the offset table
    (splat_load_thread.cc:115-135) to the same frame as the buffer of the standard layout."""
    sc = S.gaussian_cloud(3000, seed=63, pos_std=1.0, logscale=(-2.5, 0.3))
    rng = np.random.default_rng(0)
    names = list(S.PLY_PROPERTIES) + ["extra_a", "extra_b"]
    perm = rng.permutation(len(names))
    table = np.concatenate([sc.rows, rng.normal(size=(sc.n, 2)).astype(np.float32)], 1)[:, perm]
    path = os.path.join(tmp_path, "permuted.ply")
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {sc.n}"]
    header += [f"property float {names[i]}" for i in perm] + ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode() + np.ascontiguousarray(table, "<f4").tobytes())
    cam = S.orbit_camera(96, 72, 4.0)
    a = P.Renderer(sc.n)
    a.load_ply(path)
    b = P.Renderer(sc.n)
    b.load_buffer(sc.ply_buffer(), sc.n)
    ia, ib = a.render(cam).cpu().numpy(), b.render(cam).cpu().numpy()
    sa, sb = a.storage(), b.storage()
    for x, y in zip((sa.position, sa.cov3d, sa.opacity, sa.sh), (sb.position, sb.cov3d, sb.opacity, sb.sh)):
        assert torch.equal(x, y)
    a.close()
    b.close()
    np.testing.assert_array_equal(ia, ib)


def test_pygs_load_render_show_close(tmp_path):
    """pygs (binding/python/pygs/__init__.py:14-32): load raises FileNotFoundError for a missing file,
    render gives the renderer's frame, the show loop publishes frames and stops on close.  The scene's
    metre-sized splats need more tile pairs than a fresh renderer's first storage: the frame is rendered
    again with room for the whole list, never returned incomplete."""
    from vkgs_amd import pygs

    with pytest.raises(FileNotFoundError):
        pygs.load(os.path.join(tmp_path, "missing.ply"))
    sc = S.gaussian_cloud(4000, seed=64, pos_std=1.0, logscale=(0.0, 0.0))
    path = os.path.join(tmp_path, "big.ply")
    sc.write_ply(path)
    pygs._state["renderer"] = None
    pygs.load(path)
    cam = S.orbit_camera(768, 512, 4.0)
    img = pygs.render(cam).cpu().numpy()
    r = pygs._state["renderer"]
    st, v, k = r.frame_info()
    print("pygs frame: visible", v, "pairs", k, flush=True)
    assert st == 0 and k > 4 * r.capacity + (1 << 20)  # beyond the first pair storage
    ref = P.Renderer(sc.n)
    ref.load_buffer(sc.ply_buffer(), sc.n)
    for _ in range(2):  # a roomy renderer: the second frame has the grown storage
        want = ref.render(cam)
        if ref.frame_info()[0] == 0:
            break
    np.testing.assert_array_equal(img, want.cpu().numpy())
    ref.close()
    pygs.set_camera(96, 64, 3.0)
    pygs.show()
    import time
    t0 = time.time()
    while pygs.latest_frame() is None and time.time() - t0 < 30:
        time.sleep(0.01)
    pygs.close()
    assert pygs.latest_error() is None
    frame = pygs.latest_frame()
    assert frame is not None and tuple(frame.shape) == (64, 96, 4)
    assert torch.isfinite(frame).all()
    pygs._state["renderer"].close()
    pygs._state["renderer"] = None
    pygs._state["frame"] = None


@pytest.mark.parametrize("pairs", [False, True])
def test_stage_timing_sums_to_total_and_graph_replay_resumes(pairs):
    """vkgs_renderer_set_timing (the reference's per-stage timestamps, engine.cc:1031-1045): the stage
    times are positive and add up to the frame total; images equal the graph-replayed ones; with timing
    off the frame is replayed from a hipGraph again (faster than timed eager frames).  With 8x8 tiles
    two per wave forced (pairs), the timed frames run the pair raster's work-counting form and the
    replayed ones the form without counts: the images must still be equal bit for bit."""
    sc = S.gaussian_cloud(200_000, seed=65)
    cam = S.orbit_camera(800, 600, 6.0)
    r = P.Renderer(sc.n)
    r.load_buffer(sc.ply_buffer(), sc.n)
    if pairs:
        cfg = P._lib.RasterConfig()
        cfg.tile_w, cfg.tile_h, cfg.t_min, cfg.flags = 8, 8, 1.0 / 8192.0, 4  # VKGS_RASTER_PAIRS_ALWAYS
        r.set_config(config=cfg)
    for _ in range(2):
        base = r.render(cam).cpu().numpy()
    r.set_timing(True)
    for _ in range(3):
        img = r.render(cam).cpu().numpy()
        ms = r.stage_ms()
    np.testing.assert_array_equal(img, base)
    parts = [ms[k] for k in P.STAGES if k != "total"]
    print("stage ms", ms, flush=True)
    assert all(p > 0 for p in parts)
    assert abs(sum(parts) - ms["total"]) <= 0.02 * ms["total"] + 0.005
    assert r.raster_work() > 0
    r.set_timing(False)

    def per_frame(k=20):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(k):
            r.render(cam)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    replay = per_frame()
    np.testing.assert_array_equal(r.render(cam).cpu().numpy(), base)
    r.set_timing(True)
    timed = per_frame()
    r.set_timing(False)
    print("graph replay", replay, "ms/frame; timed eager", timed, flush=True)
    assert replay < timed
    r.close()


def test_async_load_then_render_and_cancel(tmp_path):
    """vkgs_renderer_load_ply_async (SplatLoadThread + Engine::LoadSplatsAsync): the load runs on the
    renderer's loader thread while frames keep the previous splats; the first frame after it is ready
    equals a synchronous load's frame.  A new request cancels the one in progress (the later file wins);
    cancel_load leaves the splats as they were; a missing file fails with VKGS_ERR_IO."""
    import time
    a = S.gaussian_cloud(300_000, seed=66, pos_std=1.0, logscale=(-3.0, 0.3))
    b = S.gaussian_cloud(20_000, seed=67, pos_std=1.0, logscale=(-2.5, 0.3))
    pa, pb = os.path.join(tmp_path, "a.ply"), os.path.join(tmp_path, "b.ply")
    a.write_ply(pa)
    b.write_ply(pb)
    cam = S.orbit_camera(128, 96, 4.0)

    def frame(rr):
        for _ in range(2):  # a frame that outgrew the pair storage is rendered again
            img = rr.render(cam)
            if rr.frame_info()[0] == 0:
                return img.cpu().numpy()
        raise AssertionError("pair overflow twice")

    def wait(r, want):
        t0 = time.time()
        while time.time() - t0 < 60:
            img = r.render(cam)  # takes a finished load
            total, loaded, state = r.load_progress()
            if state == want:
                return img, total, loaded
            time.sleep(0.005)
        raise AssertionError(f"load state {state}, wanted {want}")

    r = P.Renderer(a.n)
    ref = P.Renderer(a.n)
    ref.load_buffer(b.ply_buffer(), b.n)
    want_b = frame(ref)
    ref.load_buffer(a.ply_buffer(), a.n)
    want_a = frame(ref)
    ref.close()
    r.load_ply_async(pa)
    r.load_ply_async(pb)  # supersedes a
    img, total, loaded = wait(r, "done")
    assert r.n == b.n and total == b.n and loaded == b.n
    np.testing.assert_array_equal(frame(r), want_b)
    r.load_ply_async(pa)
    img, total, loaded = wait(r, "done")
    assert r.n == a.n
    np.testing.assert_array_equal(frame(r), want_a)
    r.load_ply_async(pb)
    r.cancel_load()
    for _ in range(20):
        r.render(cam)
    assert r.load_progress()[2] in ("canceled", "done")
    if r.load_progress()[2] == "canceled":
        assert r.n == a.n
    r.load_ply_async(os.path.join(tmp_path, "missing.ply"))
    t0 = time.time()
    while time.time() - t0 < 30:
        try:
            st = r.load_progress()[2]
        except _lib.VkgsError as e:
            assert e.status == 4  # VKGS_ERR_IO
            break
        time.sleep(0.005)
    else:
        raise AssertionError("missing file did not fail")
    assert r.n in (a.n, b.n)
    r.close()
