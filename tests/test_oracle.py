"""Pins the CPU oracle against closed forms, library routines and brute force.

None of these checks re-types the oracle's formulas: each compares against an
independent construction (scipy rotations and spherical harmonics, numpy
float16, a float64 numerical Jacobian of the NDC map, numpy's stable argsort,
per-pixel brute force) or a closed form, chosen so that a dropped term, a sign
error or a swapped index in the oracle fails at least one of them.
"""
import math

import numpy as np
import pytest

import oracle
from vkgs_amd import scene as S

scipy_rot = pytest.importorskip("scipy.spatial.transform")
special = pytest.importorskip("scipy.special")


def cam_arrays(cam):
    P, V, eye = S.camera_arrays(cam)
    return P.reshape(4, 4).T.astype(np.float64), V.reshape(4, 4).T.astype(np.float64), eye.astype(np.float64)


# ------------------------------------------------------------------ fp16 ---

def test_f16_matches_numpy():
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.normal(0, 1, 3000), rng.normal(0, 1e-5, 1000), rng.normal(0, 3e4, 500),
        np.array([0.0, -0.0, 65504.0, 65520.0, 65519.99, 1e-8, 6e-8, 5.96e-8, 2.98e-8, np.inf, -np.inf]),
        # exact halfway cases between fp16 neighbours (round to even)
        (np.arange(1, 200) * 2 + 1) * 2.0 ** -11 / 2 + 1.0,
    ]).astype(np.float32)
    for v in vals:
        got = oracle.f32_to_f16(float(v))
        want = int(np.array([v], np.float32).astype(np.float16).view(np.uint16)[0])
        assert got == want, (v, hex(got), hex(want))
    for h in list(range(0, 0x7c00, 97)) + [0x0001, 0x03ff, 0x0400, 0x7bff, 0x8001]:
        assert oracle.f16_to_f32(h) == float(np.array([h], np.uint16).view(np.float16)[0])


# ------------------------------------------------------------- parse_ply ---

def _rows(n, seed=0):
    return S.gaussian_cloud(n, seed=seed).rows


def test_parse_identity_quaternion_closed_form():
    rows = _rows(64)
    c = S.COL
    rows[:, c["rot_0"]:c["rot_0"] + 4] = [2.0, 0.0, 0.0, 0.0]  # un-normalized identity
    rows[:, c["scale_0"]:c["scale_0"] + 3] = np.log(0.3)
    pos, cov, opa, sh = oracle.parse_ply(rows, S.standard_offsets())
    s2 = np.float32(np.exp(np.float32(np.log(0.3)))) ** 2
    np.testing.assert_allclose(cov[:, [0, 3, 5]], s2, rtol=1e-6)
    np.testing.assert_array_equal(cov[:, [1, 2, 4]], 0.0)
    np.testing.assert_array_equal(pos, rows[:, 0:3])


def test_parse_cov_matches_scipy_rotation():
    rows = _rows(500, seed=3)
    pos, cov, opa, sh = oracle.parse_ply(rows, S.standard_offsets())
    c = S.COL
    wxyz = rows[:, c["rot_0"]:c["rot_0"] + 4].astype(np.float64)
    R = scipy_rot.Rotation.from_quat(wxyz[:, [1, 2, 3, 0]]).as_matrix()  # scipy takes (x, y, z, w)
    s = np.exp(rows[:, c["scale_0"]:c["scale_0"] + 3].astype(np.float64))
    full = np.einsum("nij,nj,nkj->nik", R, s * s, R)
    want = full[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    scale = np.abs(want[:, [0, 3, 5]]).max(1, keepdims=True)
    assert np.all(np.abs(cov - want) <= 2e-5 * scale)


def test_parse_quaternion_sign_invariance_and_activations():
    rows = _rows(200, seed=4)
    flipped = rows.copy()
    c = S.COL
    flipped[:, c["rot_0"]:c["rot_0"] + 4] *= -1
    _, cov_a, opa, sh = oracle.parse_ply(rows, S.standard_offsets())
    _, cov_b, _, _ = oracle.parse_ply(flipped, S.standard_offsets())
    np.testing.assert_array_equal(cov_a, cov_b)
    want = 1.0 / (1.0 + np.exp(-rows[:, c["opacity"]].astype(np.float64)))
    np.testing.assert_allclose(opa, want, rtol=1e-6)
    z = rows.copy()
    z[:, c["opacity"]] = 0.0
    assert np.all(oracle.parse_ply(z, S.standard_offsets())[2] == 0.5)
    # SH layout: [R: dc, rest0..14][G: dc, rest15..29][B: dc, rest30..44], fp16
    chan = np.concatenate([
        rows[:, [c["f_dc_0"]]], rows[:, c["f_rest_0"]:c["f_rest_0"] + 15],
        rows[:, [c["f_dc_1"]]], rows[:, c["f_rest_0"] + 15:c["f_rest_0"] + 30],
        rows[:, [c["f_dc_2"]]], rows[:, c["f_rest_0"] + 30:c["f_rest_0"] + 45]], 1)
    np.testing.assert_array_equal(sh, chan.astype(np.float16).view(np.uint16))


# ------------------------------------------------------------------ rank ---

def ndc64(cam, pts):
    P, V, _ = cam_arrays(cam)
    h = np.c_[pts.astype(np.float64), np.ones(len(pts))]
    clip = h @ (P @ V).T
    return clip[:, :3] / clip[:, 3:4]


def test_rank_against_float64_projection():
    cam = S.orbit_camera(1600, 900, 6.0)
    pts = S.gaussian_cloud(20000, seed=5).rows[:, 0:3]
    v, key, index = oracle.rank(cam, pts)
    ndc = ndc64(cam, pts)
    inside = (np.abs(ndc[:, 0]) <= 1) & (np.abs(ndc[:, 1]) <= 1) & (ndc[:, 2] >= 0) & (ndc[:, 2] <= 1)
    margin = np.minimum.reduce([1 - np.abs(ndc[:, 0]), 1 - np.abs(ndc[:, 1]), ndc[:, 2], 1 - ndc[:, 2]])
    clear = np.abs(margin) > 1e-5
    vis = np.zeros(len(pts), bool)
    vis[index] = True
    np.testing.assert_array_equal(vis[clear], inside[clear])
    assert np.all(np.diff(index.astype(np.int64)) > 0)  # ascending id order
    key_f = key.view(np.float32).astype(np.float64)
    np.testing.assert_allclose(key_f, 1.0 - ndc[index, 2], atol=2e-6)


def test_rank_closed_forms():
    # camera on +z looking at the origin: theta = 0, phi = 90 deg
    cam = S.orbit_camera(800, 800, 5.0, phi=math.pi / 2, theta=0.0)
    n, f = 0.01, 100.0
    depths = np.array([0.5, 1.0, 5.0, 50.0, 99.0])
    pts = np.stack([np.zeros(5), np.zeros(5), 5.0 - depths], 1).astype(np.float32)
    behind = np.array([[0, 0, 6.0], [0, 0, 5.0 - 101.0], [0, 0, 5.0 - 0.001]], np.float32)  # behind, past far, before near
    v, key, index = oracle.rank(cam, np.concatenate([pts, behind]))
    assert v == 5 and list(index) == [0, 1, 2, 3, 4]
    z = f / (f - n) * (1 - n / depths)  # Vulkan depth of view depth d (GL->VK conversion)
    np.testing.assert_allclose(key.view(np.float32), 1.0 - z, atol=1e-6)
    assert np.all(np.diff(key.view(np.float32)) < 0)  # farther = smaller key (drawn first)
    v0, _, _ = oracle.rank(cam, np.zeros((0, 3), np.float32))
    assert v0 == 0
    nan = np.array([[np.nan, 0, 0]], np.float32)
    assert oracle.rank(cam, nan)[0] == 0


# ------------------------------------------------------------------ sort ---

@pytest.mark.parametrize("n,distinct", [(0, 1), (1, 1), (1000, 3), (50000, 1 << 30)])
def test_sort_matches_numpy_stable(n, distinct):
    rng = np.random.default_rng(n)
    k = rng.integers(0, distinct, n, dtype=np.uint64).astype(np.uint32)
    v = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    sk, sv = oracle.sort_kv(k, v)
    o = np.argsort(k, kind="stable")
    np.testing.assert_array_equal(sk, k[o])
    np.testing.assert_array_equal(sv, v[o])


def test_inverse_is_a_partial_permutation_inverse():
    rng = np.random.default_rng(1)
    n = 1000
    idx = rng.permutation(n)[:600].astype(np.uint32)
    inv = oracle.inverse(idx, n)
    assert np.all(inv[idx] == np.arange(600))
    rest = np.setdiff1d(np.arange(n), idx)
    assert np.all(inv[rest] == -1)


# ------------------------------------------------------------ projection ---

def _project_scene(n=3000, seed=6, w=640, h=480, r=5.0):
    sc = S.gaussian_cloud(n, seed=seed, pos_std=1.0)
    cam = S.orbit_camera(w, h, r)
    pos, cov, opa, sh = oracle.parse_ply(sc.rows, S.standard_offsets())
    v, key, index = oracle.rank(cam, pos)
    sk, si = oracle.sort_kv(key, index)
    inv = oracle.inverse(si, n)
    inst, di = oracle.projection(cam, pos, cov, opa, sh, inv, v)
    return sc, cam, pos, cov, opa, sh, si, inst, di


def test_projection_cov2d_is_first_order_propagation():
    sc, cam, pos, cov, opa, sh, order, inst, di = _project_scene()
    P, V, _ = cam_arrays(cam)
    w, h = cam.screen_size[0], cam.screen_size[1]
    ids = order[:400]
    p = pos[ids].astype(np.float64)

    def ndc_xy(x):
        c = np.c_[x, np.ones(len(x))] @ (P @ V).T
        return c[:, :2] / c[:, 3:4]

    eps = 1e-4
    J = np.zeros((len(ids), 2, 3))
    for a in range(3):
        d = np.zeros(3)
        d[a] = eps
        J[:, :, a] = (ndc_xy(p + d) - ndc_xy(p - d)) / (2 * eps)
    C = cov[ids].astype(np.float64)
    Sig = np.stack([C[:, [0, 1, 2]], C[:, [1, 3, 4]], C[:, [2, 4, 5]]], 1)
    want = np.einsum("nij,njk,nlk->nil", J, Sig, J)
    want[:, 0, 0] += 1.0 / w / w
    want[:, 1, 1] += 1.0 / h / h
    rs = inst[: len(ids), 4:8].astype(np.float64)  # (c0.x, c0.y, c1.x, c1.y)
    M = np.stack([np.stack([rs[:, 0], rs[:, 2]], 1), np.stack([rs[:, 1], rs[:, 3]], 1)], 1)
    got = M @ np.transpose(M, (0, 2, 1))
    ok = np.isfinite(got).all((1, 2))
    assert ok.mean() > 0.99
    scale = np.abs(want[ok]).max((1, 2))
    err = np.abs(got[ok] - want[ok]).max((1, 2)) / scale
    assert np.percentile(err, 99) < 1e-3 and err.max() < 1e-2
    # NDC centre
    np.testing.assert_allclose(inst[: len(ids), 0:2], ndc_xy(p), atol=1e-5)
    assert list(di) == [6 * len(order), 1, 0, 0, 0, 0, 0, 0, len(order), 1, 0, 0]


def test_sh_basis_matches_scipy():
    """Colour with one unit SH coefficient equals that basis function (+0.5)."""
    rng = np.random.default_rng(9)
    n = 200
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    # closed-form real SH, standard normalisation, m = -l..l
    C0 = 0.5 / math.sqrt(math.pi)
    C1 = math.sqrt(3 / (4 * math.pi))
    ref = np.stack([
        np.full(n, C0),
        C1 * y, C1 * z, C1 * x,
        0.5 * math.sqrt(15 / math.pi) * x * y, 0.5 * math.sqrt(15 / math.pi) * y * z,
        0.25 * math.sqrt(5 / math.pi) * (3 * z * z - 1), 0.5 * math.sqrt(15 / math.pi) * x * z,
        0.25 * math.sqrt(15 / math.pi) * (x * x - y * y),
        0.25 * math.sqrt(35 / (2 * math.pi)) * y * (3 * x * x - y * y), 0.5 * math.sqrt(105 / math.pi) * x * y * z,
        0.25 * math.sqrt(21 / (2 * math.pi)) * y * (5 * z * z - 1), 0.25 * math.sqrt(7 / math.pi) * z * (5 * z * z - 3),
        0.25 * math.sqrt(21 / (2 * math.pi)) * x * (5 * z * z - 1), 0.25 * math.sqrt(105 / math.pi) * z * (x * x - y * y),
        0.25 * math.sqrt(35 / (2 * math.pi)) * x * (x * x - 3 * y * y)], 1)
    # the closed forms above are scipy's real SH without the Condon-Shortley
    # phase: Y_l0, and sqrt(2) (-1)^|m| Re / Im of Y_l^|m| for m > 0 / m < 0
    theta = np.arccos(np.clip(z, -1, 1))
    phi = np.arctan2(y, x)
    k = 0
    for l in range(4):
        for m in range(-l, l + 1):
            if hasattr(special, "sph_harm_y"):
                Y = special.sph_harm_y(l, abs(m), theta, phi)
            else:
                Y = special.sph_harm(abs(m), l, phi, theta)
            real = Y.real if m >= 0 else Y.imag
            if m != 0:
                real = real * math.sqrt(2) * (-1) ** abs(m)
            np.testing.assert_allclose(real, ref[:, k], atol=1e-9)
            k += 1
    # projection.comp:159-163 uses the 3DGS sign (-1)^m on that basis, m = -l..l
    sign = np.array([(-1) ** abs(m) for l in range(4) for m in range(-l, l + 1)], np.float64)
    expect = ref * sign
    # oracle: one splat per direction, unit coefficient on basis j, R channel
    cam = S.orbit_camera(64, 64, 3.0)
    _, _, eye = cam_arrays(cam)
    for j in range(16):
        pos = (eye[None, :] + 0.7 * d).astype(np.float32)  # dir = normalize(pos - eye) = d
        cov = np.tile(np.array([1e-4, 0, 0, 1e-4, 0, 1e-4], np.float32), (n, 1))
        opa = np.ones(n, np.float32)
        coef = np.zeros((n, 48), np.float32)
        coef[:, j] = 1.0
        sh = coef.astype(np.float16).view(np.uint16)
        inv = np.arange(n, dtype=np.int32)
        inst, _ = oracle.projection(cam, pos, cov, opa, sh, inv, n)
        got = inst[:, 8].astype(np.float64) - 0.5
        want = expect[:, j]
        keep = want + 0.5 > 1e-6  # colour is clamped at 0
        np.testing.assert_allclose(got[keep], want[keep], atol=2e-6)
        assert np.all(inst[~keep, 8] == 0.0)
        np.testing.assert_array_equal(inst[:, 9:11], 0.5)  # G, B: only the constant 0.5


# ------------------------------------------------------------------ draw ---

def _inst(cx, cy, m, rgb, o, z=0.5):
    return np.array([cx, cy, z, 0, m[0][0], m[1][0], m[0][1], m[1][1], *rgb, o], np.float32)


def test_draw_empty_is_clear_colour():
    img = oracle.draw(np.zeros((0, 12), np.float32), 16, 8)
    assert np.all(img[..., :3] == 0) and np.all(img[..., 3] == 1)


def test_draw_single_splat_closed_form():
    W, H = 32, 32
    # centre exactly on pixel (16, 16)'s centre: ndc = (2*16.5/32 - 1)
    cx = 2 * 16.5 / W - 1
    s = 2.0 / W * 2.0  # 2 px sigma in NDC
    inst = _inst(cx, cx, [[s, 0], [0, s]], (0.2, 0.4, 0.8), 0.7)[None]
    img = oracle.draw(inst, W, H)
    a = 0.7
    np.testing.assert_allclose(img[16, 16, :3], np.array([0.2, 0.4, 0.8]) * a, rtol=1e-6)
    np.testing.assert_allclose(img[16, 16, 3], a * a + (1 - a), rtol=1e-6)
    # one pixel right: p = (0.5, 0): alpha = o exp(-1/8)
    a1 = 0.7 * math.exp(-0.5 * 0.25)
    np.testing.assert_allclose(img[16, 17, 0], 0.2 * a1, rtol=1e-6)
    # outside the +-3 sigma square (7 px away): untouched
    assert img[16, 23, 3] == 1.0 and img[16, 23, 0] == 0.0
    # the square's edge at 3 sigma: with sigma = 2.05 px pixel 16+6 (u = 2.93) is covered, 16+7
    # (u = 3.41) is not (a pixel centre exactly on the edge is a tie the reference settles with its
    # rasterizer's fill rule; the analytic test decides it by the rounding of the scaled frame)
    s2 = 2.0 / W * 2.05
    img2 = oracle.draw(_inst(cx, cx, [[s2, 0], [0, s2]], (0.2, 0.4, 0.8), 0.7)[None], W, H)
    assert img2[16, 22, 0] > 0 and img2[22, 16, 0] > 0 and img2[16, 10, 0] > 0
    assert img2[16, 23, 3] == 1.0 and img2[23, 16, 3] == 1.0 and img2[16, 9, 3] == 1.0
    # the footprint is a square in the eigenframe (splat.vert:17-25), not a disc:
    # p = (2.5, 2.5) has |p| > 3 and is still covered
    a_corner = 0.7 * math.exp(-0.5 * 12.5)
    np.testing.assert_allclose(img[21, 21, 0], 0.2 * a_corner, rtol=1e-5)
    # z = 1 fails the depth test; z < 0 is clipped
    assert np.all(oracle.draw(_inst(cx, cx, [[s, 0], [0, s]], (1, 1, 1), 0.7, z=1.0)[None], W, H)[..., 3] == 1)
    assert np.all(oracle.draw(_inst(cx, cx, [[s, 0], [0, s]], (1, 1, 1), 0.7, z=-1e-6)[None], W, H)[..., 3] == 1)


def test_draw_two_splats_over_operator():
    W = H = 16
    c = 2 * 8.5 / W - 1
    s = 0.5
    far = _inst(c, c, [[s, 0], [0, s]], (1, 0, 0), 0.6)
    near = _inst(c, c, [[s, 0], [0, s]], (0, 0, 1), 0.3)
    img = oracle.draw(np.stack([far, near]), W, H)  # draw order = far then near
    a1, a2 = 0.6, 0.3
    want_rgb = np.array([a1 * (1 - a2), 0, a2])
    want_a = a2 * a2 + (1 - a2) * (a1 * a1 + (1 - a1))
    np.testing.assert_allclose(img[8, 8, :3], want_rgb, rtol=1e-6)
    np.testing.assert_allclose(img[8, 8, 3], want_a, rtol=1e-6)


def test_draw_uniform_stack_closed_form():
    W = H = 8
    c = 2 * 4.5 / W - 1
    k = 20
    a = 0.15
    inst = np.stack([_inst(c, c, [[1, 0], [0, 1]], (0.3, 0.3, 0.3), a)] * k)
    img = oracle.draw(inst, W, H)
    T = (1 - a) ** k
    np.testing.assert_allclose(img[4, 4, 0], 0.3 * (1 - T), rtol=1e-5)


def test_draw_rectangle_is_conservative_and_bins_match_bruteforce():
    sc, cam, pos, cov, opa, sh, order, inst, di = _project_scene(n=2500, seed=12, w=64, h=48, r=3.0)
    a = oracle.draw(inst, 64, 48)
    b = oracle.draw_bruteforce(inst, 64, 48)
    np.testing.assert_array_equal(a, b)
    for tw, th in [(16, 16), (8, 8), (32, 8)]:
        ranges, lst = oracle.bin_tiles(inst, 64, 48, tw, th)
        tx = (64 + tw - 1) // tw
        # every covered pixel's splat appears in its tile's list, in draw order
        for t in range(ranges.shape[0]):
            items = lst[ranges[t, 0]:ranges[t, 1]]
            assert np.all(np.diff(items.astype(np.int64)) > 0)
        # rebuild the image tile by tile from the lists only
        img = np.zeros((48, 64, 4), np.float32)
        img[..., 3] = 1
        for t in range(ranges.shape[0]):
            x0, y0 = (t % tx) * tw, (t // tx) * th
            sub = oracle.draw_bruteforce(inst[lst[ranges[t, 0]:ranges[t, 1]]], 64, 48)
            img[y0:y0 + th, x0:x0 + tw] = sub[y0:y0 + th, x0:x0 + tw]
        np.testing.assert_array_equal(img, a)


# ----------------------------------------------------------- triangle-exact draw ---

def test_triangle_draw_square_edges_follow_the_top_left_rule():
    """An axis-aligned quad whose edges fall on pixel centres (centre (32.5, 32.5), half extent 6 px, all
    vertices on the 1/256 grid): the left and top edges' pixels are covered, the right and bottom edges'
    are not (the tie rule); inside, the interpolated p is the exact affine map, so alpha matches
    o exp(-|p|^2 / 2) with p = (x - c) / sigma, sigma = 2 px."""
    W = H = 64
    c = 2 * 32.5 / W - 1
    s = 0.0625  # 3 s W / 2 = 6 px
    inst = _inst(c, c, [[s, 0], [0, s]], (0.2, 0.4, 0.8), 0.7)[None]
    img = oracle.draw_triangles(inst, W, H)
    cov = img[..., 3] != 1.0
    rows, cols = np.nonzero(cov)
    assert (cols.min(), cols.max(), rows.min(), rows.max()) == (26, 37, 26, 37)
    for y, x in [(32, 32), (30, 35), (26, 26), (37, 29)]:
        p = np.array([x + 0.5 - 32.5, y + 0.5 - 32.5]) / 2.0
        a = 0.7 * math.exp(-0.5 * float(p @ p))
        np.testing.assert_allclose(img[y, x, 0], 0.2 * a, rtol=2e-6)


def test_triangle_draw_covers_the_shared_diagonal_once():
    """Pixel centres on the diagonal the two triangles share are covered by exactly one of them: with a
    uniform alpha the diagonal matches its neighbours' single blend, not a double one."""
    W = H = 64
    c = 2 * 32.5 / W - 1
    s = 0.0625
    inst = _inst(c, c, [[s, 0], [0, s]], (1.0, 1.0, 1.0), 0.5)[None]
    img = oracle.draw_triangles(inst, W, H)
    # the vertices 1 = (-1, +1) and 2 = (+1, -1) in q: the diagonal x - 32.5 = -(y - 32.5)
    for d in range(-5, 6):
        x, y = 32 + d, 32 - d
        p = np.array([x + 0.5 - 32.5, y + 0.5 - 32.5]) / 2.0
        a = 0.5 * math.exp(-0.5 * float(p @ p))
        np.testing.assert_allclose(img[y, x, 0], a, rtol=2e-6)
        np.testing.assert_allclose(img[y, x, 3], a * a + (1 - a), rtol=2e-6)


def test_triangle_draw_rejects_depth_and_nan_cases():
    W = H = 32
    c = 2 * 16.5 / W - 1
    s = 0.2
    for inst in (_inst(c, c, [[s, 0], [0, s]], (1, 1, 1), 0.7, z=1.0),
                 _inst(c, c, [[s, 0], [0, s]], (1, 1, 1), 0.7, z=-1e-6),
                 _inst(c, c, [[float("nan")] * 2] * 2, (1, 1, 1), 0.7)):
        img = oracle.draw_triangles(inst[None], W, H)
        assert np.all(img[..., 3] == 1.0) and np.all(img[..., :3] == 0.0)


def test_triangle_draw_agrees_with_the_analytic_draw_away_from_edges():
    """On a small random scene the two rasterizations differ only where the snapped, interpolated quad
    differs from the analytic square: pixels within a small margin of a quad edge, and the last bits of
    alpha from the 1/256-pixel vertex snap."""
    from vkgs_amd import scene as S
    sc = S.gaussian_cloud(300, seed=23, pos_std=1.0, logscale=(-2.0, 0.2))
    cam = S.orbit_camera(64, 48, 4.0)
    pos, cov, opa, sh = oracle.parse_ply(sc.rows, S.standard_offsets())
    v, key, index = oracle.rank(cam, pos)
    _, order = oracle.sort_kv(key, index)
    inst, _ = oracle.projection(cam, pos, cov, opa, sh, oracle.inverse(order, sc.n), v)
    a = oracle.draw(inst[:v], 64, 48)
    t = oracle.draw_triangles(inst[:v], 64, 48)
    d = np.abs(a - t).max(-1)
    assert (d > 1e-3).mean() < 0.05 and d.max() < 0.05
