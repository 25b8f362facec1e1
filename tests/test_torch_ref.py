"""Pins the float64 torch restatement of the differentiable forward (the
gradient oracle): its image against the C oracle's fp32 frame, and its autograd
gradients against FP64 central finite differences (h = 1e-6)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from oracle import torch_ref  # noqa: E402
from vkgs_amd import scene as S  # noqa: E402

from _stack import held_rgb, stack_closed_form, stacked_scene  # noqa: E402


def small_scene(n, seed, w, h, r=4.0):
    sc = S.gaussian_cloud(n, seed=seed, pos_std=1.0, logscale=(-2.5, 0.4))
    c = S.COL
    sc.rows[:, c["opacity"]] = np.clip(sc.rows[:, c["opacity"]], -6, 4)  # alpha stays below the 0.999 clamp
    # SH representable in fp16, so the fp16 storage of the plain path is exact
    sc.rows[:, c["f_dc_0"]:c["f_rest_0"] + 45] = sc.rows[:, c["f_dc_0"]:c["f_rest_0"] + 45].astype(np.float16)
    cam = S.orbit_camera(w, h, r)
    return sc, cam


def params64(sc):
    return [torch.tensor(np.ascontiguousarray(a), dtype=torch.float64)
            for a in (sc.means(), sc.log_scales(), sc.quats_wxyz(), sc.opacity_logits(), sc.sh_channel_major())]


def cam64(cam):
    P, V, eye = S.camera_arrays(cam)
    return (torch.tensor(P.reshape(4, 4).T, dtype=torch.float64), torch.tensor(V.reshape(4, 4).T, dtype=torch.float64),
            torch.tensor(eye, dtype=torch.float64))


def draw_order(sc, cam):
    pos, cov, opa, sh = oracle.parse_ply(sc.rows, S.standard_offsets())
    v, key, index = oracle.rank(cam, pos)
    _, order = oracle.sort_kv(key, index)
    return order.astype(np.int64), (pos, cov, opa, sh)


def test_torch_ref_image_matches_oracle_frame():
    sc, cam = small_scene(1500, 21, 64, 48)
    order, (pos, cov, opa, sh) = draw_order(sc, cam)
    P, V, eye = cam64(cam)
    img = torch_ref.render(*params64(sc), P, V, eye, 64, 48, order).numpy()
    ref = oracle.frame(cam, pos, cov, opa, sh, 64, 48)["image"]
    d = np.abs(img - ref).max(-1)
    # float64 and float32 can disagree on a pixel centre lying within rounding of
    # a quad edge; such a flip moves the pixel by at most one edge alpha
    # (opacity * exp(-4.5) ~ 0.011) times its colour
    assert (d > 1e-5).mean() < 0.005
    assert d.max() < 0.02


def test_torch_ref_autograd_matches_central_differences():
    sc, cam = small_scene(250, 22, 32, 32, r=3.5)
    order, _ = draw_order(sc, cam)
    P, V, eye = cam64(cam)
    ps = params64(sc)
    rng = np.random.default_rng(0)
    wts = torch.tensor(rng.uniform(0.5, 1.5, (32, 32, 4)), dtype=torch.float64)

    def loss_and_mask(pp):
        img, mask = torch_ref.render(*pp, P, V, eye, 32, 32, order, return_mask=True)
        return (img * wts).sum(), mask

    for p in ps:
        p.requires_grad_(True)
    L, mask0 = loss_and_mask(ps)
    grads = torch.autograd.grad(L, ps)
    h = 1e-6
    checked = 0
    with torch.no_grad():
        for gi, (p, g) in enumerate(zip(ps, grads)):
            flat = p.detach().reshape(-1)
            # sample parameters of drawn splats
            rows = rng.choice(order, size=min(12, len(order)), replace=False)
            width = flat.numel() // p.shape[0]
            for r_ in rows:
                j = int(r_) * width + int(rng.integers(0, width))
                base = [q.detach().clone() for q in ps]
                plus = [q.clone() for q in base]
                minus = [q.clone() for q in base]
                plus[gi].view(-1)[j] += h
                minus[gi].view(-1)[j] -= h
                lp, mp = loss_and_mask(plus)
                lm, mm = loss_and_mask(minus)
                if not torch.equal(mp, mm):
                    continue  # a coverage edge moved across a pixel centre: not differentiable there
                fd = float((lp - lm) / (2 * h))
                an = float(g.reshape(-1)[j])
                assert abs(fd - an) <= 1e-5 + 1e-4 * abs(an), (gi, j, fd, an)
                checked += 1
    assert checked >= 40


@pytest.mark.parametrize("t_min,m_want", [(0.0, 30), (1.0 / 8192.0, 14), (2.0 ** -20, 21)])
def test_torch_ref_t_min_stop_matches_closed_form(t_min, m_want):
    # alpha 0.5 (logit 0): T_j = 2^-j hits t_min = 2^-13 and 2^-20 exactly, so the ">=" of the
    # stop rule is exercised at equality
    rgb = (0.8, 0.3, 0.55)
    sc = stacked_scene(30, 0.0, rgb)
    cam = S.orbit_camera(33, 33, 4.0)  # odd size: the origin projects onto the centre pixel's centre
    order = np.arange(30)
    P, V, eye = cam64(cam)
    img = torch_ref.render(*params64(sc), P, V, eye, 33, 33, order, t_min=t_min).numpy()
    want, m = stack_closed_form(0.5, held_rgb(sc), 30, t_min)
    assert m == m_want
    np.testing.assert_allclose(img[16, 16], want, rtol=0, atol=1e-9)
