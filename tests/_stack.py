"""Stacked-splat scenes for the compositing stop rule (test helpers)."""
import numpy as np

from oracle import torch_ref
from vkgs_amd import scene as S


def stacked_scene(k, logit, rgb, log_scale=(-2.0, -2.4, -2.2), jitter=0.0, seed=0):
    """k equal axis-aligned splats at the origin (optionally jittered by N(0, jitter) m), colour rgb
    (a (3,) colour or a (k, 3) array) from the SH DC term alone.  The default scales differ per
    axis: an isotropic splat on the optical axis of a square image has an isotropic 2D covariance,
    whose eigendecomposition is NaN in the reference's projection, so it would draw nothing."""
    rows = np.zeros((k, S.STRIDE), np.float32)
    c = S.COL
    if jitter > 0:
        rows[:, 0:3] = np.random.default_rng(seed).normal(0.0, jitter, (k, 3))
    rows[:, c["scale_0"]:c["scale_0"] + 3] = np.asarray(log_scale, np.float32)
    rows[:, c["rot_0"]] = 1.0
    rows[:, c["opacity"]] = logit
    rgb = np.broadcast_to(np.asarray(rgb, np.float64), (k, 3))
    for ch in range(3):
        rows[:, c[f"f_dc_{ch}"]] = (rgb[:, ch] - 0.5) / torch_ref.C0
    return S.Scene(rows)


def held_rgb(sc, i=0):
    """The colour splat i's float32 SH DC coefficients hold."""
    return torch_ref.C0 * sc.sh_channel_major()[i, [0, 16, 32]].astype(np.float64) + 0.5


def stack_closed_form(alpha, rgb, k, t_min):
    """Centre pixel of k equal splats of alpha and colour rgb: the j-th from the front sees
    T_j = (1 - alpha)^j and contributes iff T_j >= t_min, so with m contributors
    C = rgb (1 - (1 - alpha)^m) and A = (1 - alpha)^m + alpha (1 - (1 - alpha)^m)."""
    m = sum(1 for j in range(k) if (1.0 - alpha) ** j >= t_min)
    t = (1.0 - alpha) ** m
    return np.array([*(np.asarray(rgb) * (1.0 - t)), t + alpha * (1.0 - t)]), m
