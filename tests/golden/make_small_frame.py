"""Writes small_frame.npz: a regression fixture of one whole frame on a small scene.

The reference holds no tests, fixtures or golden vectors, so this fixture does
not pin parity with it: it freezes what the oracle computes today (whose
functions are pinned by closed forms, library routines and brute force in
test_oracle.py), so that a later change to the oracle, the camera producer or
the scene generator shows up as a diff against committed data.

Contents: the scene rows (float32, standard 3DGS PLY columns), the camera
block bytes (192 B), and the oracle's frame: visible count, sorted keys and
ids, inverse index, the 48-B instances, the DrawIndirect words and the image.

Run from the repository root: python tests/golden/make_small_frame.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from vkgs_amd import scene as S  # noqa: E402

W, H, N = 64, 48, 400


def make():
    sc = S.gaussian_cloud(N, seed=21, pos_std=1.0, logscale=(-2.5, 0.2))
    cam = S.orbit_camera(W, H, 4.0)
    pos, cov, opa, sh = oracle.parse_ply(sc.rows, S.standard_offsets())
    f = oracle.frame(cam, pos, cov, opa, sh, W, H)
    return dict(rows=sc.rows.astype(np.float32), camera=np.frombuffer(bytes(cam), np.uint8).copy(),
                visible=np.array([f["visible"]], np.int64), keys=f["keys"].astype(np.uint32),
                order=f["order"].astype(np.uint32), inverse=f["inverse"].astype(np.int32),
                instances=f["instances"].astype(np.float32), draw_indirect=f["draw_indirect"].astype(np.uint32),
                image=f["image"].astype(np.float32))


if __name__ == "__main__":
    out = os.path.join(HERE, "small_frame.npz")
    np.savez_compressed(out, **make())
    print("wrote", out, os.path.getsize(out), "bytes")
