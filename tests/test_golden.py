"""The committed small-frame fixture (tests/golden/make_small_frame.py).

The reference holds no golden vectors; this fixture freezes the oracle's frame
on a 400-splat scene at 64x48 so that a change to the oracle, the camera
producer or the scene generator is caught against committed data.  The GPU
stage path is checked against the same data in test_gpu_parity.py.
"""
import os

import numpy as np

import oracle
from vkgs_amd import _lib
from vkgs_amd import scene as S

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "small_frame.npz")


def load():
    d = np.load(FIXTURE)  # no pickles: plain arrays
    return {k: d[k] for k in d.files}


def test_generators_reproduce_fixture_inputs():
    d = load()
    sc = S.gaussian_cloud(400, seed=21, pos_std=1.0, logscale=(-2.5, 0.2))
    np.testing.assert_array_equal(sc.rows.astype(np.float32), d["rows"])
    cam = S.orbit_camera(64, 48, 4.0)
    np.testing.assert_array_equal(np.frombuffer(bytes(cam), np.uint8), d["camera"])


def test_oracle_reproduces_fixture_frame():
    d = load()
    cam = _lib.Camera.from_buffer_copy(d["camera"].tobytes())
    pos, cov, opa, sh = oracle.parse_ply(d["rows"], S.standard_offsets())
    f = oracle.frame(cam, pos, cov, opa, sh, 64, 48)
    assert f["visible"] == int(d["visible"][0])
    np.testing.assert_array_equal(f["keys"].astype(np.uint32), d["keys"])
    np.testing.assert_array_equal(f["order"].astype(np.uint32), d["order"])
    np.testing.assert_array_equal(f["inverse"].astype(np.int32), d["inverse"])
    np.testing.assert_array_equal(f["instances"].astype(np.float32).view(np.uint32), d["instances"].view(np.uint32))
    np.testing.assert_array_equal(f["draw_indirect"].astype(np.uint32), d["draw_indirect"])
    np.testing.assert_array_equal(f["image"].astype(np.float32), d["image"])
    assert (d["image"][..., 3] < 1.0).sum() > 100  # the frame covers a good part of the image
