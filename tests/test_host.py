"""Host-side logic of the C ABI (no device work): library exports, camera
producer, 4x4 inverse and the checked PLY reader.  Runs without a GPU."""
import ctypes as C
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from vkgs_amd import _lib
from vkgs_amd import scene as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "vkgs_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(vkgs_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.vkgs_abi_version() == 1
    assert set(names) <= set(_lib.SIGNATURES), sorted(set(names) - set(_lib.SIGNATURES))


def test_camera_struct_layout_matches_uniforms_h():
    # vk::shader::Camera: mat4 @0, mat4 @64, vec3 @128, alignas(16) uvec2 @144, alignas(64) -> 192
    assert C.sizeof(_lib.Camera) == 192
    assert _lib.Camera.view.offset == 64
    assert _lib.Camera.camera_position.offset == 128
    assert _lib.Camera.screen_size.offset == 144


def test_raster_config_layout_and_launch_choice_validation():
    """vkgs_raster_config keeps its 32-byte layout; the launch choices (pair form, LDS batch depth,
    resident-wave cap) are validated on the host, here through vkgs_draw_storage_bytes, which
    returns 0 for a configuration no launch accepts."""
    assert C.sizeof(_lib.RasterConfig) == 32
    assert _lib.RasterConfig.lds_batch.offset == 20 and _lib.RasterConfig.waves_per_cu.offset == 24
    lib = _lib.lib()

    def size(**kw):
        cfg = _lib.default_raster_config()
        for k, v in kw.items():
            setattr(cfg, k, v)
        return lib.vkgs_draw_storage_bytes(1000, 640, 480, 100000, C.byref(cfg))

    assert size() > 0
    for good in (dict(lds_batch=8), dict(lds_batch=32), dict(waves_per_cu=4), dict(waves_per_cu=32),
                 dict(flags=4), dict(flags=8)):
        assert size(**good) > 0, good
    for bad in (dict(lds_batch=12), dict(lds_batch=64), dict(waves_per_cu=2), dict(waves_per_cu=33),
                dict(flags=4 | 8)):
        assert size(**bad) == 0, bad


def _mats(cam):
    P, V, eye = S.camera_arrays(cam)
    return P.reshape(4, 4).T.astype(np.float64), V.reshape(4, 4).T.astype(np.float64), eye.astype(np.float64)


def test_camera_orbit_invariants():
    w, h, r = 1600, 900, 6.0
    phi, theta = math.radians(45), math.radians(45)
    cam = S.orbit_camera(w, h, r, phi=phi, theta=theta, center=(1.0, 2.0, -0.5))
    P, V, eye = _mats(cam)
    c = np.array([1.0, 2.0, -0.5])
    # camera.cc:39-45
    want_eye = c + r * np.array([math.sin(phi) * math.sin(theta), math.cos(phi), math.sin(phi) * math.cos(theta)])
    np.testing.assert_allclose(eye, want_eye, rtol=1e-6)
    # view: eye -> origin, centre -> (0, 0, -r), +y stays up, orthonormal
    np.testing.assert_allclose(V @ np.r_[eye, 1], [0, 0, 0, 1], atol=1e-5)
    np.testing.assert_allclose(V @ np.r_[c, 1], [0, 0, -r, 1], atol=1e-5)
    np.testing.assert_allclose(V[:3, :3] @ V[:3, :3].T, np.eye(3), atol=1e-6)
    assert (V[:3, :3] @ np.array([0, 1, 0]))[1] > 0
    # projection (perspective + GL->Vulkan): near -> z 0, far -> z 1, y flipped, fov and aspect
    n, f = 0.01, 100.0
    for d, z in [(n, 0.0), (f, 1.0), (6.0, f / (f - n) * (1 - n / 6.0))]:
        clip = P @ np.array([0, 0, -d, 1])
        assert abs(clip[2] / clip[3] - z) < 1e-5
    t = math.tan(math.radians(30))
    clip = P @ np.array([0, t * 5.0, -5.0, 1])  # top edge of the frustum at depth 5
    assert abs(clip[1] / clip[3] + 1.0) < 1e-5  # Vulkan: +y view is -y NDC (camera.cc:31)
    clip = P @ np.array([t * 5.0 * w / h, 0, -5.0, 1])
    assert abs(clip[0] / clip[3] - 1.0) < 1e-5
    assert list(cam.screen_size) == [w, h]


def test_camera_rejects_bad_arguments():
    cam = _lib.Camera()
    c = (C.c_float * 3)(0, 0, 0)
    lib = _lib.lib()
    assert lib.vkgs_camera_orbit(1.0, 0, 10, 0.01, 100.0, c, 2.0, 0.7, 0.7, C.byref(cam)) == 1
    assert lib.vkgs_camera_orbit(1.0, 10, 10, 1.0, 0.5, c, 2.0, 0.7, 0.7, C.byref(cam)) == 1
    assert lib.vkgs_camera_orbit(1.0, 10, 10, 0.01, 100.0, c, 2.0, 0.0, 0.7, C.byref(cam)) == 1  # looking along y


def test_mat4_inverse():
    rng = np.random.default_rng(0)
    lib = _lib.lib()
    for _ in range(20):
        m = rng.normal(size=(4, 4)).astype(np.float32)
        out = np.zeros(16, np.float32)
        assert lib.vkgs_mat4_inverse(m.T.ravel().ctypes.data, out.ctypes.data) == 0
        np.testing.assert_allclose(out.reshape(4, 4).T @ m, np.eye(4), atol=1e-4)
    eye = np.eye(4, dtype=np.float32).ravel()
    out = np.zeros(16, np.float32)
    assert lib.vkgs_mat4_inverse(eye.ctypes.data, out.ctypes.data) == 0
    np.testing.assert_array_equal(out, eye)
    z = np.zeros(16, np.float32)
    assert lib.vkgs_mat4_inverse(z.ctypes.data, out.ctypes.data) == 1


# ------------------------------------------------------------------- PLY ---

def read_header(path):
    info = _lib.PlyInfo()
    st = _lib.lib().vkgs_ply_read_header(path.encode(), C.byref(info))
    return st, info


def test_ply_roundtrip(tmp_path):
    sc = S.gaussian_cloud(1000, seed=2)
    p = str(tmp_path / "a.ply")
    sc.write_ply(p)
    st, info = read_header(p)
    assert st == 0
    assert info.point_count == 1000 and info.stride == 62
    np.testing.assert_array_equal(np.array(info.offsets), S.standard_offsets())
    nbytes = _lib.lib().vkgs_ply_buffer_bytes(C.byref(info))
    buf = np.zeros(nbytes // 4, np.uint32)
    assert _lib.lib().vkgs_ply_read(p.encode(), C.byref(info), buf.ctypes.data, nbytes) == 0
    np.testing.assert_array_equal(buf, sc.ply_buffer())
    assert _lib.lib().vkgs_ply_read(p.encode(), C.byref(info), buf.ctypes.data, nbytes - 4) == 2  # capacity


def test_ply_permuted_properties(tmp_path):
    """Any property order works through the offset table (splat_load_thread.cc:115-135)."""
    sc = S.gaussian_cloud(300, seed=3)
    rng = np.random.default_rng(0)
    perm = rng.permutation(S.STRIDE)
    names = [S.PLY_PROPERTIES[i] for i in perm] + ["extra"]
    rows = np.c_[sc.rows[:, perm], np.full(300, 7.0, np.float32)]
    header = ["ply", "format binary_little_endian 1.0", "comment made by a test", "element vertex 300"]
    header += [f"property float {n}" for n in names] + ["element face 0", "property list uchar int vertex_indices",
                                                       "end_header"]
    p = str(tmp_path / "p.ply")
    with open(p, "wb") as f:
        f.write(("\n".join(header) + "\n").encode() + rows.astype("<f4").tobytes())
    st, info = read_header(p)
    assert st == 0 and info.stride == 63
    nbytes = _lib.lib().vkgs_ply_buffer_bytes(C.byref(info))
    buf = np.zeros(nbytes // 4, np.uint32)
    assert _lib.lib().vkgs_ply_read(p.encode(), C.byref(info), buf.ctypes.data, nbytes) == 0
    off = buf[:60]
    data = buf[60:].view(np.float32).reshape(300, 63)
    std = S.standard_offsets()
    for slot in range(59):
        np.testing.assert_array_equal(data[:, off[slot]], sc.rows[:, std[slot]])


@pytest.mark.parametrize("mutate,status", [
    (lambda h, b: (h.replace("binary_little_endian", "ascii"), b), 5),
    (lambda h, b: (h.replace("binary_little_endian", "binary_big_endian"), b), 5),
    (lambda h, b: (h.replace("property float opacity", "property double opacity"), b), 5),
    (lambda h, b: (h.replace("property float rot_2\n", ""), b[:-4]), 5),
    (lambda h, b: (h.replace("property float f_rest_44", "property uchar f_rest_44"), b), 5),
    (lambda h, b: (h, b[:-1]), 5),                       # truncated data
    (lambda h, b: (h.replace("end_header\n", ""), b), 5),  # no end_header
    (lambda h, b: (h.replace("ply\n", "plx\n", 1), b), 5),
])
def test_ply_malformed(tmp_path, mutate, status):
    sc = S.gaussian_cloud(4, seed=1)
    raw = sc.ply_bytes()
    cut = raw.index(b"end_header\n") + len(b"end_header\n")
    h, b = mutate(raw[:cut].decode(), raw[cut:])
    p = str(tmp_path / "m.ply")
    with open(p, "wb") as f:
        f.write(h.encode() + b)
    st, _ = read_header(p)
    assert st == status


def test_ply_missing_file():
    st, _ = read_header("/nonexistent/x.ply")
    assert st == 4


def test_bench_refuses_multi_gpu():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stdout


def test_integer_depth_key_orders_and_ties_as_the_float_key():
    """The renderer's depth sort and the batch's view-major keys sort k = fl(1 - z) * 2^24 instead of the
    float bits of fl(1 - z) (rank.comp:39).  For z in [0, 1] (the visible range) every fl(1 - z) is a
    multiple of 2^-24 in [0, 1], so k is an exact integer <= 2^24 (25 bits); a stable sort of k gives the
    float key's order and ties, and k * 2^-24 gives its bits back.  Checked on uniform z, on z crowded
    near 0 and near 1 (where the spacing of fl(1 - z) changes), on the edge values, and on the oracle's
    keys of a scene."""
    import oracle

    rng = np.random.default_rng(3)
    f32 = np.float32
    edges = np.array([0.0, -0.0, 1.0, 0.5, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 1 - 2.0 ** -24,
                      np.nextafter(f32(0.5), f32(0)), np.nextafter(f32(0.5), f32(1)),
                      np.nextafter(f32(1), f32(0)), np.nextafter(f32(0), f32(1))], f32)
    z = np.concatenate([rng.uniform(0, 1, 400_000).astype(f32), (rng.uniform(0, 1, 100_000) ** 12).astype(f32),
                        (1 - rng.uniform(0, 1, 100_000) ** 12).astype(f32), np.repeat(edges, 3)])
    one_minus_z = (f32(1) - z).astype(f32)  # fl(1 - z), as the kernels compute it
    k = (one_minus_z * f32(2 ** 24)).astype(np.uint32)  # depth_rank_key
    assert np.array_equal(k.astype(np.float64), one_minus_z.astype(np.float64) * 2.0 ** 24)  # exact
    assert k.max() <= 1 << 24
    bits = one_minus_z.view(np.uint32)
    np.testing.assert_array_equal(np.argsort(k, kind="stable"), np.argsort(bits, kind="stable"))
    np.testing.assert_array_equal((k.astype(f32) * f32(2.0 ** -24)).view(np.uint32), bits)  # depth_key_bits
    # the oracle's rank keys (float bits of 1 - z of visible splats) of a scene with near and far splats
    sc = S.gaussian_cloud(50_000, seed=4, pos_std=2.0)
    _, okey, _ = oracle.rank(S.orbit_camera(640, 480, 3.0), np.ascontiguousarray(sc.rows[:, 0:3], f32))
    ok = (np.asarray(okey, np.uint32).view(f32) * f32(2 ** 24)).astype(np.uint32)
    np.testing.assert_array_equal(np.argsort(ok, kind="stable"), np.argsort(okey, kind="stable"))
