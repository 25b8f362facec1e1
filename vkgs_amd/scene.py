"""Synthetic 3DGS scenes, PLY bytes and camera set-up for the splat pipeline.

The PLY layout is the standard 3DGS binary one that vkgs reads
(splat_load_thread.cc:55-135): x y z nx ny nz f_dc_0..2 f_rest_0..44 opacity
scale_0..2 rot_0..3, all float, little endian; rot_0 is w.

Scene generators (seeded numpy PCG64):
  gaussian_cloud : positions ~ N(0, pos_std) iso, log-scales ~ N(-4, 0.5),
                   quaternions uniform on S^3, opacity logits ~ N(0, 1.5),
                   SH dc ~ N(0, 0.5), rest ~ N(0, 0.05)      (100k / 500k workloads)
  clustered      : anisotropic clusters with centres uniform in a box plus a
                   fraction of points on the y = 0 ground plane  (1M / 6M workloads)
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

PLY_PROPERTIES = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
                  + [f"f_rest_{i}" for i in range(45)]
                  + ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])
STRIDE = len(PLY_PROPERTIES)  # 62
COL = {name: i for i, name in enumerate(PLY_PROPERTIES)}


def standard_offsets() -> np.ndarray:
    """The 60-entry offset table of splat_load_thread.cc:115-135 for PLY_PROPERTIES."""
    off = np.zeros(60, np.uint32)
    off[0], off[1], off[2] = COL["x"], COL["y"], COL["z"]
    off[3], off[4], off[5] = COL["scale_0"], COL["scale_1"], COL["scale_2"]
    off[6], off[7], off[8], off[9] = COL["rot_1"], COL["rot_2"], COL["rot_3"], COL["rot_0"]
    off[10 + 0], off[10 + 16], off[10 + 32] = COL["f_dc_0"], COL["f_dc_1"], COL["f_dc_2"]
    for i in range(15):
        off[10 + 1 + i] = COL[f"f_rest_{i}"]
        off[10 + 17 + i] = COL[f"f_rest_{15 + i}"]
        off[10 + 33 + i] = COL[f"f_rest_{30 + i}"]
    off[58] = COL["opacity"]
    off[59] = STRIDE
    return off


@dataclass
class Scene:
    rows: np.ndarray  # (N, 62) float32 PLY rows

    @property
    def n(self) -> int:
        return int(self.rows.shape[0])

    def ply_buffer(self) -> np.ndarray:
        """The reference's ply buffer: 60 u32 offsets then N x 62 floats (splat_load_thread.cc:157-158)."""
        buf = np.empty(60 + self.rows.size, np.uint32)
        buf[:60] = standard_offsets()
        buf[60:] = np.ascontiguousarray(self.rows, np.float32).view(np.uint32).ravel()
        return buf

    def ply_bytes(self) -> bytes:
        header = ["ply", "format binary_little_endian 1.0", f"element vertex {self.n}"]
        header += [f"property float {p}" for p in PLY_PROPERTIES]
        header += ["end_header"]
        return ("\n".join(header) + "\n").encode() + np.ascontiguousarray(self.rows, "<f4").tobytes()

    def write_ply(self, path: str) -> None:
        with open(path, "wb") as f:
            f.write(self.ply_bytes())

    # raw (differentiable) parameters in PLY order
    def means(self) -> np.ndarray:
        return self.rows[:, 0:3]

    def log_scales(self) -> np.ndarray:
        return self.rows[:, COL["scale_0"]:COL["scale_0"] + 3]

    def quats_wxyz(self) -> np.ndarray:
        return self.rows[:, COL["rot_0"]:COL["rot_0"] + 4]

    def opacity_logits(self) -> np.ndarray:
        return self.rows[:, COL["opacity"]]

    def sh_channel_major(self) -> np.ndarray:
        """(N, 48) SH in the vkgs order [R: dc, rest0..14][G: dc, rest15..29][B: dc, rest30..44]."""
        off = standard_offsets()
        return self.rows[:, off[10:58]]


def _fill_common(rng, rows, logscale, logit, dc_std, rest_std):
    n = rows.shape[0]
    rows[:, COL["scale_0"]:COL["scale_0"] + 3] = rng.normal(logscale[0], logscale[1], (n, 3))
    q = rng.normal(0.0, 1.0, (n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    rows[:, COL["rot_0"]:COL["rot_0"] + 4] = q
    rows[:, COL["opacity"]] = rng.normal(logit[0], logit[1], n)
    rows[:, COL["f_dc_0"]:COL["f_dc_0"] + 3] = rng.normal(0.0, dc_std, (n, 3))
    rows[:, COL["f_rest_0"]:COL["f_rest_0"] + 45] = rng.normal(0.0, rest_std, (n, 45))


def gaussian_cloud(n: int, seed: int = 1, pos_std: float = 2.0, logscale=(-4.0, 0.5), logit=(0.0, 1.5),
                   dc_std: float = 0.5, rest_std: float = 0.05) -> Scene:
    rng = np.random.Generator(np.random.PCG64(seed))
    rows = np.zeros((n, STRIDE), np.float64)
    rows[:, 0:3] = rng.normal(0.0, pos_std, (n, 3))
    _fill_common(rng, rows, logscale, logit, dc_std, rest_std)
    return Scene(rows.astype(np.float32))


def clustered(n: int, seed: int, clusters: int, box=(40.0, 10.0, 40.0), ground_frac: float = 0.2,
              ground_jitter: float = 0.02, logscale=(-4.5, 0.8), logit=(0.5, 2.0), dc_std: float = 0.5,
              rest_std: float = 0.05) -> Scene:
    """Anisotropic clusters, centres uniform in [-bx/2,bx/2] x [0,by] x [-bz/2,bz/2]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    bx, by, bz = box
    rows = np.zeros((n, STRIDE), np.float64)
    n_ground = int(round(n * ground_frac))
    n_cl = n - n_ground
    centers = np.stack([rng.uniform(-bx / 2, bx / 2, clusters), rng.uniform(0.0, by, clusters),
                        rng.uniform(-bz / 2, bz / 2, clusters)], 1)
    stds = rng.uniform(0.2, 0.1 * max(bx, bz), (clusters, 3))
    rq = rng.normal(0.0, 1.0, (clusters, 4))
    rq /= np.linalg.norm(rq, axis=1, keepdims=True)
    w, x, y, z = rq.T
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    which = rng.integers(0, clusters, n_cl)
    local = rng.normal(0.0, 1.0, (n_cl, 3)) * stds[which]
    rows[:n_cl, 0:3] = centers[which] + np.einsum("nij,nj->ni", R[which], local)
    rows[n_cl:, 0] = rng.uniform(-bx / 2, bx / 2, n_ground)
    rows[n_cl:, 1] = rng.normal(0.0, ground_jitter, n_ground)
    rows[n_cl:, 2] = rng.uniform(-bz / 2, bz / 2, n_ground)
    _fill_common(rng, rows, logscale, logit, dc_std, rest_std)
    perm = rng.permutation(n)
    return Scene(rows[perm].astype(np.float32))


@dataclass
class Workload:
    name: str
    n: int
    width: int
    height: int
    radius: float
    make: callable


WORKLOADS = {
    "c1_100k": Workload("c1_100k", 100_000, 1600, 900, 6.0, lambda: gaussian_cloud(100_000, seed=1)),
    "c2_1m": Workload("c2_1m", 1_000_000, 1600, 900, 15.0,
                      lambda: clustered(1_000_000, seed=2, clusters=64, box=(40.0, 10.0, 40.0))),
    "c3_6m": Workload("c3_6m", 6_000_000, 1600, 900, 60.0,
                      lambda: clustered(6_000_000, seed=3, clusters=256, box=(80.0, 20.0, 80.0))),
    "c4_500k": Workload("c4_500k", 500_000, 1024, 1024, 6.0, lambda: gaussian_cloud(500_000, seed=4)),
}


def orbit_camera(width: int, height: int, r: float, phi: float = math.radians(45.0),
                 theta: float = math.radians(45.0), center=(0.0, 0.0, 0.0), fovy: float = math.radians(60.0),
                 znear: float = 0.01, zfar: float = 100.0):
    """vkgs::Camera (camera.cc:25-45) through the C ABI's camera producer."""
    from . import _lib
    cam = _lib.Camera()
    c = (C.c_float * 3)(*center)
    _lib.check(_lib.lib().vkgs_camera_orbit(fovy, width, height, znear, zfar, c, r, phi, theta, C.byref(cam)),
               "vkgs_camera_orbit")
    return cam


def camera_arrays(cam):
    """(projection, view, eye) as numpy float32 arrays (column-major 4x4 flattened)."""
    return (np.frombuffer(bytes(cam.projection), np.float32).copy(),
            np.frombuffer(bytes(cam.view), np.float32).copy(),
            np.frombuffer(bytes(cam.camera_position), np.float32).copy())
