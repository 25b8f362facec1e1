"""Float64 torch restatement of the differentiable splat forward (TEST INFRASTRUCTURE).

Used only by tests/ and __graft_entry__.smoke() as the gradient oracle: its
autograd gradients are pinned against FP64 central finite differences, and the
HIP backward is compared against them.

It restates, in float64 and vectorised over (splat, pixel):
  activation      parse_ply.comp:46-96 (scale = exp, normalized quaternion,
                  cov3d = R diag(s^2) R^T, opacity = sigmoid; SH kept in fp32)
  projection      projection.comp:80-179 (cov2d = P2 J W Sigma W^T J^T P2 with
                  the 1/W^2, 1/H^2 low-pass; eigendecomposition with atan2, as
                  the shader; SH colour clamped at 0)
  raster + blend  splat.vert / splat.frag / engine.cc:279-299: the +-3 sigma
                  square in the eigenframe, alpha = opacity * exp(-|p|^2 / 2)
                  clamped to alpha_max, composited in draw order (far to near)
                  over the clear colour (0, 0, 0, 1).
The draw order (a discrete quantity) is passed in; the CPU oracle's rank and
stable sort produce it.
"""
from __future__ import annotations

import math

import torch

C0 = 0.28209479177387814
C1 = 0.4886025119029199
C20 = 1.0925484305920792
C21 = 0.31539156525252005
C22 = 0.5462742152960396
C30 = 0.5900435899266435
C31 = 2.890611442640554
C32 = 0.4570457994644658
C33 = 0.3731763325901154
C34 = 1.445305721320277


def sh_basis(d):
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
    return torch.stack([
        torch.full_like(x, C0), -C1 * y, C1 * z, -C1 * x,
        C20 * xy, -C20 * yz, C21 * (2 * zz - xx - yy), -C20 * xz,
        C22 * (xx - yy), -C30 * y * (3 * xx - yy), C31 * xy * z, -C32 * y * (4 * zz - xx - yy),
        C33 * z * (2 * zz - 3 * xx - 3 * yy), -C32 * x * (4 * zz - xx - yy), C34 * z * (xx - yy),
        -C30 * x * (xx - 3 * yy)], 1)


def project(means, log_scales, quats, opacity_logits, sh, P, V, eye, width, height):
    """Per-splat instance quantities: ndc centre (N,2), depth (N,), rot_scale (N,2,2) with
    columns as in projection.comp:134, cov2d (N,2,2), colour (N,3), opacity (N,)."""
    s = torch.exp(log_scales)
    q = quats[:, [1, 2, 3, 0]]
    q = q / torch.linalg.norm(q, dim=1, keepdim=True)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = torch.stack([
        torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        torch.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        torch.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
    Sig = R @ torch.diag_embed(s * s) @ R.transpose(1, 2)
    W3 = V[:3, :3]
    pv = means @ W3.T + V[:3, 3]
    Sp = W3 @ Sig @ W3.T
    px, py, pz = pv[:, 0], pv[:, 1], pv[:, 2]
    zero = torch.zeros_like(pz)
    J2 = torch.stack([torch.stack([-1 / pz, zero, px / pz / pz], -1),
                      torch.stack([zero, -1 / pz, py / pz / pz], -1)], 1)
    S2 = J2 @ Sp @ J2.transpose(1, 2)
    Ps = P[:2, :2]
    cov2 = Ps @ S2 @ Ps
    lp = torch.tensor([[1.0 / width / width, 0.0], [0.0, 1.0 / height / height]], dtype=means.dtype)
    cov2 = cov2 + lp
    a, b, c = cov2[:, 0, 0], cov2[:, 1, 1], cov2[:, 0, 1]
    D = torch.sqrt((a - b) ** 2 + 4 * c * c)
    s0 = torch.sqrt(0.5 * (a + b + D))
    s1 = torch.sqrt(0.5 * (a + b - D))
    theta = torch.atan2(2 * c / D, (a - b) / D) / 2
    ct, st = torch.cos(theta), torch.sin(theta)
    rs = torch.stack([torch.stack([s0 * ct, -s1 * st], -1), torch.stack([s0 * st, s1 * ct], -1)], 1)
    clip = pv @ P[:3, :3].T + P[:3, 3]
    w4 = pv @ P[3, :3] + P[3, 3]
    ndc = clip[:, :2] / w4[:, None]
    depth = clip[:, 2] / w4
    d = means - eye
    d = d / torch.linalg.norm(d, dim=1, keepdim=True)
    B = sh_basis(d)
    shc = sh.reshape(-1, 3, 16)
    color = torch.clamp_min(torch.einsum("nj,ncj->nc", B, shc) + 0.5, 0.0)
    opacity = torch.sigmoid(opacity_logits)
    return ndc, depth, rs, cov2, color, opacity


def render(means, log_scales, quats, opacity_logits, sh, P, V, eye, width, height, order, alpha_max=0.999,
           return_mask=False, t_min=0.0, return_edge=False):
    """(H, W, 4) image; order = splat ids in draw order (far to near).

    t_min: a pixel stops compositing once its transmittance falls below t_min -- a splat
    contributes only if the transmittance in front of it is >= t_min, and the alpha channel's
    T term is the transmittance after the last contributing splat (the rule of the GPU raster;
    0 composites every splat).  The stop is a discrete choice, held constant under autograd.

    return_edge: also return, per pixel (H, W), the smallest distance of max(|u_x|, |u_y|) to the
    footprint's edge 3 over the drawn splats (in front of the far plane), in eigenframe units: a
    float32 renderer can decide coverage differently from this float64 one only where it is
    within rounding of 0."""
    ndc, depth, rs, cov2, color, opacity = project(means, log_scales, quats, opacity_logits, sh, P, V, eye, width,
                                                   height)
    idx = torch.as_tensor(order, dtype=torch.long)
    ndc, depth, rs, color, opacity = ndc[idx], depth[idx], rs[idx], color[idx], opacity[idx]
    hw, hh = 0.5 * width, 0.5 * height
    cpx = torch.stack([(ndc[:, 0] + 1) * hw, (ndc[:, 1] + 1) * hh], 1)
    A = rs * torch.tensor([[hw], [hh]], dtype=rs.dtype)  # rows scaled: pixel-space rot_scale
    Ainv = torch.linalg.inv(A)
    ys, xs = torch.meshgrid(torch.arange(height, dtype=means.dtype) + 0.5,
                            torch.arange(width, dtype=means.dtype) + 0.5, indexing="ij")
    pix = torch.stack([xs.reshape(-1), ys.reshape(-1)], 1)  # (P, 2)
    dvec = pix[None, :, :] - cpx[:, None, :]  # (K, P, 2)
    u = torch.einsum("kij,kpj->kpi", Ainv, dvec)
    cover = (u.abs() <= 3).all(-1) & (depth >= 0)[:, None] & (depth < 1)[:, None]
    alpha = opacity[:, None] * torch.exp(-0.5 * (u * u).sum(-1))
    alpha = torch.where(cover, torch.clamp_max(alpha, alpha_max), torch.zeros_like(alpha))
    # front to back: reverse draw order
    af = alpha.flip(0)
    cf = color.flip(0)
    if t_min > 0:
        with torch.no_grad():  # T only decreases, so "T in front >= t_min" holds up to the stop
            t_front = torch.cumprod(torch.cat([torch.ones_like(af[:1]), 1 - af[:-1]], 0), 0)
            keep = t_front >= t_min
        af = torch.where(keep, af, torch.zeros_like(af))
    one_minus = 1 - af
    T = torch.cumprod(torch.cat([torch.ones_like(af[:1]), one_minus[:-1]], 0), 0)  # T in front of each
    wgt = af * T
    rgb = torch.einsum("kp,kc->pc", wgt, cf)
    a_out = (wgt * af).sum(0) + T[-1] * one_minus[-1]
    img = torch.cat([rgb, a_out[:, None]], 1).reshape(height, width, 4)
    out = (img,)
    if return_mask:
        out += (cover,)
    if return_edge:
        with torch.no_grad():
            dist = (u.abs().amax(-1) - 3).abs()
            dist = torch.where(((depth >= 0) & (depth < 1))[:, None], dist, torch.full_like(dist, math.inf))
            edge = dist.amin(0) if dist.shape[0] else torch.full((height * width,), math.inf, dtype=dist.dtype)
        out += (edge.reshape(height, width),)
    return out[0] if len(out) == 1 else out
