"""FP64 autodiff of the brute-force splat renderer: the error oracle of the
backward. Test infrastructure, not product: only tests/, the smoke check and
the benchmark's CPU baseline import it, and only as the checker.

The forward restates lrt_oracle.c's or_diff_render_f64 in torch float64 ops,
so torch.autograd gives its exact FP64 gradient:

- projection: reference_project.slang:31-105 (EWA with the Jacobian clamped to
  1.3x the frustum, the dilation and its opacity compensation, the conic, the
  cutoff min(2 ln(max(255 alpha, 1.0001)), 9)) from FP32 parameters: scale
  exp(log s), opacity sigmoid(logit), unit quaternion q / |q|, colour
  max(0.5 + C0 dc + rest(dir), 0) with sh.slang:25-50's basis;
- walk: reference_blend.slang:24-119 over every splat in the rasteriser's
  keyed order (bits(float z) >> 7, then index): skip past the cutoff, alpha =
  min(0.99, opacity exp(-power / 2)), skip below 1/255, stop once T < 1e-4,
  then linearise and add the background.

The skip and stop tests and the order are piecewise constant, so they are
taken on detached values; the clamp at 0.99 passes no gradient, as the
rasteriser's backward.
"""
from __future__ import annotations

import numpy as np
import torch

C0 = 0.28209479177387814
C1 = 0.4886025119029199
C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396)
C3 = (-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
      1.445305721320277, -0.5900435899266435)


def sh_basis(d: torch.Tensor, degree: int) -> torch.Tensor:
    """The rest basis (sh.slang:25-50), [N, 15], zero past the degree."""
    x, y, z = d.unbind(1)
    zero = torch.zeros_like(x)
    b = [zero] * 15
    if degree >= 1:
        b[0], b[1], b[2] = -C1 * y, C1 * z, -C1 * x
    if degree >= 2:
        xx, yy, zz = x * x, y * y, z * z
        b[3], b[4] = C2[0] * x * y, C2[1] * y * z
        b[5] = C2[2] * (2.0 * zz - xx - yy)
        b[6], b[7] = C2[3] * x * z, C2[4] * (xx - yy)
    if degree >= 3:
        b[8] = C3[0] * y * (3.0 * xx - yy)
        b[9] = C3[1] * x * y * z
        b[10] = C3[2] * y * (4.0 * zz - xx - yy)
        b[11] = C3[3] * z * (2.0 * zz - 3.0 * xx - 3.0 * yy)
        b[12] = C3[4] * x * (4.0 * zz - xx - yy)
        b[13] = C3[5] * z * (xx - yy)
        b[14] = C3[6] * x * (xx - 3.0 * yy)
    return torch.stack(b, dim=1)


def degree_of(rest: int) -> int:
    return 3 if rest >= 15 else 2 if rest >= 8 else 1 if rest >= 3 else 0


def project(frame, means, log_scales, quats, logits, sh, filter3d=None):
    """Per-splat screen records in FP64: (valid, cx, cy, z, cutoff, a, b, c, alpha, colour[N, 3])."""
    d = torch.float64
    m = torch.tensor(list(frame.m), dtype=d).reshape(3, 4)
    wm, w0 = m[:, :3], m[:, 3]
    p = means.to(d)
    s = torch.exp(log_scales.to(d))
    op = torch.sigmoid(logits.to(d))
    if frame.use_filter3d:
        f2 = filter3d.to(d)[:, None] ** 2
        s2 = s * s
        op = op * torch.sqrt(s2.prod(1) / (s2 + f2).prod(1))
        s = torch.sqrt(s2 + f2)
    q = quats.to(d)
    q = q / q.norm(dim=1, keepdim=True)
    x, y, z, w = q.unbind(1)
    r = torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                     2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                     2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], dim=1).reshape(-1, 3, 3)
    mm = r * s[:, None, :]
    v = p @ wm.T + w0
    tz = v[:, 2]
    valid = (tz > frame.near_z) & (tz <= frame.far_z)
    fx, fy, ccx, ccy = float(frame.focal_x), float(frame.focal_y), float(frame.centre_x), float(frame.centre_y)
    if frame.orthographic:
        cx = fx * v[:, 0] + ccx
        cy = fy * v[:, 1] + ccy
        zero = torch.zeros_like(tz)
        j00, j02, j11, j12 = zero + fx, zero, zero + fy, zero
    else:
        safe = torch.where(valid, tz, torch.ones_like(tz))
        cx = fx * v[:, 0] / safe + ccx
        cy = fy * v[:, 1] / safe + ccy
        lox = (-0.15 * frame.width - ccx) / fx
        hix = (1.15 * frame.width - ccx) / fx
        loy = (-0.15 * frame.height - ccy) / fy
        hiy = (1.15 * frame.height - ccy) / fy
        tx = torch.clamp(v[:, 0] / safe, lox, hix)
        ty = torch.clamp(v[:, 1] / safe, loy, hiy)
        j00 = fx / safe
        j02 = -fx * (tx * safe) / (safe * safe)
        j11 = fy / safe
        j12 = -fy * (ty * safe) / (safe * safe)
    t3x = j00[:, None] * wm[0][None, :] + j02[:, None] * wm[2][None, :]
    t3y = j11[:, None] * wm[1][None, :] + j12[:, None] * wm[2][None, :]
    ax = torch.einsum("ni,nij->nj", t3x, mm)
    ay = torch.einsum("ni,nij->nj", t3y, mm)
    a0, b, c0 = (ax * ax).sum(1), (ax * ay).sum(1), (ay * ay).sum(1)
    before = a0 * c0 - b * b
    a1, c1 = a0 + float(frame.dilation), c0 + float(frame.dilation)
    det = a1 * c1 - b * b
    valid = valid & (det > 1e-12)
    det_safe = torch.where(valid, det, torch.ones_like(det))
    alpha = op * torch.sqrt(torch.clamp(before, min=0.0) / det_safe) if frame.antialias else op
    valid = valid & (alpha >= 1.0 / 255.0)
    with torch.no_grad():
        cutoff = torch.clamp(2.0 * torch.log(torch.clamp(255.0 * alpha, min=1.0001)), max=9.0)
    eye = torch.tensor(list(frame.eye), dtype=d)
    rel = p - eye
    dirs = rel / rel.norm(dim=1, keepdim=True)
    rest = int(sh.shape[1]) - 1
    degree = min(degree_of(rest), int(frame.sh_limit))
    shd = sh.to(d)
    coeff = torch.zeros((shd.shape[0], 15, 3), dtype=d)
    if rest:
        coeff = torch.cat([shd[:, 1:, :], torch.zeros((shd.shape[0], 15 - rest, 3), dtype=d)], dim=1)
    col = 0.5 + C0 * shd[:, 0, :] + torch.einsum("nb,nbc->nc", sh_basis(dirs, degree), coeff)
    col = torch.clamp(col, min=0.0)
    return dict(valid=valid, cx=cx, cy=cy, z=tz, cutoff=cutoff, a=c1 / det_safe, b=-b / det_safe,
                c=a1 / det_safe, alpha=alpha, colour=col)


def keyed_order(rec) -> torch.Tensor:
    """Visible splats in the rasteriser's order: bits(float z) >> 7, then index."""
    valid = rec["valid"].numpy()
    idx = np.nonzero(valid)[0]
    key = rec["z"].detach().numpy()[idx].astype(np.float32).view(np.uint32) >> 7
    order = idx[np.lexsort((idx, key))]
    return torch.from_numpy(order.astype(np.int64))


def _srgb_to_linear(x: torch.Tensor) -> torch.Tensor:
    x = torch.clamp(x, min=0.0)
    lo = x / 12.92
    hi = torch.pow((torch.clamp(x, min=0.04045) + 0.055) / 1.055, 2.4)
    return torch.where(x <= 0.04045, lo, hi)


def shade(frame, rec, order, xs: torch.Tensor, ys: torch.Tensor) -> torch.Tensor:
    """Colour [P, 4] of pixels (xs, ys) (integer coordinates, bottom row first)."""
    d = torch.float64
    px = xs.to(d)[:, None] + 0.5
    py = ys.to(d)[:, None] + 0.5
    dx = px - rec["cx"][order][None, :]
    dy = py - rec["cy"][order][None, :]
    power = rec["a"][order] * dx * dx + 2.0 * rec["b"][order] * dx * dy + rec["c"][order] * dy * dy
    a_pix = torch.clamp(rec["alpha"][order][None, :] * torch.exp(-0.5 * power), max=0.99)
    with torch.no_grad():
        ok = (power <= rec["cutoff"][order][None, :]) & (a_pix >= 1.0 / 255.0)
    a_eff = torch.where(ok, a_pix, torch.zeros_like(a_pix))
    with torch.no_grad():
        ones = torch.ones((a_eff.shape[0], 1), dtype=d)
        t_before = torch.cat([ones, torch.cumprod(1.0 - a_eff, dim=1)[:, :-1]], dim=1)
        processed = t_before >= 1e-4
    a_proc = torch.where(processed, a_eff, torch.zeros_like(a_eff))
    one = 1.0 - a_proc
    ones = torch.ones((one.shape[0], 1), dtype=d)
    tb = torch.cat([ones, torch.cumprod(one, dim=1)[:, :-1]], dim=1)
    weight = tb * a_proc
    rgb = weight @ rec["colour"][order]
    t = torch.prod(one, dim=1)
    coverage = 1.0 - t
    if frame.linearise:
        cov = torch.where(coverage > 1e-6, coverage, torch.ones_like(coverage))[:, None]
        lin = _srgb_to_linear(rgb / cov) * cov
        rgb = torch.where((coverage > 1e-6)[:, None], lin, rgb)
    bg = torch.tensor(list(frame.bg), dtype=d)
    return torch.cat([rgb + t[:, None] * bg[None, :3], (coverage + t * bg[3])[:, None]], dim=1)


def render(frame, means, log_scales, quats, logits, sh, filter3d=None, chunk: int = 2048) -> torch.Tensor:
    """The FP64 image [H, W, 4], bottom row first."""
    with torch.no_grad():
        rec = project(frame, means, log_scales, quats, logits, sh, filter3d)
        order = keyed_order(rec)
        ys, xs = torch.meshgrid(torch.arange(frame.height), torch.arange(frame.width), indexing="ij")
        xs, ys = xs.reshape(-1), ys.reshape(-1)
        out = [shade(frame, rec, order, xs[k:k + chunk], ys[k:k + chunk]) for k in range(0, xs.numel(), chunk)]
    return torch.cat(out).reshape(frame.height, frame.width, 4)


def decisions(frame, params: dict, filter3d=None, chunk: int = 2048):
    """The walk's discrete choices for float32 parameters: (the visible splats'
    keyed order, a [pixels, visible] bool of the pairs that pass the cutoff and
    alpha-floor tests and come before the T < 1e-4 stop). The image is smooth
    in the parameters wherever these stay the same; a finite difference whose
    two ends differ here straddles a step of the forward."""
    names = ("means", "log_scales", "quats", "opacity_logits", "sh")
    with torch.no_grad():
        t = [torch.from_numpy(np.asarray(params[k], dtype=np.float32)) for k in names]
        f3 = torch.from_numpy(np.asarray(filter3d, dtype=np.float32)) if filter3d is not None else None
        rec = project(frame, *t, f3)
        order = keyed_order(rec)
        ys, xs = torch.meshgrid(torch.arange(frame.height), torch.arange(frame.width), indexing="ij")
        xs, ys = xs.reshape(-1), ys.reshape(-1)
        masks = []
        for k in range(0, xs.numel(), chunk):
            px = xs[k:k + chunk].to(torch.float64)[:, None] + 0.5
            py = ys[k:k + chunk].to(torch.float64)[:, None] + 0.5
            dx = px - rec["cx"][order][None, :]
            dy = py - rec["cy"][order][None, :]
            power = rec["a"][order] * dx * dx + 2.0 * rec["b"][order] * dx * dy + rec["c"][order] * dy * dy
            a_pix = torch.clamp(rec["alpha"][order][None, :] * torch.exp(-0.5 * power), max=0.99)
            ok = (power <= rec["cutoff"][order][None, :]) & (a_pix >= 1.0 / 255.0)
            a_eff = torch.where(ok, a_pix, torch.zeros_like(a_pix))
            ones = torch.ones((a_eff.shape[0], 1), dtype=torch.float64)
            t_before = torch.cat([ones, torch.cumprod(1.0 - a_eff, dim=1)[:, :-1]], dim=1)
            masks.append((ok & (t_before >= 1e-4)).numpy())
    return order.numpy(), np.concatenate(masks)


def same_decisions(a, b) -> bool:
    return a[0].shape == b[0].shape and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def gradients(frame, params: dict, grad_colour: np.ndarray, filter3d=None, chunk: int = 1024) -> dict:
    """d(sum(colour * grad_colour))/d(parameter) in FP64 by autodiff, per parameter name.

    params: float32 numpy arrays means, log_scales, quats, opacity_logits, sh.
    """
    names = ("means", "log_scales", "quats", "opacity_logits", "sh")
    leaves = {k: torch.tensor(np.asarray(params[k], dtype=np.float64), requires_grad=True) for k in names}
    g = torch.from_numpy(np.asarray(grad_colour, dtype=np.float64).reshape(-1, 4))
    f3 = torch.from_numpy(np.asarray(filter3d, dtype=np.float64)) if filter3d is not None else None
    ys, xs = torch.meshgrid(torch.arange(frame.height), torch.arange(frame.width), indexing="ij")
    xs, ys = xs.reshape(-1), ys.reshape(-1)
    order = None
    for k in range(0, xs.numel(), chunk):
        rec = project(frame, leaves["means"], leaves["log_scales"], leaves["quats"], leaves["opacity_logits"],
                      leaves["sh"], f3)
        if order is None:
            order = keyed_order(rec)
        out = shade(frame, rec, order, xs[k:k + chunk], ys[k:k + chunk])
        (out * g[k:k + chunk]).sum().backward()
    return {k: leaves[k].grad.numpy() for k in names}
