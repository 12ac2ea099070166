"""
The 3D-astigmatism workload generated on the GPU with torch (same model and
camera as synth.frame_z, torch's generator instead of numpy's), so that
hundreds of 512 x 512 frames are ready in seconds and already resident in HBM.
"""
import math

import torch

from .synth import WX_Z, WY_Z


def _width(z, p):
    X = (z - p[1]) / p[2]
    return p[0] * torch.sqrt(1.0 + X * X + p[3] * X ** 3 + p[4] * X ** 4)


def frames_z(n_frames, size=512, n=1000, seed=20261204, device="cuda:0", chunk=25, roi=18):
    """
    Returns (images (F, size, size) float64, background (F, size, size)
    float64, rows (F*n, 7) float64 "hdf5" start peaks, counts (F,) int32,
    truth dict of (F, n) tensors).
    """
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    m = roi // 2 + 2
    f64 = dict(dtype=torch.float64, device=device)
    x = m + (size - 2 * m) * torch.rand((n_frames, n), generator=g, **f64)
    y = m + (size - 2 * m) * torch.rand((n_frames, n), generator=g, **f64)
    z = -0.5 + torch.rand((n_frames, n), generator=g, **f64)
    h = torch.exp(math.log(1500.0) + 0.5 * torch.randn((n_frames, n), generator=g, **f64))
    sx = 0.5 * _width(z, WX_Z)
    sy = 0.5 * _width(z, WY_Z)

    jj = torch.arange(size, **f64)
    bg1 = 125.0 + 75.0 * torch.sin(2 * math.pi * jj / size)[None, :] * torch.cos(2 * math.pi * jj / size)[:, None]
    background = bg1.expand(n_frames, size, size).contiguous()

    half = int(math.ceil(5.0 * float(max(sx.max(), sy.max())))) + 1
    k = torch.arange(-half, half + 1, device=device)
    expected = background.clone().reshape(-1)
    for f0 in range(0, n_frames, chunk):
        f1 = min(n_frames, f0 + chunk)
        xs, ys = x[f0:f1], y[f0:f1]
        cx, cy = torch.floor(xs).long(), torch.floor(ys).long()
        cols = cx[..., None] + k                          # (c, n, K)
        rows = cy[..., None] + k
        gx = torch.exp(-((cols - xs[..., None]) ** 2) / (2.0 * sx[f0:f1, :, None] ** 2))
        gy = torch.exp(-((rows - ys[..., None]) ** 2) / (2.0 * sy[f0:f1, :, None] ** 2))
        vals = h[f0:f1, :, None, None] * gy[..., :, None] * gx[..., None, :]
        rr = rows[..., :, None].expand_as(vals)
        cc = cols[..., None, :].expand_as(vals)
        fr = torch.arange(f0, f1, device=device)[:, None, None, None].expand_as(vals)
        ok = (rr >= 0) & (rr < size) & (cc >= 0) & (cc < size)
        idx = (fr * size + rr) * size + cc
        expected.index_add_(0, idx[ok], vals[ok])
    expected = expected.reshape(n_frames, size, size)
    photons = torch.poisson(expected, generator=g)
    images = torch.clamp(photons, min=1.0)          # offset 100 ADU, gain 1: (ADU - 100) / 1, >= 1

    xs = x + (2.0 * torch.rand((n_frames, n), generator=g, **f64) - 1.0)
    ys = y + (2.0 * torch.rand((n_frames, n), generator=g, **f64) - 1.0)
    ci = torch.clamp(torch.round(xs).long(), 1, size - 2)
    ri = torch.clamp(torch.round(ys).long(), 1, size - 2)
    fidx = torch.arange(n_frames, device=device)[:, None].expand_as(ci)
    best = torch.full_like(xs, -float("inf"))
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            best = torch.maximum(best, images[fidx, ri + di, ci + dj])
    bg_at = 125.0 + 75.0 * torch.sin(2 * math.pi * x / size) * torch.cos(2 * math.pi * y / size)
    hgt = torch.clamp(best - bg_at, min=1.0)
    s0 = torch.full_like(xs, 1.5)
    rows7 = torch.stack((xs, ys, torch.zeros_like(xs), bg_at, hgt, s0, s0), dim=-1).reshape(-1, 7).contiguous()
    counts = torch.full((n_frames,), n, dtype=torch.int32)
    truth = {"x": x, "y": y, "z": z, "height": h}
    return images.contiguous(), background, rows7, counts, truth
