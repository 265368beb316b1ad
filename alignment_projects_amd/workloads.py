"""Seeded synthetic workloads of the benchmark, generated on the device.

The band-limited noise is the one of `synth` (white noise, 13-tap Gaussian
blur with sigma 2 px on a field with a margin, affine rescale to the stated
mean and sd, round-half-even, clip), computed with torch on the GPU so that
the larger workloads (thousands of tiles up to 1024 x 1024, 2048 x 2048
fields) are produced in seconds.  Every generator takes an explicit seed and
uses its own torch.Generator, so the same inputs come back on every run on
the same device type.

Tiles are returned as device tensors (uint8 / uint16 / float32 and uint8
masks) that the caller adopts with `Context.tile_put(..., device_ptr=...)`;
keep the tensors alive while the tiles are in use.
"""
import math

import numpy as np
import torch
import torch.nn.functional as F

_DEV = "cuda"


def _gen(seed):
    g = torch.Generator(device=_DEV)
    g.manual_seed(int(seed))
    return g


def _taps(sigma=2.0, radius=6):
    x = torch.arange(-radius, radius + 1, dtype=torch.float64, device=_DEV)
    k = torch.exp(-0.5 * (x / sigma) ** 2)
    return (k / k.sum()).to(torch.float32)


def field(gen, h, w, mean, sd):
    """Band-limited float32 field of h x w on the device."""
    m = 6
    z = torch.randn((1, 1, h + 2 * m, w + 2 * m), generator=gen, device=_DEV, dtype=torch.float32)
    k = _taps()
    z = F.conv2d(z, k.view(1, 1, 1, -1))
    z = F.conv2d(z, k.view(1, 1, -1, 1))[0, 0]
    z = (z - z.mean()) / z.std()
    return z * sd + mean


def quantize(t, hi):
    return torch.clamp(torch.round(t), 0, hi)


def _bilinear(F2, xs, ys):
    """Sample F2 (H x W) at float pixel coordinates (x along columns)."""
    H, W = F2.shape
    gx = 2.0 * xs / (W - 1) - 1.0
    gy = 2.0 * ys / (H - 1) - 1.0
    grid = torch.stack([gx, gy], dim=-1)[None].to(torch.float32)
    return F.grid_sample(F2[None, None], grid, mode="bilinear", padding_mode="border", align_corners=True)[0, 0]


def _holes(gen, h, w, max_holes=3, rmin=10.0, rmax=80.0):
    """Overlap-rectangle mask minus 0..max_holes elliptical holes (uint8 on the device)."""
    m = torch.ones((h, w), dtype=torch.uint8, device=_DEV)
    k = int(torch.randint(0, max_holes + 1, (1,), generator=gen, device=_DEV).item())
    if k == 0:
        return m
    yy = torch.arange(h, device=_DEV, dtype=torch.float32)[:, None]
    xx = torch.arange(w, device=_DEV, dtype=torch.float32)[None, :]
    p = torch.rand((k, 5), generator=gen, device=_DEV)
    for i in range(k):
        cx, cy = float(p[i, 0]) * w, float(p[i, 1]) * h
        a, b = rmin + float(p[i, 2]) * (rmax - rmin), rmin + float(p[i, 3]) * (rmax - rmin)
        th = float(p[i, 4]) * math.pi
        c, s = math.cos(th), math.sin(th)
        u = (xx - cx) * c + (yy - cy) * s
        v = -(xx - cx) * s + (yy - cy) * c
        m[(u / a) ** 2 + (v / b) ** 2 <= 1.0] = 0
    return m


# ----------------------------------------------------------------------------
# Cross-layer thumbnail rotation sweep (CThmScan): 2048 x 2048 uint16 pairs,
# B rotated by U[-2.5, 2.5] deg and shifted by U[-60, 60] px, both decimated
# 8x (DecimateVector's box means) to 256 x 256 and normalised by the library's
# thumbnail kernel (MakeThumbs on the device).
# ----------------------------------------------------------------------------

def thumb_pairs(ctx, n, seed0=3000, size=2048, dec=8, mean=30000.0, sd=6000.0, max_deg=2.5, max_shift=60.0):
    """Returns (A list, B list) of float64 numpy thumbnails and the true (deg, sx, sy) per pair.  ctx (an
    xcorr.Context) makes the thumbnails from the 16-bit pictures on the device."""
    from . import abi
    A, B, truth = [], [], []
    mg = 160
    n_f = size + 2 * mg
    c0 = n_f / 2.0
    yy, xx = torch.meshgrid(torch.arange(size, device=_DEV, dtype=torch.float64),
                            torch.arange(size, device=_DEV, dtype=torch.float64), indexing="ij")
    for i in range(n):
        g = _gen(seed0 + i)
        Fd = field(g, n_f, n_f, mean, sd)
        p = torch.rand(3, generator=g, device=_DEV, dtype=torch.float64)
        deg = float((2 * p[0] - 1) * max_deg)
        sx, sy = float((2 * p[1] - 1) * max_shift), float((2 * p[2] - 1) * max_shift)
        th = math.radians(deg)
        # content at A position q appears at B position R(th)(q - c) + c + s
        u, v = xx + mg - c0 - sx, yy + mg - c0 - sy
        xs = math.cos(th) * u + math.sin(th) * v + c0
        ys = -math.sin(th) * u + math.cos(th) * v + c0
        a16 = quantize(Fd[mg:mg + size, mg:mg + size], 65535)
        b16 = quantize(_bilinear(Fd, xs, ys), 65535)
        out = []
        for im in (a16, b16):
            pic = im.to(torch.float32).contiguous()  # 16-bit values, exact in FP32
            t = torch.empty((size // dec, size // dec), dtype=torch.float64, device=_DEV)
            torch.cuda.synchronize()  # the library works on its own stream: pic must be written first
            ctx.thumb_raster_dev(pic.data_ptr(), abi.XC_F32, size, size, size, dec, t.data_ptr())
            out.append(t.cpu().numpy())
        A.append(out[0])
        B.append(out[1])
        truth.append((deg, sx / dec, sy / dec))
    return A, B, truth


# ----------------------------------------------------------------------------
# Uneven batch with irregular masks: w, h ~ U{128..1024}, equal per pair,
# uint8, shift U{-50..50}^2, N(0, 12) on B, 0-3 elliptical holes per mask.
# ----------------------------------------------------------------------------

def _pitched(t):
    """t (h x w, uint8 on the device) as a view of an h x P tensor, P the next multiple of 16 bytes: the row
    pitch the library's table sweep reads fastest (pass t.stride(0) as the tile's stride)."""
    h, w = t.shape
    p = torch.zeros((h, (w + 15) // 16 * 16), dtype=t.dtype, device=t.device)
    p[:, :w] = t
    return p[:, :w]


def uneven_pairs(n, seed0=4000, lo=128, hi=1024, smax=50, noise=12.0, fsize=6144):
    """Returns a list of (A, mA, B, mB, (sx, sy)) device tensors, rows 16-byte pitched (A and its mask share
    one stride, as do B and its mask).  Pairs are cut from shared fields (one per 64 pairs) at random places."""
    out = []
    g = _gen(seed0)
    Fd = None
    for i in range(n):
        if i % 64 == 0:
            Fd = field(g, fsize, fsize, 128.0, 40.0)
        w, h = (int(v) for v in torch.randint(lo, hi + 1, (2,), generator=g, device=_DEV))
        sx, sy = (int(v) for v in torch.randint(-smax, smax + 1, (2,), generator=g, device=_DEV))
        x0 = int(torch.randint(smax + 1, fsize - w - smax - 1, (1,), generator=g, device=_DEV))
        y0 = int(torch.randint(smax + 1, fsize - h - smax - 1, (1,), generator=g, device=_DEV))
        a = quantize(Fd[y0:y0 + h, x0:x0 + w], 255).to(torch.uint8).contiguous()
        nb = torch.randn((h, w), generator=g, device=_DEV) * noise
        b = quantize(Fd[y0 - sy:y0 - sy + h, x0 - sx:x0 - sx + w] + nb, 255).to(torch.uint8).contiguous()
        out.append((_pitched(a), _pitched(_holes(g, h, w)), _pitched(b), _pitched(_holes(g, h, w)), (sx, sy)))
    return out


def uneven_pair_sized(seed, w, h, smax=50, noise=12.0):
    """One pair of the uneven batch's kind at a chosen size (w x h), with its own field: (A, mA, B, mB,
    (sx, sy)) device tensors as uneven_pairs gives them."""
    g = _gen(seed)
    mg = smax + 8
    Fd = field(g, h + 2 * mg, w + 2 * mg, 128.0, 40.0)
    sx, sy = (int(v) for v in torch.randint(-smax, smax + 1, (2,), generator=g, device=_DEV))
    a = quantize(Fd[mg:mg + h, mg:mg + w], 255).to(torch.uint8).contiguous()
    nb = torch.randn((h, w), generator=g, device=_DEV) * noise
    b = quantize(Fd[mg - sy:mg - sy + h, mg - sx:mg - sx + w] + nb, 255).to(torch.uint8).contiguous()
    return a, _holes(g, h, w), b, _holes(g, h, w), (sx, sy)


# ----------------------------------------------------------------------------
# Serving stream: tiles on small tori of neighbours (each tile referenced by
# four requests: its right, left, lower and upper neighbour pairs).
# ----------------------------------------------------------------------------

def lattice(seed, w, h, k=4, jitter=25, noise=8.0):
    """k x k tiles of w x h cut from one field at jittered offsets (uint8 device tensors) and the
    requests (i, j) of right and down neighbours on the torus, with their true shifts."""
    g = _gen(seed)
    mg = jitter + 2
    Fd = field(g, h + 2 * mg, w + 2 * mg, 128.0, 40.0)
    off = torch.randint(-jitter, jitter + 1, (k * k, 2), generator=g, device=_DEV).cpu().numpy()
    tiles = []
    for t in range(k * k):
        ox, oy = int(off[t, 0]), int(off[t, 1])
        crop = Fd[mg + oy:mg + oy + h, mg + ox:mg + ox + w] + torch.randn((h, w), generator=g, device=_DEV) * noise
        tiles.append(quantize(crop, 255).to(torch.uint8).contiguous())
    reqs = []
    for r in range(k):
        for c in range(k):
            t = r * k + c
            for u in (r * k + (c + 1) % k, ((r + 1) % k) * k + c):
                # content at A(p) sits at field p + off_a; in B at p + off_a - off_b
                reqs.append((t, u, (int(off[t, 0] - off[u, 0]), int(off[t, 1] - off[u, 1]))))
    return tiles, reqs


def poisson_times(rate, n, seed):
    """Arrival times (s) of n requests at `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))
