"""Seeded synthetic workloads for the tile-pair correlation path.

Band-limited noise: N(0,1) white noise on a field with a 6 px margin, a
separable Gaussian blur (sigma 2 px, 13 taps, normalised), the margin cropped,
an affine rescale to the requested mean/sd over the field, round-half-even and
clip to the integer type.  Shifts follow the correlator's convention: lag
(x, y) means B-coordinate = A-coordinate + (x, y), i.e. content at A(x, y)
appears at B(x + sx, y + sy).

Everything is numpy/scipy with explicit seeds, so the same inputs are produced
on every machine with this image; `digest()` gives a SHA-256 per array.
"""
import hashlib

import numpy as np
from scipy import ndimage

_TAPS = None


def gauss_taps(sigma=2.0, radius=6):
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def bandlimited(rng, h, w, mean, sd, sigma=2.0):
    """Band-limited float field of h x w with the given mean and sd."""
    m = 6
    z = rng.standard_normal((h + 2 * m, w + 2 * m))
    k = gauss_taps(sigma)
    z = ndimage.correlate1d(z, k, axis=0, mode="constant")
    z = ndimage.correlate1d(z, k, axis=1, mode="constant")
    z = z[m:m + h, m:m + w]
    z = (z - z.mean()) / z.std()
    return z * sd + mean


def quantize(f, dtype):
    info = np.iinfo(dtype)
    return np.clip(np.rint(f), info.min, info.max).astype(dtype)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def bilinear(field, xs, ys):
    """Sample `field` at float coordinates (x along columns, y along rows)."""
    return ndimage.map_coordinates(field, [ys, xs], order=1, mode="nearest")


# ----------------------------------------------------------------------------
# Config 1: one 512 x 512 uint8 pair cropped from a 1024 x 1024 field.
# ----------------------------------------------------------------------------

def plain_pair(seed=1, size=512, field=1024, shift=(37, -21), noise=8.0):
    """Two size x size uint8 crops; B holds A's content translated by `shift` plus N(0, noise)."""
    rng = np.random.default_rng(seed)
    F = bandlimited(rng, field, field, 128.0, 40.0)
    sx, sy = shift
    x0 = (field - size) // 2
    y0 = (field - size) // 2
    A = F[y0:y0 + size, x0:x0 + size]
    B = F[y0 - sy:y0 - sy + size, x0 - sx:x0 - sx + size] + rng.normal(0.0, noise, (size, size))
    return quantize(A, np.uint8), quantize(B, np.uint8)


# ----------------------------------------------------------------------------
# Config 2: strip pairs with fractional shifts.
# ----------------------------------------------------------------------------

def strip_pair(seed, w=2048, h=256, kmax=40, frac=0.3, noise=10.0):
    """One w x h uint8 strip pair; true shift (kx + frac*sign, ky + frac*sign) by bilinear resampling."""
    rng = np.random.default_rng(seed)
    kx, ky = rng.integers(-kmax, kmax + 1, size=2)
    sx = kx + (frac if kx >= 0 else -frac)
    sy = ky + (frac if ky >= 0 else -frac)
    mg = kmax + 4
    F = bandlimited(rng, h + 2 * mg, w + 2 * mg, 128.0, 40.0)
    A = F[mg:mg + h, mg:mg + w]
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    B = bilinear(F, xx + mg - sx, yy + mg - sy) + rng.normal(0.0, noise, (h, w))
    return quantize(A, np.uint8), quantize(B, np.uint8), (float(sx), float(sy))


def strip_batch(n, seed0=1000, **kw):
    A, B, S = [], [], []
    for i in range(n):
        a, b, s = strip_pair(seed0 + i, **kw)
        A.append(a)
        B.append(b)
        S.append(s)
    return np.stack(A), np.stack(B), np.array(S)


# ----------------------------------------------------------------------------
# Config 4: uneven pairs with elliptical holes.
# ----------------------------------------------------------------------------

def holes_mask(rng, h, w, max_holes=3, rmin=10, rmax=80):
    m = np.ones((h, w), np.uint8)
    yy, xx = np.mgrid[0:h, 0:w]
    for _ in range(int(rng.integers(0, max_holes + 1))):
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        a, b = rng.uniform(rmin, rmax, size=2)
        th = rng.uniform(0, np.pi)
        c, s = np.cos(th), np.sin(th)
        u = (xx - cx) * c + (yy - cy) * s
        v = -(xx - cx) * s + (yy - cy) * c
        m[(u / a) ** 2 + (v / b) ** 2 <= 1.0] = 0
    return m


def uneven_pair(seed, lo=128, hi=1024, smax=50, noise=12.0):
    """One pair of equal size w x h (w, h uniform in [lo, hi]) with hole masks; integer shift."""
    rng = np.random.default_rng(seed)
    w, h = (int(v) for v in rng.integers(lo, hi + 1, size=2))
    sx, sy = (int(v) for v in rng.integers(-smax, smax + 1, size=2))
    mg = smax + 2
    F = bandlimited(rng, h + 2 * mg, w + 2 * mg, 128.0, 40.0)
    A = F[mg:mg + h, mg:mg + w]
    B = F[mg - sy:mg - sy + h, mg - sx:mg - sx + w] + rng.normal(0.0, noise, (h, w))
    return (quantize(A, np.uint8), holes_mask(rng, h, w), quantize(B, np.uint8), holes_mask(rng, h, w),
            (sx, sy))


# ----------------------------------------------------------------------------
# Point-list views (the reference's vector<Point> + vector<double> form).
# ----------------------------------------------------------------------------

def raster_points(img, mask=None, x0=0, y0=0):
    """Point list (x, y) and values of the present pixels of `img`, in raster order."""
    h, w = img.shape
    yy, xx = np.mgrid[0:h, 0:w]
    sel = np.ones((h, w), bool) if mask is None else mask.astype(bool)
    pts = np.stack([xx[sel] + x0, yy[sel] + y0], axis=1).astype(np.float64)
    return pts, img[sel].astype(np.float64)


def normalize(v):
    """Normalize (Maths.cpp:192-212): mean 0, sd 1 with the n-1 deviation."""
    v = np.asarray(v, dtype=np.float64)
    n = v.size
    s, s2 = v.sum(), (v * v).sum()
    avg = s / n
    sd = np.sqrt(max(s2 - n * avg * avg, 0.0) / (n - 1.0))
    return (v - avg) / sd


# ----------------------------------------------------------------------------
# Config 3: thumbnail rotation sweeps.
# ----------------------------------------------------------------------------

def sweep_angles(lo=-5.0, hi=5.0, step=0.5):
    """Angle list built as CThmScan builds it: for (a = lo; a <= hi; a += step) (CThmScan.cpp:540)."""
    out, a = [], lo
    while a <= hi:
        out.append(a)
        a += step
    return out


def rotated_pair(seed, size=96, margin=40, deg=2.0, shift=(3.0, -2.0), mean=0.0, sd=1.0):
    """A: size x size crop of a band-limited field; B: the field rotated by `deg` about the crop centre and
    shifted by `shift` (bilinear), same crop.  Values as float64 (thumbnail-like, not quantised)."""
    rng = np.random.default_rng(seed)
    n = size + 2 * margin
    F = bandlimited(rng, n, n, mean, sd)
    c0 = n / 2.0
    yy, xx = np.mgrid[0:size, 0:size].astype(np.float64)
    X, Y = xx + margin, yy + margin
    th = np.deg2rad(deg)
    # content at A-position p appears at B-position R(th) (p - c) + c + shift
    u, v = X - c0 - shift[0], Y - c0 - shift[1]
    xs = np.cos(th) * u + np.sin(th) * v + c0
    ys = -np.sin(th) * u + np.cos(th) * v + c0
    A = F[margin:margin + size, margin:margin + size].copy()
    B = bilinear(F, xs, ys)
    return A, B
