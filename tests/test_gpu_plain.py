"""Plain FP64 form on the GPU against the CPU oracle (CorrPatches / CorrImagesR / CorrImagesS).

Bar: integer peak lag and status identical, r within 1e-9 relative, dx/dy within 1e-9 px.
"""
import numpy as np
import pytest

import oracle
from alignment_projects_amd import abi, synth

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def _same(g, o, what, ties=None):
    """GPU result g against oracle result o.  `ties` (set of signed lags) lists the lags whose
    oracle value lies within RTOL of the maximum: a near-tie is decided by FFT rounding, so the
    GPU may pick any of them (then r must still agree and dx, dy are not compared)."""
    gst, gr, gdx, gdy, gres = g
    ost, orr, odx, ody, ores = o
    assert gst == ost, f"{what}: status {gst} vs {ost}"
    assert (gres.nx, gres.ny) == (ores.nx, ores.ny)
    assert abs(gr - orr) <= RTOL * max(1.0, abs(orr)), f"{what}: r {gr!r} vs {orr!r}"
    glag, olag = (gres.lag_x, gres.lag_y), (ores.lag_x, ores.lag_y)
    if ties is not None and len(ties) > 1:
        assert glag in ties, f"{what}: lag {glag} not among near-ties {sorted(ties)[:8]}"
        return
    assert glag == olag, f"{what}: lag {glag} vs {olag}"
    assert abs(gdx - odx) <= 1e-9 and abs(gdy - ody) <= 1e-9, f"{what}: d {(gdx, gdy)} vs {(odx, ody)}"


def _ncc_ties(rslt, best, res, nnegx, nnegy):
    """Signed lags of the oracle's scanned R image within RTOL of the maximum."""
    Ny, Nx = rslt.shape
    iy, ix = np.nonzero(rslt >= best - RTOL * max(1.0, abs(best)))
    lx = np.where(ix >= Nx - nnegx, ix - Nx, ix)
    ly = np.where(iy >= Ny - nnegy, iy - Ny, iy)
    return set(zip(lx.tolist(), ly.tolist()))


def test_testnormcorr_kat(ctx):
    # Retired_Proj/4_TestNormCorr/testnormcorr.cpp:20-63: exact match at dx=-4, dy=+2, r=1
    pb = [(0, 3), (2, 3), (3, 2), (3, 3), (3, 4), (4, 3), (5, 4)]
    vb = [1, 2, 1, 3, 1, 2, 1]
    pa = [(7, 2), (8, 2), (9, 2), (7, 1), (8, 1)]
    va = [1, 0, 1, 3, 2]
    g = ctx.corr_patches(pa, va, pb, vb, 0, 0, 4096, rgn=abi.rgn_some_area())
    o = oracle.corr_patches(pa, va, pb, vb, 0, 0, 4096, rgn=abi.rgn_some_area())
    _same(g, o, "testnormcorr")
    assert round(g[2]) == -4 and round(g[3]) == 2 and abs(g[1] - 1.0) < 1e-12


@pytest.mark.parametrize("seed", range(24))
def test_corr_patches_random(ctx, seed):
    rng = np.random.default_rng(seed)
    w1, h1, w2, h2 = (int(v) for v in rng.integers(2, 48, size=4))
    A = rng.standard_normal((h1, w1))
    B = rng.standard_normal((h2, w2))
    pa, va = synth.raster_points(A, x0=int(rng.integers(-5, 5)), y0=int(rng.integers(-5, 5)))
    pb, vb = synth.raster_points(B, x0=int(rng.integers(-5, 5)), y0=int(rng.integers(-5, 5)))
    rules = [(None, None), (abi.rgn_some_area(), None), (abi.rgn_min(3, 3, 12), abi.cnt_enough_points(5))]
    rgn, cnt = rules[seed % 3]
    radius = [4096, 10, 25][seed % 3]
    g = ctx.corr_patches(pa, va, pb, vb, 1, -1, radius, rgn=rgn, cnt=cnt)
    o = oracle.corr_patches(pa, va, pb, vb, 1, -1, radius, rgn=rgn, cnt=cnt, want_rslt=True)
    ties = _ncc_ties(o[5], o[1], o[4], w1 - 1, h1 - 1) if o[0] == abi.XC_OK else None
    _same(g, o[:5], f"seed {seed} {(w1, h1, w2, h2)}", ties)


def test_corr_patches_no_peak_leaves_dxdy(ctx):
    pa, va = synth.raster_points(np.arange(12.0).reshape(3, 4))
    pb, vb = synth.raster_points(np.arange(20.0).reshape(4, 5))
    g = ctx.corr_patches(pa, va, pb, vb, rgn=abi.rgn_exact(99, 99), dx0=7.5, dy0=-3.25)
    assert g[0] == abi.XC_NO_PEAK and g[1] == 0.0 and (g[2], g[3]) == (7.5, -3.25)


def test_cfg1_plain_pair(ctx):
    A, B = synth.plain_pair(seed=1)
    pa, va = synth.raster_points(A)
    pb, vb = synth.raster_points(B)
    # With LegalRgn = NULL the scan's maximum is a 1 x 2 corner overlap (r = 1 for any two
    # pixels); a BigEnough-style rule (CThmScan.cpp:28-33) restricts it to real overlaps.
    rgn = abi.rgn_big_enough(128, 512 * 512 // 4)
    g = ctx.corr_patches(pa, va, pb, vb, 0, 0, 4096, rgn=rgn)
    o = oracle.corr_patches(pa, va, pb, vb, 0, 0, 4096, rgn=rgn)
    _same(g, o, "cfg1")
    assert (g[4].lag_x, g[4].lag_y) == (37, -21)


@pytest.mark.parametrize("mode", [abi.XC_MODE_R, abi.XC_MODE_S])
@pytest.mark.parametrize("seed", range(8))
def test_corr_images_random(ctx, mode, seed):
    rng = np.random.default_rng(100 + seed)
    w1, h1, w2, h2 = (int(v) for v in rng.integers(7, 60, size=4))
    A = rng.standard_normal((h1, w1))
    B = rng.standard_normal((h2, w2))
    pa, va = synth.raster_points(A)
    pb, vb = synth.raster_points(B, x0=3, y0=-2)
    req = int(0.25 * min(w1 * h1, w2 * h2))
    rgn = abi.rgn_big_enough(4, req)
    cnt = abi.cnt_enough_points(req)
    Rx, Ry = (-1, -1) if seed % 2 == 0 else (12, 7)
    g = ctx.corr_images(mode, pa, va, pb, vb, rgn=rgn, cnt=cnt, Ox=1, Oy=2, Rx=Rx, Ry=Ry)
    o = oracle.corr_images(mode, pa, va, pb, vb, rgn=rgn, cnt=cnt, Ox=1, Oy=2, Rx=Rx, Ry=Ry, want_images=True)
    ties = None
    if o[0] == abi.XC_OK:
        F = o[6] if mode == abi.XC_MODE_S else o[5]
        A = o[7].astype(bool)
        peak = F[A].max()
        yy, xx = np.nonzero(A & (F >= peak - RTOL * max(1.0, abs(peak))))
        ties = set(zip((xx - (w1 - 1)).tolist(), (yy - (h1 - 1)).tolist()))
    _same(g, o[:5], f"mode {mode} seed {seed}", ties)


@pytest.mark.parametrize("shape", [(1, 2), (8, 16), (64, 32), (512, 1024), (256, 4096)])
def test_fft_seam_vs_oracle_and_numpy(ctx, shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape)
    f = ctx.fft_2d(a)
    ref = np.fft.rfft2(a)
    assert np.abs(f - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(f - oracle.fft2d_r2c(a)).max() <= 1e-12 * np.abs(ref).max()
    b = ctx.ift_2d(f, shape[1])
    assert np.abs(b - oracle.ift2d_c2r(f, shape[1])).max() <= 1e-11 * a.size


@pytest.mark.parametrize("bfilter", [False, True])
@pytest.mark.parametrize("seed", range(4))
def test_corr_patch_to_image(ctx, seed, bfilter):
    rng = np.random.default_rng(40 + seed)
    N = [64, 128, 256, 64][seed]
    img = np.zeros((N, N))
    h2, w2 = int(rng.integers(N // 4, N // 2)), int(rng.integers(N // 4, N // 2))
    img[:h2, :w2] = rng.standard_normal((h2, w2)) + 2.0   # raw (unnormalised) correlation: the peak need
                                                           # not sit at the patch's origin
    ph, pw = int(rng.integers(4, h2 // 2)), int(rng.integers(4, w2 // 2))
    x0, y0 = int(rng.integers(0, w2 - pw)), int(rng.integers(0, h2 - ph))
    patch = img[y0:y0 + ph, x0:x0 + pw] + 0.05 * rng.standard_normal((ph, pw))
    pa, va = synth.raster_points(patch)
    g = ctx.corr_patch_to_image(pa, va, img, 0, 0, N, bfilter)
    o = oracle.corr_patch_to_image(pa, va, img, 0, 0, N, bfilter)
    assert g[0] == o[0]
    assert (g[4].lag_x, g[4].lag_y) == (o[4].lag_x, o[4].lag_y)
    assert abs(g[1] - o[1]) <= 1e-9 * max(1.0, abs(o[1]))
    assert abs(g[2] - o[2]) <= 1e-9 and abs(g[3] - o[3]) <= 1e-9


@pytest.mark.parametrize("seed", range(12))
def test_corr_images_f(ctx, seed):
    """CorrImagesF: R image, MakeF, OrderF, FPeak, ReturnR on the device against the oracle."""
    from tests import bruteforce as bf
    A, B, _ = bf.f_pair(seed, lo=12, hi=64, field=96) if seed >= 6 else bf.f_pair(seed)
    pa, va = synth.raster_points(A, x0=2)
    pb, vb = synth.raster_points(B, y0=-3)
    h1, w1 = A.shape
    h2, w2 = B.shape
    req = int(0.2 * min(w1 * h1, w2 * h2))
    rgn = abi.rgn_big_enough(3, req)
    cnt = abi.cnt_enough_points(req) if seed % 3 == 0 else None
    mincor = 0.0 if seed % 2 else 0.2
    Rx, Ry = (9, 6) if seed % 4 == 1 else (-1, -1)
    kw = dict(rgn=rgn, cnt=cnt, mincor=mincor, nbmaxht=0.5, Ox=1, Oy=0, Rx=Rx, Ry=Ry)
    g = ctx.corr_images(abi.XC_MODE_F, pa, va, pb, vb, **kw)
    o = oracle.corr_images(abi.XC_MODE_F, pa, va, pb, vb, **kw)
    _same(g, o, f"F mode seed {seed}")


def test_corr_images_f_thumbnails(ctx):
    """CorrImagesF on a 128 x 96 pair with a known shift (the R image is 255 x 191)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(77)
    base = gaussian_filter(rng.standard_normal((160, 200)), 1.5)
    A = base[20:116, 30:158]
    B = base[25:121, 22:150] + 0.05 * rng.standard_normal((96, 128))
    pa, va = synth.raster_points(A)
    pb, vb = synth.raster_points(B)
    rgn = abi.rgn_big_enough(32, 64 * 48)
    g = ctx.corr_images(abi.XC_MODE_F, pa, va, pb, vb, rgn=rgn, nbmaxht=0.5)
    o = oracle.corr_images(abi.XC_MODE_F, pa, va, pb, vb, rgn=rgn, nbmaxht=0.5)
    _same(g, o, "F thumbnails")
    assert g[0] == abi.XC_OK and (g[4].lag_x, g[4].lag_y) == (8, -5)
