"""CThmUtil::Crop and Crop_NoCR (1_DMesh/CThmUtil.cpp:352-464): the overlap boxes from a's offset in b's
system, the points of each connected region inside its box, their values from the picture.

The oracle is pinned by what the geometry fixes: when B is A moved by (dx, dy), the two boxes hold the same
pixels, so the two value lists are equal; the offset is truncated toward zero; whole pictures are used for
xyconf 0 or a box thinner than max(olap1d, 8); the kept points are numpy's boolean selection in list order, their
values numpy's fancy indexing at the truncated coordinates, shifted by the floor / ceil bounding box.  The device
forms (values gathered from device pictures) are held to the oracle bit for bit."""
import numpy as np
import pytest

import oracle
from alignment_projects_amd import abi


def _pair(seed, ws, hs, dx, dy):
    """A and B = A moved by (dx, dy): B(x + dx, y + dy) = A(x, y); fresh noise where A does not reach."""
    rng = np.random.default_rng(seed)
    big = rng.normal(size=(hs + 2 * abs(dy) + 2, ws + 2 * abs(dx) + 2))
    oy, ox = abs(dy) + 1, abs(dx) + 1
    A = big[oy:oy + hs, ox:ox + ws].copy()
    B = big[oy - dy:oy - dy + hs, ox - dx:ox - dx + ws].copy()
    return A, B


def _region(seed, ws, hs, frac=False):
    """A connected-region point list: raster points of a disc-shaped blob, shuffled, optionally off the grid."""
    yy, xx = np.mgrid[0:hs, 0:ws]
    keep = (xx - 0.45 * ws) ** 2 + (yy - 0.55 * hs) ** 2 < (0.4 * min(ws, hs)) ** 2
    pts = np.stack([xx[keep], yy[keep]], axis=1).astype(np.float64)
    pts = pts[np.random.default_rng(seed).permutation(len(pts))]
    if frac:
        pts = pts + np.random.default_rng(seed + 1).uniform(0.0, 0.999, size=pts.shape)
    return pts


def test_nocr_boxes_of_a_moved_picture_hold_the_same_pixels():
    for dx, dy in ((13, -7), (-20, 0), (0, 31), (-9, -11)):
        A, B = _pair(1, 120, 90, dx, dy)
        tx, ty = dx + 0.6 * np.sign(dx), dy + 0.6 * np.sign(dy)  # truncated toward zero: (dx, dy)
        st, whole, p, (av, a), (bv, b) = oracle.crop_nocr(A, B, tx, ty, 1.0, 1.0, 8)
        assert st == abi.XC_OK and whole == 0
        assert np.array_equal(av, bv)
        assert (a.w, a.h) == (120 - abs(dx), 90 - abs(dy)) and a.n == a.w * a.h
        assert (b.Ox - a.Ox, b.Oy - a.Oy) == (dx, dy)
        assert np.array_equal(p[:, 0], np.tile(np.arange(a.w), a.h)) and np.array_equal(p[:, 1],
                                                                                      np.repeat(np.arange(a.h), a.w))


def test_offset_truncates_toward_zero_and_scales():
    A, B = _pair(2, 64, 64, -5, 3)
    # xyconf 0.5, scl 2: -23.9 * 0.5 / 2 = -5.975 -> -5;  14.2 * 0.5 / 2 = 3.55 -> 3
    st, whole, _, (av, a), (bv, b) = oracle.crop_nocr(A, B, -23.9, 14.2, 0.5, 2.0, 8)
    assert whole == 0 and (a.Ox, a.Oy, b.Ox, b.Oy) == (5, 0, 0, 3) and np.array_equal(av, bv)


def test_whole_pictures_when_no_confidence_or_thin_overlap():
    A, B = _pair(3, 50, 40, 0, 0)
    for args in ((0.0, 0.0, 0.0), (45.0, 0.0, 1.0), (0.0, 33.5, 1.0)):  # xyconf 0; 5 wide < 8; 6 high < 8
        st, whole, p, (av, a), (bv, b) = oracle.crop_nocr(A, B, args[0], args[1], args[2], 1.0, 4)
        assert whole == 1 and np.array_equal(av, A.ravel()) and np.array_equal(bv, B.ravel())
        assert (a.Ox, a.Oy, a.w, a.h) == (0, 0, 50, 40)
    # olap1d above 8 sets the bar: a 20-wide box passes at 20, not at 21
    assert oracle.crop_nocr(A, B, 30.0, 0.0, 1.0, 1.0, 20)[1] == 0
    assert oracle.crop_nocr(A, B, 30.0, 0.0, 1.0, 1.0, 21)[1] == 1


@pytest.mark.parametrize("frac", [False, True])
def test_crop_selection_and_values_match_numpy(frac):
    ws, hs, dx, dy = 150, 110, 37, -21
    A, B = _pair(4, ws, hs, dx, dy)
    ap, bp = _region(5, ws, hs, frac), _region(6, ws, hs, frac)
    st, whole, (pa, va, a), (pb, vb, b) = oracle.crop(A, B, ap, bp, dx, dy, 1.0, 1.0, 8, 100)
    assert st == abi.XC_OK and whole == 0
    for P, V, S, pic, (L, R, Bt, T) in ((ap, va, a, A, (0, ws - 1 - dx, -dy, hs - 1)),
                                       (bp, vb, b, B, (dx, ws - 1, 0, hs - 1 + dy))):
        sel = P[(P[:, 0] >= L) & (P[:, 0] <= R) & (P[:, 1] >= Bt) & (P[:, 1] <= T)]
        assert S.n == len(sel) and np.array_equal(V, pic[sel[:, 1].astype(int), sel[:, 0].astype(int)])
        lo = np.floor(sel.min(axis=0))
        hi = np.ceil(sel.max(axis=0))
        assert (S.Ox, S.Oy, S.w, S.h) == (lo[0], lo[1], hi[0] - lo[0] + 1, hi[1] - lo[1] + 1)
    np.testing.assert_array_equal(pa + [a.Ox, a.Oy], ap[(ap[:, 0] <= ws - 1 - dx) & (ap[:, 1] >= -dy)])
    # whole pictures: every point, in list order
    st, whole, (pa, va, a), _ = oracle.crop(A, B, ap, bp, dx, dy, 0.0, 1.0, 8, 100)
    assert whole == 1 and a.n == len(ap) and np.array_equal(va, A[ap[:, 1].astype(int), ap[:, 0].astype(int)])


def test_crop_too_small_is_the_references_false():
    A, B = _pair(7, 80, 80, 40, 40)
    ap, bp = _region(8, 80, 80), _region(9, 80, 80)
    n_in = int(((ap[:, 0] <= 39) & (ap[:, 1] <= 39)).sum())
    assert oracle.crop(A, B, ap, bp, 40, 40, 1.0, 1.0, 8, n_in - 1)[0] == abi.XC_OK
    assert oracle.crop(A, B, ap, bp, 40, 40, 1.0, 1.0, 8, n_in)[0] == abi.XC_TOO_SMALL


# ---------------------------------------------------------------------------------------------- device


class _DevPics:
    """Two FP64 pictures in device memory (HIP runtime calls; no other framework)."""

    def __init__(self, *pics):
        import ctypes as C
        from tests.test_thumbs import _hip
        self.C, self.hip = C, _hip()
        self.ptrs = []
        for pic in pics:
            host = np.ascontiguousarray(pic, dtype=np.float64)
            d = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(d), host.nbytes) == 0
            self.ptrs.append(d)
            assert self.hip.hipMemcpy(d, host.ctypes.data, host.nbytes, 1) == 0

    def __enter__(self):
        return self.ptrs

    def __exit__(self, *exc):
        for d in self.ptrs:
            self.hip.hipFree(d)


@pytest.mark.gpu
@pytest.mark.parametrize("frac", [False, True])
def test_crop_device_matches_oracle(ctx, frac):
    ws, hs, dx, dy = 1000, 700, -130, 95
    A, B = _pair(10, ws, hs, dx, dy)
    ap, bp = _region(11, ws, hs, frac), _region(12, ws, hs, frac)
    with _DevPics(A, B) as (da, db):
        for xyconf, olap2d in ((1.0, 1000), (0.0, 1000), (1.0, 10 ** 9)):
            g = ctx.crop(da, db, ws, hs, ap, bp, dx - 0.3, dy + 0.7, xyconf, 1.0, 40, olap2d)
            o = oracle.crop(A, B, ap, bp, dx - 0.3, dy + 0.7, xyconf, 1.0, 40, olap2d)
            assert g[:2] == o[:2], (xyconf, olap2d)
            if g[0] != abi.XC_OK:
                continue
            for (gp, gv, gs), (op, ov, os_) in zip(g[2:], o[2:]):
                assert np.array_equal(gp, op) and np.array_equal(gv, ov)
                assert (gs.Ox, gs.Oy, gs.w, gs.h, gs.n) == (os_.Ox, os_.Oy, os_.w, os_.h, os_.n)


@pytest.mark.gpu
def test_crop_nocr_device_matches_oracle(ctx):
    ws, hs = 1024, 768
    A, B = _pair(13, ws, hs, 77, -41)
    with _DevPics(A, B) as (da, db):
        for tx, ty, xyconf in ((77.2, -41.9, 1.0), (0.0, 0.0, 0.0), (2000.0, 0.0, 1.0), (-154.0, 83.0, 0.5)):
            whole, p, (av, a), (bv, b) = ctx.crop_nocr(da, db, ws, hs, tx, ty, xyconf, 1.0, 16)
            _, ow, op, (oav, oa), (obv, ob) = oracle.crop_nocr(A, B, tx, ty, xyconf, 1.0, 16)
            assert whole == ow and np.array_equal(p, op)
            assert np.array_equal(av, oav) and np.array_equal(bv, obv)
            assert [(s.Ox, s.Oy, s.w, s.h, s.n) for s in (a, b)] == [(s.Ox, s.Oy, s.w, s.h, s.n) for s in (oa, ob)]


@pytest.mark.gpu
def test_crop_refuses_points_outside_the_picture(ctx):
    from alignment_projects_amd.xcorr import XcorrError
    A, B = _pair(14, 64, 64, 0, 0)
    good = _region(15, 64, 64)
    with _DevPics(A, B) as (da, db):
        for bad in ([[64.0, 3.0]], [[3.0, 64.0]], [[-1.5, 2.0]], [[np.nan, 1.0]]):
            with pytest.raises(XcorrError):  # refused on the host, before any device access
                ctx.crop(da, db, 64, 64, np.vstack([good, bad]), good, 0, 0, 0.0, 1.0, 8, 10)
        with pytest.raises(XcorrError):
            ctx.crop(da, db, 64, 64, good, good, 1e300, 0, 1.0, 1.0, 8, 10)
