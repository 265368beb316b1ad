// xc_anglescan.hip -- CThmScan's angle search around the rotation sweep:
// RFromAngle on point lists (FP64), AngleScanMaxR / AngleScanConstXY and
// PeakHunt, evaluated either with the FP64 plain correlator (the
// reference's arithmetic, one angle per call) or with the batched FP32
// sweep over resident tiles (every angle of a scan in one launch).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <vector>

#include "xc_host.hpp"

using namespace xc;

namespace {

constexpr double REF_PI = 3.14159265358;  // the reference's truncated pi (GenDefs.h)

struct Rules {
    xc_rgn_rule rgn;
    xc_cnt_rule cnt;
};

// BigEnough(olap1D, reqArea) and EnoughPoints(reqArea), as RFromAngle passes them
// (CThmScan.cpp:28-33, 360-365): both sides >= olap1D, area > reqArea; both counts > reqArea.
Rules scan_rules(int olap1D, int64_t req)
{
    Rules r;
    memset(&r, 0, sizeof r);
    r.rgn.enabled = 1;
    r.rgn.min_w = r.rgn.min_h = olap1D;
    r.rgn.min_area = req;
    r.cnt.enabled = 1;
    r.cnt.k1 = r.cnt.k2 = req;
    return r;
}

// LineFit's correlation coefficient (Maths.cpp:457-507); NaN or Inf for a zero variance.
#pragma clang fp contract(off)
double lincor(const double *x, const double *y, int n)
{
    double nn = 0.0, sx = 0.0, sy = 0.0, sxx = 0.0, syy = 0.0, sxy = 0.0;
    for (int i = 0; i < n; ++i) {
        const double X = x[i], Y = y[i];
        sx += X;
        sy += Y;
        sxx += X * X;
        syy += Y * Y;
        sxy += X * Y;
        ++nn;
    }
    const double r = nn * sxy - sx * sy, vx = nn * sxx - sx * sx, vy = nn * syy - sy * sy;
    return r / sqrt(vx * vy);
}
#pragma clang fp contract(on)

// ROUND (Maths.h:19)
double round_ref(double x) { return (double)(x >= 0.0 ? (int)(x + 0.5) : (int)(x - 0.5)); }

// The angle lists of AngleScanMaxR (CThmScan.cpp:444-470) and AngleScanConstXY (:543-544).
std::vector<double> scan_angles(int mode, double center, double hlfwid, double step, bool favor)
{
    std::vector<double> a;
    if (mode == XC_SCAN_MAXR) {
        a.push_back(center);
        if (favor) {
            const double cwide = 2.0 * step, cstep = step / 3.0;
            for (double v = center - cwide; v <= center + cwide; v += cstep) a.push_back(v);
            if (hlfwid >= 180) {
                const double opc = center + 180;
                for (double v = opc - cwide; v <= opc + cwide; v += cstep) a.push_back(v);
            }
        }
    }
    for (double v = center - hlfwid; v <= center + hlfwid; v += step) a.push_back(v);
    return a;
}

// The selections of AngleScanMaxR (:476-487) and AngleScanConstXY (:554-613).
xc_angle_rec scan_select(int mode, const std::vector<xc_angle_rec> &vC)
{
    xc_angle_rec best;
    memset(&best, 0, sizeof best);
    const int nC = (int)vC.size();
    if (mode == XC_SCAN_MAXR) {
        if (!nC) return best;
        best = vC[0];
        for (int ic = 1; ic < nC; ++ic)
            if (vC[ic].R > best.R) best = vC[ic];
        return best;
    }
    std::vector<int> order(nC);
    for (int i = 0; i < nC; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return vC[a].R > vC[b].R; });
    const double rthresh = 0.7;
    const int m = 3;
    for (int ic : order) {
        if (ic < m || nC - 1 - ic < m) continue;
        double A[2 * m + 1], X[2 * m + 1], Y[2 * m + 1];
        int n = 0;
        for (int j = ic - m; j <= ic + m; ++j, ++n) {
            A[n] = vC[j].A;
            X[n] = round_ref(vC[j].X);
            Y[n] = round_ref(vC[j].Y);
        }
        double lc = lincor(A, X, n);
        if (std::isfinite(lc) && fabs(lc) < rthresh) continue;
        lc = lincor(A, Y, n);
        if (std::isfinite(lc) && fabs(lc) < rthresh) continue;
        return vC[ic];
    }
    return best;
}

// TAffine (t0 t1 X t3 t4 Y) and its product (TAffine.cpp:220-232).
struct Aff {
    double t[6];
};
constexpr Aff AFF_ID = {{1, 0, 0, 0, 1, 0}};

Aff aff_mul(const Aff &L, const Aff &R)
{
#pragma clang fp contract(off)
    Aff r;
    r.t[0] = L.t[0] * R.t[0] + L.t[1] * R.t[3];
    r.t[1] = L.t[0] * R.t[1] + L.t[1] * R.t[4];
    r.t[2] = L.t[0] * R.t[2] + L.t[1] * R.t[5] + L.t[2];
    r.t[3] = L.t[3] * R.t[0] + L.t[4] * R.t[3];
    r.t[4] = L.t[3] * R.t[1] + L.t[4] * R.t[4];
    r.t[5] = L.t[3] * R.t[2] + L.t[4] * R.t[5] + L.t[5];
    return r;
}

Aff aff_of(const double *t) { return t ? Aff{{t[0], t[1], t[2], t[3], t[4], t[5]}} : AFF_ID; }

// TAffine::NUSelect (TAffine.cpp:76-135): 0 scale, 1 x-scale, 2 y-scale, 3 x-skew, 4 y-skew, 5 rotation.
Aff nu_select(int sel, double a)
{
    switch (sel) {
    case 1: return Aff{{a, 0, 0, 0, 1, 0}};
    case 2: return Aff{{1, 0, 0, 0, a, 0}};
    case 3: return Aff{{1, a, 0, 0, 1, 0}};
    case 4: return Aff{{1, 0, 0, a, 1, 0}};
    case 5: {
        const double c = cos(a), s = sin(a);
        return Aff{{c, -s, 0, s, c, 0}};
    }
    default: return Aff{{a, 0, 0, 0, a, 0}};
    }
}

// RFromAngle on point lists (CThmScan.cpp:329-366): A's points by the 2x2 part of
// R(deg) * (Tdfm * Ttw) (identity for both when null), then CorrImagesS.
int rfa64_aff(xc_ctx *ctx, const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv, int nb,
              double deg, const Aff *Tdfm, const Aff *Ttw, const Rules &ru, int Ox, int Oy, int Rx, int Ry,
              std::vector<xc_point> &buf, xc_angle_rec *rec)
{
#pragma clang fp contract(off)
    const Aff T = aff_mul(nu_select(5, deg * REF_PI / 180.0), aff_mul(Tdfm ? *Tdfm : AFF_ID, Ttw ? *Ttw : AFF_ID));
    const double t0 = T.t[0], t1 = T.t[1], t3 = T.t[3], t4 = T.t[4];
    buf.resize(na);
    for (int i = 0; i < na; ++i) {
        buf[i].x = ap[i].x * t0 + ap[i].y * t1;
        buf[i].y = ap[i].x * t3 + ap[i].y * t4;
    }
    double X = 0.0, Y = 0.0, R = 0.0;
    const int st = xc_corr_images_f64(ctx, XC_MODE_S, buf.data(), av, na, bp, bv, nb, &ru.rgn, &ru.cnt, 0.0, 0.0, Ox,
                                      Oy, Rx, Ry, &X, &Y, &R, nullptr);
    if (st < 0) return st;
    rec->A = deg;
    rec->R = R;
    rec->X = X;
    rec->Y = Y;
    return XC_OK;
}

// RFromAngle on point lists (CThmScan.cpp:329-366, identity deformation and tweak).
int rfa64(xc_ctx *ctx, const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv, int nb,
          double deg, const Rules &ru, int Ox, int Oy, int Rx, int Ry, std::vector<xc_point> &buf, xc_angle_rec *rec)
{
#pragma clang fp contract(off)
    const double c = cos(deg * REF_PI / 180.0), s = sin(deg * REF_PI / 180.0);
    const double t0 = c, t1 = -s, t3 = s, t4 = c;
    buf.resize(na);
    for (int i = 0; i < na; ++i) {
        buf[i].x = ap[i].x * t0 + ap[i].y * t1;
        buf[i].y = ap[i].x * t3 + ap[i].y * t4;
    }
    double X = 0.0, Y = 0.0, R = 0.0;
    const int st = xc_corr_images_f64(ctx, XC_MODE_S, buf.data(), av, na, bp, bv, nb, &ru.rgn, &ru.cnt, 0.0, 0.0, Ox,
                                      Oy, Rx, Ry, &X, &Y, &R, nullptr);
    if (st < 0) return st;
    rec->A = deg;
    rec->R = R;
    rec->X = X;
    rec->Y = Y;
    return XC_OK;
}

// A batch of angles through the FP32 sweep of tiles.
int rfa32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, const std::vector<double> &angles, const Rules &ru, int Ox,
          int Oy, int Rx, int Ry, std::vector<xc_angle_rec> &out)
{
    const int n = (int)angles.size();
    out.resize(n);
    if (!n) return XC_OK;
    std::vector<xc_result> res(n);
    const int st = xc_sweep_angles_f32(ctx, a, b, angles.data(), n, &ru.rgn, &ru.cnt, Ox, Oy, Rx, Ry, res.data());
    if (st < 0) return st;
    for (int i = 0; i < n; ++i) {
        const bool ok = res[i].status == XC_OK;
        out[i].A = angles[i];
        out[i].R = ok ? res[i].r : 0.0;  // CorrImagesS: 0 and (0, 0) without a peak
        out[i].X = ok ? res[i].dx : 0.0;
        out[i].Y = ok ? res[i].dy : 0.0;
    }
    return XC_OK;
}

// PeakHunt (CThmScan.cpp:676-732) over an evaluator of one angle.
int peak_hunt(double hlfwid, xc_angle_rec *best, int *n_evals, const std::function<int(double, xc_angle_rec *)> &eval)
{
    xc_angle_rec C, B0 = *best;
    double L = best->A - hlfwid, R = best->A + hlfwid, M;
    int k = 1, st;
    if ((st = eval(M = (L + R) / 2.0, best))) return st;
    while (R - L > 0.0001) {
        double a;
        ++k;
        if ((st = eval(a = (L + M) / 2.0, &C))) return st;
        if (C.R >= best->R) {
            R = M;
            *best = C;
            M = a;
        } else {
            L = a;
        }
        if ((st = eval(a = (M + R) / 2.0, &C))) return st;
        if (C.R >= best->R) {
            L = M;
            *best = C;
            M = a;
        } else {
            R = a;
        }
    }
    if (B0.R > best->R) *best = B0;
    if (n_evals) *n_evals = k;
    return XC_OK;
}

int finish_scan(int mode, const std::vector<xc_angle_rec> &vC, xc_angle_rec *best, xc_angle_rec *recs, int cap,
                int *n_recs)
{
    *best = scan_select(mode, vC);
    if (recs && cap > 0) memcpy(recs, vC.data(), sizeof(xc_angle_rec) * std::min<size_t>(cap, vC.size()));
    if (n_recs) *n_recs = (int)vC.size();
    return XC_OK;
}

bool bad_scan(int mode, double step, const xc_angle_rec *best)
{
    return !best || (mode != XC_SCAN_MAXR && mode != XC_SCAN_CONSTXY) || !(step > 0.0);
}

}  // namespace

extern "C" int xc_rfromangle_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na, const xc_point *bp,
                                 const double *bv, int nb, double deg, int olap1D, int64_t req_area, int Ox, int Oy,
                                 int Rx, int Ry, xc_angle_rec *rec)
{
    if (!ctx || !rec) return XC_E_ARG;
    std::vector<xc_point> buf;
    return rfa64(ctx, ap, av, na, bp, bv, nb, deg, scan_rules(olap1D, req_area), Ox, Oy, Rx, Ry, buf, rec);
}

extern "C" int xc_angle_scan_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na, const xc_point *bp,
                                 const double *bv, int nb, int mode, double center, double hlfwid, double step,
                                 int favor_center, int olap1D, int64_t req_area, int Ox, int Oy, int Rx, int Ry,
                                 xc_angle_rec *best, xc_angle_rec *recs, int cap, int *n_recs)
{
    if (!ctx || bad_scan(mode, step, best)) return XC_E_ARG;
    const Rules ru = scan_rules(olap1D, req_area);
    const std::vector<double> ang = scan_angles(mode, center, hlfwid, step, favor_center != 0);
    std::vector<xc_angle_rec> vC(ang.size());
    std::vector<xc_point> buf;
    for (size_t i = 0; i < ang.size(); ++i) {
        const int st = rfa64(ctx, ap, av, na, bp, bv, nb, ang[i], ru, Ox, Oy, Rx, Ry, buf, &vC[i]);
        if (st) return st;
    }
    return finish_scan(mode, vC, best, recs, cap, n_recs);
}

extern "C" int xc_angle_scan_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, int mode, double center,
                                 double hlfwid, double step, int favor_center, int olap1D, int64_t req_area, int Ox,
                                 int Oy, int Rx, int Ry, xc_angle_rec *best, xc_angle_rec *recs, int cap, int *n_recs)
{
    if (!ctx || !a || !b || bad_scan(mode, step, best)) return XC_E_ARG;
    const Rules ru = scan_rules(olap1D, req_area);
    const std::vector<double> ang = scan_angles(mode, center, hlfwid, step, favor_center != 0);
    std::vector<xc_angle_rec> vC;
    const int st = rfa32(ctx, a, b, ang, ru, Ox, Oy, Rx, Ry, vC);
    if (st) return st;
    return finish_scan(mode, vC, best, recs, cap, n_recs);
}

extern "C" int xc_peak_hunt_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na, const xc_point *bp,
                                const double *bv, int nb, double hlfwid, int olap1D, int64_t req_area, int Ox, int Oy,
                                int Rx, int Ry, xc_angle_rec *best, int *n_evals)
{
    if (!ctx || !best) return XC_E_ARG;
    const Rules ru = scan_rules(olap1D, req_area);
    std::vector<xc_point> buf;
    return peak_hunt(hlfwid, best, n_evals, [&](double deg, xc_angle_rec *rec) {
        return rfa64(ctx, ap, av, na, bp, bv, nb, deg, ru, Ox, Oy, Rx, Ry, buf, rec);
    });
}

extern "C" int xc_peak_hunt_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, double hlfwid, int olap1D,
                                int64_t req_area, int Ox, int Oy, int Rx, int Ry, xc_angle_rec *best, int *n_evals)
{
    if (!ctx || !a || !b || !best) return XC_E_ARG;
    const Rules ru = scan_rules(olap1D, req_area);
    std::vector<xc_angle_rec> one;
    return peak_hunt(hlfwid, best, n_evals, [&](double deg, xc_angle_rec *rec) {
        const int st = rfa32(ctx, a, b, std::vector<double>(1, deg), ru, Ox, Oy, Rx, Ry, one);
        if (!st) *rec = one[0];
        return st;
    });
}

extern "C" int xc_rfromangle_affine_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na, const xc_point *bp,
                                        const double *bv, int nb, double deg, const double *Tdfm, const double *Ttw,
                                        int olap1D, int64_t req_area, int Ox, int Oy, int Rx, int Ry,
                                        xc_angle_rec *rec)
{
    if (!ctx || !rec) return XC_E_ARG;
    std::vector<xc_point> buf;
    const Aff D = aff_of(Tdfm), W = aff_of(Ttw);
    return rfa64_aff(ctx, ap, av, na, bp, bv, nb, deg, &D, &W, scan_rules(olap1D, req_area), Ox, Oy, Rx, Ry, buf, rec);
}

namespace {

// R of each candidate pretweak M(sel[i], mag[i]) * tw at the round's angle.
using PtwEval = std::function<int(const std::vector<int> &sel, const std::vector<double> &mag, const Aff &tw,
                                  std::vector<double> &R)>;

// Pretweaks (CThmScan.cpp:205-323) with PTWInterp (:119-160) and NewXFromParabola (:48-61):
// tw (the pretweak) and bR (the best R) in/out.  Each round's magnitude sweep is one
// evaluation batch, then the parabola's two neighbours, then its vertex.
int pretweaks_core(const PtwEval &eval, Aff &tw, double &bR, bool *changed, int *n_evals)
{
    int evals = 0, st;
    const double astep = 0.02;
    const int nstep = 5;
    bool any = false;
    int vused[5] = {0, 0, 0, 0, 0};
    std::vector<double> Rv;
    for (int itype = 0; itype < 5; ++itype) {
        // sweep every unused type's magnitude about its identity value
        std::vector<int> vsel;
        std::vector<double> va;
        for (int sel = 0; sel < 5; ++sel) {
            if (vused[sel]) continue;
            const double abase = (sel >= 0 && sel <= 2) ? 1.0 : 0.0;
            for (int j = -nstep; j <= nstep; ++j) {
                vsel.push_back(sel);
                va.push_back(abase + j * astep);
            }
        }
        if ((st = eval(vsel, va, tw, Rv))) return st;
        evals += (int)vsel.size();
        double rbest = 0;
        int ibest = -1;
        for (size_t ic = 0; ic < vsel.size(); ++ic)
            if (Rv[ic] > rbest) {
                rbest = Rv[ic];
                ibest = (int)ic;
            }
        if (ibest == -1) break;
        // PTWInterp: parabola through the neighbours at +-astep
        const int sel = vsel[ibest];
        const double x1 = va[ibest], y1 = rbest, d = astep;
        if ((st = eval({sel, sel}, {x1 - d, x1 + d}, tw, Rv))) return st;
        evals += 2;
        const double y0 = Rv[0], y2 = Rv[1];
        double xnew = x1;
        {
#pragma clang fp contract(off)
            const double Dp = d * (y0 - y2) / (2 * (y0 + y2 - y1 - y1));
            if (fabs(Dp) <= 2 * d) xnew += Dp;
        }
        if ((st = eval({sel}, {xnew}, tw, Rv))) return st;
        evals += 1;
        double ynew = Rv[0];
        if (ynew < y1) {
            xnew = x1;
            ynew = y1;
        }
        rbest = ynew;
        if (rbest > bR) {
            tw = aff_mul(nu_select(sel, xnew), tw);
            vused[sel] = 1;
            bR = rbest;
            any = true;
        } else {
            break;
        }
    }
    if (changed) *changed = any;
    if (n_evals) *n_evals = evals;
    return XC_OK;
}

// Pretweaks on point lists in FP64, one device correlation per candidate.
int pretweaks_run(xc_ctx *ctx, const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv,
                  int nb, double deg, const Aff &D, Aff &tw, double &bR, const Rules &ru, int Ox, int Oy, int Rx,
                  int Ry, std::vector<xc_point> &buf, bool *changed, int *n_evals)
{
    const PtwEval eval = [&](const std::vector<int> &sel, const std::vector<double> &mag, const Aff &t,
                             std::vector<double> &R) {
        R.resize(sel.size());
        for (size_t i = 0; i < sel.size(); ++i) {
            const Aff W = aff_mul(nu_select(sel[i], mag[i]), t);  // M * Tptwk (RFromAngle's iaux branch)
            xc_angle_rec C;
            const int st = rfa64_aff(ctx, ap, av, na, bp, bv, nb, deg, &D, &W, ru, Ox, Oy, Rx, Ry, buf, &C);
            if (st) return st;
            R[i] = C.R;
        }
        return (int)XC_OK;
    };
    return pretweaks_core(eval, tw, bR, changed, n_evals);
}

// The search oval of one evaluation: CorrImagesS's (Ox, Oy, Rx, Ry).
struct Oval {
    int Ox, Oy, Rx, Ry;
    bool operator==(const Oval &o) const { return Ox == o.Ox && Oy == o.Oy && Rx == o.Rx && Ry == o.Ry; }
};

// RFromAngle at each angle with the deformation D, the pretweak W and that angle's oval ov[i].
using AngEval = std::function<int(const std::vector<double> &deg, const Aff &D, const Aff &W,
                                  const std::vector<Oval> &ov, std::vector<xc_angle_rec> &out)>;
// R of each candidate pretweak M(sel[i], mag[i]) * tw at angle deg, with D and the oval.
using TwkEval = std::function<int(double deg, const Aff &D, const Oval &ov, const std::vector<int> &sel,
                                  const std::vector<double> &mag, const Aff &tw, std::vector<double> &R)>;

// The two evaluators on point lists in FP64: one device correlation per probe.
struct PointEvals {
    xc_ctx *ctx;
    const xc_point *ap;
    const double *av;
    int na;
    const xc_point *bp;
    const double *bv;
    int nb;
    Rules ru;
    std::vector<xc_point> buf;
    AngEval ang()
    {
        return [this](const std::vector<double> &deg, const Aff &D, const Aff &W, const std::vector<Oval> &ov,
                      std::vector<xc_angle_rec> &out) {
            out.resize(deg.size());
            for (size_t i = 0; i < deg.size(); ++i) {
                const Oval &o = ov[i];
                const int st =
                    rfa64_aff(ctx, ap, av, na, bp, bv, nb, deg[i], &D, &W, ru, o.Ox, o.Oy, o.Rx, o.Ry, buf, &out[i]);
                if (st) return st;
            }
            return (int)XC_OK;
        };
    }
    TwkEval twk()
    {
        return [this](double deg, const Aff &D, const Oval &o, const std::vector<int> &sel,
                      const std::vector<double> &mag, const Aff &tw, std::vector<double> &R) {
            R.resize(sel.size());
            for (size_t i = 0; i < sel.size(); ++i) {
                const Aff W = aff_mul(nu_select(sel[i], mag[i]), tw);  // M * Tptwk (RFromAngle's iaux branch)
                xc_angle_rec C;
                const int st = rfa64_aff(ctx, ap, av, na, bp, bv, nb, deg, &D, &W, ru, o.Ox, o.Oy, o.Rx, o.Ry, buf, &C);
                if (st) return st;
                R[i] = C.R;
            }
            return (int)XC_OK;
        };
    }
};

// The two evaluators over resident tiles in FP32: every probe list is one batched sweep of
// the maps R(deg) * (D * W) (xc_sweep_affine_f32); no peak gives R = 0 at (0, 0) as CorrImagesS does.
struct TileEvals {
    xc_ctx *ctx;
    const xc_tile *a, *b;
    Rules ru;
    std::vector<double> maps;
    std::vector<xc_result> res;
    // one batched sweep of the maps T[i], each with its oval ov[i]
    int sweep(const std::vector<Aff> &T, const std::vector<Oval> &ov)
    {
        const size_t n = T.size();
        maps.resize(4 * n);
        res.resize(n);
        std::vector<int> ovals(4 * n);
        for (size_t i = 0; i < n; ++i) {
            maps[4 * i + 0] = T[i].t[0];
            maps[4 * i + 1] = T[i].t[1];
            maps[4 * i + 2] = T[i].t[3];
            maps[4 * i + 3] = T[i].t[4];
            ovals[4 * i + 0] = ov[i].Ox;
            ovals[4 * i + 1] = ov[i].Oy;
            ovals[4 * i + 2] = ov[i].Rx;
            ovals[4 * i + 3] = ov[i].Ry;
        }
        const int st = sweep_affine_ovals(ctx, a, b, maps.data(), ovals.data(), (int)n, &ru.rgn, &ru.cnt, res.data());
        return st < 0 ? st : (int)XC_OK;
    }
    AngEval ang()
    {
        return [this](const std::vector<double> &deg, const Aff &D, const Aff &W, const std::vector<Oval> &ov,
                      std::vector<xc_angle_rec> &out) {
            std::vector<Aff> T(deg.size());
            const Aff DW = aff_mul(D, W);
            for (size_t i = 0; i < deg.size(); ++i) T[i] = aff_mul(nu_select(5, deg[i] * REF_PI / 180.0), DW);
            const int st = sweep(T, ov);
            if (st) return st;
            out.resize(deg.size());
            for (size_t i = 0; i < deg.size(); ++i) {
                const bool ok = res[i].status == XC_OK;
                out[i].A = deg[i];
                out[i].R = ok ? res[i].r : 0.0;
                out[i].X = ok ? res[i].dx : 0.0;
                out[i].Y = ok ? res[i].dy : 0.0;
            }
            return (int)XC_OK;
        };
    }
    TwkEval twk()
    {
        return [this](double deg, const Aff &D, const Oval &o, const std::vector<int> &sel,
                      const std::vector<double> &mag, const Aff &tw, std::vector<double> &R) {
            const Aff Rot = nu_select(5, deg * REF_PI / 180.0);
            std::vector<Aff> T(sel.size());
            for (size_t i = 0; i < sel.size(); ++i) T[i] = aff_mul(Rot, aff_mul(D, aff_mul(nu_select(sel[i], mag[i]), tw)));
            const int st = sweep(T, std::vector<Oval>(sel.size(), o));
            if (st) return st;
            R.resize(sel.size());
            for (size_t i = 0; i < sel.size(); ++i) R[i] = res[i].status == XC_OK ? res[i].r : 0.0;
            return (int)XC_OK;
        };
    }
};

// CThmScan's matching state for one pair: the parameter block (thresholds, sweep switches,
// disc, Tdfm and the running Tptwk) and the evaluation count, over either evaluator pair.
struct Thm {
    xc_thm_params *p;
    AngEval eval_ang;
    TwkEval eval_twk;

    // The oval at angle deg: newAngProc's when the hook is set, else the disc (CThmScan.cpp:351-355)
    Oval oval(double deg) const
    {
        Oval o{p->Ox, p->Oy, p->Rx, p->Ry};
        if (p->ang_proc) p->ang_proc(&o.Ox, &o.Oy, &o.Rx, &o.Ry, deg, p->ang_arg);
        return o;
    }
    // RFromAngle (CThmScan.cpp:329-366) with the current Tdfm, Tptwk and ovals, over a list of angles
    int rfa_list(const std::vector<double> &deg, std::vector<xc_angle_rec> &out)
    {
        p->n_evals += (int)deg.size();
        std::vector<Oval> ov(deg.size());
        for (size_t i = 0; i < deg.size(); ++i) ov[i] = oval(deg[i]);
        return eval_ang(deg, aff_of(p->Tdfm), aff_of(p->Tptwk), ov, out);
    }
    int rfa(double deg, xc_angle_rec *C)
    {
        std::vector<xc_angle_rec> o;
        const int st = rfa_list(std::vector<double>(1, deg), o);
        if (!st) *C = o[0];
        return st;
    }
    // AngleScanSel (CThmScan.cpp:627-641): one list of angles
    int scan_sel(xc_angle_rec *best, double center, double hlfwid, double step, bool favor)
    {
        const int mode = p->const_xy ? XC_SCAN_CONSTXY : XC_SCAN_MAXR;
        const std::vector<double> ang = scan_angles(mode, center, hlfwid, step, favor);
        std::vector<xc_angle_rec> vC;
        const int st = rfa_list(ang, vC);
        if (st) return st;
        *best = scan_select(mode, vC);
        return XC_OK;
    }
    // Pretweaks (CThmScan.cpp:205-323) folding into p->Tptwk; each round's candidates one list
    int pretweaks(double bestR, double deg, bool *changed)
    {
        Aff tw = aff_of(p->Tptwk);
        double bR = bestR;
        int ev = 0;
        const Aff D = aff_of(p->Tdfm);
        const Oval o = oval(deg);  // every candidate of a round is evaluated at deg
        const PtwEval e = [&](const std::vector<int> &sel, const std::vector<double> &mag, const Aff &t,
                              std::vector<double> &R) { return eval_twk(deg, D, o, sel, mag, t, R); };
        const int st = pretweaks_core(e, tw, bR, changed, &ev);
        p->n_evals += ev;
        if (!st) memcpy(p->Tptwk, tw.t, sizeof tw.t);
        return st;
    }
    // AngleScanWithTweaks (CThmScan.cpp:647-672)
    int with_tweaks(xc_angle_rec *best, double center, double hlfwid, double step)
    {
        int st = scan_sel(best, center, hlfwid, step, true);
        if (st || !(best->R < p->rthresh) || !p->pretweak) return st;
        bool changed = false;
        if ((st = pretweaks(best->R, best->R > 0.0 ? best->A : center, &changed))) return st;
        return changed ? scan_sel(best, center, hlfwid, step, true) : XC_OK;
    }
    // PeakHunt (CThmScan.cpp:676-732): each probe depends on the last decision
    int hunt(xc_angle_rec *best, double hlfwid)
    {
        return peak_hunt(hlfwid, best, nullptr, [&](double deg, xc_angle_rec *rec) { return rfa(deg, rec); });
    }
    // UsePriorAngles (CThmScan.cpp:734-758)
    int use_prior(double ang0, double hfangpr, xc_angle_rec *best, int *ok)
    {
        int st = with_tweaks(best, ang0, hfangpr, 0.1);
        bool good = !st && !(best->R < p->rthresh);
        if (good) {
            st = hunt(best, 0.1 * 1.5);
            good = !st && !(best->R < p->rthresh);
        }
        if (ok) *ok = good ? 1 : 0;
        return st;
    }
    // DenovoBestAngle (CThmScan.cpp:764-790)
    int denovo(double ang0, double hfangdn, double step, xc_angle_rec *best, int *ok)
    {
        int st = with_tweaks(best, ang0, hfangdn, step);
        bool good = !st && !(best->R < p->rthresh);
        if (good) {
            st = scan_sel(best, best->A, step * 2.0, step * 0.05, false);
            good = !st && !(best->R < p->rthresh);
        }
        if (good) {
            st = hunt(best, step * 0.05 * 1.5);
            good = !st && !(best->R < p->rthresh);
        }
        if (ok) *ok = good ? 1 : 0;
        return st;
    }
    // PostTweaks (CThmScan.cpp:796-802)
    int post_tweaks(xc_angle_rec *best)
    {
        p->Ox = (int)best->X;  // SetDisc( (int)best.X, (int)best.Y, 40, 40 )
        p->Oy = (int)best->Y;
        p->Rx = p->Ry = 40;
        bool changed = false;
        const int st = pretweaks(best->R, best->A, &changed);
        if (st) return st;
        return rfa(best->A, best);
    }
};

bool bad_thm(xc_ctx *ctx, const xc_thm_params *p, const xc_angle_rec *best)
{
    return !ctx || !p || !best;
}

}  // namespace

extern "C" int xc_pretweaks_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na, const xc_point *bp,
                                const double *bv, int nb, double deg, const double *Tdfm, double *Tptwk,
                                double *bestR, int olap1D, int64_t req_area, int Ox, int Oy, int Rx, int Ry,
                                int *changed, int *n_evals)
{
    if (!ctx || !Tptwk || !bestR) return XC_E_ARG;
    std::vector<xc_point> buf;
    Aff tw = aff_of(Tptwk);
    double bR = *bestR;
    bool any = false;
    const int st = pretweaks_run(ctx, ap, av, na, bp, bv, nb, deg, aff_of(Tdfm), tw, bR,
                                 scan_rules(olap1D, req_area), Ox, Oy, Rx, Ry, buf, &any, n_evals);
    if (st) return st;
    memcpy(Tptwk, tw.t, sizeof tw.t);
    *bestR = bR;
    if (changed) *changed = any ? 1 : 0;
    return XC_OK;
}

extern "C" int xc_finish_at_full_res_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                                         const xc_point *bp, const double *bv, int nb, const double *Tdfm,
                                         const double *Tptwk, int olap1D, int64_t req_area, xc_angle_rec *best,
                                         int *used)
{
    if (!ctx || !best) return XC_E_ARG;
    const xc_angle_rec b0 = *best;
    const Aff D = aff_of(Tdfm), W = aff_of(Tptwk);
    std::vector<xc_point> buf;
    xc_angle_rec C;
    // SetDisc( (int)best.X, (int)best.Y, 80, 80 ), then RFromAngle( best, best.A )
    const int st = rfa64_aff(ctx, ap, av, na, bp, bv, nb, b0.A, &D, &W, scan_rules(olap1D, req_area), (int)b0.X,
                             (int)b0.Y, 80, 80, buf, &C);
    if (st) return st;
    const bool ok = fabs(C.X - b0.X) <= 20 && fabs(C.Y - b0.Y) <= 20;
    if (ok) {
        *best = C;
        best->R = b0.R;  // the thumbnail R, comparable across the table
    } else {
        *best = b0;
    }
    if (used) *used = ok ? 1 : 0;
    return XC_OK;
}

extern "C" void xc_thm_params_default(xc_thm_params *p)
{
    if (!p) return;
    memset(p, 0, sizeof *p);
    p->rthresh = 0.30;
    p->const_xy = 1;
    p->pretweak = 1;
    p->Rx = p->Ry = -1;
    p->Tdfm[0] = p->Tdfm[4] = 1.0;
    p->Tptwk[0] = p->Tptwk[4] = 1.0;
}

extern "C" int xc_thm_scan_with_tweaks_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                                           const xc_point *bp, const double *bv, int nb, xc_thm_params *p,
                                           double center, double hlfwid, double step, xc_angle_rec *best)
{
    if (bad_thm(ctx, p, best) || !(step > 0.0)) return XC_E_ARG;
    PointEvals E{ctx, ap, av, na, bp, bv, nb, scan_rules(p->olap1D, p->req_area), {}};
    Thm t{p, E.ang(), E.twk()};
    return t.with_tweaks(best, center, hlfwid, step);
}

extern "C" int xc_thm_use_prior_angles_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                                           const xc_point *bp, const double *bv, int nb, xc_thm_params *p,
                                           double ang0, double hfangpr, xc_angle_rec *best, int *ok)
{
    if (bad_thm(ctx, p, best)) return XC_E_ARG;
    PointEvals E{ctx, ap, av, na, bp, bv, nb, scan_rules(p->olap1D, p->req_area), {}};
    Thm t{p, E.ang(), E.twk()};
    return t.use_prior(ang0, hfangpr, best, ok);
}

extern "C" int xc_thm_denovo_best_angle_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                                            const xc_point *bp, const double *bv, int nb, xc_thm_params *p,
                                            double ang0, double hfangdn, double step, xc_angle_rec *best, int *ok)
{
    if (bad_thm(ctx, p, best) || !(step > 0.0)) return XC_E_ARG;
    PointEvals E{ctx, ap, av, na, bp, bv, nb, scan_rules(p->olap1D, p->req_area), {}};
    Thm t{p, E.ang(), E.twk()};
    return t.denovo(ang0, hfangdn, step, best, ok);
}

extern "C" int xc_thm_post_tweaks_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                                      const xc_point *bp, const double *bv, int nb, xc_thm_params *p,
                                      xc_angle_rec *best)
{
    if (bad_thm(ctx, p, best)) return XC_E_ARG;
    PointEvals E{ctx, ap, av, na, bp, bv, nb, scan_rules(p->olap1D, p->req_area), {}};
    Thm t{p, E.ang(), E.twk()};
    return t.post_tweaks(best);
}

extern "C" int xc_thm_scan_with_tweaks_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, xc_thm_params *p,
                                           double center, double hlfwid, double step, xc_angle_rec *best)
{
    if (bad_thm(ctx, p, best) || !a || !b || !(step > 0.0)) return XC_E_ARG;
    TileEvals E{ctx, a, b, scan_rules(p->olap1D, p->req_area), {}, {}};
    Thm t{p, E.ang(), E.twk()};
    return t.with_tweaks(best, center, hlfwid, step);
}

extern "C" int xc_thm_use_prior_angles_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, xc_thm_params *p,
                                           double ang0, double hfangpr, xc_angle_rec *best, int *ok)
{
    if (bad_thm(ctx, p, best) || !a || !b) return XC_E_ARG;
    TileEvals E{ctx, a, b, scan_rules(p->olap1D, p->req_area), {}, {}};
    Thm t{p, E.ang(), E.twk()};
    return t.use_prior(ang0, hfangpr, best, ok);
}

extern "C" int xc_thm_denovo_best_angle_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, xc_thm_params *p,
                                            double ang0, double hfangdn, double step, xc_angle_rec *best, int *ok)
{
    if (bad_thm(ctx, p, best) || !a || !b || !(step > 0.0)) return XC_E_ARG;
    TileEvals E{ctx, a, b, scan_rules(p->olap1D, p->req_area), {}, {}};
    Thm t{p, E.ang(), E.twk()};
    return t.denovo(ang0, hfangdn, step, best, ok);
}

extern "C" int xc_thm_post_tweaks_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, xc_thm_params *p,
                                      xc_angle_rec *best)
{
    if (bad_thm(ctx, p, best) || !a || !b) return XC_E_ARG;
    TileEvals E{ctx, a, b, scan_rules(p->olap1D, p->req_area), {}, {}};
    Thm t{p, E.ang(), E.twk()};
    return t.post_tweaks(best);
}

extern "C" int xc_pretweaks_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, double deg, const double *Tdfm,
                                double *Tptwk, double *bestR, int olap1D, int64_t req_area, int Ox, int Oy, int Rx,
                                int Ry, int *changed, int *n_evals)
{
    if (!ctx || !a || !b || !Tptwk || !bestR) return XC_E_ARG;
    TileEvals E{ctx, a, b, scan_rules(olap1D, req_area), {}, {}};
    const TwkEval twk = E.twk();
    const Aff D = aff_of(Tdfm);
    Aff tw = aff_of(Tptwk);
    double bR = *bestR;
    bool any = false;
    const PtwEval e = [&](const std::vector<int> &sel, const std::vector<double> &mag, const Aff &t,
                          std::vector<double> &R) { return twk(deg, D, Oval{Ox, Oy, Rx, Ry}, sel, mag, t, R); };
    const int st = pretweaks_core(e, tw, bR, &any, n_evals);
    if (st) return st;
    memcpy(Tptwk, tw.t, sizeof tw.t);
    *bestR = bR;
    if (changed) *changed = any ? 1 : 0;
    return XC_OK;
}

extern "C" int xc_finish_at_full_res_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, const double *Tdfm,
                                         const double *Tptwk, int olap1D, int64_t req_area, xc_angle_rec *best,
                                         int *used)
{
    if (!ctx || !a || !b || !best) return XC_E_ARG;
    TileEvals E{ctx, a, b, scan_rules(olap1D, req_area), {}, {}};
    const AngEval ang = E.ang();
    const xc_angle_rec b0 = *best;
    std::vector<xc_angle_rec> out;
    // SetDisc( (int)best.X, (int)best.Y, 80, 80 ), then RFromAngle( best, best.A )
    const int st = ang(std::vector<double>(1, b0.A), aff_of(Tdfm), aff_of(Tptwk),
                       std::vector<Oval>(1, Oval{(int)b0.X, (int)b0.Y, 80, 80}), out);
    if (st) return st;
    const xc_angle_rec &C = out[0];
    const bool ok = fabs(C.X - b0.X) <= 20 && fabs(C.Y - b0.Y) <= 20;
    if (ok) {
        *best = C;
        best->R = b0.R;  // the thumbnail R, comparable across the table
    } else {
        *best = b0;
    }
    if (used) *used = ok ? 1 : 0;
    return XC_OK;
}
