// xc_thmpair.hip -- ThmPair_za^zb.txt tables, the thumbnail matcher's result
// file: header, appended entries in the reference's printf format, readers,
// and the mode-Y starting angle (median of prior angles).  Host code only.
#include <fcntl.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <vector>

#include "../../include/xcorr.h"

namespace {

// One table line as ReadThmPair's sscanf reads it (PipeFiles.cpp:1163-1170).
bool parse_line(const char *line, xc_thmpair *e)
{
    xc_thmpair p;
    memset(&p, 0, sizeof p);
    const int n = sscanf(line, "%d\t%d\t%d\t%d\t%d\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf", &p.atl, &p.acr, &p.btl,
                         &p.bcr, &p.err, &p.A, &p.R, &p.T[0], &p.T[1], &p.T[2], &p.T[3], &p.T[4], &p.T[5]);
    *e = p;  // sscanf leaves unmatched fields; the reference keeps whatever they held (here 0)
    return n > 0;
}

// Lines after the header; false when the file cannot be opened or is empty.
bool read_lines(const char *path, std::vector<xc_thmpair> &out, bool stop_at_match, int atl, int acr, int btl,
                int bcr, bool *found)
{
    FILE *f = fopen(path, "r");
    if (!f) return false;
    char *line = nullptr;
    size_t cap = 0;
    bool ok = getline(&line, &cap, f) > 0;  // header
    if (found) *found = false;
    while (ok && getline(&line, &cap, f) > 0) {
        xc_thmpair e;
        parse_line(line, &e);
        if (stop_at_match) {
            if (e.atl != atl || e.btl != btl || e.acr != acr || e.bcr != bcr) continue;
            out.push_back(e);
            *found = true;
            break;
        }
        out.push_back(e);
    }
    free(line);
    fclose(f);
    return ok;
}

}  // namespace

extern "C" int xc_thmpair_create(const char *path)
{
    if (!path) return XC_E_ARG;
    FILE *f = fopen(path, "w");
    if (!f) return XC_E_ARG;
    fprintf(f, "Atl\tAcr\tBtl\tBcr\tErr\tDeg\tR\tT0\tT1\tX\tT3\tT4\tY\n");
    fclose(f);
    return XC_OK;
}

extern "C" int xc_thmpair_append(const char *path, const xc_thmpair *e)
{
    if (!path || !e) return XC_E_ARG;
    if (e->atl < 0 || e->btl < 0) return XC_OK;  // "Table writing bypassed if invalid tile ID"
    // the line at whatever length fprintf would give it (a %f of a diverged entry near 1e300 runs to
    // ~310 characters), formatted once into a buffer of that size
    const char *fmt = "%d\t%d\t%d\t%d\t%d\t%f\t%f\t%f\t%f\t%f\t%f\t%f\t%f\n";
    const int n = snprintf(nullptr, 0, fmt, e->atl, e->acr, e->btl, e->bcr, e->err, e->A, e->R, e->T[0], e->T[1],
                           e->T[2], e->T[3], e->T[4], e->T[5]);
    if (n <= 0) return XC_E_ARG;
    std::vector<char> buf((size_t)n + 1);
    if (snprintf(buf.data(), buf.size(), fmt, e->atl, e->acr, e->btl, e->bcr, e->err, e->A, e->R, e->T[0], e->T[1],
                 e->T[2], e->T[3], e->T[4], e->T[5]) != n)
        return XC_E_ARG;
    // one append write: concurrent writers (one per pair job) do not interleave within a line
    const int fd = open(path, O_WRONLY | O_APPEND | O_CREAT, 0644);
    if (fd < 0) return XC_E_ARG;
    const ssize_t w = write(fd, buf.data(), (size_t)n);
    close(fd);
    return w == n ? XC_OK : XC_E_ARG;
}

extern "C" int xc_thmpair_read_all(const char *path, xc_thmpair *out, int cap)
{
    if (!path) return XC_E_ARG;
    std::vector<xc_thmpair> v;
    if (!read_lines(path, v, false, 0, 0, 0, 0, nullptr)) return XC_E_ARG;
    if (out && cap > 0 && !v.empty()) memcpy(out, v.data(), sizeof(xc_thmpair) * std::min<size_t>(cap, v.size()));
    return (int)v.size();
}

extern "C" int xc_thmpair_find(const char *path, int atl, int acr, int btl, int bcr, xc_thmpair *out)
{
    if (!path || !out) return XC_E_ARG;
    std::vector<xc_thmpair> v;
    bool found = false;
    if (!read_lines(path, v, true, atl, acr, btl, bcr, &found)) return XC_E_ARG;
    if (!found) return XC_NO_PEAK;
    *out = v[0];
    return XC_OK;
}

extern "C" int xc_thmpair_prior_angle(const xc_thmpair *e, int n, double *ang0)
{
    if ((!e && n > 0) || !ang0) return XC_E_ARG;
    std::vector<double> A;
    for (int i = 0; i < n; ++i)
        if (!e[i].err) A.push_back(e[i].A);
    const int np = (int)A.size();
    if (np < 4) return 0;
    std::sort(A.begin(), A.end());
    double m = (np & 1) ? A[np / 2] : (A[np / 2 - 1] + A[np / 2]) / 2.0;
    if (fabs(m) < 0.001) m = 0.0;
    *ang0 = m;
    return np;
}
