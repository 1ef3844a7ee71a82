// Result and mask files of the reference's NEMOLite2D driver, byte for byte:
//
//   nemo_write_results   common/gocean2d_io_mod.f90:98-149 (model_write):
//                        go2d_<istp>_<rank>.dat, one '(6e16.7)' record per
//                        internal T point, x fastest, and an empty record
//                        after every row
//   nemo_dump_tmask      common/model_mod.f90:251-291 (dump_tmask):
//                        tmask_<rank>.dat, '(2e16.7,1x,I3)' per internal T
//                        point, or '(3(I4,1x))' per point of the full mask
//
// Fortran's Ew.d edit descriptor is restated here from the standard: the
// value rounded to d significant digits as 0.ddddddd times 10^e, written
// right-justified in w characters as [-]0.dddddddE+ee, or [-]0.ddddddd+eee
// when |e| > 99; zero keeps exponent 0; NaN and Inf right-justified.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "nemolite2d_setup.h"

namespace {

// One Ew.d field (w <= 32) into out[0..w).
void fortran_e(double x, int w, int d, char *out)
{
    char buf[64];
    std::memset(out, ' ', (size_t)w);
    if (std::isnan(x) || std::isinf(x)) {
        const char *s = std::isnan(x) ? "NaN" : (x > 0 ? "Inf" : "-Inf");
        const size_t n = std::strlen(s);
        std::memcpy(out + w - n, s, n);
        return;
    }
    int e10 = 0;
    std::string digits;
    if (x == 0.0) {
        digits.assign((size_t)d, '0');
    } else {
        // d significant digits, correctly rounded: d.ddd...e+XX
        std::snprintf(buf, sizeof buf, "%.*e", d - 1, std::fabs(x));
        const char *ep = std::strchr(buf, 'e');
        digits.push_back(buf[0]);
        for (const char *p = buf + 2; p < ep; p++) digits.push_back(*p);
        e10 = std::atoi(ep + 1) + 1;  // 0.ddd form
    }
    std::string s = std::signbit(x) ? "-0." : "0.";
    s += digits;
    if (std::abs(e10) <= 99) {
        std::snprintf(buf, sizeof buf, "E%c%02d", e10 < 0 ? '-' : '+', std::abs(e10));
    } else {
        std::snprintf(buf, sizeof buf, "%c%03d", e10 < 0 ? '-' : '+', std::abs(e10));
    }
    s += buf;
    if ((int)s.size() > w) {
        std::memset(out, '*', (size_t)w);
        return;
    }
    std::memcpy(out + w - s.size(), s.data(), s.size());
}

}  // namespace

extern "C" {

void nemo_fortran_e(double x, int w, int d, char *out)
{
    if (w <= 0 || w > 32 || d <= 0 || d > 17) {
        std::memset(out, '*', w > 0 ? (size_t)w : 0);
        return;
    }
    fortran_e(x, w, d, out);
}

int nemo_write_results(const char *path, const double *ht, const double *sshn,
                       const double *un, const double *vn, int width, int xs, int xe,
                       int ys, int ye, double dx, double dy)
{
    FILE *f = std::fopen(path, "w");
    if (!f) return -1;
    char line[6 * 16 + 2];
    line[6 * 16] = '\n';
    for (int jj = ys; jj <= ye; jj++) {
        for (int ji = xs; ji <= xe; ji++) {
            const size_t k = (size_t)jj * width + ji;
            // x, y of the T point (1-based Fortran index minus one, times the
            // spacing), depth, ssh, u and v averaged onto the T point
            const double v[6] = {ji * dx, jj * dy, ht[k], sshn[k], 0.5 * (un[k - 1] + un[k]),
                                 0.5 * (vn[k - width] + vn[k])};
            for (int c = 0; c < 6; c++) fortran_e(v[c], 16, 7, line + 16 * c);
            std::fwrite(line, 1, sizeof line - 1, f);
        }
        std::fputc('\n', f);
    }
    return std::fclose(f) == 0 ? 0 : -1;
}

int nemo_dump_tmask(const char *path, const int *tmask, int width, int M, int N, int xs,
                    int xe, int ys, int ye, double dx, double dy, int raw)
{
    FILE *f = std::fopen(path, "w");
    if (!f) return -1;
    char buf[64];
    if (raw) {
        for (int jj = 0; jj < N; jj++) {
            for (int ji = 0; ji < M; ji++) {
                // the trailing 1x of the record writes nothing
                std::fprintf(f, "%4d %4d %4d\n", ji + 1, jj + 1, tmask[(size_t)jj * width + ji]);
            }
            std::fputc('\n', f);
        }
    } else {
        for (int jj = ys; jj <= ye; jj++) {
            for (int ji = xs; ji <= xe; ji++) {
                fortran_e(ji * dx, 16, 7, buf);
                fortran_e(jj * dy, 16, 7, buf + 16);
                std::snprintf(buf + 32, sizeof buf - 32, " %3d\n", tmask[(size_t)jj * width + ji]);
                std::fputs(buf, f);
            }
            std::fputc('\n', f);
        }
    }
    return std::fclose(f) == 0 ? 0 : -1;
}

}  // extern "C"
