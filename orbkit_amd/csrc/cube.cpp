// Text body of a Gaussian cube file as orbkit writes it
// (orbkit/output/cube.py:88-96): for each x, for each y, the values of all
// data sets at z = 0, 1, ... each formatted '%13.5E' (Python's
// '%0.5E'.rjust(13)), a line break after every sixth value of the (x, y)
// run and one more at its end.  Host code; data is [nsets][nx][ny][nz]
// (C order, x slowest).
#include <cstdint>
#include <cstdio>
#include <cstring>

extern "C" size_t okmi_cube_format(const double* data, int nsets, int64_t nx, int64_t ny, int64_t nz, char* out,
                                   size_t cap)
{
    const int64_t npts = nx * ny * nz;
    // worst case per value: 13 characters (wider for 3-digit exponents) + a line break
    const size_t need = (size_t)nsets * npts * 15 + (size_t)nx * ny + 1;
    if (!out) return need;
    if (cap < need) return 0;
    char* w = out;
    char buf[32];
    for (int64_t rr = 0; rr < nx; ++rr)
        for (int64_t ss = 0; ss < ny; ++ss) {
            int64_t c = 0;
            for (int64_t tt = 0; tt < nz; ++tt) {
                const int64_t p = (rr * ny + ss) * nz + tt;
                for (int d = 0; d < nsets; ++d) {
                    const int n = std::snprintf(buf, sizeof buf, "%13.5E", data[(int64_t)d * npts + p]);
                    std::memcpy(w, buf, (size_t)n);
                    w += n;
                    if (c % 6 == 5) *w++ = '\n';
                    ++c;
                }
            }
            *w++ = '\n';
        }
    *w = '\0';
    return (size_t)(w - out);
}
