// The consumer side of the maps on the host: AutoDock 4.2.1's .fld reader
// (readfield) and .map reader (readmap) into AutoDock's float map-set layout
// [z][y][x][slot].  AD = autogrid_original_4_2_1_ref/autodock in the
// reference.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "capi_common.hpp"
#include "fgmaps.h"

namespace {

constexpr int kLineLen = 256;  // AD/autocomm.h:85

// The float sscanf("%f") stores for the number at the start of `s`, and
// whether one was read.  Plain decimals with at most 8 digits after the
// point and 15 in all are m / 10^d: the double division is the nearest
// double to the decimal, and no float rounding boundary lies within 2^-53
// (relative) of such a decimal unless the decimal is that boundary, so
// rounding the double to float gives the float nearest to the decimal, which
// is what strtof returns.  Other text goes to strtof.
bool readFloat(const char *s, float &out)
{
    static const double kPow10[9] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8};
    const char *q = s;
    while (*q == ' ' || *q == '\t' || *q == '\r' || *q == '\n' || *q == '\v' || *q == '\f') ++q;
    bool neg = false;
    if (*q == '-' || *q == '+') neg = *q++ == '-';
    uint64_t m = 0;
    int digits = 0, frac = -1;
    bool fast = true;
    for (;; ++q) {
        if (*q >= '0' && *q <= '9') {
            m = m * 10 + uint64_t(*q - '0');
            if (frac >= 0) ++frac;
            if (++digits > 15 || frac > 8) {
                fast = false;
                break;
            }
        } else if (*q == '.' && frac < 0) {
            frac = 0;
        } else {
            break;
        }
    }
    if (fast && digits > 0 && (*q == 0 || std::strchr(" \t\r\n\v\f", *q))) {
        const double v = frac > 0 ? double(m) / kPow10[frac] : double(m);
        out = float(neg ? -v : v);
        return true;
    }
    char *end = nullptr;
    const float v = std::strtof(s, &end);
    if (end == s) return false;
    out = v;
    return true;
}

bool startsWith(const char *s, const char *prefix) { return std::strncmp(s, prefix, std::strlen(prefix)) == 0; }

}  // namespace

extern "C" {

int fg_read_fld(const char *path, fg_map_grid *g)
{
    // AD/readfield.cc:55-140
    if (!path || !g) return fgc::fail(FG_E_ARG, "fg_read_fld: NULL argument");
    std::memset(g, 0, sizeof *g);
    FILE *f = std::fopen(path, "r");
    if (!f) return fgc::fail(FG_E_IO, std::string("can't find or open grid data file ") + path);
    char line[kLineLen], word[kLineLen];
    while (std::fgets(line, kLineLen, f)) {
        word[0] = 0;
        (void)std::sscanf(line, "%255s", word);
        if (startsWith(word, "#SPACING")) {
            double s = 0;
            (void)std::sscanf(line, "%*s %lf", &s);
            g->spacing = s;
            break;
        }
    }
    g->inv_spacing = 1. / g->spacing;
    if (!std::fgets(line, kLineLen, f)) line[0] = 0;
    (void)std::sscanf(line, "%*s %d %d %d", &g->num_points[0], &g->num_points[1], &g->num_points[2]);
    for (int i = 0; i < 3; ++i)
        if (g->num_points[i] % 2 != 0 || g->num_points[i] <= 0) {
            std::fclose(f);
            return fgc::fail(FG_E_INPUT, "the number of user-specified grid points must be even and positive in the "
                                         "\"#NELEMENTS\" line of " + std::string(path));
        }
    for (int i = 0; i < 3; ++i) g->num_points1[i] = g->num_points[i] + 1;
    if (!std::fgets(line, kLineLen, f)) line[0] = 0;
    (void)std::sscanf(line, "%*s %lf %lf %lf", &g->center[0], &g->center[1], &g->center[2]);
    std::fclose(f);
    for (int i = 0; i < 3; ++i) {
        g->lo[i] = g->center[i] - (g->num_points[i] / 2) * g->spacing;
        g->hi[i] = g->center[i] + (g->num_points[i] / 2) * g->spacing;
    }
    return FG_OK;
}

int fg_read_map(const char *path, const fg_map_grid *g, int slot, int nslots, float *values)
{
    // AD/readmap.cc:110-300
    if (!path || !g || !values) return fgc::fail(FG_E_ARG, "fg_read_map: NULL argument");
    if (nslots <= 0 || slot < 0 || slot >= nslots) return fgc::fail(FG_E_ARG, "fg_read_map: slot out of range");
    FILE *f = std::fopen(path, "r");
    if (!f) return fgc::fail(FG_E_IO, std::string("can't find or open grid map ") + path);
    char line[kLineLen];
    std::string bad;
    for (int h = 0; h < 6; ++h) {
        if (!std::fgets(line, kLineLen, f)) continue;  // a warning in the reference
        if (h == 3) {
            float s = 0.f;
            (void)std::sscanf(line, "%*s %f", &s);
            if (s != float(g->spacing)) bad = "grid point spacing";
        } else if (h == 4) {
            int nel[3] = {0, 0, 0};
            (void)std::sscanf(line, "%*s %d %d %d", &nel[0], &nel[1], &nel[2]);
            for (int k = 0; k < 3; ++k)
                if (nel[k] != g->num_points[k]) bad = "number of grid points";
        } else if (h == 5) {
            float c[3] = {0.f, 0.f, 0.f};
            (void)std::sscanf(line, "%*s %f %f %f", &c[0], &c[1], &c[2]);
            for (int k = 0; k < 3; ++k)
                if (c[k] != float(g->center[k])) bad = "grid-map center";
        }
        // GRID_PARAMETER_FILE, GRID_DATA_FILE, MACROMOLECULE: names are only logged
    }
    if (!bad.empty()) {
        std::fclose(f);
        return fgc::fail(FG_E_INPUT, "Wrong " + bad + " in grid-map file \"" + path + "\".");
    }
    const int64_t nx = g->num_points1[0], ny = g->num_points1[1], nz = g->num_points1[2];
    int64_t nv = 0;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                if (!std::fgets(line, kLineLen, f)) continue;
                float v;
                if (!readFloat(line, v)) continue;
                values[((k * ny + j) * nx + i) * nslots + slot] = v;
                ++nv;
            }
    std::fclose(f);
    if (nv != nx * ny * nz)
        return fgc::fail(FG_E_INPUT, std::string("wrong number of values read in ") + path + ": " +
                                         std::to_string(nv) + " of " + std::to_string(nx * ny * nz));
    return FG_OK;
}

}  // extern "C"
