/*
 * fginterp.c — AutoDock 4.2.1's grid-map reader and trilinear interpolation,
 * restated in C (test infrastructure; see fginterp.h).
 * AD = autogrid_original_4_2_1_ref/autodock.
 *
 * Compiled with -ffp-contract=off: the double operations are the reference's,
 * in its order, as x86-64 SSE2 computes them.  The reference's
 * `p1u = 1.0L - p0u` is evaluated in long double and stored to a double; p0u
 * is a double in [0, 1) with at least 2^-46 between nonzero values for grids
 * of up to 1024 points, so 1 - p0u is exact in long double and the store
 * rounds once, as the double subtraction below does.
 */
#include "fginterp.h"

#include <stdio.h>
#include <string.h>

#define LINE_LEN 256             /* AD/autocomm.h:85 */
#define ENERGYPENALTY 500.0      /* AD/trilinterp.h:30 */

int oi_read_fld(const char *path, oi_grid *g)
{
    /* AD/readfield.cc:55-140 */
    char line[LINE_LEN], rec9[LINE_LEN];
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    memset(g, 0, sizeof *g);
    while (fgets(line, LINE_LEN, f) != NULL) {
        rec9[0] = 0;
        (void)sscanf(line, "%255s", rec9);
        if (strncmp(rec9, "#SPACING", 8) == 0) {
            double s = 0;
            (void)sscanf(line, "%*s %lf", &s);
            g->spacing = s;
            break;
        }
    }
    g->inv_spacing = 1. / (g->spacing);
    if (fgets(line, LINE_LEN, f) == NULL) line[0] = 0;
    (void)sscanf(line, "%*s %d %d %d", &g->num_points[0], &g->num_points[1], &g->num_points[2]);
    for (int i = 0; i < 3; ++i)
        if (g->num_points[i] % 2 != 0 || g->num_points[i] <= 0) {
            fclose(f);
            return -2;
        }
    for (int i = 0; i < 3; ++i) g->num_points1[i] = g->num_points[i] + 1;
    if (fgets(line, LINE_LEN, f) == NULL) line[0] = 0;
    (void)sscanf(line, "%*s %lf %lf %lf", &g->center[0], &g->center[1], &g->center[2]);
    fclose(f);
    for (int i = 0; i < 3; ++i) {
        g->lo[i] = g->center[i] - (g->num_points[i] / 2) * (g->spacing);
        g->hi[i] = g->center[i] + (g->num_points[i] / 2) * (g->spacing);
    }
    return 0;
}

int oi_read_map(const char *path, const oi_grid *g, int slot, int nslots, float *map)
{
    /* AD/readmap.cc:110-300 */
    char line[LINE_LEN];
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    int bad = 0;
    for (int h = 0; h < 6; ++h) {
        if (fgets(line, LINE_LEN, f) == NULL) continue;  /* warn_bad_file: a warning */
        if (h == 3) {  /* SPACING, compared as Real (float) */
            float s = 0.f;
            (void)sscanf(line, "%*s %f", &s);
            bad |= s != (float)g->spacing;
        } else if (h == 4) {
            int nel[3] = {0, 0, 0};
            (void)sscanf(line, "%*s %d %d %d", &nel[0], &nel[1], &nel[2]);
            for (int k = 0; k < 3; ++k) bad |= nel[k] != g->num_points[k];
        } else if (h == 5) {
            float c[3] = {0.f, 0.f, 0.f};
            (void)sscanf(line, "%*s %f %f %f", &c[0], &c[1], &c[2]);
            for (int k = 0; k < 3; ++k) bad |= c[k] != (float)g->center[k];
        }
        /* GRID_PARAMETER_FILE, GRID_DATA_FILE, MACROMOLECULE: names only logged */
    }
    if (bad) {
        fclose(f);
        return -3;
    }
    const int nx = g->num_points1[0], ny = g->num_points1[1], nz = g->num_points1[2];
    long nv = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (fgets(line, LINE_LEN, f) == NULL) continue;
                float v;
                if (sscanf(line, "%f", &v) != 1) continue;
                map[(((long)k * ny + j) * nx + i) * nslots + slot] = v;
                nv++;
            }
    fclose(f);
    return nv == (long)nx * ny * nz ? 0 : -4;
}

float oi_trilinterp(const oi_grid *g, const float *map, int nslots, int elec_slot, int desolv_slot, int natoms,
                    const float *xyz, const float *charge, const float *abs_charge, const int *type,
                    int outside, const int *ignore, float *elec, float *emap, double *elec_total,
                    double *emap_total)
{
    /* AD/trilinterp.cc:87-185 */
    const long nx = g->num_points1[0], ny = g->num_points1[1], nz = g->num_points1[2];
    /* A corner one past the grid's last point (an atom just inside hi whose u
     * rounds to num_points) reads 0: AutoDock's map array is a zeroed static
     * [128][128][128][MAX_MAPS] (AD/map_declare.h) and holds 0 past the
     * points a map file fills; that corner's weight is 0 or ~0. */
#define MAP(w, v, u, m) ((u) < nx && (v) < ny && (w) < nz ? (double)map[(((long)(w) * ny + (v)) * nx + (u)) * nslots + (m)] : 0.0)
    double et = 0, mt = 0;
    for (int i = 0; i < natoms; ++i) {
        if (ignore && ignore[i]) {
            if (elec) elec[i] = 0;
            if (emap) emap[i] = 0;
            continue;
        }
        double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        /* is_out_grid_info (AD/constants.h:315).  The reference tests it only
         * when the caller says some atoms may be outside, and reads past the
         * grid otherwise (undefined); here every atom is tested. */
        (void)outside;
        if (x <= g->lo[0] || x >= g->hi[0] || y <= g->lo[1] || y >= g->hi[1] || z <= g->lo[2] || z >= g->hi[2]) {
            x -= g->center[0];
            y -= g->center[1];
            z -= g->center[2];
            const double ep = (x * x + y * y + z * z) * ENERGYPENALTY;  /* sqhypotenuse, AD/constants.h:278 */
            if (elec) elec[i] = (float)ep;
            if (emap) emap[i] = (float)ep;
            et += ep;
            mt += ep;
            continue;
        }
        const int t = type[i];
        double u, v, w, p0u, p0v, p0w, p1u, p1v, p1w;
        int u0, v0, w0, u1, v1, w1;
        u = ((double)xyz[3 * i] - (double)g->lo[0]) * (double)g->inv_spacing;
        u0 = (int)u;
        u1 = u0 + 1;
        p0u = u - (double)u0;
        p1u = 1.0 - p0u;
        v = ((double)xyz[3 * i + 1] - (double)g->lo[1]) * (double)g->inv_spacing;
        v0 = (int)v;
        v1 = v0 + 1;
        p0v = v - (double)v0;
        p1v = 1.0 - p0v;
        w = ((double)xyz[3 * i + 2] - (double)g->lo[2]) * (double)g->inv_spacing;
        w0 = (int)w;
        w1 = w0 + 1;
        p0w = w - (double)w0;
        p1w = 1.0 - p0w;
        double e = 0, m = 0, d = 0;
        e += p1u * p1v * p1w * MAP(w0, v0, u0, elec_slot);
        m += p1u * p1v * p1w * MAP(w0, v0, u0, t);
        d += p1u * p1v * p1w * MAP(w0, v0, u0, desolv_slot);

        d += p0u * p1v * p1w * MAP(w0, v0, u1, desolv_slot);
        m += p0u * p1v * p1w * MAP(w0, v0, u1, t);
        e += p0u * p1v * p1w * MAP(w0, v0, u1, elec_slot);

        e += p1u * p0v * p1w * MAP(w0, v1, u0, elec_slot);
        m += p1u * p0v * p1w * MAP(w0, v1, u0, t);
        d += p1u * p0v * p1w * MAP(w0, v1, u0, desolv_slot);

        d += p0u * p0v * p1w * MAP(w0, v1, u1, desolv_slot);
        m += p0u * p0v * p1w * MAP(w0, v1, u1, t);
        e += p0u * p0v * p1w * MAP(w0, v1, u1, elec_slot);

        e += p1u * p1v * p0w * MAP(w1, v0, u0, elec_slot);
        m += p1u * p1v * p0w * MAP(w1, v0, u0, t);
        d += p1u * p1v * p0w * MAP(w1, v0, u0, desolv_slot);

        d += p0u * p1v * p0w * MAP(w1, v0, u1, desolv_slot);
        m += p0u * p1v * p0w * MAP(w1, v0, u1, t);
        e += p0u * p1v * p0w * MAP(w1, v0, u1, elec_slot);

        e += p1u * p0v * p0w * MAP(w1, v1, u0, elec_slot);
        m += p1u * p0v * p0w * MAP(w1, v1, u0, t);
        d += p1u * p0v * p0w * MAP(w1, v1, u0, desolv_slot);

        d += p0u * p0v * p0w * MAP(w1, v1, u1, desolv_slot);
        m += p0u * p0v * p0w * MAP(w1, v1, u1, t);
        e += p0u * p0v * p0w * MAP(w1, v1, u1, elec_slot);

        et += e * charge[i];
        mt += m + d * abs_charge[i];
        if (elec) elec[i] = (float)(e * charge[i]);
        if (emap) emap[i] = (float)(m + d * abs_charge[i]);
    }
#undef MAP
    if (elec_total) *elec_total = et;
    if (emap_total) *emap_total = mt;
    return (float)((double)(float)et + mt);
}

void oi_trilinterp_poses(const oi_grid *g, const float *map, int nslots, int elec_slot, int desolv_slot,
                         int natoms, int npose, const float *xyz, const float *charge, const float *abs_charge,
                         const int *type, int outside, const int *ignore, float *elec, float *emap,
                         double *elec_total, double *emap_total, float *total, int threads)
{
    if (threads <= 0) threads = 1;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int p = 0; p < npose; ++p) {
        double et, mt;
        const float r = oi_trilinterp(g, map, nslots, elec_slot, desolv_slot, natoms, xyz + (long)p * natoms * 3,
                                      charge, abs_charge, type, outside, ignore,
                                      elec ? elec + (long)p * natoms : NULL, emap ? emap + (long)p * natoms : NULL,
                                      &et, &mt);
        if (elec_total) elec_total[p] = et;
        if (emap_total) emap_total[p] = mt;
        if (total) total[p] = r;
    }
}
