/*
 * fginterp — CPU restatement of AutoDock 4.2.1's consumer of grid maps: the
 * .fld reader, the .map reader and trilinear interpolation of a ligand's
 * atoms in the maps.
 *
 * TEST INFRASTRUCTURE ONLY, like the rest of oracle/: the checker of the HIP
 * product path, never linked by it.
 *
 * AD = autogrid_original_4_2_1_ref/autodock in the reference.  Types follow
 * AutoDock's default build (Real = float, typedefs.h:56-66): map values,
 * ligand coordinates and charges are floats, the interpolation arithmetic is
 * double (trilinterp.cc:87-185).
 */
#ifndef FGINTERP_H
#define FGINTERP_H

#ifdef __cplusplus
extern "C" {
#endif

/* GridMapSetInfo as readfield leaves it (AD/grid.h:36-54). */
typedef struct {
    double spacing, inv_spacing;
    int num_points[3];   /* NELEMENTS: points per axis minus 1, even */
    int num_points1[3];  /* points per axis */
    double lo[3], hi[3], center[3];
} oi_grid;

/* readfield (AD/readfield.cc:55-140).  0, or -1 when the file cannot be
 * opened, -2 for an odd or non-positive NELEMENTS. */
int oi_read_fld(const char *path, oi_grid *g);

/* readmap (AD/readmap.cc:49-300): header checks against g (SPACING, NELEMENTS
 * and CENTER must agree, as floats), then one value per line read with
 * sscanf "%f" into map[((k * ny1 + j) * nx1 + i) * nslots + slot].  0, -1
 * when the file cannot be opened, -3 for a header mismatch, -4 when the
 * number of values read is not the number of grid points. */
int oi_read_map(const char *path, const oi_grid *g, int slot, int nslots, float *map);

/* trilinterp (AD/trilinterp.cc:46-185) over atoms [0, natoms) of one pose:
 * xyz[3 * i + k] (float), charge, abs_charge (float), type = map slot of the
 * atom's affinity map, ignore (may be NULL).  elec_slot and desolv_slot are
 * the slots of the electrostatic and desolvation maps (the reference's
 * ElecMap / DesolvMap).  Atoms outside the grid get the distance-squared
 * penalty (the reference tests this only when `outside` is set and reads
 * past the grid otherwise; here every atom is tested, so `outside` changes
 * nothing).  elec/emap (may be NULL) receive the
 * per-atom floats; the totals are double; returns (float)elec_total +
 * emap_total as the reference's Real. */
float oi_trilinterp(const oi_grid *g, const float *map, int nslots, int elec_slot, int desolv_slot, int natoms,
                    const float *xyz, const float *charge, const float *abs_charge, const int *type,
                    int outside, const int *ignore, float *elec, float *emap, double *elec_total,
                    double *emap_total);

/* The same over npose poses sharing the atom list (xyz is [pose][atom][3]),
 * parallel over poses; per-pose totals and returned values. */
void oi_trilinterp_poses(const oi_grid *g, const float *map, int nslots, int elec_slot, int desolv_slot,
                         int natoms, int npose, const float *xyz, const float *charge, const float *abs_charge,
                         const int *type, int outside, const int *ignore, float *elec, float *emap,
                         double *elec_total, double *emap_total, float *total, int threads);

#ifdef __cplusplus
}
#endif

#endif
