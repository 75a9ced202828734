/*
 * fgoracle.c — FP64 CPU restatement of FastGrid 4.2.1 (test infrastructure).
 * See fgoracle.h.  FG = fastgrid_master/fastgrid, AC = fastgrid_master/autodock.
 *
 * Semantics pinned where the reference has undefined behaviour:
 *  - BondVectors leaves rexp/rvector/rvector2 uninitialised for unbonded
 *    atoms (FG/BondVectors.cpp:34).  Here they are zero, which is what the
 *    reference computes with zero-filled allocations (MALLOC_PERTURB_=255)
 *    and what AutoGrid 4.2.1 computes (zeroed statics).
 *  - DesolvExpFunc leaves func[0] uninitialised (FG/DesolvExpFunc.cpp:29,37);
 *    here it is 0 for the same reason.
 *  - The nearest donor hydrogen (FG/CalculateGridmaps.cpp:65-82, 248) breaks
 *    ties toward the lowest atom index (the reference's --no-nns semantics).
 */
#include "fgoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#include <omp.h>

#define APPROX_ZERO 1.0E-6      /* AC/autocomm.h:74 */
#define BIG 1.0E12              /* AC/autocomm.h:75 */
#define PI_ 3.14159265358979323846
#define EINTCLAMP 100000.       /* AC/autocomm.h:98 */
#define LINE_LEN 256            /* AC/autocomm.h:85 */

/* ------------------------------------------------------------------ */
/* Built-in force-field parameters (AC/default_parameters.h:1-82).     */
/* ------------------------------------------------------------------ */
typedef struct { const char *t; double R, eps, vol, sol, RHB, epsHB; int hb, bond; } or_defpar;

static const or_defpar kAtomPar[] = {
    {"H",  2.00, 0.020,  0.0000, 0.00051, 0.0, 0.0, 0, 3}, {"HD", 2.00, 0.020,  0.0000, 0.00051, 0.0, 0.0, 2, 3},
    {"HS", 2.00, 0.020,  0.0000, 0.00051, 0.0, 0.0, 1, 3}, {"C",  4.00, 0.150, 33.5103, -0.00143, 0.0, 0.0, 0, 0},
    {"A",  4.00, 0.150, 33.5103, -0.00052, 0.0, 0.0, 0, 0}, {"N",  3.50, 0.160, 22.4493, -0.00162, 0.0, 0.0, 0, 1},
    {"NA", 3.50, 0.160, 22.4493, -0.00162, 1.9, 5.0, 4, 1}, {"NS", 3.50, 0.160, 22.4493, -0.00162, 1.9, 5.0, 3, 1},
    {"OA", 3.20, 0.200, 17.1573, -0.00251, 1.9, 5.0, 5, 2}, {"OS", 3.20, 0.200, 17.1573, -0.00251, 1.9, 5.0, 3, 2},
    {"F",  3.09, 0.080, 15.4480, -0.00110, 0.0, 0.0, 0, 4}, {"Mg", 1.30, 0.875,  1.5600, -0.00110, 0.0, 0.0, 0, 4},
    {"MG", 1.30, 0.875,  1.5600, -0.00110, 0.0, 0.0, 0, 4}, {"P",  4.20, 0.200, 38.7924, -0.00110, 0.0, 0.0, 0, 5},
    {"SA", 4.00, 0.200, 33.5103, -0.00214, 2.5, 1.0, 5, 6}, {"S",  4.00, 0.200, 33.5103, -0.00214, 0.0, 0.0, 0, 6},
    {"Cl", 4.09, 0.276, 35.8235, -0.00110, 0.0, 0.0, 0, 4}, {"CL", 4.09, 0.276, 35.8235, -0.00110, 0.0, 0.0, 0, 4},
    {"Ca", 1.98, 0.550,  2.7700, -0.00110, 0.0, 0.0, 0, 4}, {"CA", 1.98, 0.550,  2.7700, -0.00110, 0.0, 0.0, 0, 4},
    {"Mn", 1.30, 0.875,  2.1400, -0.00110, 0.0, 0.0, 0, 4}, {"MN", 1.30, 0.875,  2.1400, -0.00110, 0.0, 0.0, 0, 4},
    {"Fe", 1.30, 0.010,  1.8400, -0.00110, 0.0, 0.0, 0, 4}, {"FE", 1.30, 0.010,  1.8400, -0.00110, 0.0, 0.0, 0, 4},
    {"Zn", 1.48, 0.550,  1.7000, -0.00110, 0.0, 0.0, 0, 4}, {"ZN", 1.48, 0.550,  1.7000, -0.00110, 0.0, 0.0, 0, 4},
    {"Br", 4.33, 0.389, 42.5661, -0.00110, 0.0, 0.0, 0, 4}, {"BR", 4.33, 0.389, 42.5661, -0.00110, 0.0, 0.0, 0, 4},
    {"I",  4.72, 0.550, 55.0585, -0.00110, 0.0, 0.0, 0, 4}, {"Z",  4.00, 0.150, 33.5103, -0.00143, 0.0, 0.0, 0, 0},
    {"G",  4.00, 0.150, 33.5103, -0.00143, 0.0, 0.0, 0, 0}, {"GA", 4.00, 0.150, 33.5103, -0.00052, 0.0, 0.0, 0, 0},
    {"J",  4.00, 0.150, 33.5103, -0.00143, 0.0, 0.0, 0, 0}, {"Q",  4.00, 0.150, 33.5103, -0.00143, 0.0, 0.0, 0, 0},
};
/* FE coefficients: AD4.0 (AC/default_parameters.h:2-6), AD4.1 (:43-47). */
static const double kFE40[5] = {0.1560, 0.0974, 0.1465, 0.1159, 0.2744};
static const double kFE41[5] = {0.1662, 0.1209, 0.1406, 0.1322, 0.2983};

/* FG/ParameterLibrary.cpp:67-78 */
static unsigned lib_hash(const char *k)
{
    size_t l = strlen(k);
    if (l == 0) return 0;
    if (l == 1) return (unsigned char)k[0];
    return (unsigned char)k[0] + 256u * (unsigned char)k[1];
}

static or_param *lib_find(or_lib *L, const char *k)
{
    int s = L->slot[lib_hash(k)];
    return s ? &L->e[s - 1] : NULL;
}

/* FG/ParameterLibrary.cpp:80-86: insert overwrites the whole entry. */
static void lib_insert(or_lib *L, const or_param *p)
{
    unsigned h = lib_hash(p->type);
    if (!L->slot[h]) {
        if (L->n >= 256) return;
        L->slot[h] = ++L->n;
    }
    L->e[L->slot[h] - 1] = *p;
}

/* FG/ParameterLibrary.cpp:196-218: epsij/epsijHB are weighted by the
 * coefficients current when the line is read. */
static void lib_insert_atom(or_lib *L, const char *t, double R, double eps, double vol, double sol,
                            double RHB, double epsHB, int hb, int rec, int map, int bond)
{
    or_param p;
    memset(&p, 0, sizeof p);
    strncpy(p.type, t, 7);
    p.Rij = R; p.epsij = eps * L->coeff_vdW; p.vol = vol; p.solpar = sol;
    p.RijHB = RHB; p.epsijHB = epsHB * L->coeff_hbond;
    p.hbond = hb <= OR_A2 ? hb : OR_NON;
    p.recIndex = rec; p.mapIndex = map; p.bondIndex = bond;
    lib_insert(L, &p);
}

/* FG/ParameterLibrary.cpp:34-58 */
static void lib_init(or_lib *L, int v4)
{
    memset(L, 0, sizeof *L);
    const double *fe = v4 ? kFE40 : kFE41;
    L->coeff_vdW = fe[0]; L->coeff_hbond = fe[1]; L->coeff_estat = fe[2];
    L->coeff_desolv = fe[3]; L->coeff_tors = fe[4];
    for (size_t i = 0; i < sizeof kAtomPar / sizeof kAtomPar[0]; i++) {
        const or_defpar *d = &kAtomPar[i];
        lib_insert_atom(L, d->t, d->R, d->eps, d->vol, d->sol, d->RHB, d->epsHB, d->hb, -1, -1, d->bond);
    }
}

/* FG/ParameterLibrary.cpp:124-284: one line of an AD4 parameter file. */
static void lib_read_line(or_lib *L, const char *line)
{
    char kw[LINE_LEN];
    int j = 0;
    for (; line[j] && line[j] != ' ' && line[j] != '\t' && line[j] != '\n' && j < LINE_LEN - 1; j++)
        kw[j] = (char)tolower((unsigned char)line[j]);
    kw[j] = 0;
    if (kw[0] == 0 || kw[0] == '#') return;
    static const char *tok[6] = {"FE_coeff_vdW", "FE_coeff_hbond", "FE_coeff_estat", "FE_coeff_desolv", "FE_coeff_tors", "atom_par"};
    int t = -1;
    for (int i = 0; i < 6 && t < 0; i++)
        if (strncasecmp(tok[i], kw, (size_t)j) == 0) t = i;
    double v;
    switch (t) {
    case 0: if (sscanf(line, "%*s %lf", &v) == 1) L->coeff_vdW = v; break;
    case 1: if (sscanf(line, "%*s %lf", &v) == 1) L->coeff_hbond = v; break;
    case 2: if (sscanf(line, "%*s %lf", &v) == 1) L->coeff_estat = v; break;
    case 3: if (sscanf(line, "%*s %lf", &v) == 1) L->coeff_desolv = v; break;
    case 4: if (sscanf(line, "%*s %lf", &v) == 1) L->coeff_tors = v; break;
    case 5: {
        char ty[LINE_LEN];
        double R = 0, eps = 0, vol = 0, sol = 0, RHB = 0, epsHB = 0;
        int hb = 0, rec = 0, map = 0, bond = 0;
        int nf = sscanf(line, "%*s %255s %lf %lf %lf %lf %lf %lf %d %d %d %d", ty, &R, &eps, &vol, &sol,
                        &RHB, &epsHB, &hb, &rec, &map, &bond);
        if (nf < 2) return;
        ty[7] = 0;
        lib_insert_atom(L, ty, R, eps, vol, sol, RHB, epsHB, hb, rec, map, bond);
        break;
    }
    default: break;
    }
}

static int lib_load_file(or_lib *L, const char *fn)
{
    FILE *f = fopen(fn, "r");
    if (!f) return -1;
    char line[LINE_LEN];
    while (fgets(line, sizeof line, f)) lib_read_line(L, line);
    fclose(f);
    return 0;
}

static int lib_rec_index(or_lib *L, const char *k)
{
    or_param *p = lib_find(L, k);
    return p ? p->recIndex : -1;
}

/* ------------------------------------------------------------------ */
/* Numeric helpers (FG/autogrid.h:106-131, FG/math/Math.h:51-98).      */
/* ------------------------------------------------------------------ */
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double sat(double f) { return dmin(dmax(f, 0), 1); }
static inline double acos_c(double f) { return f > 1 ? 0 : (f < -1 ? PI_ : acos(f)); }
static inline int bin_index(double r)
{
    int i = (int)(r * 100.0);
    return i < OR_MAX_DIST - 1 ? i : OR_MAX_DIST - 1;
}

double or_round3(double a)
{
    return fabs(a) < 0.0005 ? 0 : floor(a * 1000 + 0.5) * 0.001;
}

void or_ddd_table(double *eps)
{
    for (int i = 0; i < OR_MAX_DIST; i++) {
        double d = (double)i * (1 / 100.0);
        double E = exp(-0.3153767175 * d);
        eps[i] = (332.0 + (332 * 7.7839) * E) / (78.4 + -66.57180475 * E);
    }
}

/* FG/DesolvExpFunc.cpp:27-43; S[0] = 0 (see header). */
void or_desolv_table(double coeffDesolv, double *S)
{
    double sigma = 3.6;
    double m = -1 / (2 * sigma * sigma);
    S[0] = 0;
    for (int i = 1; i < OR_MAX_DIST; i++) {
        double r = (double)i * (1 / 100.0);
        S[i] = exp((r * r) * m);
        S[i] *= coeffDesolv;
    }
}

/* ------------------------------------------------------------------ */
/* GPF / PDBQT parsing (FG/InputDataLoader.cpp:100-938).               */
/* ------------------------------------------------------------------ */
enum { G_NULL = 0, G_COMMENT, G_RECEPTOR, G_GRIDFLD, G_NPTS, G_SPACING, G_GRIDCENTER, G_LIGAND_TYPES,
       G_MAP, G_NBP_COEFFS, G_NBP_R_EPS, G_ELECMAP, G_DIEL, G_FMAP, G_DISORDER, G_SMOOTH, G_SOL_PAR,
       G_CONSTANT, G_COVALENTMAP, G_RECEPTOR_TYPES, G_PARAM_FILE, G_DSOLVMAP, G_QASP };

/* FG/InputDataLoader.cpp:799-865 */
static int gpf_keyword(const char *line)
{
    const char *sp = strchr(line, ' ');
    size_t l;
    if (sp) l = (size_t)(sp - line);
    else {
        const char *tb = strchr(line, '\t');
        l = tb ? (size_t)(tb - line) : strlen(line);
    }
    char c[LINE_LEN];
    if (l >= LINE_LEN) l = LINE_LEN - 1;
    for (size_t i = 0; i < l; i++) c[i] = (char)tolower((unsigned char)line[i]);
    c[l] = 0;
    if (c[0] == '\n' || c[0] == 0) return G_NULL;
    if (c[0] == '#') return G_COMMENT;
    static const struct { const char *k; int t; } tab[] = {
        {"receptor_types", G_RECEPTOR_TYPES}, {"receptor", G_RECEPTOR}, {"gridfld", G_GRIDFLD},
        {"npts", G_NPTS}, {"spacing", G_SPACING}, {"gridcenter", G_GRIDCENTER}, {"types", G_LIGAND_TYPES},
        {"ligand_types", G_LIGAND_TYPES}, {"map", G_MAP}, {"elecmap", G_ELECMAP}, {"dsolvmap", G_DSOLVMAP},
        {"covalentmap", G_COVALENTMAP}, {"nbp_coeffs", G_NBP_COEFFS}, {"nbp_r_eps", G_NBP_R_EPS},
        {"dielectric", G_DIEL}, {"qasp", G_QASP}, {"fmap", G_FMAP}, {"disorder_h", G_DISORDER},
        {"smooth", G_SMOOTH}, {"sol_par", G_SOL_PAR}, {"constant", G_CONSTANT}, {"parameter_file", G_PARAM_FILE},
    };
    for (size_t i = 0; i < sizeof tab / sizeof tab[0]; i++)
        if (strncmp(c, tab[i].k, strlen(tab[i].k)) == 0) return tab[i].t;
    return -1;
}

/* FG/InputDataLoader.cpp:890-938: words after the keyword. */
static int parse_types(char *line, char **words, int maxw)
{
    char *p = line;
    int n = 0, kw = 0;
    for (;;) {
        while (isspace((unsigned char)*p)) p++;
        if (!*p) return n;
        if (!kw) kw = 1;
        else words[n++] = p;
        while (!isspace((unsigned char)*p) && *p) p++;
        if (!*p) return n;
        *p++ = 0;
        if (n >= maxw) return n;
    }
}

/* FG/InputDataLoader.cpp:868-887 */
static int check_size(or_input *in, int nm1, int *fatal)
{
    if (nm1 < 0) { *fatal = 1; snprintf(in->err, sizeof in->err, "Negative number of grid elements"); }
    else if (nm1 > OR_MAX_GRID_PTS) nm1 = OR_MAX_GRID_PTS;
    else if (nm1 % 2 == 1) nm1 -= 1;
    return nm1;
}

#define FAIL(code, ...) do { snprintf(in->err, sizeof in->err, __VA_ARGS__); ret = (code); goto done; } while (0)

static void apply_param_file(or_input *in)
{
    if (in->paramFile[0] && lib_load_file(&in->lib, in->paramFile) != 0)
        snprintf(in->err, sizeof in->err, "cannot open parameter file %s", in->paramFile);
}

int or_load(or_input *in, const char *gpf, int v4, int agParamOrder, FILE *log)
{
    memset(in, 0, sizeof *in);
    in->log = log;
    lib_init(&in->lib, v4);
    in->solparQ = .01097;           /* FG/InputDataLoader.cpp:74 */
    in->spacing = 0.375;
    in->covHalfWidthSqInv = 1.0;
    in->covBarrier = 1000.0;
    in->P = -1;
    strncpy(in->gpfName, gpf ? gpf : "", OR_NAME - 1);

    int ret = 0, mapIndex = -1, recTypesGpf = 0, hasRecTypes = 0, fatal = 0;
    double cmean[3] = {0, 0, 0};
    FILE *xyzf = NULL;
    FILE *G = (gpf && gpf[0]) ? fopen(gpf, "r") : stdin;
    if (!G) { snprintf(in->err, sizeof in->err, "can't open GPF %s", gpf); return 911; }

    char L[LINE_LEN];
    while (fgets(L, LINE_LEN, G)) {
        int kw = gpf_keyword(L);
        if (log) fprintf(log, "GPF> %s", L);
        if (kw == -1) continue;
        if (kw != G_NULL && kw != G_COMMENT) {
            char *h = strchr(L, '#');
            if (h) *h = 0;
        }
        switch (kw) {
        case G_RECEPTOR: {
            sscanf(L, "%*s %127s", in->receptor);
            FILE *R = fopen(in->receptor, "r");
            if (!R) FAIL(-2, "can't find or open receptor PDBQT file \"%s\"", in->receptor);
            int cap = 1, ch;
            while ((ch = fgetc(R)) != EOF) if (ch == '\n') cap++;
            rewind(R);
            in->x = malloc(sizeof(double) * cap); in->y = malloc(sizeof(double) * cap);
            in->z = malloc(sizeof(double) * cap); in->w = malloc(sizeof(double) * cap);
            in->charge = malloc(sizeof(double) * cap); in->vol = malloc(sizeof(double) * cap);
            in->solpar = malloc(sizeof(double) * cap);
            in->type = malloc(sizeof(int) * cap); in->hbond = malloc(sizeof(int) * cap);
            char line[LINE_LEN], rec[LINE_LEN] = "";
            double csum[3] = {0, 0, 0}, qmax = -BIG, qmin = BIG;
            int ia = 0;
            while (fgets(line, LINE_LEN, R)) {
                sscanf(line, "%6s", rec);
                if (strncmp(rec, "ATOM", 4) && strncmp(rec, "HETA", 4) && strncmp(rec, "CHAR", 4)) continue;
                size_t ll = strlen(line);
                char ty[LINE_LEN] = "";
                in->x[ia] = in->y[ia] = in->z[ia] = in->charge[ia] = 0;
                if (ll > 30) sscanf(&line[30], "%lf", &in->x[ia]);
                if (ll > 38) sscanf(&line[38], "%lf", &in->y[ia]);
                if (ll > 46) sscanf(&line[46], "%lf", &in->z[ia]);
                if (ll > 70) sscanf(&line[70], "%lf", &in->charge[ia]);
                if (ll > 77) sscanf(&line[77], "%255s", ty);
                ty[7] = 0;
                or_param *p = lib_find(&in->lib, ty);
                if (!p) { fclose(R); FAIL(-1, "receptor file contains unknown type: '%s'", ty); }
                if (p->recIndex < 0) {
                    strncpy(in->recTypes[in->numRecTypes], p->type, 3);
                    in->recTypes[in->numRecTypes][2] = 0;
                    p->recIndex = in->numRecTypes++;
                }
                in->type[ia] = p->recIndex;
                in->solpar[ia] = p->solpar;
                in->vol[ia] = p->vol;
                in->hbond[ia] = p->hbond;
                qmax = dmax(qmax, in->charge[ia]);
                qmin = dmin(qmin, in->charge[ia]);
                csum[0] += in->x[ia]; csum[1] += in->y[ia]; csum[2] += in->z[ia];
                ia++;
            }
            fclose(R);
            if (hasRecTypes && in->numRecTypes != recTypesGpf)
                FAIL(-2, "The number of atom types found in the receptor PDBQT (%d) does not match receptor_types (%d)",
                     in->numRecTypes, recTypesGpf);
            in->N = ia;
            if (qmax == 0 && qmin == 0) FAIL(-2, "No partial atomic charges were found in %s", in->receptor);
            for (int k = 0; k < 3; k++) cmean[k] = csum[k] / in->N;
            break;
        }
        case G_GRIDFLD: {
            sscanf(L, "%*s %127s", in->fld);
            char *f = strstr(in->fld, ".fld");
            if (!f) FAIL(-2, "Grid data file needs the extension \".fld\" for AVS input");
            strncpy(in->xyz, in->fld, OR_NAME - 1);
            char *g = strstr(in->xyz, "fld");
            g[0] = 'x'; g[1] = 'y'; g[2] = 'z';
            xyzf = fopen(in->xyz, "w");
            if (!xyzf) FAIL(-2, "can't create grid extrema data file %s", in->xyz);
            break;
        }
        case G_NPTS: {
            int nm1[3] = {0, 0, 0};
            sscanf(L, "%*s %d %d %d", &nm1[0], &nm1[1], &nm1[2]);
            for (int k = 0; k < 3; k++) nm1[k] = check_size(in, nm1[k], &fatal);
            if (fatal) { ret = -2; goto done; }
            for (int k = 0; k < 3; k++) { in->div2[k] = nm1[k] / 2; in->n[k] = nm1[k] + 1; }
            in->P = in->n[0] * in->n[1] * in->n[2];
            break;
        }
        case G_SPACING: sscanf(L, "%*s %lf", &in->spacing); break;
        case G_GRIDCENTER: {
            char tok[LINE_LEN] = "";
            sscanf(L, "%*s %255s", tok);
            if (strncmp(tok, "auto", 4) == 0) memcpy(in->center, cmean, sizeof cmean);
            else sscanf(L, "%*s %lf %lf %lf", &in->center[0], &in->center[1], &in->center[2]);
            for (int ia = 0; ia < in->N; ia++) {
                in->x[ia] -= in->center[0]; in->y[ia] -= in->center[1]; in->z[ia] -= in->center[2];
            }
            if (xyzf) {
                for (int k = 0; k < 3; k++) {
                    double ext = (double)in->div2[k] * in->spacing;
                    fprintf(xyzf, "%.3lf %.3lf\n", in->center[k] - ext, in->center[k] + ext);
                }
                fclose(xyzf);
                xyzf = NULL;
            }
            break;
        }
        case G_LIGAND_TYPES: {
            char *words[OR_MAX_MAPS];
            int na = parse_types(L, words, OR_MAX_ATOM_TYPES);
            char lt[OR_MAX_MAPS][3];
            memset(lt, 0, sizeof lt);
            for (int i = 0; i < na; i++) strncpy(lt[i], words[i], 2);
            for (int i = 0; i < na; i++) {
                or_param *p = lib_find(&in->lib, lt[i]);
                if (!p) FAIL(-1, "unknown ligand atom type %s", lt[i]);
                p->mapIndex = i;
            }
            for (int i = 0; i < OR_MAX_MAPS + 2; i++) free(in->maps[i].energies);
            memset(in->maps, 0, sizeof in->maps);
            in->numAtomMaps = na;
            in->numMaps = na + 2;
            if (in->P < 0) FAIL(-2, "You need to set the number of grid points using \"npts\" before \"ligand_types\"");
            for (int i = 0; i < na; i++) {
                or_map *m = &in->maps[i];
                or_param *p = lib_find(&in->lib, lt[i]);
                memcpy(m->type, lt[i], 3);
                if (strcmp(lt[i], "Z") == 0) m->isCovalent = 1;
                m->atomType = p->mapIndex;
                m->solparProbe = p->solpar; m->volProbe = p->vol;
                m->Rij = p->Rij; m->epsij = p->epsij; m->hbond = p->hbond;
                m->RijHB = p->RijHB; m->epsijHB = p->epsijHB;
                if (m->hbond > 0) m->isHBonder = 1;
                for (int j = 0; j < in->numRecTypes; j++) {
                    or_param *r = lib_find(&in->lib, in->recTypes[j]);
                    if (!r) FAIL(-2, "unknown receptor type %s", in->recTypes[j]);
                    m->nbpR[j] = (m->Rij + r->Rij) / 2;
                    m->nbpEps[j] = sqrt(m->epsij * r->epsij);
                    m->xA[j] = 12; m->xB[j] = 6; m->hbonder[j] = 0;
                    if (m->hbond > OR_D1 && (r->hbond == OR_DS || r->hbond == OR_D1)) {
                        m->xB[j] = 10; m->hbonder[j] = 1; m->isHBonder = 1;
                        m->nbpR[j] = m->RijHB; m->nbpEps[j] = m->epsijHB;
                    } else if ((m->hbond == OR_DS || m->hbond == OR_D1) && r->hbond > OR_D1) {
                        m->xB[j] = 10; m->hbonder[j] = 1; m->isHBonder = 1;
                        m->nbpR[j] = r->RijHB; m->nbpEps[j] = r->epsijHB;
                    }
                }
            }
            break;
        }
        case G_RECEPTOR_TYPES: {
            char *words[OR_NUM_TYPES];
            in->numRecTypes = parse_types(L, words, OR_MAX_ATOM_TYPES);
            recTypesGpf = in->numRecTypes;
            hasRecTypes = 1;
            for (int i = 0; i < in->numRecTypes; i++) { strncpy(in->recTypes[i], words[i], 2); in->recTypes[i][2] = 0; }
            for (int i = 0; i < in->numRecTypes; i++) {
                or_param *p = lib_find(&in->lib, words[i]);
                if (!p) FAIL(-2, "Unknown receptor type: \"%s\"", words[i]);
                p->recIndex = i;
            }
            break;
        }
        case G_SOL_PAR: {
            char ty[LINE_LEN] = "";
            double tv = 0, ts = 0;
            sscanf(L, "%*s %255s %lf %lf", ty, &tv, &ts);
            ty[7] = 0;
            or_param *p = lib_find(&in->lib, ty);
            if (p) {
                p->vol = tv; p->solpar = ts;
                if (p->mapIndex >= 0) in->maps[p->mapIndex].solparProbe = ts;
            }
            break;
        }
        case G_MAP:
            ++mapIndex;
            if (mapIndex >= in->numAtomMaps) FAIL(-2, "Too many \"map\" keywords");
            sscanf(L, "%*s %127s", in->maps[mapIndex].filename);
            break;
        case G_ELECMAP:
            if (in->numMaps < 2) FAIL(-2, "elecmap before ligand_types");
            sscanf(L, "%*s %127s", in->maps[in->numMaps - 2].filename);
            break;
        case G_DSOLVMAP:
            if (in->numMaps < 2) FAIL(-2, "dsolvmap before ligand_types");
            sscanf(L, "%*s %127s", in->maps[in->numMaps - 1].filename);
            break;
        case G_COVALENTMAP: {
            double hw = 0;
            sscanf(L, "%*s %lf %lf %lf %lf %lf", &hw, &in->covBarrier, &in->covPoint[0], &in->covPoint[1], &in->covPoint[2]);
            in->covHalfWidthSqInv = 1 / (hw * hw);
            for (int k = 0; k < 3; k++) in->covPoint[k] -= in->center[k];
            break;
        }
        case G_DISORDER: in->disorderH = 1; break;
        case G_SMOOTH: sscanf(L, "%*s %lf", &in->rSmooth); break;
        case G_QASP: sscanf(L, "%*s %lf", &in->solparQ); break;
        case G_DIEL: {
            double diel = 0;
            sscanf(L, "%*s %lf", &diel);
            if (diel < 0) {
                in->distDepDiel = 1;
                or_ddd_table(in->eps);
            } else {
                in->distDepDiel = 0;
                if (diel <= APPROX_ZERO) diel = 40;
                in->invDielCal = 332 / diel;
            }
            break;
        }
        case G_FMAP: sscanf(L, "%*s %127s", in->fmap); break;
        case G_PARAM_FILE:
            sscanf(L, "%*s %127s", in->paramFile);
            /* AutoGrid 4.2.1 reads the library here; FastGrid after the GPF. */
            if (agParamOrder) apply_param_file(in);
            break;
        default: break;
        }
    }
    if (mapIndex < in->numAtomMaps - 1) FAIL(-2, "Too few \"map\" keywords");
    if (in->numMaps < 2 || !in->maps[in->numMaps - 1].filename[0]) FAIL(-2, "No desolvation map file defined.");
    if (!in->maps[in->numMaps - 2].filename[0]) FAIL(-2, "No electrostatic map file defined.");

    /* FG/InputDataLoader.cpp:789-796 */
    for (int ia = 0; ia < in->N; ia++)
        in->w[ia] = in->distDepDiel ? in->charge[ia] * in->lib.coeff_estat
                                    : in->charge[ia] * in->lib.coeff_estat * in->invDielCal;
    /* FG/main.cpp:170-171: FastGrid loads the parameter file after parsing. */
    if (!agParamOrder) apply_param_file(in);
done:
    if (xyzf) fclose(xyzf);
    if (G != stdin) fclose(G);
    return ret;
}

/* ------------------------------------------------------------------ */
/* Bond vectors (FG/BondVectors.cpp:45-393), zero-initialised.         */
/* ------------------------------------------------------------------ */
static inline void v_sub(double *o, const double *a, const double *b) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static inline double v_msq(const double *a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
static inline double v_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void v_scale(double *a, double f) { a[0] *= f; a[1] *= f; a[2] *= f; }
static inline void v_cross(double *o, const double *a, const double *b)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
/* Normalisation with the reference's clamp: rd2 < 1e-6 -> 1e-6, inv = 1/sqrt. */
static inline void v_norm_clamped(double *a)
{
    double rd2 = v_msq(a);
    if (rd2 < APPROX_ZERO) rd2 = APPROX_ZERO;
    v_scale(a, 1 / sqrt(rd2));
}

static void atom_pos(const or_input *in, int i, double *p) { p[0] = in->x[i]; p[1] = in->y[i]; p[2] = in->z[i]; }

static void bond_vectors(or_input *in)
{
    int N = in->N;
    in->rvec = calloc((size_t)3 * (N ? N : 1), sizeof(double));
    in->rvec2 = calloc((size_t)3 * (N ? N : 1), sizeof(double));
    in->rexp = calloc((size_t)(N ? N : 1), sizeof(int));
    in->disorder = calloc((size_t)(N ? N : 1), 1);
    int hydrogen = lib_rec_index(&in->lib, "HD");
    int nonHBh = lib_rec_index(&in->lib, "H");
    int carbon = lib_rec_index(&in->lib, "C");
    int aromC = lib_rec_index(&in->lib, "A");
    int oxygen = lib_rec_index(&in->lib, "OA");
    int sulphur = lib_rec_index(&in->lib, "SA");
    int i1 = 0, i2 = 0, i3 = 0;
    const int *T = in->type;
    for (int ia = 0; ia < N; ia++) {
        double pa[3], pb[3], d[3], dc[3], rd2;
        double *rv = &in->rvec[3 * ia], *rv2 = &in->rvec2[3 * ia];
        int from = ia - 20 > 0 ? ia - 20 : 0;
        int to = ia + 20 < N - 1 ? ia + 20 : N - 1;
        atom_pos(in, ia, pa);
        if (in->hbond[ia] == OR_D1) {
            for (int ib = from; ib <= to; ib++) {
                if (ib == ia) continue;
                atom_pos(in, ib, pb);
                v_sub(d, pa, pb);
                rd2 = v_msq(d);
                if (rd2 < 1.90) {
                    if (rd2 < APPROX_ZERO) rd2 = APPROX_ZERO;
                    double inv = 1 / sqrt(rd2);
                    if (T[ib] != oxygen && T[ib] != sulphur) in->rexp[ia] = 2;
                    if (T[ib] == oxygen || T[ib] == sulphur) {
                        in->rexp[ia] = 4;
                        if (in->disorderH) in->disorder[ia] = 1;
                    }
                    rv[0] = d[0] * inv; rv[1] = d[1] * inv; rv[2] = d[2] * inv;
                    break;
                }
            }
        } else if (in->hbond[ia] == OR_A2) {
            int nbond = 0, ib = from;
            for (; ib <= to; ib++) {
                if (ib == ia) continue;
                atom_pos(in, ib, pb);
                v_sub(dc, pa, pb);
                rd2 = v_msq(dc);
                int isH = T[ib] == hydrogen || T[ib] == nonHBh;
                if ((rd2 < 3.61 && !isH) || (rd2 < 1.69 && isH)) {
                    if (nbond == 1) { nbond = 2; i2 = ib; }
                    if (nbond == 0) { nbond = 1; i1 = ib; }
                }
            }
            if (nbond == 1) {
                double p1[3];
                atom_pos(in, i1, p1);
                v_sub(rv, pa, p1);
                v_norm_clamped(rv);
                for (int j2 = from; j2 <= to; j2++) {
                    if (j2 == i1 || j2 == ia) continue;
                    double p2[3];
                    atom_pos(in, j2, p2);
                    v_sub(dc, p1, p2);
                    rd2 = v_msq(dc);
                    if ((rd2 < 2.89 && T[j2] != hydrogen) || (rd2 < 1.69 && T[j2] == hydrogen)) {
                        v_sub(d, p2, p1);
                        v_norm_clamped(d);
                        v_cross(rv2, rv, d);
                        v_norm_clamped(rv2);
                    }
                }
            }
            if (nbond == 2) {
                double p1[3], p2[3];
                atom_pos(in, i1, p1);
                atom_pos(in, i2, p2);
                if ((T[i1] == hydrogen || T[i2] == hydrogen) && T[i1] != T[i2] && in->disorderH) {
                    /* FG/BondVectors.cpp:238-241 including its i1/i2 typo; ib may
                     * be left at to+1 by the scan above (out of range at the end
                     * of the receptor: treated as the origin here). */
                    if (T[i1] == carbon || T[i1] == aromC) ib = i1;
                    if (T[i2] == carbon || T[i1] == aromC) ib = i2;
                    in->disorder[ia] = 1;
                    if (ib < N) atom_pos(in, ib, pb); else { pb[0] = pb[1] = pb[2] = 0; }
                    v_sub(rv, pa, pb);
                    v_norm_clamped(rv);
                } else {
                    v_sub(rv2, p2, p1);
                    v_norm_clamped(rv2);
                    double t[3];
                    v_sub(t, pa, p1);
                    double rdot = v_dot(t, rv2);
                    double q[3] = {rv2[0] * rdot + p1[0], rv2[1] * rdot + p1[1], rv2[2] * rdot + p1[2]};
                    v_sub(rv, pa, q);
                    v_norm_clamped(rv);
                }
            }
        } else if (in->hbond[ia] == OR_A1) {
            int nbond = 0;
            for (int ib = from; ib <= to; ib++) {
                if (ib == ia) continue;
                atom_pos(in, ib, pb);
                v_sub(dc, pa, pb);
                rd2 = v_msq(dc);
                int isH = T[ib] == hydrogen || T[ib] == nonHBh;
                if ((rd2 < 2.89 && !isH) || (rd2 < 1.69 && isH)) {
                    if (nbond == 2) { nbond = 3; i3 = ib; }
                    if (nbond == 1) { nbond = 2; i2 = ib; }
                    if (nbond == 0) { nbond = 1; i1 = ib; }
                }
            }
            double p1[3], p2[3], p3[3];
            if (nbond == 1) {
                atom_pos(in, i1, p1);
                v_sub(rv, pa, p1);
                v_norm_clamped(rv);
            } else if (nbond == 2) {
                atom_pos(in, i1, p1); atom_pos(in, i2, p2);
                double mid[3] = {(p2[0] + p1[0]) / 2, (p2[1] + p1[1]) / 2, (p2[2] + p1[2]) / 2};
                v_sub(rv, pa, mid);
                v_norm_clamped(rv);
            } else if (nbond == 3) {
                atom_pos(in, i1, p1); atom_pos(in, i2, p2); atom_pos(in, i3, p3);
                double mid[3] = {(p1[0] + p2[0] + p3[0]) / 3, (p1[1] + p2[1] + p3[1]) / 3, (p1[2] + p2[2] + p3[2]) / 3};
                v_sub(rv, pa, mid);
                v_norm_clamped(rv);
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Pair-energy tables (FG/PairwiseInteractionEnergies.cpp:36-115).     */
/* ------------------------------------------------------------------ */
/* pow(double, int) as the reference computes it when built as C++98
 * (libstdc++ maps it to __builtin_powi: binary exponentiation, the libgcc
 * __powidf2 algorithm).  Used for R^xA and R^xB at
 * FG/PairwiseInteractionEnergies.cpp:72-73, where the exponent is an int. */
static double powi_(double x, int m)
{
    unsigned n = m < 0 ? 0u - (unsigned)m : (unsigned)m;
    double y = (n % 2) ? x : 1;
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? 1 / y : y;
}

static int pair_tables(or_input *in)
{
    /* rows past the declared types stay zero (the reference's table is a
     * zero-initialised [32][8192][16]); atoms may carry such an index */
    int T = in->numRecTypes, M = in->numAtomMaps, rows = T > 0 ? T : 1;
    for (int ia = 0; ia < in->N; ia++) if (in->type[ia] + 1 > rows) rows = in->type[ia] + 1;
    in->tableTypes = rows;
    in->pair = calloc((size_t)rows * OR_MAX_DIST * (M ? M : 1), sizeof(double));
    int iSmooth = (int)(in->rSmooth * 100 / 2);
    static double sm[OR_MAX_DIST];
    for (int m = 0; m < M; m++) {
        const or_map *g = &in->maps[m];
        if (g->isCovalent) continue;
        for (int i = 0; i < T; i++) {
            int xA = g->xA[i], xB = g->xB[i];
            double Rij = g->nbpR[i], epsij = g->nbpEps[i], tmp;
            double cA = (tmp = epsij / (xA - xB)) * powi_(Rij, xA) * xB;
            double cB = tmp * powi_(Rij, xB) * xA;
            if (isnan(cA) || isnan(cB)) { snprintf(in->err, sizeof in->err, "Van der Waals coefficient is not a number"); return -2; }
            if (xA == 0 || xB == 0) { snprintf(in->err, sizeof in->err, "Van der Waals exponent is 0"); return -2; }
            double dxA = xA, dxB = xB;
#define PT(k) in->pair[((size_t)i * OR_MAX_DIST + (k)) * M + m]
            for (int k = 1; k < OR_MAX_DIST; k++) {
                double r = (double)k * (1 / 100.0);
                double rA = pow(r, dxA), rB = pow(r, dxB);
                PT(k) = dmin(EINTCLAMP, cA / rA - cB / rB);
            }
            PT(0) = EINTCLAMP;
            PT(OR_MAX_DIST - 1) = 0;
            if (iSmooth > 0) {
                for (int k = 1; k < OR_MAX_DIST; k++) {
                    sm[k] = 100000;
                    int lo = k - iSmooth > 0 ? k - iSmooth : 0;
                    int hi = k + iSmooth < OR_MAX_DIST ? k + iSmooth : OR_MAX_DIST;
                    for (int j = lo; j < hi; j++) sm[k] = dmin(sm[k], PT(j));
                }
                for (int k = 1; k < OR_MAX_DIST; k++) PT(k) = sm[k];
            }
#undef PT
        }
    }
    return 0;
}

int or_precompute(or_input *in)
{
    for (int i = 0; i < in->numMaps; i++)
        if (in->maps[i].filename[0]) in->maps[i].energies = calloc((size_t)in->P, sizeof(double));
    int r = pair_tables(in);
    if (r) return r;
    or_desolv_table(in->lib.coeff_desolv, in->desolv);
    bond_vectors(in);
    in->hydrogen = lib_rec_index(&in->lib, "HD");
    return 0;
}

/* ------------------------------------------------------------------ */
/* Electrostatic map (FG/electrostatics/Cpu.cpp:27-88).                */
/* ------------------------------------------------------------------ */
void or_elec_map(const int n[3], const int div2[3], double spacing,
                 int N, const double *x, const double *y, const double *z, const double *w,
                 const double *eps, int ddd, double *e, int round3, int threads)
{
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
    {
        double *dz2 = malloc(sizeof(double) * 2 * (N ? N : 1));
        double *dyz2 = dz2 + N;
#pragma omp for schedule(dynamic, 1)
        for (int iz = 0; iz < n[2]; iz++) {
            double gz = (iz - div2[2]) * spacing;
            for (int ia = 0; ia < N; ia++) { double t = z[ia] - gz; dz2[ia] = t * t; }
            for (int iy = 0; iy < n[1]; iy++) {
                double gy = (iy - div2[1]) * spacing;
                for (int ia = 0; ia < N; ia++) { double t = y[ia] - gy; dyz2[ia] = t * t + dz2[ia]; }
                size_t row = ((size_t)iz * n[1] + iy) * n[0];
                for (int ix = 0; ix < n[0]; ix++) {
                    double gx = (ix - div2[0]) * spacing;
                    double en = e[row + ix];
                    for (int ia = 0; ia < N; ia++) {
                        double t = x[ia] - gx;
                        double r = sqrt(t * t + dyz2[ia]);
                        double inv = 1 / r;
                        if (ddd) en += w[ia] * dmin(inv, 2) * eps[bin_index(r)];
                        else en += w[ia] * dmin(inv, 2);
                    }
                    e[row + ix] = round3 ? or_round3(en) : en;
                }
            }
        }
        free(dz2);
    }
}

/* ------------------------------------------------------------------ */
/* Affinity + desolvation maps (FG/CalculateGridmaps.cpp:32-316).      */
/* ------------------------------------------------------------------ */
typedef struct {
    int N, M, T, P;
    const double *x, *y, *z, *charge, *vol, *solpar, *rvec, *rvec2, *pair, *S;
    const int *type, *hbond, *rexp, *mHb, *mIsHB, *mCov;
    const unsigned char *dis;
    const double *mSol, *mVol;
    double solparQ;
    int hydrogen;
} or_aff;

/* FG/CalculateGridmaps.cpp:84-197 */
static void angular(const or_aff *A, int ia, int cH, const double *d, double *racc, double *rdon, double *Hramp)
{
    *racc = 1; *rdon = 1; *Hramp = 1;
    const double *rv = &A->rvec[3 * ia];
    double cosT = -v_dot(d, rv);
    int hb = A->hbond[ia];
    if (hb == OR_D1) {
        *racc = 0;
        if (cosT > 0) {
            double ra = cosT;
            switch (A->rexp[ia]) {
            case 4: ra = ra * ra; /* fall through */
            case 2: ra = ra * ra;
            }
            *racc = ra;
            if (ia != cH) {
                double th = acos_c(v_dot(&A->rvec[3 * cH], rv));
                *Hramp = 0.5 - 0.5 * cos(th * (120.0 / 90.0));
            }
        }
    } else if (hb == OR_A1) {
        *rdon = 0;
        if (cosT > 0) *rdon = cosT * cosT;
    } else if (hb == OR_A2) {
        *rdon = 0;
        const double *rv2 = &A->rvec2[3 * ia];
        if (A->dis[ia]) {
            *racc = 0;
            double th = acos_c(cosT);
            if (th <= 1.24791 + (PI_ / 2)) {
                double c = cos(th - 1.24791);
                c = c * c;
                *rdon = c * c;
                *racc = *rdon;
            }
        } else {
            double rd = 0;
            if (cosT >= 0) {
                double cr[3];
                v_cross(cr, d, rv2);
                double rd2 = v_msq(cr);
                if (rd2 < APPROX_ZERO) rd2 = APPROX_ZERO;
                double inv = 1 / sqrt(rd2);
                double ti = fabs(acos_c(inv * v_dot(cr, rv)) - (PI_ / 2));
                rd = 0.9 + 0.1 * sin(ti + ti);
            } else if (cosT >= -0.34202) {
                double t = 0.116978 - cosT * cosT;
                rd = 562.25 * (t * t * t);
            }
            double t0 = (PI_ / 2) - acos_c(v_dot(d, rv2));
            rd *= cos(t0);
            *rdon = rd;
        }
    }
}

typedef struct { double mn, mx; int flag; } or_hb;

/* FG/CalculateGridmaps.cpp:238-316 for one point; cand[] lists candidate
 * atoms in ascending index (a superset of those within 8 A). */
static void grid_point(const or_aff *A, const int *cand, int nc, double gx, double gy, double gz,
                       double *maps, double *dmap, size_t p, int round3)
{
    or_hb hb[OR_MAX_MAPS];
    for (int m = 0; m < A->M; m++) { hb[m].mn = 999999; hb[m].mx = -999999; hb[m].flag = 0; }
    /* nearest DS/D1 atom, lowest index on ties (FG/CalculateGridmaps.cpp:65-82) */
    int cH = 0;
    double rmin = 999999.0 * 999999.0;
    for (int c = 0; c < nc; c++) {
        int ia = cand[c];
        if (A->hbond[ia] != OR_DS && A->hbond[ia] != OR_D1) continue;
        double dx = A->x[ia] - gx, dy = A->y[ia] - gy, dz = A->z[ia] - gz;
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < rmin) { rmin = r2; cH = ia; }
    }
    for (int c = 0; c < nc; c++) {
        int ia = cand[c];
        double d[3] = {A->x[ia] - gx, A->y[ia] - gy, A->z[ia] - gz};
        double rSq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (rSq > 8.0 * 8.0) continue;
        if (A->type[ia] == A->hydrogen && A->dis[ia]) continue;
        if (rSq < (APPROX_ZERO * APPROX_ZERO) * (APPROX_ZERO * APPROX_ZERO))
            rSq = (APPROX_ZERO * APPROX_ZERO) * (APPROX_ZERO * APPROX_ZERO);
        double invR = 1 / sqrt(rSq);
        v_scale(d, invR);
        int idx = bin_index(1 / invR);
        double racc, rdon, Hramp;
        angular(A, ia, cH, d, &racc, &rdon, &Hramp);
        int ahb = A->hbond[ia];
        int aDon = ahb == OR_DS || ahb == OR_D1;
        double S = A->S[idx];
        const double *Trow = &A->pair[((size_t)A->type[ia] * OR_MAX_DIST + idx) * A->M];
        for (int m = 0; m < A->M; m++) {
            if (A->mCov[m]) continue;
            double *e = &maps[(size_t)m * A->P + p];
            double E = Trow[m];
            int mhb = A->mHb[m];
            if (A->mIsHB[m]) {
                double rsph = sat(E / 100);
                if ((mhb == OR_AS || mhb == OR_A2) && aDon) {
                    /* a negative HD index is undefined in the reference; E is used */
                    double Ep = (A->dis[ia] && A->hydrogen >= 0)
                                    ? A->pair[((size_t)A->hydrogen * OR_MAX_DIST + (idx - 110 > 0 ? idx - 110 : 0)) * A->M + m]
                                    : E;
                    *e += Ep * Hramp * (racc + (1 - racc) * rsph);
                } else if (mhb == OR_A1 && aDon) {
                    double v = E * (racc + (1 - racc) * rsph);
                    hb[m].mn = dmin(hb[m].mn, v); hb[m].mx = dmax(hb[m].mx, v); hb[m].flag = 1;
                } else if ((mhb == OR_DS || mhb == OR_D1) && ahb >= OR_AS) {
                    double v = E * (rdon + (1 - rdon) * rsph);
                    hb[m].mn = dmin(hb[m].mn, v); hb[m].mx = dmax(hb[m].mx, v); hb[m].flag = 1;
                } else {
                    *e += E;
                }
            } else {
                *e += E;
            }
            *e += A->mSol[m] * A->vol[ia] * S + (A->solpar[ia] + A->solparQ * fabs(A->charge[ia])) * A->mVol[m] * S;
        }
        dmap[p] += A->solparQ * A->vol[ia] * S;
    }
    for (int m = 0; m < A->M; m++) {
        double *e = &maps[(size_t)m * A->P + p];
        if (hb[m].flag) *e += hb[m].mn + hb[m].mx;
        if (round3) *e = or_round3(*e);
    }
    if (round3) dmap[p] = or_round3(dmap[p]);
}

/* Cell list over blocks of B^3 grid points: each block lists, in ascending
 * index, every atom whose distance to the block's box is <= 8 A (plus a
 * margin), a superset of the atoms within 8 A of any point in the block. */
#define OR_B 8
typedef struct { int nb[3]; int *start, *list; } or_cells;

static void build_cells(or_cells *C, const int n[3], const int div2[3], double s, int N,
                        const double *x, const double *y, const double *z)
{
    for (int k = 0; k < 3; k++) C->nb[k] = (n[k] + OR_B - 1) / OR_B;
    int NB = C->nb[0] * C->nb[1] * C->nb[2];
    int *cnt = calloc((size_t)NB + 1, sizeof(int));
    const double *xs[3] = {x, y, z};
    double cut = 8.0 + 1e-6;
    /* two passes: count, then fill (ascending atom order preserved) */
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            C->start = malloc(sizeof(int) * ((size_t)NB + 1));
            C->start[0] = 0;
            for (int b = 0; b < NB; b++) C->start[b + 1] = C->start[b] + cnt[b];
            C->list = malloc(sizeof(int) * ((size_t)C->start[NB] + 1));
            memset(cnt, 0, sizeof(int) * (size_t)NB);
        }
        for (int ia = 0; ia < N; ia++) {
            int lo[3], hi[3];
            for (int k = 0; k < 3; k++) {
                double c = xs[k][ia];
                /* grid index range within 8 A of the atom along this axis */
                double gl = (c - cut) / s + div2[k], gh = (c + cut) / s + div2[k];
                int il = (int)floor(gl), ih = (int)ceil(gh);
                if (il < 0) il = 0;
                if (ih > n[k] - 1) ih = n[k] - 1;
                if (il > ih) { lo[0] = 1; hi[0] = 0; break; }
                lo[k] = il / OR_B; hi[k] = ih / OR_B;
            }
            if (lo[0] > hi[0]) continue;
            for (int bz = lo[2]; bz <= hi[2]; bz++)
                for (int by = lo[1]; by <= hi[1]; by++)
                    for (int bx = lo[0]; bx <= hi[0]; bx++) {
                        int bi[3] = {bx, by, bz};
                        double d2 = 0;
                        for (int k = 0; k < 3; k++) {
                            int i0 = bi[k] * OR_B, i1 = bi[k] * OR_B + OR_B - 1;
                            if (i1 > n[k] - 1) i1 = n[k] - 1;
                            double g0 = (i0 - div2[k]) * s, g1 = (i1 - div2[k]) * s;
                            double c = xs[k][ia], t = 0;
                            if (c < g0) t = g0 - c; else if (c > g1) t = c - g1;
                            d2 += t * t;
                        }
                        if (d2 > cut * cut) continue;
                        int b = (bz * C->nb[1] + by) * C->nb[0] + bx;
                        if (pass == 1) C->list[C->start[b] + cnt[b]] = ia;
                        cnt[b]++;
                    }
        }
    }
    free(cnt);
}

void or_affinity_maps(const int n[3], const int div2[3], double spacing,
                      int N, const double *x, const double *y, const double *z,
                      const double *charge, const double *vol, const double *solpar,
                      const int *type, const int *hbond,
                      const double *rvec, const double *rvec2, const int *rexp,
                      const unsigned char *disorder,
                      int numRecTypes, int numAtomMaps, const double *pair, const double *desolv,
                      const int *mapHbond, const int *mapIsHBonder, const int *mapIsCovalent,
                      const double *mapSolpar, const double *mapVol,
                      double solparQ, int hydrogen, double *maps, double *d, int round3, int threads)
{
    or_aff A = {N, numAtomMaps, numRecTypes, n[0] * n[1] * n[2], x, y, z, charge, vol, solpar, rvec, rvec2,
                pair, desolv, type, hbond, rexp, mapHbond, mapIsHBonder, mapIsCovalent, disorder,
                mapSolpar, mapVol, solparQ, hydrogen};
    or_cells C;
    build_cells(&C, n, div2, spacing, N, x, y, z);
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int iz = 0; iz < n[2]; iz++) {
        double gz = (iz - div2[2]) * spacing;
        for (int iy = 0; iy < n[1]; iy++) {
            double gy = (iy - div2[1]) * spacing;
            for (int ix = 0; ix < n[0]; ix++) {
                double gx = (ix - div2[0]) * spacing;
                int b = ((iz / OR_B) * C.nb[1] + iy / OR_B) * C.nb[0] + ix / OR_B;
                size_t p = ((size_t)iz * n[1] + iy) * n[0] + ix;
                grid_point(&A, &C.list[C.start[b]], C.start[b + 1] - C.start[b], gx, gy, gz, maps, d, p, round3);
            }
        }
    }
    free(C.start);
    free(C.list);
}

long long or_count_cutoff_pairs(const int n[3], const int div2[3], double spacing,
                                int N, const double *x, const double *y, const double *z, int threads)
{
    or_cells C;
    build_cells(&C, n, div2, spacing, N, x, y, z);
    long long total = 0;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int iz = 0; iz < n[2]; iz++) {
        double gz = (iz - div2[2]) * spacing;
        for (int iy = 0; iy < n[1]; iy++) {
            double gy = (iy - div2[1]) * spacing;
            for (int ix = 0; ix < n[0]; ix++) {
                double gx = (ix - div2[0]) * spacing;
                int b = ((iz / OR_B) * C.nb[1] + iy / OR_B) * C.nb[0] + ix / OR_B;
                for (int c = C.start[b]; c < C.start[b + 1]; c++) {
                    int ia = C.list[c];
                    double dx = x[ia] - gx, dy = y[ia] - gy, dz = z[ia] - gz;
                    if (dx * dx + dy * dy + dz * dz <= 64.0) total++;
                }
            }
        }
    }
    free(C.start);
    free(C.list);
    return total;
}

/* FG/main.cpp:32-52 */
void or_covalent_map(const int n[3], const int div2[3], double spacing, const double covPoint[3],
                     double covHalfWidthSqInv, double covBarrier, double *out)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int iz = 0; iz < n[2]; iz++)
        for (int iy = 0; iy < n[1]; iy++)
            for (int ix = 0; ix < n[0]; ix++) {
                double g[3] = {(ix - div2[0]) * spacing, (iy - div2[1]) * spacing, (iz - div2[2]) * spacing};
                double d[3];
                v_sub(d, covPoint, g);
                double r = v_msq(d) * covHalfWidthSqInv;
                if (r < APPROX_ZERO * APPROX_ZERO) r = APPROX_ZERO * APPROX_ZERO;
                out[((size_t)iz * n[1] + iy) * n[0] + ix] = covBarrier * (1 - exp(-0.69314718055994529 * r));
            }
}

static void covalent_maps(or_input *in)
{
    for (int m = 0; m < in->numAtomMaps; m++)
        if (in->maps[m].isCovalent)
            or_covalent_map(in->n, in->div2, in->spacing, in->covPoint, in->covHalfWidthSqInv, in->covBarrier,
                            in->maps[m].energies);
}

/* FG/main.cpp:54-79 */
void or_floating_grid(const int n[3], const int div2[3], double spacing, int N, const double *x, const double *y,
                      const double *z, float *out, int threads)
{
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int iz = 0; iz < n[2]; iz++)
        for (int iy = 0; iy < n[1]; iy++)
            for (int ix = 0; ix < n[0]; ix++) {
                double g[3] = {(ix - div2[0]) * spacing, (iy - div2[1]) * spacing, (iz - div2[2]) * spacing};
                double minDistSq = 1.0E12 * 1.0E12;
                for (int ia = 0; ia < N; ia++) {
                    double a[3] = {x[ia], y[ia], z[ia]}, d[3];
                    v_sub(d, a, g);
                    double distSq = v_msq(d);
                    if (distSq < minDistSq) minDistSq = distSq;
                }
                out[((size_t)iz * n[1] + iy) * n[0] + ix] = (float)sqrt(minDistSq);
            }
}

void or_compute(or_input *in, int mode, int threads)
{
    int P = in->P, M = in->numAtomMaps;
    if (mode & 1)
        or_elec_map(in->n, in->div2, in->spacing, in->N, in->x, in->y, in->z, in->w, in->eps, in->distDepDiel,
                    in->maps[in->numMaps - 2].energies, 1, threads);
    if (mode & 2) {
        covalent_maps(in);
        double *maps = calloc((size_t)(M ? M : 1) * P, sizeof(double));
        int mHb[OR_MAX_MAPS], mIsHB[OR_MAX_MAPS], mCov[OR_MAX_MAPS];
        double mSol[OR_MAX_MAPS], mVol[OR_MAX_MAPS];
        for (int m = 0; m < M; m++) {
            mHb[m] = in->maps[m].hbond; mIsHB[m] = in->maps[m].isHBonder; mCov[m] = in->maps[m].isCovalent;
            mSol[m] = in->maps[m].solparProbe; mVol[m] = in->maps[m].volProbe;
            if (in->maps[m].energies) memcpy(&maps[(size_t)m * P], in->maps[m].energies, sizeof(double) * P);
        }
        or_affinity_maps(in->n, in->div2, in->spacing, in->N, in->x, in->y, in->z, in->charge, in->vol, in->solpar,
                         in->type, in->hbond, in->rvec, in->rvec2, in->rexp, in->disorder, in->numRecTypes, M,
                         in->pair, in->desolv, mHb, mIsHB, mCov, mSol, mVol, in->solparQ, in->hydrogen, maps,
                         in->maps[in->numMaps - 1].energies, 1, threads);
        for (int m = 0; m < M; m++)
            if (in->maps[m].energies) memcpy(in->maps[m].energies, &maps[(size_t)m * P], sizeof(double) * P);
        free(maps);
    }
    if ((mode & 4) && in->fmap[0]) {
        if (!in->fmapValues) in->fmapValues = malloc(sizeof(float) * (size_t)P);
        or_floating_grid(in->n, in->div2, in->spacing, in->N, in->x, in->y, in->z, in->fmapValues, threads);
    }
}

/* ------------------------------------------------------------------ */
/* Writers (FG/GridMap.cpp:127-212, FG/Utils.cpp:40-84).               */
/* ------------------------------------------------------------------ */
int or_write_outputs(or_input *in, const char *programName)
{
    char hdr[4096];
    int hl = snprintf(hdr, sizeof hdr,
                      "GRID_PARAMETER_FILE %s\nGRID_DATA_FILE %s\nMACROMOLECULE %s\nSPACING %.3lf\n"
                      "NELEMENTS %d %d %d\nCENTER %.3lf %.3lf %.3lf\n",
                      in->gpfName, in->fld, in->receptor, in->spacing, in->n[0] - 1, in->n[1] - 1, in->n[2] - 1,
                      in->center[0], in->center[1], in->center[2]);
    FILE *f = fopen(in->fld, "w");
    if (!f) return -2;
    int nm = in->numMaps;
    /* the floating grid is one more variable of the field (FG/Utils.cpp:51-81) */
    int fm = in->fmap[0] != 0, nmF = nm + fm;
    fprintf(f, "# AVS field file\n#\n# AutoDock Atomic Affinity and Electrostatic Grids\n#\n# Created by %s.\n#\n", programName);
    fprintf(f, "#SPACING %.3f\n", (double)(float)in->spacing);
    fprintf(f, "#NELEMENTS %d %d %d\n", in->n[0] - 1, in->n[1] - 1, in->n[2] - 1);
    fprintf(f, "#CENTER %.3lf %.3lf %.3lf\n", in->center[0], in->center[1], in->center[2]);
    fprintf(f, "#MACROMOLECULE %s\n#GRID_PARAMETER_FILE %s\n#\n", in->receptor, in->gpfName);
    fprintf(f, "ndim=3\t\t\t# number of dimensions in the field\n");
    fprintf(f, "dim1=%d\t\t\t# number of x-elements\n", in->n[0]);
    fprintf(f, "dim2=%d\t\t\t# number of y-elements\n", in->n[1]);
    fprintf(f, "dim3=%d\t\t\t# number of z-elements\n", in->n[2]);
    fprintf(f, "nspace=3\t\t# number of physical coordinates per point\n");
    fprintf(f, "veclen=%d\t\t# number of affinity values at each point\n", nmF);
    fprintf(f, "data=float\t\t# data type (byte, integer, float, double)\n");
    fprintf(f, "field=uniform\t\t# field type (uniform, rectilinear, irregular)\n");
    for (int i = 0; i < 3; i++) fprintf(f, "coord %d file=%s filetype=ascii offset=%d\n", i + 1, in->xyz, i * 2);
    for (int i = 0; i < in->numAtomMaps; i++)
        fprintf(f, "label=%s-affinity\t# component label for variable %d\n", in->maps[i].type, i + 1);
    fprintf(f, "label=Electrostatics\t# component label for variable %d\n", nmF - 2);
    fprintf(f, "label=Desolvation\t# component label for variable %d\n", nmF - 1);
    if (fm) fprintf(f, "label=Floating_Grid\t# component label for variable %d\n", nmF);
    fprintf(f, "#\n# location of affinity grid files and how to read them\n#\n");
    for (int i = 0; i < nm; i++)
        fprintf(f, "variable %d file=%s filetype=ascii skip=6\n", i + 1, in->maps[i].filename);
    if (fm) fprintf(f, "variable %d file=%s filetype=ascii skip=6\n", nmF, in->fmap);
    fclose(f);
    for (int i = 0; i < nm; i++) {
        if (!in->maps[i].filename[0]) continue;
        FILE *mf = fopen(in->maps[i].filename, "w");
        if (!mf) return -2;
        fwrite(hdr, (size_t)hl, 1, mf);
        const double *E = in->maps[i].energies;
        for (int j = 0; j < in->P; j++) {
            if (E[j] == 0) fputs("0\n", mf);
            else fprintf(mf, "%.3f\n", E[j]);
        }
        fclose(mf);
    }
    if (fm && in->fmapValues) {
        /* FG/GridMap.cpp:160-181: "%.3f" of each float */
        FILE *mf = fopen(in->fmap, "w");
        if (!mf) return -2;
        fwrite(hdr, (size_t)hl, 1, mf);
        for (int j = 0; j < in->P; j++) fprintf(mf, "%.3f\n", in->fmapValues[j]);
        fclose(mf);
    }
    return 0;
}

void or_free(or_input *in)
{
    free(in->fmapValues);
    in->fmapValues = NULL;
    free(in->x); free(in->y); free(in->z); free(in->w); free(in->charge); free(in->vol); free(in->solpar);
    free(in->type); free(in->hbond); free(in->rvec); free(in->rvec2); free(in->rexp); free(in->disorder);
    free(in->pair);
    for (int i = 0; i < OR_MAX_MAPS + 2; i++) free(in->maps[i].energies);
    memset(in, 0, sizeof *in);
}

/* ------------------------------------------------------------------ */
/* Accessors for the test harness (ctypes).                            */
/* ------------------------------------------------------------------ */
or_input *or_open(const char *gpf, int v4, int agParamOrder, int *rc)
{
    or_input *in = calloc(1, sizeof *in);
    int r = or_load(in, gpf, v4, agParamOrder, NULL);
    if (r == 0) r = or_precompute(in);
    if (rc) *rc = r;
    if (r) { or_free(in); free(in); return NULL; }
    return in;
}

void or_close(or_input *in)
{
    if (!in) return;
    or_free(in);
    free(in);
}

/* N, numAtomMaps, tableTypes, n[3], div2[3], distDepDiel, hydrogen, P */
void or_dims(const or_input *in, int out[12])
{
    int v[12] = {in->N, in->numAtomMaps, in->tableTypes, in->n[0], in->n[1], in->n[2],
                 in->div2[0], in->div2[1], in->div2[2], in->distDepDiel, in->hydrogen, in->P};
    memcpy(out, v, sizeof v);
}

/* spacing, solparQ, center xyz */
void or_scalars(const or_input *in, double out[5])
{
    out[0] = in->spacing; out[1] = in->solparQ;
    out[2] = in->center[0]; out[3] = in->center[1]; out[4] = in->center[2];
}

const double *or_darray(const or_input *in, const char *name)
{
    if (!strcmp(name, "x")) return in->x;
    if (!strcmp(name, "y")) return in->y;
    if (!strcmp(name, "z")) return in->z;
    if (!strcmp(name, "w")) return in->w;
    if (!strcmp(name, "charge")) return in->charge;
    if (!strcmp(name, "vol")) return in->vol;
    if (!strcmp(name, "solpar")) return in->solpar;
    if (!strcmp(name, "rvector")) return in->rvec;
    if (!strcmp(name, "rvector2")) return in->rvec2;
    if (!strcmp(name, "pair")) return in->pair;
    if (!strcmp(name, "desolv")) return in->desolv;
    if (!strcmp(name, "eps")) return in->eps;
    return NULL;
}

const int *or_iarray(const or_input *in, const char *name)
{
    if (!strcmp(name, "type")) return in->type;
    if (!strcmp(name, "hbond")) return in->hbond;
    if (!strcmp(name, "rexp")) return in->rexp;
    return NULL;
}

const unsigned char *or_disorder(const or_input *in) { return in->disorder; }

/* hbond, isHBonder, isCovalent; solparProbe, volProbe */
void or_map_params(const or_input *in, int m, int iv[3], double dv[2])
{
    iv[0] = in->maps[m].hbond; iv[1] = in->maps[m].isHBonder; iv[2] = in->maps[m].isCovalent;
    dv[0] = in->maps[m].solparProbe; dv[1] = in->maps[m].volProbe;
}

const char *or_error(const or_input *in) { return in ? in->err : ""; }
