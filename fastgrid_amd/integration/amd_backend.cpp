// The binding a maintainer adds to the reference FastGrid 4.2.1 tree
// (fastgrid_master/fastgrid) to compute its maps with libfgmaps: this file
// is compiled in place of the reference's two map backends,
// electrostatics/Cpu.cpp (or Cuda.cpp) and CalculateGridmaps.cpp, and keeps
// their interfaces:
//   calculateElectrostaticMapAsync / synchronizeCalculation
//       (FG/electrostatics/Electrostatics.h:28-36)
//   calculateGridmaps (FG/CalculateGridmaps.h:30-31)
// Everything else (parsing, tables, bond vectors, covalent maps, floating
// grid, writers, log) stays the reference's own code.  C++98, as the
// reference is built.  The reference's own build never sees this file; the
// recipe that links it with the reference's objects is oracle/ref.mk.
#include "electrostatics/Electrostatics.h"
#include "CalculateGridmaps.h"
#include "Exceptions.h"

#include <fgmaps.h>

#include <cstdio>
#include <vector>

namespace {

// the bond vectors are passed as [n][3] doubles
typedef char Vec3dIsThreeDoubles[sizeof(Vec3d) == 3 * sizeof(double) ? 1 : -1];

// InputData's atoms (FG/InputData.h:50-58) as the structure of arrays the ABI
// takes.
struct Soa {
    std::vector<double> x, y, z, w;
    std::vector<int32_t> hb;
    fg_atoms a;
    explicit Soa(const InputData *in)
    {
        const int n = in->numReceptorAtoms;
        x.resize(n + 1);
        y.resize(n + 1);
        z.resize(n + 1);
        w.resize(n + 1);
        hb.resize(n + 1);
        for (int i = 0; i < n; i++) {
            x[i] = in->receptorAtom[i].x;
            y[i] = in->receptorAtom[i].y;
            z[i] = in->receptorAtom[i].z;
            w[i] = in->receptorAtom[i].w;
            hb[i] = int32_t(in->hbond[i]);  // HBondType: NON, DS, D1, AS, A1, A2
        }
        a.n = n;
        a.x = &x[0];
        a.y = &y[0];
        a.z = &z[0];
        a.w = &w[0];
        a.charge = in->charge;
        a.vol = in->vol;
        a.solpar = in->solpar;
        a.type = in->atomType;
        a.hbond = &hb[0];
    }
};

fg_grid gridOf(const InputData *in)
{
    fg_grid g;
    g.nx = in->numGridPoints.x;
    g.ny = in->numGridPoints.y;
    g.nz = in->numGridPoints.z;
    g.div2x = in->numGridPointsDiv2.x;
    g.div2y = in->numGridPointsDiv2.y;
    g.div2z = in->numGridPointsDiv2.z;
    g.spacing = in->gridSpacing;
    return g;
}

void check(int rc)
{
    if (rc != FG_OK) {
        fprintf(stderr, "libfgmaps: %s\n", fg_last_error());
        throw ExitProgram(0xbad);  // as the reference's CUDA path on a launch failure
    }
}

// One context for the process, on the device --cuda-dev names.
fg_ctx *context(const ProgramParameters &p)
{
    static fg_ctx *ctx = 0;
    if (!ctx) check(fg_create(p.getDeviceIDCUDA(), &ctx));
    return ctx;
}

// The reference's backends store rounded values (roundOutput,
// FG/autogrid.h:128-131; FG/electrostatics/Cpu.cpp:82,
// FG/CalculateGridmaps.cpp:307-315); the ABI returns them unrounded.
void round3(double *v, int n)
{
    for (int i = 0; i < n; i++) v[i] = fg_round3(v[i]);
}

struct ElecHandle {
    fg_handle *h;
    GridMapList *maps;
    int points;
};

}  // namespace

void *calculateElectrostaticMapAsync(const InputData *input, const ProgramParameters &programParams,
                                     GridMapList &gridmaps)
{
    Soa s(input);
    fg_grid g = gridOf(input);
    ElecHandle *eh = new ElecHandle;
    eh->maps = &gridmaps;
    eh->points = input->numGridPointsPerMap;
    // the inputs are copied before the call returns; the map is written by fg_wait
    check(fg_elec_map_async(context(programParams), &g, &s.a, input->epsilon, input->distDepDiel ? 1 : 0,
                            FG_PREC_COMPENSATED, gridmaps.getElectrostaticMap().energies, &eh->h));
    return eh;
}

void synchronizeCalculation(void *handle)
{
    if (!handle) return;
    ElecHandle *eh = static_cast<ElecHandle *>(handle);
    check(fg_wait(eh->h));
    round3(eh->maps->getElectrostaticMap().energies, eh->points);
    // the e-map backend saves its own map (FG/electrostatics/Cuda.cpp:194)
    eh->maps->saveElectrostaticMap();
    delete eh;
}

void calculateGridmaps(const InputData *input, const ProgramParameters &programParams, GridMapList &gridmaps,
                       const ParameterLibrary &parameterLibrary, const PairwiseInteractionEnergies &energyLookup,
                       const DesolvExpFunc &desolvExpFunc, const BondVectors *bondVectors)
{
    Soa s(input);
    fg_grid g = gridOf(input);
    const int M = gridmaps.getNumAtomMaps(), N = input->numReceptorAtoms, P = input->numGridPointsPerMap;
    // the pair table as [type][bin][map], every row: a receptor_types line
    // read after the receptor renames types without retyping the atoms, so
    // an atom's type index may exceed numReceptorTypes - 1
    std::vector<double> pair(size_t(NUM_RECEPTOR_TYPES) * MAX_DIST * (M > 0 ? M : 1));
    for (int t = 0; t < NUM_RECEPTOR_TYPES; t++)
        for (int k = 0; k < MAX_DIST; k++)
            for (int m = 0; m < M; m++) pair[(size_t(t) * MAX_DIST + k) * M + m] = energyLookup(t, k, m);
    std::vector<double> desolv(MAX_DIST);
    for (int k = 0; k < MAX_DIST; k++) desolv[k] = desolvExpFunc(k);
    std::vector<uint8_t> disorder(N + 1);
    for (int i = 0; i < N; i++) disorder[i] = bondVectors->disorder[i] ? 1 : 0;
    std::vector<fg_map_params> mp(M + 1);
    std::vector<double *> out(M + 1);
    for (int m = 0; m < M; m++) {
        mp[m].hbond = int32_t(gridmaps[m].hbond);
        mp[m].is_hbonder = gridmaps[m].isHBonder;
        mp[m].is_covalent = gridmaps[m].isCovalent ? 1 : 0;
        mp[m].solpar_probe = gridmaps[m].solparProbe;
        mp[m].vol_probe = gridmaps[m].volProbe;
        out[m] = gridmaps[m].energies;  // covalent maps hold initCovalentMaps' values
    }
    fg_hbond_geom geom;
    geom.rvector = &bondVectors->rvector[0].x;
    geom.rvector2 = &bondVectors->rvector2[0].x;
    geom.rexp = bondVectors->rexp;
    geom.disorder = &disorder[0];
    fg_tables tab;
    tab.n_types = NUM_RECEPTOR_TYPES;
    tab.n_maps = M;
    tab.pair = &pair[0];
    tab.desolv = &desolv[0];
    tab.eps = input->epsilon;
    double *d = gridmaps.getDesolvationMap().energies;
    check(fg_affinity_d_maps(context(programParams), &g, &s.a, &geom, &tab, &mp[0], M, input->solparQ,
                             parameterLibrary.getAtomParameterRecIndex("HD"), &out[0], d));
    for (int m = 0; m < M; m++) round3(gridmaps[m].energies, P);
    round3(d, P);
}
