// Host front end of libfgmaps: parameter library, GPF/PDBQT loader,
// precomputed tables and bond vectors, writers and the synthetic generator.
// Plain C++17, no device code.  FG = fastgrid_master/fastgrid in the
// reference.
#pragma once

#include <array>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "fgmaps.h"

namespace fgh {

constexpr int kTableLen = FG_TABLE_LEN;  // MAX_DIST, FG/autogrid.h:88
constexpr int kNumTypes = 32;            // NUM_RECEPTOR_TYPES, FG/autogrid.h:93-94
constexpr int kMaxAtomTypes = 14;        // MAX_ATOM_TYPES, autocomm.h:102
constexpr int kMaxGridPts = 1024;        // MAX_GRID_PTS, autocomm.h:94
constexpr double kApproxZero = 1.0e-6;   // autocomm.h:74
constexpr double kPi = 3.14159265358979323846;

enum HBond : int32_t { NON = 0, DS = 1, D1 = 2, AS = 3, A1 = 4, A2 = 5 };

struct AtomParam {
    std::string type;
    double Rij = 0, epsij = 0, vol = 0, solpar = 0, RijHB = 0, epsijHB = 0;
    int32_t hbond = NON, recIndex = -1, mapIndex = -1, bondIndex = 0;
};

// AD4 parameter library (FG/ParameterLibrary.cpp:34-284): a dictionary keyed
// by the first two characters of the type name, weighted epsilons, FE
// coefficients.
class ParamLibrary {
public:
    explicit ParamLibrary(bool v4);
    AtomParam *find(const std::string &key);
    int32_t recIndexOf(const std::string &key);
    bool loadFile(const std::string &path);
    void readLine(const char *line);

    double coeffVdW = 0, coeffHbond = 0, coeffEstat = 0, coeffDesolv = 0, coeffTors = 0;

private:
    void insertAtom(const std::string &type, double R, double eps, double vol, double sol,
                    double RHB, double epsHB, int hb, int rec, int map, int bond);
    static uint32_t keyOf(const std::string &key);
    std::vector<AtomParam> entries_;
    std::vector<int32_t> slot_;  // 65536 buckets -> entry index + 1
};

// One output map (FG/GridMap.h:27-55) without its energies.
struct MapDef {
    std::string filename, type;
    int32_t atomType = 0;
    bool isCovalent = false, isHBonder = false;
    int32_t hbond = NON;
    double solparProbe = 0, volProbe = 0, Rij = 0, epsij = 0, RijHB = 0, epsijHB = 0;
    std::array<double, kNumTypes> nbpR{}, nbpEps{};
    std::array<int32_t, kNumTypes> xA{}, xB{}, hbonder{};
};

// Everything the map kernels need, after parsing and precomputation.
struct Prepared {
    std::string gpfPath, fldPath, xyzPath, receptorPath, fmapPath, paramFile;
    std::array<int32_t, 3> n{0, 0, 0}, div2{0, 0, 0};
    int64_t P = -1;
    double spacing = 0.375;
    std::array<double, 3> center{0, 0, 0};

    std::vector<std::string> recTypes;
    std::vector<double> x, y, z, w, charge, vol, solpar;
    std::vector<int32_t> type, hbond;

    int numAtomMaps = 0;
    std::vector<MapDef> maps;  // numAtomMaps + e + d

    std::vector<double> eps, desolv, pair;  // pair: [tableTypes][8192][numAtomMaps]
    int32_t tableTypes = 0;
    std::vector<double> rvec, rvec2;        // [N][3]
    std::vector<int32_t> rexp;
    std::vector<uint8_t> disorder;
    std::vector<fg_map_params> mapParams;

    double solparQ = .01097, invDielCal = 0, rSmooth = 0;
    double covHalfWidthSqInv = 1.0, covBarrier = 1000.0;
    std::array<double, 3> covPoint{0, 0, 0};
    bool distDepDiel = false, disorderH = false;
    int32_t hydrogen = -1;

    // summary of written maps: min/max per map (FG/GridMap.cpp:71-101)
    std::vector<double> mapMin, mapMax;
    std::string log;  // GPF echo and notes, for the .glg writer

    int32_t numAtoms() const { return int32_t(x.size()); }
    fg_job_desc desc() const;
};

// FG/InputDataLoader.cpp:100-797 plus FG/main.cpp:150-193.  On failure returns
// the reference's exit code (non-zero) and sets err.
int loadGpf(const std::string &gpf, const fg_load_opts &opts, Prepared &out, std::string &err);
int precompute(Prepared &p, ParamLibrary &lib, std::string &err);

double round3(double a);
// Writes the .fld and every .map file; fills p.mapMin/mapMax.
// maps: unrounded FP64 maps [n_maps + 2 (+ floating grid)], or milli: the
// rounded maps in integer thousandths [n_maps + 2] with the floating grid in
// fmapValues.
int writeOutputs(Prepared &p, double *const *maps, const int32_t *const *milli, const double *fmapValues,
                 const std::string &programName, int threads,
                 std::string &err);
// The .glg log: GPF echo, map summary, completion line, times.
int writeLog(const Prepared &p, FILE *f, const std::string &programName, double realSeconds);

int synthWrite(const fg_synth_spec &spec, const std::string &dir, std::string &err);

}  // namespace fgh
