// AD4 parameter library (FG/ParameterLibrary.cpp:34-284).
// The built-in values are the AutoDock 4.0 / 4.1 force-field constants of
// fastgrid_master/autodock/default_parameters.h:1-82.
#include <cctype>
#include <cstring>
#include <strings.h>

#include "host.hpp"

namespace fgh {
namespace {

struct BuiltinAtom {
    const char *type;
    double R, eps, vol, sol, RHB, epsHB;
    int hb, bond;
};

// type, Rii, epsii, vol, solpar, Rij_hb, epsij_hb, hbond class, bond index
constexpr BuiltinAtom kBuiltin[] = {
    {"H", 2.00, 0.020, 0.0000, 0.00051, 0.0, 0.0, 0, 3},    {"HD", 2.00, 0.020, 0.0000, 0.00051, 0.0, 0.0, 2, 3},
    {"HS", 2.00, 0.020, 0.0000, 0.00051, 0.0, 0.0, 1, 3},   {"C", 4.00, 0.150, 33.5103, -0.00143, 0.0, 0.0, 0, 0},
    {"A", 4.00, 0.150, 33.5103, -0.00052, 0.0, 0.0, 0, 0},  {"N", 3.50, 0.160, 22.4493, -0.00162, 0.0, 0.0, 0, 1},
    {"NA", 3.50, 0.160, 22.4493, -0.00162, 1.9, 5.0, 4, 1}, {"NS", 3.50, 0.160, 22.4493, -0.00162, 1.9, 5.0, 3, 1},
    {"OA", 3.20, 0.200, 17.1573, -0.00251, 1.9, 5.0, 5, 2}, {"OS", 3.20, 0.200, 17.1573, -0.00251, 1.9, 5.0, 3, 2},
    {"F", 3.09, 0.080, 15.4480, -0.00110, 0.0, 0.0, 0, 4},  {"Mg", 1.30, 0.875, 1.5600, -0.00110, 0.0, 0.0, 0, 4},
    {"MG", 1.30, 0.875, 1.5600, -0.00110, 0.0, 0.0, 0, 4},  {"P", 4.20, 0.200, 38.7924, -0.00110, 0.0, 0.0, 0, 5},
    {"SA", 4.00, 0.200, 33.5103, -0.00214, 2.5, 1.0, 5, 6}, {"S", 4.00, 0.200, 33.5103, -0.00214, 0.0, 0.0, 0, 6},
    {"Cl", 4.09, 0.276, 35.8235, -0.00110, 0.0, 0.0, 0, 4}, {"CL", 4.09, 0.276, 35.8235, -0.00110, 0.0, 0.0, 0, 4},
    {"Ca", 1.98, 0.550, 2.7700, -0.00110, 0.0, 0.0, 0, 4},  {"CA", 1.98, 0.550, 2.7700, -0.00110, 0.0, 0.0, 0, 4},
    {"Mn", 1.30, 0.875, 2.1400, -0.00110, 0.0, 0.0, 0, 4},  {"MN", 1.30, 0.875, 2.1400, -0.00110, 0.0, 0.0, 0, 4},
    {"Fe", 1.30, 0.010, 1.8400, -0.00110, 0.0, 0.0, 0, 4},  {"FE", 1.30, 0.010, 1.8400, -0.00110, 0.0, 0.0, 0, 4},
    {"Zn", 1.48, 0.550, 1.7000, -0.00110, 0.0, 0.0, 0, 4},  {"ZN", 1.48, 0.550, 1.7000, -0.00110, 0.0, 0.0, 0, 4},
    {"Br", 4.33, 0.389, 42.5661, -0.00110, 0.0, 0.0, 0, 4}, {"BR", 4.33, 0.389, 42.5661, -0.00110, 0.0, 0.0, 0, 4},
    {"I", 4.72, 0.550, 55.0585, -0.00110, 0.0, 0.0, 0, 4},  {"Z", 4.00, 0.150, 33.5103, -0.00143, 0.0, 0.0, 0, 0},
    {"G", 4.00, 0.150, 33.5103, -0.00143, 0.0, 0.0, 0, 0},  {"GA", 4.00, 0.150, 33.5103, -0.00052, 0.0, 0.0, 0, 0},
    {"J", 4.00, 0.150, 33.5103, -0.00143, 0.0, 0.0, 0, 0},  {"Q", 4.00, 0.150, 33.5103, -0.00143, 0.0, 0.0, 0, 0},
};

// Free-energy coefficients vdW, hbond, estat, desolv, tors.
constexpr double kCoeffAD40[5] = {0.1560, 0.0974, 0.1465, 0.1159, 0.2744};
constexpr double kCoeffAD41[5] = {0.1662, 0.1209, 0.1406, 0.1322, 0.2983};

}  // namespace

ParamLibrary::ParamLibrary(bool v4) : slot_(65536, 0)
{
    const double *c = v4 ? kCoeffAD40 : kCoeffAD41;
    coeffVdW = c[0];
    coeffHbond = c[1];
    coeffEstat = c[2];
    coeffDesolv = c[3];
    coeffTors = c[4];
    for (const BuiltinAtom &a : kBuiltin)
        insertAtom(a.type, a.R, a.eps, a.vol, a.sol, a.RHB, a.epsHB, a.hb, -1, -1, a.bond);
}

// Bucket of a type name: first character plus 256 x second character
// (FG/ParameterLibrary.cpp:67-78), so "NA" and "NAx" share an entry.
uint32_t ParamLibrary::keyOf(const std::string &key)
{
    if (key.empty()) return 0;
    if (key.size() == 1) return uint8_t(key[0]);
    return uint32_t(uint8_t(key[0])) + 256u * uint32_t(uint8_t(key[1]));
}

AtomParam *ParamLibrary::find(const std::string &key)
{
    int32_t s = slot_[keyOf(key)];
    return s ? &entries_[size_t(s - 1)] : nullptr;
}

int32_t ParamLibrary::recIndexOf(const std::string &key)
{
    const AtomParam *p = find(key);
    return p ? p->recIndex : -1;
}

// Epsilons are weighted by the coefficients in force when the entry is read
// (FG/ParameterLibrary.cpp:214-216); a later entry replaces the whole record.
void ParamLibrary::insertAtom(const std::string &type, double R, double eps, double vol, double sol,
                              double RHB, double epsHB, int hb, int rec, int map, int bond)
{
    AtomParam a;
    a.type = type.substr(0, 7);
    a.Rij = R;
    a.epsij = eps * coeffVdW;
    a.vol = vol;
    a.solpar = sol;
    a.RijHB = RHB;
    a.epsijHB = epsHB * coeffHbond;
    a.hbond = hb <= A2 ? hb : NON;
    a.recIndex = rec;
    a.mapIndex = map;
    a.bondIndex = bond;
    int32_t &s = slot_[keyOf(a.type)];
    if (!s) {
        entries_.push_back(a);
        s = int32_t(entries_.size());
    } else {
        entries_[size_t(s - 1)] = a;
    }
}

// One line of an AD4 parameter file: keyword matched case-insensitively as a
// prefix of the known keywords (FG/ParameterLibrary.cpp:236-284).
void ParamLibrary::readLine(const char *line)
{
    size_t j = 0;
    char word[256];
    while (line[j] && line[j] != ' ' && line[j] != '\t' && line[j] != '\n' && j < sizeof word - 1) {
        word[j] = char(std::tolower(uint8_t(line[j])));
        ++j;
    }
    word[j] = 0;
    if (!word[0] || word[0] == '#') return;
    static const char *const kKeys[] = {"FE_coeff_vdW", "FE_coeff_hbond", "FE_coeff_estat",
                                        "FE_coeff_desolv", "FE_coeff_tors", "atom_par"};
    int key = -1;
    for (int i = 0; i < 6 && key < 0; ++i)
        if (strncasecmp(kKeys[i], word, j) == 0) key = i;
    double v = 0;
    double *coeff[5] = {&coeffVdW, &coeffHbond, &coeffEstat, &coeffDesolv, &coeffTors};
    if (key >= 0 && key < 5) {
        if (std::sscanf(line, "%*s %lf", &v) == 1) *coeff[key] = v;
    } else if (key == 5) {
        char t[256] = "";
        double R = 0, eps = 0, vol = 0, sol = 0, RHB = 0, epsHB = 0;
        int hb = 0, rec = 0, map = 0, bond = 0;
        int nf = std::sscanf(line, "%*s %255s %lf %lf %lf %lf %lf %lf %d %d %d %d", t, &R, &eps, &vol, &sol,
                             &RHB, &epsHB, &hb, &rec, &map, &bond);
        if (nf >= 2) insertAtom(t, R, eps, vol, sol, RHB, epsHB, hb, rec, map, bond);
    }
}

bool ParamLibrary::loadFile(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    char line[256];
    while (std::fgets(line, sizeof line, f)) readLine(line);
    std::fclose(f);
    return true;
}

}  // namespace fgh
