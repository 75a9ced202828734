// GPF + PDBQT loader (FG/InputDataLoader.cpp:100-938) and the host-side
// precompute of FG/main.cpp:163-193: energies bookkeeping, parameter file,
// pair-energy / desolvation / dielectric tables and bond vectors.
//
// The loader keeps the reference's order-dependent semantics, which the
// reference's own golden maps encode:
//  - receptor type indices are handed out in first-appearance order when the
//    receptor is read; a later receptor_types line renames the indices but
//    does not retype atoms already read (:262-276 vs :615-627);
//  - per-map pair parameters are fixed when ligand_types is read, from the
//    receptor types known at that moment (:527-585);
//  - gridcenter centres the atoms read so far and writes the .xyz file opened
//    by gridfld (:452-491);
//  - w = q * coeff_estat [* 332/diel] uses the coefficient in force at the end
//    of the GPF, before FastGrid reads parameter_file (:789-796,
//    FG/main.cpp:170-171).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "host.hpp"

namespace fgh {
namespace {

enum Keyword {
    K_NONE = -1, K_NULL = 0, K_COMMENT, K_RECEPTOR, K_GRIDFLD, K_NPTS, K_SPACING, K_GRIDCENTER,
    K_LIGAND_TYPES, K_MAP, K_NBP_COEFFS, K_NBP_R_EPS, K_ELECMAP, K_DIEL, K_FMAP, K_DISORDER,
    K_SMOOTH, K_SOL_PAR, K_CONSTANT, K_COVALENTMAP, K_RECEPTOR_TYPES, K_PARAM_FILE, K_DSOLVMAP, K_QASP
};

// The keyword is the text up to the first space (else tab), lower-cased and
// matched as a prefix in the reference's order (FG/InputDataLoader.cpp:799-865).
Keyword classify(const char *line)
{
    const char *end = std::strchr(line, ' ');
    if (!end) end = std::strchr(line, '\t');
    size_t len = end ? size_t(end - line) : std::strlen(line);
    std::string w(line, len);
    for (char &c : w) c = char(std::tolower(uint8_t(c)));
    if (w.empty() || w[0] == '\n') return K_NULL;
    if (w[0] == '#') return K_COMMENT;
    static const std::pair<const char *, Keyword> kTable[] = {
        {"receptor_types", K_RECEPTOR_TYPES}, {"receptor", K_RECEPTOR}, {"gridfld", K_GRIDFLD},
        {"npts", K_NPTS}, {"spacing", K_SPACING}, {"gridcenter", K_GRIDCENTER},
        {"types", K_LIGAND_TYPES}, {"ligand_types", K_LIGAND_TYPES}, {"map", K_MAP},
        {"elecmap", K_ELECMAP}, {"dsolvmap", K_DSOLVMAP}, {"covalentmap", K_COVALENTMAP},
        {"nbp_coeffs", K_NBP_COEFFS}, {"nbp_r_eps", K_NBP_R_EPS}, {"dielectric", K_DIEL},
        {"qasp", K_QASP}, {"fmap", K_FMAP}, {"disorder_h", K_DISORDER}, {"smooth", K_SMOOTH},
        {"sol_par", K_SOL_PAR}, {"constant", K_CONSTANT}, {"parameter_file", K_PARAM_FILE},
    };
    for (const auto &kv : kTable)
        if (w.compare(0, std::strlen(kv.first), kv.first) == 0) return kv.second;
    return K_NONE;
}

// Whitespace-separated words after the keyword, at most maxWords
// (FG/InputDataLoader.cpp:890-938).
std::vector<std::string> wordsAfterKeyword(const char *line, size_t maxWords)
{
    std::vector<std::string> out;
    const char *p = line;
    bool keyword = true;
    while (true) {
        while (std::isspace(uint8_t(*p))) ++p;
        if (!*p) break;
        const char *s = p;
        while (*p && !std::isspace(uint8_t(*p))) ++p;
        if (keyword) keyword = false;
        else out.emplace_back(s, size_t(p - s));
        if (!*p || out.size() >= maxWords) break;
    }
    return out;
}

std::string firstArg(const char *line)
{
    char buf[256] = "";
    std::sscanf(line, "%*s %255s", buf);
    return buf;
}

// First whitespace-delimited word of `s`, at most `maxLen` characters, into
// out (as sscanf "%<maxLen>s" stores it); returns its length, leaving out
// unchanged (and returning 0) when `s` holds no word.
size_t word(const char *s, char *out, size_t maxLen)
{
    while (*s && std::isspace(static_cast<unsigned char>(*s))) ++s;
    size_t n = 0;
    while (s[n] && !std::isspace(static_cast<unsigned char>(s[n])) && n < maxLen) ++n;
    if (n == 0) return 0;
    std::memcpy(out, s, n);
    out[n] = 0;
    return n;
}

// Value of the number at the start of `s` as strtod gives it.  Plain
// decimals (optional sign, at most 15 digits around one point, then a blank
// or the end) are m / 10^d with m and 10^d exact doubles, so the one
// correctly rounded division is the nearest double to the decimal, which is
// strtod's result; any other text goes to strtod.
double readNumber(const char *s)
{
    static const double kPow10[16] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10, 1e11, 1e12, 1e13, 1e14, 1e15};
    const char *q = s;
    while (*q == ' ' || *q == '\t') ++q;
    bool neg = false;
    if (*q == '-' || *q == '+') neg = *q++ == '-';
    uint64_t m = 0;
    int digits = 0, frac = -1;
    for (;; ++q) {
        if (*q >= '0' && *q <= '9') {
            m = m * 10 + uint64_t(*q - '0');
            if (frac >= 0) ++frac;
            if (++digits > 15) return std::strtod(s, nullptr);
        } else if (*q == '.' && frac < 0) {
            frac = 0;
        } else {
            break;
        }
    }
    const bool endOk = *q == 0 || std::isspace(static_cast<unsigned char>(*q));
    if (digits == 0 || !endOk) return std::strtod(s, nullptr);
    const double v = frac > 0 ? double(m) / kPow10[frac] : double(m);
    return neg ? -v : v;
}

// libstdc++ evaluates pow(double, int) in C++98 by binary exponentiation
// (__builtin_powi); the reference's R^xA / R^xB (int exponents,
// FG/PairwiseInteractionEnergies.cpp:72-73) are computed that way.
double powInt(double x, int m)
{
    unsigned n = m < 0 ? 0u - unsigned(m) : unsigned(m);
    double y = (n & 1u) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n & 1u) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

struct Vec3 {
    double x, y, z;
};
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator*(Vec3 a, double f) { return {a.x * f, a.y * f, a.z * f}; }
inline Vec3 operator/(Vec3 a, double f) { return {a.x / f, a.y / f, a.z / f}; }
inline double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm2(Vec3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// Unit vector with the reference's guard: |v|^2 below 1e-6 is taken as 1e-6.
inline Vec3 unitGuarded(Vec3 v)
{
    double r2 = norm2(v);
    if (r2 < kApproxZero) r2 = kApproxZero;
    return v * (1 / std::sqrt(r2));
}

}  // namespace

double round3(double a) { return std::fabs(a) < 0.0005 ? 0 : std::floor(a * 1000 + 0.5) * 0.001; }

int loadGpf(const std::string &gpfPath, const fg_load_opts &opts, Prepared &p, std::string &err)
{
    p = Prepared();
    p.gpfPath = gpfPath;
    ParamLibrary lib(opts.v4 != 0);
    int mapCursor = -1, recTypesDeclared = 0;
    bool hasRecTypes = false;
    std::array<double, 3> mean{0, 0, 0};
    FILE *xyz = nullptr;
    FILE *gpf = gpfPath.empty() ? stdin : std::fopen(gpfPath.c_str(), "r");
    if (!gpf) {
        err = "Sorry, I can't find or open Grid Parameter File \"" + gpfPath + "\"";
        return 911;
    }
    int rc = 0;
    auto fail = [&](int code, const std::string &msg) {
        err = msg;
        rc = code;
    };
    auto readParamFile = [&]() {
        if (!p.paramFile.empty() && !lib.loadFile(p.paramFile))
            fail(-1, "Sorry, I can't find or open " + p.paramFile);
    };

    char line[256];
    while (rc == 0 && std::fgets(line, sizeof line, gpf)) {
        Keyword k = classify(line);
        p.log += "GPF> ";
        p.log += line;
        if (k == K_NONE) continue;
        if (k != K_NULL && k != K_COMMENT)
            if (char *h = std::strchr(line, '#')) *h = 0;
        switch (k) {
        case K_RECEPTOR: {
            p.receptorPath = firstArg(line);
            FILE *rf = std::fopen(p.receptorPath.c_str(), "r");
            if (!rf) {
                fail(-2, "can't find or open receptor PDBQT file \"" + p.receptorPath + "\".");
                break;
            }
            std::array<double, 3> sum{0, 0, 0};
            double qmax = -1e12, qmin = 1e12;
            char rec[8] = "", row[256];
            while (std::fgets(row, sizeof row, rf)) {
                // the record name is the first word, at most 6 characters
                // (sscanf "%6s"); a blank line keeps the previous one
                word(row, rec, 6);
                if (std::strncmp(rec, "ATOM", 4) && std::strncmp(rec, "HETA", 4) && std::strncmp(rec, "CHAR", 4))
                    continue;
                // fixed columns: x 30, y 38, z 46, charge 70, type 77; each
                // value is read as sscanf "%lf" reads it (strtod of the text
                // from that column, 0 when no number starts there)
                size_t len = std::strlen(row);
                double c[4] = {0, 0, 0, 0};
                const size_t col[4] = {30, 38, 46, 70};
                for (int i = 0; i < 4; ++i)
                    if (len > col[i]) c[i] = readNumber(row + col[i]);
                char t[8] = "";
                if (len > 77) word(row + 77, t, 7);
                std::string tname(t);
                AtomParam *a = lib.find(tname);
                if (!a) {
                    fail(-1, "receptor file contains unknown type: '" + tname + "'");
                    break;
                }
                if (a->recIndex < 0) {
                    a->recIndex = int32_t(p.recTypes.size());
                    p.recTypes.push_back(a->type.substr(0, 2));
                }
                p.x.push_back(c[0]);
                p.y.push_back(c[1]);
                p.z.push_back(c[2]);
                p.charge.push_back(c[3]);
                p.type.push_back(a->recIndex);
                p.solpar.push_back(a->solpar);
                p.vol.push_back(a->vol);
                p.hbond.push_back(a->hbond);
                qmax = std::max(qmax, c[3]);
                qmin = std::min(qmin, c[3]);
                sum[0] += c[0];
                sum[1] += c[1];
                sum[2] += c[2];
            }
            std::fclose(rf);
            if (rc) break;
            if (hasRecTypes && int(p.recTypes.size()) != recTypesDeclared) {
                fail(-2, "The number of atom types found in the receptor PDBQT does not match receptor_types");
                break;
            }
            if (qmax == 0 && qmin == 0) {
                fail(-2, "No partial atomic charges were found in the receptor PDBQT file " + p.receptorPath);
                break;
            }
            for (int i = 0; i < 3; ++i) mean[size_t(i)] = sum[size_t(i)] / double(p.x.size());
            break;
        }
        case K_GRIDFLD: {
            p.fldPath = firstArg(line);
            if (p.fldPath.find(".fld") == std::string::npos) {
                fail(-2, "Grid data file needs the extension \".fld\" for AVS input");
                break;
            }
            p.xyzPath = p.fldPath;
            p.xyzPath.replace(p.xyzPath.find("fld"), 3, "xyz");
            xyz = std::fopen(p.xyzPath.c_str(), "w");
            if (!xyz) fail(-2, "can't create grid extrema data file " + p.xyzPath);
            break;
        }
        case K_NPTS: {
            int v[3] = {0, 0, 0};
            std::sscanf(line, "%*s %d %d %d", &v[0], &v[1], &v[2]);
            for (int i = 0; i < 3 && rc == 0; ++i) {
                // checkSize, FG/InputDataLoader.cpp:868-887
                if (v[i] < 0) fail(-2, "Negative number of grid elements!  Aborting.");
                else if (v[i] > kMaxGridPts) v[i] = kMaxGridPts;
                else if (v[i] % 2 == 1) v[i] -= 1;
                p.div2[size_t(i)] = v[i] / 2;
                p.n[size_t(i)] = v[i] + 1;
            }
            p.P = int64_t(p.n[0]) * p.n[1] * p.n[2];
            break;
        }
        case K_SPACING:
            std::sscanf(line, "%*s %lf", &p.spacing);
            break;
        case K_GRIDCENTER: {
            if (firstArg(line).compare(0, 4, "auto") == 0) p.center = mean;
            else std::sscanf(line, "%*s %lf %lf %lf", &p.center[0], &p.center[1], &p.center[2]);
            for (size_t i = 0; i < p.x.size(); ++i) {
                p.x[i] -= p.center[0];
                p.y[i] -= p.center[1];
                p.z[i] -= p.center[2];
            }
            if (xyz) {
                for (size_t i = 0; i < 3; ++i) {
                    double ext = double(p.div2[i]) * p.spacing;
                    std::fprintf(xyz, "%.3lf %.3lf\n", p.center[i] - ext, p.center[i] + ext);
                }
                std::fclose(xyz);
                xyz = nullptr;
            }
            break;
        }
        case K_LIGAND_TYPES: {
            std::vector<std::string> types = wordsAfterKeyword(line, kMaxAtomTypes);
            for (std::string &t : types) t = t.substr(0, 2);
            for (size_t i = 0; i < types.size() && rc == 0; ++i) {
                AtomParam *a = lib.find(types[i]);
                if (!a) fail(-1, "unknown ligand atom type " + types[i]);
                else a->mapIndex = int32_t(i);
            }
            if (rc) break;
            p.numAtomMaps = int(types.size());
            p.maps.assign(types.size() + 2, MapDef());
            if (p.P < 0) {
                fail(-2, "You need to set the number of grid points using \"npts\" before setting the ligand atom types");
                break;
            }
            for (size_t i = 0; i < types.size() && rc == 0; ++i) {
                MapDef &m = p.maps[i];
                const AtomParam &a = *lib.find(types[i]);
                m.type = types[i];
                m.isCovalent = types[i] == "Z";
                m.atomType = a.mapIndex;
                m.solparProbe = a.solpar;
                m.volProbe = a.vol;
                m.Rij = a.Rij;
                m.epsij = a.epsij;
                m.hbond = a.hbond;
                m.RijHB = a.RijHB;
                m.epsijHB = a.epsijHB;
                m.isHBonder = a.hbond > 0;
                for (size_t j = 0; j < p.recTypes.size(); ++j) {
                    const AtomParam *r = lib.find(p.recTypes[j]);
                    if (!r) {
                        fail(-2, "unknown receptor type " + p.recTypes[j]);
                        break;
                    }
                    m.nbpR[j] = (m.Rij + r->Rij) / 2;
                    m.nbpEps[j] = std::sqrt(m.epsij * r->epsij);
                    m.xA[j] = 12;
                    m.xB[j] = 6;
                    m.hbonder[j] = 0;
                    bool recDonor = r->hbond == DS || r->hbond == D1;
                    bool mapDonor = m.hbond == DS || m.hbond == D1;
                    if (m.hbond > D1 && recDonor) {  // acceptor probe vs donor receptor: 12-10
                        m.xB[j] = 10;
                        m.hbonder[j] = 1;
                        m.isHBonder = true;
                        m.nbpR[j] = m.RijHB;
                        m.nbpEps[j] = m.epsijHB;
                    } else if (mapDonor && r->hbond > D1) {  // donor probe vs acceptor receptor
                        m.xB[j] = 10;
                        m.hbonder[j] = 1;
                        m.isHBonder = true;
                        m.nbpR[j] = r->RijHB;
                        m.nbpEps[j] = r->epsijHB;
                    }
                }
            }
            break;
        }
        case K_RECEPTOR_TYPES: {
            std::vector<std::string> types = wordsAfterKeyword(line, kMaxAtomTypes);
            recTypesDeclared = int(types.size());
            hasRecTypes = true;
            p.recTypes.clear();
            for (const std::string &t : types) p.recTypes.push_back(t.substr(0, 2));
            for (size_t i = 0; i < types.size(); ++i) {
                AtomParam *a = lib.find(types[i]);
                if (!a) {
                    fail(-2, "Unknown receptor type: \"" + types[i] + "\"");
                    break;
                }
                a->recIndex = int32_t(i);
            }
            break;
        }
        case K_SOL_PAR: {
            char t[256] = "";
            double v = 0, s = 0;
            std::sscanf(line, "%*s %255s %lf %lf", t, &v, &s);
            if (AtomParam *a = lib.find(std::string(t).substr(0, 7))) {
                a->vol = v;
                a->solpar = s;
                if (a->mapIndex >= 0 && size_t(a->mapIndex) < p.maps.size()) p.maps[size_t(a->mapIndex)].solparProbe = s;
            }
            break;
        }
        case K_MAP:
            ++mapCursor;
            if (mapCursor >= p.numAtomMaps) fail(-2, "Too many \"map\" keywords");
            else p.maps[size_t(mapCursor)].filename = firstArg(line);
            break;
        case K_ELECMAP:
        case K_DSOLVMAP:
            if (p.maps.size() < 2) fail(-2, "elecmap/dsolvmap given before ligand_types");
            else p.maps[p.maps.size() - (k == K_ELECMAP ? 2 : 1)].filename = firstArg(line);
            break;
        case K_COVALENTMAP: {
            double hw = 0;
            std::sscanf(line, "%*s %lf %lf %lf %lf %lf", &hw, &p.covBarrier, &p.covPoint[0], &p.covPoint[1],
                        &p.covPoint[2]);
            p.covHalfWidthSqInv = 1 / (hw * hw);
            for (size_t i = 0; i < 3; ++i) p.covPoint[i] -= p.center[i];
            break;
        }
        case K_DISORDER:
            p.disorderH = true;
            break;
        case K_SMOOTH:
            std::sscanf(line, "%*s %lf", &p.rSmooth);
            break;
        case K_QASP:
            std::sscanf(line, "%*s %lf", &p.solparQ);
            break;
        case K_DIEL: {
            double diel = 0;
            std::sscanf(line, "%*s %lf", &diel);
            if (diel < 0) {
                p.distDepDiel = true;
            } else {
                p.distDepDiel = false;
                if (diel <= kApproxZero) diel = 40;
                p.invDielCal = 332 / diel;
            }
            break;
        }
        case K_FMAP:
            p.fmapPath = firstArg(line);
            break;
        case K_PARAM_FILE:
            p.paramFile = firstArg(line);
            if (opts.ag421_param_order) readParamFile();
            break;
        default:
            break;
        }
    }
    if (xyz) std::fclose(xyz);
    if (gpf != stdin) std::fclose(gpf);
    if (rc) return rc;
    if (mapCursor < p.numAtomMaps - 1) {
        err = "Too few \"map\" keywords";
        return -2;
    }
    if (p.maps.size() < 2 || p.maps.back().filename.empty()) {
        err = "No desolvation map file defined.";
        return -2;
    }
    if (p.maps[p.maps.size() - 2].filename.empty()) {
        err = "No electrostatic map file defined.";
        return -2;
    }
    p.maps[p.maps.size() - 2].type.clear();
    p.maps.back().type.clear();
    size_t N = p.x.size();
    p.w.resize(N);
    for (size_t i = 0; i < N; ++i)
        p.w[i] = p.distDepDiel ? p.charge[i] * lib.coeffEstat : p.charge[i] * lib.coeffEstat * p.invDielCal;
    if (!opts.ag421_param_order) readParamFile();
    if (rc) return rc;
    return precompute(p, lib, err);
}

namespace {

// r^x at r = k/100, k = 0..8191, with libm's pow(double, double) as the
// reference calls it (FG/PairwiseInteractionEnergies.cpp:93-97).  Every row
// of every table reads the same few exponents, so each exponent's powers are
// computed once per process; the values are those of the direct calls.
const double *powersOf(double x)
{
    static std::mutex mu;
    static std::map<double, std::unique_ptr<double[]>> cache;
    std::lock_guard<std::mutex> lock(mu);
    std::unique_ptr<double[]> &slot = cache[x];
    if (!slot) {
        slot.reset(new double[kTableLen]);
        for (int k = 0; k < kTableLen; ++k) slot[size_t(k)] = std::pow(double(k) * (1 / 100.0), x);
    }
    return slot.get();
}

// One row of the pair-energy table (FG/PairwiseInteractionEnergies.cpp:93-114):
// min(1e5, cA / r^xA - cB / r^xB) at r = k/100, entry 0 = 1e5 and entry 8191
// = 0, then for k >= 1 the minimum over entries [k - half, k + half) (the
// reference's running minimum, here with a monotonic queue: the same values).
void pairRow(double cA, double cB, int xA, int xB, int half, double *row)
{
    const double *rA = powersOf(double(xA)), *rB = powersOf(double(xB));
    for (int k = 1; k < kTableLen; ++k) row[k] = std::min(100000.0, cA / rA[k] - cB / rB[k]);
    row[0] = 100000.0;
    row[kTableLen - 1] = 0;
    if (half <= 0) return;
    std::vector<double> smooth(kTableLen);
    std::deque<int> q;  // window indices, values increasing from the front
    int next = 0;       // next index to enter the window
    for (int k = 1; k < kTableLen; ++k) {
        const int lo = std::max(0, k - half), hi = std::min(kTableLen, k + half);
        for (; next < hi; ++next) {
            while (!q.empty() && row[q.back()] >= row[next]) q.pop_back();
            q.push_back(next);
        }
        while (!q.empty() && q.front() < lo) q.pop_front();
        smooth[size_t(k)] = q.empty() ? 100000.0 : std::min(100000.0, row[q.front()]);
    }
    for (int k = 1; k < kTableLen; ++k) row[k] = smooth[size_t(k)];
}

// Rows depend only on (cA, cB, xA, xB, half); preparations that share a
// parameter set (a batch of pockets) share their rows.  At most kRowCache
// rows (64 KiB each) are kept per process.
constexpr size_t kRowCache = 1024;

void cachedPairRow(double cA, double cB, int xA, int xB, int half, double *row)
{
    using Key = std::tuple<double, double, int, int, int>;
    static std::mutex mu;
    static std::map<Key, std::unique_ptr<double[]>> cache;
    const Key key{cA, cB, xA, xB, half};
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            std::copy(it->second.get(), it->second.get() + kTableLen, row);
            return;
        }
    }
    pairRow(cA, cB, xA, xB, half, row);
    std::lock_guard<std::mutex> lock(mu);
    if (cache.size() < kRowCache && !cache.count(key)) {
        std::unique_ptr<double[]> v(new double[kTableLen]);
        std::copy(row, row + kTableLen, v.get());
        cache.emplace(key, std::move(v));
    }
}

// Pair-energy table T[type][k][map] (FG/PairwiseInteractionEnergies.cpp:59-115):
// clamped 12-6 / 12-10 potential at r = k/100, then a running minimum over
// +-rSmooth/2.
int pairTables(Prepared &p, std::string &err)
{
    // Rows past the declared receptor types stay zero, as in the reference's
    // zero-initialised [32][8192][16] table; atoms can carry such an index
    // when receptor_types follows receptor and lists fewer types.
    const size_t T = p.recTypes.size(), M = size_t(p.numAtomMaps);
    size_t rows = std::max<size_t>(T, 1);
    for (int32_t t : p.type) rows = std::max(rows, size_t(t) + 1);
    p.tableTypes = int32_t(rows);
    p.pair.assign(rows * kTableLen * std::max<size_t>(M, 1), 0.0);
    const int half = int(p.rSmooth * 100 / 2);
    std::vector<double> row(kTableLen);
    for (size_t m = 0; m < M; ++m) {
        const MapDef &md = p.maps[m];
        if (md.isCovalent) continue;
        for (size_t t = 0; t < T; ++t) {
            const int xA = md.xA[t], xB = md.xB[t];
            const double scale = md.nbpEps[t] / (xA - xB);
            const double cA = scale * powInt(md.nbpR[t], xA) * xB;
            const double cB = scale * powInt(md.nbpR[t], xB) * xA;
            if (std::isnan(cA) || std::isnan(cB)) {
                err = "Van der Waals coefficient cA/cB is not a number.";
                return -2;
            }
            if (xA == 0 || xB == 0) {
                err = "Van der Waals exponent is 0.";
                return -2;
            }
            cachedPairRow(cA, cB, xA, xB, half, row.data());
            for (int k = 0; k < kTableLen; ++k) p.pair[(t * kTableLen + size_t(k)) * M + m] = row[size_t(k)];
        }
    }
    return 0;
}

// Bond vectors for directional H-bonds (FG/BondVectors.cpp:45-393).  Partners
// are searched among the 20 atoms on either side in file order.  Atoms with no
// partner keep zero vectors and rexp = 0.
void bondVectors(Prepared &p, ParamLibrary &lib)
{
    const int N = p.numAtoms();
    p.rvec.assign(size_t(3) * size_t(N), 0.0);
    p.rvec2.assign(size_t(3) * size_t(N), 0.0);
    p.rexp.assign(size_t(N), 0);
    p.disorder.assign(size_t(N), 0);
    const int32_t H = lib.recIndexOf("HD"), Hnon = lib.recIndexOf("H"), C = lib.recIndexOf("C"),
                  Car = lib.recIndexOf("A"), O = lib.recIndexOf("OA"), S = lib.recIndexOf("SA");
    auto pos = [&](int i) { return Vec3{p.x[size_t(i)], p.y[size_t(i)], p.z[size_t(i)]}; };
    auto store = [](std::vector<double> &v, int i, Vec3 a) {
        v[size_t(3 * i)] = a.x;
        v[size_t(3 * i + 1)] = a.y;
        v[size_t(3 * i + 2)] = a.z;
    };
    auto load = [](const std::vector<double> &v, int i) {
        return Vec3{v[size_t(3 * i)], v[size_t(3 * i + 1)], v[size_t(3 * i + 2)]};
    };
    const int32_t *T = p.type.data();
    int b1 = 0, b2 = 0, b3 = 0;  // bonded partners; persist across atoms as in the reference
    for (int ia = 0; ia < N; ++ia) {
        const int from = std::max(ia - 20, 0), to = std::min(ia + 20, N - 1);
        const Vec3 pa = pos(ia);
        const int32_t hb = p.hbond[size_t(ia)];
        if (hb == D1) {
            for (int ib = from; ib <= to; ++ib) {
                if (ib == ia) continue;
                Vec3 d = pa - pos(ib);
                double rd2 = norm2(d);
                if (rd2 >= 1.90) continue;
                if (rd2 < kApproxZero) rd2 = kApproxZero;
                bool toOS = T[ib] == O || T[ib] == S;
                p.rexp[size_t(ia)] = toOS ? 4 : 2;
                if (toOS && p.disorderH) p.disorder[size_t(ia)] = 1;
                store(p.rvec, ia, d * (1 / std::sqrt(rd2)));
                break;
            }
        } else if (hb == A2) {
            int nb = 0, ib = from;
            for (; ib <= to; ++ib) {
                if (ib == ia) continue;
                double rd2 = norm2(pa - pos(ib));
                bool isH = T[ib] == H || T[ib] == Hnon;
                if ((rd2 < 3.61 && !isH) || (rd2 < 1.69 && isH)) {
                    if (nb == 1) {
                        nb = 2;
                        b2 = ib;
                    }
                    if (nb == 0) {
                        nb = 1;
                        b1 = ib;
                    }
                }
            }
            if (nb == 1) {  // carbonyl O=C-X
                Vec3 rv = unitGuarded(pa - pos(b1));
                store(p.rvec, ia, rv);
                for (int j = from; j <= to; ++j) {
                    if (j == b1 || j == ia) continue;
                    double rd2 = norm2(pos(b1) - pos(j));
                    if ((rd2 < 2.89 && T[j] != H) || (rd2 < 1.69 && T[j] == H))
                        store(p.rvec2, ia, unitGuarded(cross(rv, unitGuarded(pos(j) - pos(b1)))));
                }
            } else if (nb == 2) {
                if ((T[b1] == H || T[b2] == H) && T[b1] != T[b2] && p.disorderH) {
                    // disordered hydroxyl; the second test reads T[b1] as the
                    // reference does (FG/BondVectors.cpp:240).  The scan leaves
                    // ib one past its range when neither test matches; past the
                    // last atom that reads the origin here.
                    if (T[b1] == C || T[b1] == Car) ib = b1;
                    if (T[b2] == C || T[b1] == Car) ib = b2;
                    p.disorder[size_t(ia)] = 1;
                    Vec3 pb = ib < N ? pos(ib) : Vec3{0, 0, 0};
                    store(p.rvec, ia, unitGuarded(pa - pb));
                } else {  // hydroxyl / ether X1-O-X2
                    Vec3 rv2 = unitGuarded(pos(b2) - pos(b1));
                    store(p.rvec2, ia, rv2);
                    double rdot = dot(pa - pos(b1), rv2);
                    store(p.rvec, ia, unitGuarded(pa - (rv2 * rdot + pos(b1))));
                }
            }
        } else if (hb == A1) {
            int nb = 0;
            for (int ib = from; ib <= to; ++ib) {
                if (ib == ia) continue;
                double rd2 = norm2(pa - pos(ib));
                bool isH = T[ib] == H || T[ib] == Hnon;
                if ((rd2 < 2.89 && !isH) || (rd2 < 1.69 && isH)) {
                    if (nb == 2) {
                        nb = 3;
                        b3 = ib;
                    }
                    if (nb == 1) {
                        nb = 2;
                        b2 = ib;
                    }
                    if (nb == 0) {
                        nb = 1;
                        b1 = ib;
                    }
                }
            }
            if (nb == 1) store(p.rvec, ia, unitGuarded(pa - pos(b1)));
            else if (nb == 2) store(p.rvec, ia, unitGuarded(pa - (pos(b2) + pos(b1)) / 2));
            else if (nb == 3) store(p.rvec, ia, unitGuarded(pa - (pos(b1) + pos(b2) + pos(b3)) / 3));
        }
    }
    (void)load;
}

}  // namespace

int precompute(Prepared &p, ParamLibrary &lib, std::string &err)
{
    // inverse distance-dependent dielectric of Mehler & Solmajer at r = k/100
    // (FG/autogrid.h:106-112, FG/InputDataLoader.cpp:722-723), and the
    // desolvation Gaussian, sigma 3.6 A, times coeff_desolv
    // (FG/DesolvExpFunc.cpp:27-43; entry 0 is only read for r < 0.01 A and is
    // 0 here).  Both are computed once per process (per coefficient).
    static std::mutex mu;
    static std::vector<double> epsTable;
    static std::map<double, std::vector<double>> desolvTables;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (epsTable.empty()) {
            epsTable.assign(kTableLen, 0.0);
            for (int k = 0; k < kTableLen; ++k) {
                double r = double(k) * (1 / 100.0);
                double E = std::exp(-0.3153767175 * r);
                epsTable[size_t(k)] = (332.0 + (332 * 7.7839) * E) / (78.4 + -66.57180475 * E);
            }
        }
        std::vector<double> &dt = desolvTables[lib.coeffDesolv];
        if (dt.empty()) {
            dt.assign(kTableLen, 0.0);
            const double sigma = 3.6, f = -1 / (2 * sigma * sigma);
            for (int k = 1; k < kTableLen; ++k) {
                double r = double(k) * (1 / 100.0);
                dt[size_t(k)] = std::exp((r * r) * f);
                dt[size_t(k)] *= lib.coeffDesolv;
            }
        }
        if (p.distDepDiel) p.eps = epsTable;
        else p.eps.assign(kTableLen, 0.0);
        p.desolv = dt;
    }
    if (int rc = pairTables(p, err)) return rc;
    bondVectors(p, lib);
    p.hydrogen = lib.recIndexOf("HD");
    p.mapParams.resize(size_t(p.numAtomMaps));
    for (size_t m = 0; m < size_t(p.numAtomMaps); ++m) {
        const MapDef &d = p.maps[m];
        p.mapParams[m] = fg_map_params{d.hbond, d.isHBonder ? 1 : 0, d.isCovalent ? 1 : 0, d.solparProbe, d.volProbe};
    }
    return 0;
}

fg_job_desc Prepared::desc() const
{
    fg_job_desc d{};
    d.grid = fg_grid{n[0], n[1], n[2], div2[0], div2[1], div2[2], spacing};
    d.atoms.n = numAtoms();
    d.atoms.x = x.data();
    d.atoms.y = y.data();
    d.atoms.z = z.data();
    d.atoms.w = w.data();
    d.atoms.charge = charge.data();
    d.atoms.vol = vol.data();
    d.atoms.solpar = solpar.data();
    d.atoms.type = type.data();
    d.atoms.hbond = hbond.data();
    d.geom = fg_hbond_geom{rvec.data(), rvec2.data(), rexp.data(), disorder.data()};
    d.tables = fg_tables{tableTypes, int32_t(numAtomMaps), pair.data(), desolv.data(), eps.data()};
    d.n_maps = numAtomMaps;
    d.maps = mapParams.data();
    // covalent maps are computed on the device from their parameters
    d.covalent = nullptr;
    for (int m = 0; m < numAtomMaps; ++m) d.cov_on = d.cov_on || maps[size_t(m)].isCovalent;
    for (int k = 0; k < 3; ++k) d.cov_point[k] = covPoint[size_t(k)];
    d.cov_hw_sq_inv = covHalfWidthSqInv;
    d.cov_barrier = covBarrier;
    d.floating_grid = fmapPath.empty() ? 0 : 1;
    d.solpar_q = solparQ;
    d.hydrogen = hydrogen;
    d.dist_dep_diel = distDepDiel ? 1 : 0;
    return d;
}

}  // namespace fgh
