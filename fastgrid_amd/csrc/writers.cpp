// Output writers: .map files (FG/GridMap.cpp:127-145, 183-212), the AVS .fld
// descriptor (FG/Utils.cpp:40-84) and the .glg log summary
// (FG/GridMap.cpp:71-101).
//
// A map file is six header lines followed by one value per grid point, x
// fastest.  Each value is the 3-decimal rounding of the FP64 sum, printed as
// "0" when it is exactly zero and as printf("%.3f") otherwise.  Because a
// rounded value v equals k * 0.001 for an integer k, "%.3f" of v prints k/1000
// exactly; the formatter below writes that integer directly, in parallel
// chunks, instead of calling printf 20 million times.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "host.hpp"

namespace fgh {
namespace {

// Appends the text of one rounded value; returns the number of bytes.
inline size_t formatValue(double v, char *out)
{
    if (v == 0) {
        out[0] = '0';
        out[1] = '\n';
        return 2;
    }
    long long k = std::llround(v * 1000.0);
    char *p = out;
    if (k < 0 || (k == 0 && std::signbit(v))) *p++ = '-';
    unsigned long long a = k < 0 ? 0ull - (unsigned long long)k : (unsigned long long)k;
    unsigned long long ip = a / 1000, fp = a % 1000;
    char tmp[24];
    int n = 0;
    do {
        tmp[n++] = char('0' + ip % 10);
        ip /= 10;
    } while (ip);
    while (n) *p++ = tmp[--n];
    *p++ = '.';
    *p++ = char('0' + fp / 100);
    *p++ = char('0' + fp / 10 % 10);
    *p++ = char('0' + fp % 10);
    *p++ = '\n';
    return size_t(p - out);
}

// A map as unrounded FP64 values (rounded here) or as the rounded values'
// integer thousandths.
struct MapValues {
    const double *v;
    const int32_t *milli;
    double rounded(int64_t i) const { return milli ? double(milli[i]) * 0.001 : round3(v[i]); }
};

// Formats values [lo, hi) after rounding; the buffer holds at most 24 bytes
// per value.  A rounded value is k * 0.001 for the integer k, so the
// thousandths give the same text and the same extremes.
void formatRange(MapValues v, int64_t lo, int64_t hi, std::string &buf, double &mn, double &mx)
{
    buf.resize(size_t(hi - lo) * 24);
    char *p = &buf[0];
    double a = 1e12, b = -1e12;
    for (int64_t i = lo; i < hi; ++i) {
        double r = v.rounded(i);
        a = std::min(a, r);
        b = std::max(b, r);
        // values beyond the fixed-point range fall back to printf
        if (std::fabs(r) < 9.0e15) p += formatValue(r, p);
        else p += std::snprintf(p, 24, "%.3f\n", r);
    }
    buf.resize(size_t(p - &buf[0]));
    mn = a;
    mx = b;
}

// Thousandths [lo, hi) of a map rounded on the device: the text formatValue
// gives their values k * 0.001 (0 exactly when k is 0), written from the
// integers directly; mn and mx as formatRange gives them.
void formatMilliRange(const int32_t *k, int64_t lo, int64_t hi, std::string &buf, double &mn, double &mx)
{
    buf.resize(size_t(hi - lo) * 16);
    char *p = &buf[0];
    int32_t a = INT32_MAX, b = INT32_MIN;
    for (int64_t i = lo; i < hi; ++i) {
        const int32_t v = k[i];
        a = std::min(a, v);
        b = std::max(b, v);
        if (v == 0) {
            *p++ = '0';
            *p++ = '\n';
            continue;
        }
        if (v < 0) *p++ = '-';
        uint32_t u = v < 0 ? 0u - uint32_t(v) : uint32_t(v);
        uint32_t ip = u / 1000, fp = u % 1000;
        char tmp[12];
        int n = 0;
        do {
            tmp[n++] = char('0' + ip % 10);
            ip /= 10;
        } while (ip);
        while (n) *p++ = tmp[--n];
        *p++ = '.';
        *p++ = char('0' + fp / 100);
        *p++ = char('0' + fp / 10 % 10);
        *p++ = char('0' + fp % 10);
        *p++ = '\n';
    }
    buf.resize(size_t(p - &buf[0]));
    mn = hi > lo ? double(a) * 0.001 : 1e12;
    mx = hi > lo ? double(b) * 0.001 : -1e12;
}

// Floating-grid values [lo, hi) (floats held as doubles), printed as the
// reference prints them: "%.3f" of the value, no rounding step, no "0" form.
void formatFloatRange(const double *v, int64_t lo, int64_t hi, std::string &buf)
{
    buf.resize(size_t(hi - lo) * 32);
    char *p = &buf[0];
    for (int64_t i = lo; i < hi; ++i) p += std::snprintf(p, 32, "%.3f\n", v[i]);
    buf.resize(size_t(p - &buf[0]));
}

}  // namespace

int writeOutputs(Prepared &p, double *const *maps, const int32_t *const *milli, const double *fmapValues,
                 const std::string &programName, int threads, std::string &err)
{
    const int nm = int(p.maps.size());
    // the floating grid is one more variable of the field (FG/Utils.cpp:51-81)
    const bool fmap = !p.fmapPath.empty();
    const int nmF = nm + (fmap ? 1 : 0);
    // AVS field file
    FILE *f = std::fopen(p.fldPath.c_str(), "w");
    if (!f) {
        err = "can't create grid dimensions data file " + p.fldPath;
        return -2;
    }
    std::fprintf(f, "# AVS field file\n#\n# AutoDock Atomic Affinity and Electrostatic Grids\n#\n# Created by %s.\n#\n",
                 programName.c_str());
    std::fprintf(f, "#SPACING %.3f\n", double(float(p.spacing)));
    std::fprintf(f, "#NELEMENTS %d %d %d\n", p.n[0] - 1, p.n[1] - 1, p.n[2] - 1);
    std::fprintf(f, "#CENTER %.3lf %.3lf %.3lf\n", p.center[0], p.center[1], p.center[2]);
    std::fprintf(f, "#MACROMOLECULE %s\n#GRID_PARAMETER_FILE %s\n#\n", p.receptorPath.c_str(), p.gpfPath.c_str());
    std::fprintf(f, "ndim=3\t\t\t# number of dimensions in the field\n");
    std::fprintf(f, "dim1=%d\t\t\t# number of x-elements\n", p.n[0]);
    std::fprintf(f, "dim2=%d\t\t\t# number of y-elements\n", p.n[1]);
    std::fprintf(f, "dim3=%d\t\t\t# number of z-elements\n", p.n[2]);
    std::fprintf(f, "nspace=3\t\t# number of physical coordinates per point\n");
    std::fprintf(f, "veclen=%d\t\t# number of affinity values at each point\n", nmF);
    std::fprintf(f, "data=float\t\t# data type (byte, integer, float, double)\n");
    std::fprintf(f, "field=uniform\t\t# field type (uniform, rectilinear, irregular)\n");
    for (int i = 0; i < 3; ++i)
        std::fprintf(f, "coord %d file=%s filetype=ascii offset=%d\n", i + 1, p.xyzPath.c_str(), i * 2);
    for (int i = 0; i < p.numAtomMaps; ++i)
        std::fprintf(f, "label=%s-affinity\t# component label for variable %d\n", p.maps[size_t(i)].type.c_str(), i + 1);
    std::fprintf(f, "label=Electrostatics\t# component label for variable %d\n", nmF - 2);
    std::fprintf(f, "label=Desolvation\t# component label for variable %d\n", nmF - 1);
    if (fmap) std::fprintf(f, "label=Floating_Grid\t# component label for variable %d\n", nmF);
    std::fprintf(f, "#\n# location of affinity grid files and how to read them\n#\n");
    for (int i = 0; i < nm; ++i)
        std::fprintf(f, "variable %d file=%s filetype=ascii skip=6\n", i + 1, p.maps[size_t(i)].filename.c_str());
    if (fmap) std::fprintf(f, "variable %d file=%s filetype=ascii skip=6\n", nmF, p.fmapPath.c_str());
    std::fclose(f);

    char hdr[2048];
    int hl = std::snprintf(hdr, sizeof hdr,
                           "GRID_PARAMETER_FILE %s\nGRID_DATA_FILE %s\nMACROMOLECULE %s\nSPACING %.3lf\n"
                           "NELEMENTS %d %d %d\nCENTER %.3lf %.3lf %.3lf\n",
                           p.gpfPath.c_str(), p.fldPath.c_str(), p.receptorPath.c_str(), p.spacing, p.n[0] - 1,
                           p.n[1] - 1, p.n[2] - 1, p.center[0], p.center[1], p.center[2]);
    if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
    threads = std::min(threads, 64);
    const int64_t P = p.P;
    p.mapMin.assign(size_t(nm), 0.0);
    p.mapMax.assign(size_t(nm), 0.0);
    // The maps are formatted in parallel chunks into one of two buffer sets
    // while a writer thread writes the previous map from the other.
    std::vector<std::string> chunkSets[2] = {std::vector<std::string>(size_t(threads)),
                                             std::vector<std::string>(size_t(threads))};
    std::vector<double> mn(static_cast<size_t>(threads)), mx(static_cast<size_t>(threads));
    auto writeChunks = [&hdr, hl](FILE *mf, const std::vector<std::string> &chunks) {
        bool ok = std::fwrite(hdr, size_t(hl), 1, mf) == 1;
        for (const std::string &c : chunks)
            if (!c.empty()) ok = ok && std::fwrite(c.data(), c.size(), 1, mf) == 1;
        return (std::fclose(mf) == 0) && ok;
    };
    std::thread writer;
    bool writeOk = true;  // set by the writer thread, read after joining it
    auto joinWriter = [&]() {
        if (writer.joinable()) writer.join();
    };
    struct JoinAtExit {  // every way out of this function joins the writer
        std::thread &t;
        ~JoinAtExit()
        {
            if (t.joinable()) t.join();
        }
    } joinAtExit{writer};
    int set = 0;
    for (int m = 0; m < nm; ++m) {
        const MapValues mv{maps ? maps[m] : nullptr, milli ? milli[m] : nullptr};
        if (p.maps[size_t(m)].filename.empty() || (!mv.v && !mv.milli)) continue;
        std::vector<std::string> &chunks = chunkSets[set];
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t) {
            int64_t lo = P * t / threads, hi = P * (t + 1) / threads;
            if (mv.milli)
                pool.emplace_back(formatMilliRange, mv.milli, lo, hi, std::ref(chunks[size_t(t)]),
                                  std::ref(mn[size_t(t)]), std::ref(mx[size_t(t)]));
            else
                pool.emplace_back(formatRange, mv, lo, hi, std::ref(chunks[size_t(t)]), std::ref(mn[size_t(t)]),
                                  std::ref(mx[size_t(t)]));
        }
        for (auto &th : pool) th.join();
        double a = 1e12, b = -1e12;
        for (int t = 0; t < threads; ++t) {
            a = std::min(a, mn[size_t(t)]);
            b = std::max(b, mx[size_t(t)]);
        }
        p.mapMin[size_t(m)] = a;
        p.mapMax[size_t(m)] = b;
        // the previous map is written (its buffer set is free, and a GPF that
        // names one file twice sees it rewritten, not cut short under the writer)
        joinWriter();
        FILE *mf = std::fopen(p.maps[size_t(m)].filename.c_str(), "w");
        if (!mf) {
            err = "Cannot open grid map \"" + p.maps[size_t(m)].filename + "\" for writing.";
            return -2;
        }
        if (!writeOk) {
            std::fclose(mf);
            err = "Not enough disk space.";
            return -2;
        }
        writer = std::thread([&writeOk, &writeChunks, mf, &chunks]() {
            if (!writeChunks(mf, chunks)) writeOk = false;
        });
        set ^= 1;
    }
    joinWriter();
    if (!writeOk) {
        err = "Not enough disk space.";
        return -2;
    }
    std::vector<std::string> &chunks = chunkSets[0];
    const double *fv = maps ? maps[nm] : fmapValues;
    if (fmap && fv) {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t) {
            int64_t lo = P * t / threads, hi = P * (t + 1) / threads;
            pool.emplace_back(formatFloatRange, fv, lo, hi, std::ref(chunks[size_t(t)]));
        }
        for (auto &th : pool) th.join();
        FILE *mf = std::fopen(p.fmapPath.c_str(), "w");
        if (!mf) {
            err = "can't open grid map \"" + p.fmapPath + "\" for writing.";
            return -2;
        }
        if (!writeChunks(mf, chunks)) {
            err = "Not enough disk space.";
            return -2;
        }
    }
    return 0;
}

int writeLog(const Prepared &p, FILE *f, const std::string &programName, double realSeconds)
{
    std::fprintf(f, "\n       FastGrid for MI355X (AutoGrid 4.2.1 compatible grid maps)\n\n");
    std::fputs(p.log.c_str(), f);
    std::fprintf(f, "\nCalculating %d grids over %lld elements, around %d receptor atoms.\n\n", int(p.maps.size()),
                 (long long)p.P, p.numAtoms());
    std::fprintf(f, "\nGrid\tAtom\tMinimum   \tMaximum\nMap \tType\tEnergy    \tEnergy \n\t\t(kcal/mol)\t(kcal/mol)\n"
                    "____\t____\t_____________\t_____________\n");
    const int nm = int(p.maps.size());
    for (int i = 0; i < nm && size_t(i) < p.mapMin.size(); ++i)
        std::fprintf(f, " %d\t %s\t  %6.2lf\t%9.2le\n", i + 1, p.maps[size_t(i)].type.c_str(), p.mapMin[size_t(i)],
                     p.mapMax[size_t(i)]);
    if (nm >= 2 && p.mapMin.size() == size_t(nm))
        std::fprintf(f, " %d\t %c\t  %6.2lf\t%9.2le\tElectrostatic Potential\n %d\t %c\t  %6.2lf\t%9.2le\tDesolvation Potential\n"
                        "\n\n * Note:  Every pairwise-atomic interaction was clamped at %.2f\n\n\n",
                     nm - 1, 'e', p.mapMin[size_t(nm - 2)], p.mapMax[size_t(nm - 2)], nm, 'd', p.mapMin[size_t(nm - 1)],
                     p.mapMax[size_t(nm - 1)], 100000.0);
    std::fprintf(f, "%s: Successful Completion.\n", programName.c_str());
    std::fprintf(f, "Real= %.2fs\n", realSeconds);
    return 0;
}

}  // namespace fgh
