// C ABI of the host front end (include/fgmaps.h): errors, prepared inputs,
// writers and the synthetic generator.  No device code.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "capi_common.hpp"
#include "host.hpp"

struct fg_prep {
    fgh::Prepared p;
};

namespace {
thread_local std::string g_error;
thread_local int g_exitCode = 0;
}  // namespace

namespace fgc {
int fail(int code, const std::string &msg)
{
    g_error = msg;
    return code;
}
void clearError() { g_error.clear(); }
}  // namespace fgc

extern "C" {

const char *fg_last_error(void) { return g_error.c_str(); }

double fg_round3(double a) { return fgh::round3(a); }

int fg_prep_exit_code(void) { return g_exitCode; }

int fg_prep_load(const char *gpf, const fg_load_opts *opts, fg_prep **out)
{
    if (!out) return fgc::fail(FG_E_ARG, "fg_prep_load: out is NULL");
    *out = nullptr;
    fg_load_opts o{0, 0};
    if (opts) o = *opts;
    fg_prep *prep = new (std::nothrow) fg_prep();
    if (!prep) return fgc::fail(FG_E_NOMEM, "fg_prep_load: out of memory");
    std::string err;
    int rc = 0;
    try {
        rc = fgh::loadGpf(gpf ? gpf : "", o, prep->p, err);
    } catch (const std::bad_alloc &) {
        rc = int(0xBADA110C);
        err = "Not enough memory!";
    }
    if (rc != 0) {
        g_exitCode = rc;
        delete prep;
        return fgc::fail(rc == 911 ? FG_E_IO : FG_E_INPUT, err);
    }
    g_exitCode = 0;
    *out = prep;
    return FG_OK;
}

int fg_prep_load_many(const char *const *gpfs, int n, const fg_load_opts *opts, int threads, fg_prep **out)
{
    if (!out || n < 0 || (!gpfs && n > 0)) return fgc::fail(FG_E_ARG, "fg_prep_load_many: NULL argument");
    for (int i = 0; i < n; ++i) out[i] = nullptr;
    if (threads <= 0) threads = int(std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
    threads = std::min(threads, std::max(n, 1));
    // every worker keeps the status, message and exit code of its own loads
    std::vector<int> rcs(size_t(n), FG_OK), exits(size_t(n), 0);
    std::vector<std::string> errs(static_cast<size_t>(n));
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i; (i = next.fetch_add(1)) < n;) {
            rcs[size_t(i)] = fg_prep_load(gpfs[i], opts, &out[i]);
            if (rcs[size_t(i)] != FG_OK) {
                errs[size_t(i)] = g_error;
                exits[size_t(i)] = g_exitCode;
            }
        }
    };
    std::vector<std::thread> pool;
    try {
        for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    } catch (const std::system_error &) {
        // fewer threads than asked for: the ones running (and this one) share the work
    }
    work();
    for (std::thread &t : pool) t.join();
    for (int i = 0; i < n; ++i)
        if (rcs[size_t(i)] != FG_OK) {
            for (int k = 0; k < n; ++k) {
                fg_prep_free(out[k]);
                out[k] = nullptr;
            }
            g_exitCode = exits[size_t(i)];
            return fgc::fail(rcs[size_t(i)], "GPF " + std::to_string(i) + ": " + errs[size_t(i)]);
        }
    g_exitCode = 0;
    return FG_OK;
}

void fg_prep_free(fg_prep *prep) { delete prep; }

int fg_prep_desc(const fg_prep *prep, fg_job_desc *out)
{
    if (!prep || !out) return fgc::fail(FG_E_ARG, "fg_prep_desc: NULL argument");
    *out = prep->p.desc();
    return FG_OK;
}

int fg_prep_num_maps(const fg_prep *prep) { return prep ? prep->p.numAtomMaps : 0; }

const char *fg_prep_map_type(const fg_prep *prep, int m)
{
    if (!prep || m < 0 || m >= int(prep->p.maps.size())) return "";
    return prep->p.maps[size_t(m)].type.c_str();
}

int fg_prep_num_atoms(const fg_prep *prep) { return prep ? prep->p.numAtoms() : 0; }

int64_t fg_prep_num_points(const fg_prep *prep) { return prep ? prep->p.P : 0; }

const char *fg_prep_gpf_echo(const fg_prep *prep) { return prep ? prep->p.log.c_str() : ""; }

int fg_prep_summary(const fg_prep *prep, int m, double *mn, double *mx)
{
    if (!prep || !mn || !mx) return fgc::fail(FG_E_ARG, "fg_prep_summary: NULL argument");
    if (m < 0 || size_t(m) >= prep->p.mapMin.size()) return fgc::fail(FG_E_ARG, "fg_prep_summary: no such map");
    *mn = prep->p.mapMin[size_t(m)];
    *mx = prep->p.mapMax[size_t(m)];
    return FG_OK;
}

int fg_prep_map_grid(const fg_prep *prep, fg_map_grid *g)
{
    if (!prep || !g) return fgc::fail(FG_E_ARG, "fg_prep_map_grid: NULL argument");
    const fgh::Prepared &p = prep->p;
    // the .fld's "#SPACING %.3f" of the float spacing and "#CENTER %.3lf",
    // read back with "%lf" (AD/readfield.cc:72-114)
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.3f", double(float(p.spacing)));
    g->spacing = std::strtod(buf, nullptr);
    g->inv_spacing = 1. / g->spacing;
    for (int k = 0; k < 3; ++k) {
        std::snprintf(buf, sizeof buf, "%.3lf", p.center[k]);
        g->center[k] = std::strtod(buf, nullptr);
        g->num_points[k] = p.n[k] - 1;
        g->num_points1[k] = p.n[k];
        if (g->num_points[k] % 2 != 0 || g->num_points[k] <= 0)
            return fgc::fail(FG_E_INPUT, "fg_prep_map_grid: AutoDock needs an even NELEMENTS on every axis");
        g->lo[k] = g->center[k] - (g->num_points[k] / 2) * g->spacing;
        g->hi[k] = g->center[k] + (g->num_points[k] / 2) * g->spacing;
    }
    return FG_OK;
}

int fg_write_outputs(fg_prep *prep, double *const *maps, const char *program_name, int threads)
{
    if (!prep || !maps) return fgc::fail(FG_E_ARG, "fg_write_outputs: NULL argument");
    std::string err;
    int rc = fgh::writeOutputs(prep->p, maps, nullptr, nullptr, program_name ? program_name : "fastgrid4", threads,
                               err);
    return rc ? fgc::fail(FG_E_IO, err) : FG_OK;
}

int fg_write_outputs_milli(fg_prep *prep, const int32_t *const *maps, const double *fmap, const char *program_name,
                           int threads)
{
    if (!prep || !maps) return fgc::fail(FG_E_ARG, "fg_write_outputs_milli: NULL argument");
    std::string err;
    int rc = fgh::writeOutputs(prep->p, nullptr, maps, fmap, program_name ? program_name : "fastgrid4", threads, err);
    return rc ? fgc::fail(FG_E_IO, err) : FG_OK;
}

int fg_synth_write(const fg_synth_spec *spec, const char *dir)
{
    if (!spec) return fgc::fail(FG_E_ARG, "fg_synth_write: NULL spec");
    std::string err;
    int rc = fgh::synthWrite(*spec, dir ? dir : ".", err);
    if (rc == -3) return fgc::fail(FG_E_IO, err);
    return rc ? fgc::fail(FG_E_ARG, err) : FG_OK;
}

}  // extern "C"
