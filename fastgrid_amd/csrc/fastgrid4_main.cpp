// fastgrid4 — drop-in command line of FastGrid 4.2.1 on MI355X.
//
// Same switches as the reference (FG/ProgramParameters.cpp:63-112), same
// outputs (.map/.fld/.xyz and a .glg log ending in "Successful Completion"),
// same orchestration (FG/main.cpp:115-254): parse -> precompute -> maps ->
// write.  The maps are computed by libfgmaps on the GPU; there is no CPU map
// path, so --no-cuda is refused.  --cuda-ddd=g|c|i|t picks the dielectric
// variant of the plain form (--precision=plain), as it picks the reference's
// CUDA kernel variant; without it the plain form makes the reference's own
// choice.  The other switches that tune the reference's CUDA kernels
// (--cuda-unroll, --cuda-thread, --cuda-slices) and its CPU search structures
// (--no-cogrid, --no-nns, --cogrid-mem) are accepted and have no effect on
// the maps.
//
// Additional switches:
//   --precision=plain|compensated|fp64|fp32   computation form (default compensated)
//   --ag421-param-order                  read parameter_file where the GPF names it
#include <sys/wait.h>
#include <unistd.h>

#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fgmaps.h"

namespace {

struct Options {
    std::string program, gpf, log;
    bool benchmark = false, v4 = false, agOrder = false, noCuda = false;
    int device = 0, threads = 0, debug = 0, precision = FG_PREC_COMPENSATED;
    int plainDiel = FG_PLAIN_DIEL_AUTO;  // --cuda-ddd=: the plain form's dielectric variant
};

[[noreturn]] void usage(const char *prog)
{
    std::fprintf(stderr,
                 "Usage: %s [OPTIONS]\n\n"
                 "Stream control:\n"
                 "  -l, --log FILE    log to FILE, default: standard output\n"
                 "  -p, --gpf FILE    read grid parameters from FILE, default: standard input\n\n"
                 "Miscellaneous:\n"
                 "      --benchmark   print execution times to standard error output\n"
                 "      --cogrid-mem N  accepted for compatibility (no effect)\n"
                 "      --cuda-enum   enumerate GPUs and exit\n"
                 "      --cuda-dev N  use GPU number N, default: 0\n"
                 "  -d, --debug       increment debug level\n"
                 "  -u, --help        display this help and exit\n"
                 "      --no-cuda     not supported: this build computes maps on the GPU only\n"
                 "      --omp N       use N host threads for writing maps\n"
                 "      --timeout N   terminate if calculations do not finish in N seconds\n"
                 "                    (POSIX only), this must be the first parameter if used\n"
                 "      --v4          set the AutoGrid 4.0 default parameter library\n\n"
                 "Advanced:\n"
                 "      --precision=plain|compensated|fp64|fp32  computation form, default: compensated\n"
                 "      --ag421-param-order  read parameter_file where the GPF names it (AutoGrid 4.2.1)\n"
                 "      --cuda-ddd=g|c|t|i  dielectric variant of --precision=plain (default: the reference's choice)\n"
                 "      --cuda-slices=y|n, --cuda-thread=y|n, --cuda-unroll=y|n,\n"
                 "      --no-cogrid, --no-nns   accepted for compatibility (no effect on the maps)\n\n",
                 prog);
    std::exit(0);
}

int readInt(const char *prog, const char *s)
{
    int n;
    if (!s || std::sscanf(s, "%i", &n) != 1) {
        std::fprintf(stderr, "%s: '%s' is not a number\n", prog, s ? s : "");
        std::exit(1);
    }
    return n;
}

// --timeout: the parent waits for the child and kills it after N seconds
// (FG/ProgramParameters.cpp:199-220).  Runs before any GPU call.
pid_t g_child = 0;
void onAlarm(int) { if (g_child > 0) kill(g_child, SIGKILL); }
void setTimeout(int seconds)
{
    pid_t pid = fork();
    if (pid < 0) return;
    if (pid == 0) return;  // child carries on
    g_child = pid;
    std::signal(SIGALRM, onAlarm);
    alarm(unsigned(seconds));
    int status = 0;
    while (waitpid(pid, &status, 0) < 0) {
    }
    if (WIFSIGNALED(status) && WTERMSIG(status) == SIGKILL) std::fprintf(stderr, "Timeout! Process terminated.\n");
    std::exit(0);
}

Options parse(int argc, char **argv)
{
    Options o;
    o.program = argv[0];
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
        if (a == "-l" || a == "--log") { const char *v = next(); o.log = v ? v : ""; }
        else if (a == "-p" || a == "--gpf") { const char *v = next(); o.gpf = v ? v : ""; }
        else if (a == "--benchmark") o.benchmark = true;
        else if (a == "--cogrid-mem") readInt(argv[0], next());
        else if (a == "--cuda-enum") {
            int n = fg_device_count();
            std::fprintf(stderr, "Found %i GPU devices.\n", n);
            std::exit(0);
        }
        else if (a == "--cuda-dev") o.device = readInt(argv[0], next());
        else if (a == "-d" || a == "--debug") ++o.debug;
        else if (a == "-u" || a == "--help") usage(argv[0]);
        else if (a == "--no-cuda") o.noCuda = true;
        else if (a == "--omp") o.threads = readInt(argv[0], next());
        else if (a == "--timeout") setTimeout(readInt(argv[0], next()));
        else if (a == "--v4") o.v4 = true;
        else if (a == "--ag421-param-order") o.agOrder = true;
        else if (a == "--precision=plain") o.precision = FG_PREC_PLAIN;
        else if (a == "--precision=compensated") o.precision = FG_PREC_COMPENSATED;
        else if (a == "--precision=fp64") o.precision = FG_PREC_FP64;
        else if (a == "--precision=fp32") o.precision = FG_PREC_FP32;
        else if (a == "--cuda-ddd=g" || a == "--cuda-ddd=c") o.plainDiel = FG_PLAIN_DIEL_TABLE;
        else if (a == "--cuda-ddd=i") o.plainDiel = FG_PLAIN_DIEL_INPLACE;
        else if (a == "--cuda-ddd=t") o.plainDiel = FG_PLAIN_DIEL_TEXTURE;
        else if (a.rfind("--cuda-ddd=", 0) == 0 || a.rfind("--cuda-unroll=", 0) == 0 ||
                 a.rfind("--cuda-thread=", 0) == 0 || a.rfind("--cuda-slices=", 0) == 0 || a == "--no-cogrid" ||
                 a == "--no-nns") {
        } else {
            std::fprintf(stderr, "%s: unknown switch '%s'\n", argv[0], argv[i]);
            std::exit(1);
        }
    }
    return o;
}

double seconds(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b)
{
    return std::chrono::duration<double>(b - a).count();
}

}  // namespace

int main(int argc, char **argv)
{
    const auto t0 = std::chrono::steady_clock::now();
    Options o = parse(argc, argv);
    const char *prog = o.program.c_str();
    FILE *log = o.log.empty() ? stdout : std::fopen(o.log.c_str(), "w");
    if (!log) {
        std::fprintf(stderr, "\n%s: Sorry, I can't create the log file \"%s\"\n", prog, o.log.c_str());
        std::fprintf(stderr, "\n%s: Unsuccessful Completion.\n\n", prog);
        return 911;
    }
    auto fatal = [&](int code, const char *what) {
        std::fprintf(log, "\n%s: ERROR:  %s\n", prog, what);
        std::fprintf(stderr, "\n%s: ERROR:  %s\n", prog, what);
        if (log != stdout) std::fclose(log);
        return code;
    };
    if (o.noCuda) return fatal(1, "--no-cuda: this build has no CPU map path");
    std::fprintf(log, "\n       FastGrid 4.2.1 for MI355X (HIP, gfx950)\n\n");

    fg_load_opts lo{o.v4 ? 1 : 0, o.agOrder ? 1 : 0};
    fg_prep *prep = nullptr;
    if (fg_prep_load(o.gpf.c_str(), &lo, &prep) != FG_OK) {
        const int code = fg_prep_exit_code();
        const std::string what = fg_last_error();
        if (code && code != -2) return fatal(code, what.c_str());
        // a fatal GPF error: the message, then the reference's closing line
        // (FG/InputDataLoader.cpp:757-780, FG/LogFile.cpp:147-149)
        std::fprintf(log, "\n%s: ERROR:  %s\n", prog, what.c_str());
        std::fprintf(stderr, "\n%s: ERROR:  %s\n", prog, what.c_str());
        return fatal(-2, "Unsuccessful completion.");
    }
    std::fputs(fg_prep_gpf_echo(prep), log);
    const auto t1 = std::chrono::steady_clock::now();

    fg_ctx *ctx = nullptr;
    if (fg_create(o.device, &ctx) != FG_OK) return fatal(0xbad, fg_last_error());
    fg_job_desc desc;
    fg_prep_desc(prep, &desc);
    const int M = desc.n_maps;
    const int64_t P = fg_prep_num_points(prep);
    const int nOut = M + 2 + (desc.floating_grid ? 1 : 0);  // maps incl. the floating grid
    std::fprintf(log, "Beginning grid calculations.\n\nCalculating %d grids over %lld elements, around %d receptor atoms.\n\n",
                 nOut, (long long)P, desc.atoms.n);
    fg_job *job = nullptr;
    if (fg_job_create(ctx, &desc, &job) != FG_OK || fg_job_set_plain_dielectric(job, o.plainDiel) != FG_OK)
        return fatal(0xbad, fg_last_error());
    const auto t2 = std::chrono::steady_clock::now();
    if (fg_job_run(ctx, job, FG_RUN_ALL | (o.benchmark ? FG_RUN_PHASE_TIMING : 0u), o.precision) != FG_OK)
        return fatal(0xbad, fg_last_error());
    // the maps as the files print them: integer thousandths rounded on the
    // device (half the bytes of FP64); a job with values past 32-bit
    // thousandths comes back in FP64.  The host buffers are made while the
    // GPU computes.
    std::vector<std::vector<int32_t>> milli(static_cast<size_t>(M + 2), std::vector<int32_t>(static_cast<size_t>(P)));
    std::vector<int32_t *> mptrs(static_cast<size_t>(M + 2));
    for (size_t m = 0; m < milli.size(); ++m) mptrs[m] = milli[m].data();
    if (fg_sync(ctx) != FG_OK) return fatal(0xbad, fg_last_error());
    fg_timing tm{};
    fg_last_timing(ctx, &tm);
    // the log's map extremes and total charge, computed on the GPU
    fg_summary sm{};
    if (fg_job_summary(ctx, job, &sm) != FG_OK) return fatal(0xbad, fg_last_error());
    std::vector<std::vector<double>> host;
    std::vector<double *> ptrs(static_cast<size_t>(nOut), nullptr);
    const bool asMilli = fg_job_fetch_milli(ctx, job, mptrs.data()) == FG_OK;
    if (asMilli) {
        if (desc.floating_grid) {
            host.assign(1, std::vector<double>(static_cast<size_t>(P)));
            ptrs[size_t(M + 2)] = host[0].data();
        }
    } else {
        host.assign(static_cast<size_t>(nOut), std::vector<double>(static_cast<size_t>(P)));
        for (size_t m = 0; m < host.size(); ++m) ptrs[m] = host[m].data();
    }
    if ((!asMilli || desc.floating_grid) && fg_job_fetch(ctx, job, ptrs.data()) != FG_OK)
        return fatal(0xbad, fg_last_error());
    fg_job_destroy(job);
    fg_destroy(ctx);
    const auto t3 = std::chrono::steady_clock::now();
    const int wrc = asMilli ? fg_write_outputs_milli(prep, mptrs.data(), desc.floating_grid ? ptrs[size_t(M + 2)] : nullptr,
                                                     prog, o.threads)
                            : fg_write_outputs(prep, ptrs.data(), prog, o.threads);
    if (wrc != FG_OK) return fatal(-2, fg_last_error());
    const auto t4 = std::chrono::steady_clock::now();

    // summary (FG/GridMap.cpp:71-101)
    std::fprintf(log, "\nGrid\tAtom\tMinimum   \tMaximum\nMap \tType\tEnergy    \tEnergy \n\t\t(kcal/mol)\t(kcal/mol)\n"
                      "____\t____\t_____________\t_____________\n");
    const double *mn = sm.min, *mx = sm.max;
    for (int m = 0; m < M + 2; ++m)
        std::fprintf(log, " %d\t %s\t  %6.2lf\t%9.2le\n", m + 1, fg_prep_map_type(prep, m), mn[m], mx[m]);
    std::fprintf(log, " %d\t %c\t  %6.2lf\t%9.2le\tElectrostatic Potential\n %d\t %c\t  %6.2lf\t%9.2le\tDesolvation Potential\n"
                      "\n\n * Note:  Every pairwise-atomic interaction was clamped at %.2f\n\n\n",
                 M + 1, 'e', mn[M], mx[M], M + 2, 'd', mn[M + 1], mx[M + 1], 100000.0);
    if (sm.has_charge)
        std::fprintf(log, "Total charge :\t\t\t%.2lf e%s\n\n", sm.total_charge,
                     sm.charge_integral ? "" : "  (not an integral net charge)");
    std::fprintf(stderr, "\n%s: Successful Completion.\n", prog);
    std::fprintf(log, "%s: Successful Completion.\n", prog);
    std::fprintf(log, "Real= %.2fs\n", seconds(t0, t4));
    if (o.benchmark) {
        std::fprintf(stderr, "Parse + precompute       Real: %.0f ms\n", 1e3 * seconds(t0, t1));
        std::fprintf(stderr, "Device setup + upload    Real: %.0f ms\n", 1e3 * seconds(t1, t2));
        std::fprintf(stderr, "Atom & desolvation maps  GPU: %.3f ms\n", tm.affinity_ms);
        std::fprintf(stderr, "Electrostatic map        GPU: %.3f ms\n", tm.elec_ms);
        std::fprintf(stderr, "Maps (run + fetch)       Real: %.0f ms\n", 1e3 * seconds(t2, t3));
        std::fprintf(stderr, "Write maps               Real: %.0f ms\n", 1e3 * seconds(t3, t4));
    }
    fg_prep_free(prep);
    if (log != stdout) std::fclose(log);
    return 0;
}
