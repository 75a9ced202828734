/*
 * fgoracle — command-line front of the FP64 CPU oracle (test infrastructure).
 * Mirrors the fastgrid4 command line (FG/ProgramParameters.cpp:63-112) for
 * the flags that change results: -p/--gpf, -l/--log, --v4, --omp N, plus
 * --ag421-param-order (read parameter_file where it appears in the GPF, as
 * AutoGrid 4.2.1 does) and --benchmark (phase times on stderr).
 */
#include "fgoracle.h"

#include <stdlib.h>
#include <string.h>
#include <omp.h>

int main(int argc, char **argv)
{
    const char *gpf = "", *logName = "";
    int v4 = 0, ag = 0, threads = 0, bench = 0;
    for (int i = 1; i < argc; i++) {
        if ((!strcmp(argv[i], "-p") || !strcmp(argv[i], "--gpf")) && i + 1 < argc) gpf = argv[++i];
        else if ((!strcmp(argv[i], "-l") || !strcmp(argv[i], "--log")) && i + 1 < argc) logName = argv[++i];
        else if (!strcmp(argv[i], "--v4")) v4 = 1;
        else if (!strcmp(argv[i], "--ag421-param-order")) ag = 1;
        else if (!strcmp(argv[i], "--omp") && i + 1 < argc) threads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--benchmark")) bench = 1;
        else { fprintf(stderr, "%s: unknown switch '%s'\n", argv[0], argv[i]); return 1; }
    }
    FILE *log = logName[0] ? fopen(logName, "w") : stdout;
    if (!log) { fprintf(stderr, "%s: cannot create log %s\n", argv[0], logName); return 911; }
    static or_input in;
    double t0 = omp_get_wtime();
    int rc = or_load(&in, gpf, v4, ag, log);
    if (rc == 0) rc = or_precompute(&in);
    if (rc) {
        fprintf(stderr, "%s: ERROR: %s\n", argv[0], in.err);
        if (log != stdout) fclose(log);
        return rc;
    }
    double t1 = omp_get_wtime();
    or_compute(&in, 1, threads);
    double t2 = omp_get_wtime();
    or_compute(&in, 2 | 4, threads);
    double t3 = omp_get_wtime();
    rc = or_write_outputs(&in, argv[0]);
    double t4 = omp_get_wtime();
    if (bench)
        fprintf(stderr, "oracle: load+precompute %.3f s, e-map %.3f s, affinity+d maps %.3f s, write %.3f s\n",
                t1 - t0, t2 - t1, t3 - t2, t4 - t3);
    fprintf(log, "\n%s: Successful Completion.\n", argv[0]);
    if (log != stdout) fclose(log);
    or_free(&in);
    return rc;
}
