/*
 * ps_oracle_cli.c -- TEST INFRASTRUCTURE ONLY.  Command-line driver of the oracle restatement.
 *
 *   ps_oracle_cli scan  <period_search_in> <n_from> <n_to> [conw_r]
 *       prints period_search_out lines (reference format) for periods n_from..n_to; when conw_r
 *       is omitted the convexity-weight search runs first (period_search_BOINC.cpp:487-715).
 *   ps_oracle_cli fits  <period_search_in> <n_from> <n_to> <conw_r> [npoles]
 *       runs the (period, pole) fits of periods n_from..n_to (poles 1..npoles, default 10) and
 *       prints one line with the fit count and the wall time in seconds (CPU-baseline timing).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ps_oracle.h"

static double now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s scan|fits <period_search_in> <n_from> <n_to> [conw_r [npoles]]\n", argv[0]);
        return 2;
    }
    int err = 0;
    pso_ctx *c = pso_ctx_from_file(argv[2], &err);
    if (!c) {
        fprintf(stderr, "cannot load %s (error %d)\n", argv[2], err);
        return 2;
    }
    int n_from = atoi(argv[3]), n_to = atoi(argv[4]);
    double conw_r;
    if (argc > 5)
        conw_r = atof(argv[5]);
    else
        pso_conw_search(c, &conw_r, NULL);
    if (strcmp(argv[1], "fits") == 0) {
        double f0, df;
        pso_ctx_freq(c, &f0, NULL, &df, NULL);
        pso_ctx_set_conw(c, conw_r);
        int npoles = argc > 6 ? atoi(argv[6]) : PSO_N_POLES;
        if (npoles < 1 || npoles > PSO_N_POLES) npoles = PSO_N_POLES;
        double t0 = now();
        int nfits = 0;
        for (int n = n_from; n <= n_to; n++)
            for (int m = 1; m <= npoles; m++) {
                pso_fit_result r;
                pso_fit(c, f0 - (n - 1) * df, m, &r);
                nfits++;
            }
        printf("fits %d seconds %.6f\n", nfits, now() - t0);
    } else {
        int nmax;
        pso_ctx_freq(c, NULL, NULL, NULL, &nmax);
        if (n_to <= 0 || n_to > nmax) n_to = nmax;
        pso_line *lines = (pso_line *)malloc(sizeof(pso_line) * (n_to - n_from + 1));
        int cnt = pso_scan(c, n_from, n_to, conw_r, lines);
        char buf[256];
        for (int i = 0; i < cnt; i++) {
            pso_format_line(&lines[i], buf, sizeof buf);
            fputs(buf, stdout);
        }
        free(lines);
    }
    pso_ctx_free(c);
    return 0;
}
