/*
 * ps_ref_cli.cpp -- TEST INFRASTRUCTURE ONLY.  Times the reference's own plain-C core (compiled
 * from the reference sources into _ref/, driven by ref_harness.cpp) as the CPU baseline.
 *
 *   ps_ref_cli fits <period_search_in> <n_from> <n_to> <conw_r> [npoles]
 *       runs the (period, pole) fits of periods n_from..n_to (poles 1..npoles, default 10) on the
 *       reference core and prints "fits <count> seconds <wall>".
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>

extern "C" {
int psr_load_file(const char *path);
void psr_set_conw(double conw_r);
int psr_fit(double freq, int m, double *out7);
int psr_freq(double *f0, double *df, int *nmax);
}

static double now() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv) {
    if (argc < 6 || strcmp(argv[1], "fits") != 0) {
        fprintf(stderr, "usage: %s fits <period_search_in> <n_from> <n_to> <conw_r> [npoles]\n", argv[0]);
        return 2;
    }
    if (psr_load_file(argv[2]) != 0) {
        fprintf(stderr, "cannot load %s\n", argv[2]);
        return 2;
    }
    const int n_from = atoi(argv[3]), n_to = atoi(argv[4]);
    int npoles = argc > 6 ? atoi(argv[6]) : 10;
    if (npoles < 1 || npoles > 10) npoles = 10;
    double f0, df;
    psr_freq(&f0, &df, nullptr);
    psr_set_conw(atof(argv[5]));
    const double t0 = now();
    int nfits = 0;
    for (int n = n_from; n <= n_to; n++)
        for (int m = 1; m <= npoles; m++) {
            double r[7];
            psr_fit(f0 - (n - 1) * df, m, r);
            nfits++;
        }
    printf("fits %d seconds %.6f\n", nfits, now() - t0);
    return 0;
}
