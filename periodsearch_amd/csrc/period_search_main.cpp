// period_search_amd -- command-line replacement for the reference period_search executable
// (period_search/period_search_BOINC.cpp:147-954): reads period_search_in, scans the period interval
// on the GPU and writes period_search_out (with period_search_state checkpoints) in the reference
// format.  Optional flags: --device N, --precision fp64|mixed, and explicit file names.  Exit status:
// 0 on success, 1 when a file cannot be read or written, 2 on any other failure (usage, a malformed
// input, a reference limit, an unsupported shape, no device).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/periodsearch_amd.h"

int main(int argc, char **argv) {
    const char *in = "period_search_in", *out = "period_search_out", *state = "period_search_state";
    int device = 0, precision = PS_FP64, pos = 0;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--device") && i + 1 < argc) {
            device = std::atoi(argv[++i]);
        } else if (!std::strcmp(argv[i], "--precision") && i + 1 < argc) {
            ++i;
            precision = !std::strcmp(argv[i], "mixed") ? PS_MIXED : PS_FP64;
        } else if (argv[i][0] == '-') {
            std::fprintf(stderr, "usage: %s [--device N] [--precision fp64|mixed] [in [out [state]]]\n", argv[0]);
            return 2;
        } else {
            (pos == 0 ? in : pos == 1 ? out : state) = argv[i];
            ++pos;
        }
    }
    int rc = ps_run_files(in, out, state, device, precision);
    if (rc != PS_OK) {
        std::fprintf(stderr, "period_search_amd: %s (code %d)\n", ps_last_error(), rc);
        // the reference exits with 1 when its output cannot be opened (period_search_BOINC.cpp:252-261) and
        // with 2 on an input limit (:327, :353, :361) or a failed device setup (CUDAPrepare, :608-613)
        return rc == PS_ERR_IO ? 1 : 2;
    }
    return 0;
}
