// sdpa_amd — command-line front end with the reference's flags (sdpa_main.cpp:501-677):
//   sdpa_amd -ds problem.dat-s | -dd problem.dat  [-o out|-] [-p param.sdpa] [-pt 0|1|2]
//            [-is init.ini-s | -id init.ini]  [-prec bits] [-gemm plain|ozaki] [-slices s] [-dev d]
//   sdpa_amd problem out [init] [-pt n]     (positional form: param.sdpa from the working directory,
//                                            formats chosen by the "-s" name suffix)
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sdpa_amd.h"
#include "sdpg.h"

static int usage() {
    fprintf(stderr,
            "usage: sdpa_amd -ds data.dat-s | -dd data.dat [-o out] [-p param.sdpa] [-pt 0|1|2]\n"
            "                [-is init.ini-s | -id init.ini] [-prec bits] [-gemm plain|ozaki] [-slices s] [-dev d]\n"
            "       sdpa_amd data out [init] [-pt n]\n");
    return 2;
}

int main(int argc, char **argv) {
    sdpa_amd_options o;
    memset(&o, 0, sizeof o);
    o.data_sparse = -1;
    o.init_sparse = -1;
    o.log_path = "-";
    o.gemm_mode = SDPG_GEMM_PLAIN;
    if (argc < 2) return usage();
    if (argv[1][0] != '-') {  // SDPA argument form
        if (argc < 3) return usage();
        o.data_path = argv[1];
        o.log_path = argv[2];
        o.param_path = "./param.sdpa";
        for (int i = 3; i < argc; i++) {
            if (!strcmp(argv[i], "-pt") && i + 1 < argc) {
                o.preset = atoi(argv[++i]);
                o.param_path = nullptr;
            } else {
                o.init_path = argv[i];
            }
        }
    } else {
        for (int i = 1; i < argc; i++) {
            const char *a = argv[i];
            const bool more = i + 1 < argc;
            if (!strcmp(a, "-ds") && more) o.data_path = argv[++i], o.data_sparse = 1;
            else if (!strcmp(a, "-dd") && more) o.data_path = argv[++i], o.data_sparse = 0;
            else if (!strcmp(a, "-is") && more) o.init_path = argv[++i], o.init_sparse = 1;
            else if (!strcmp(a, "-id") && more) o.init_path = argv[++i], o.init_sparse = 0;
            else if (!strcmp(a, "-o") && more) o.log_path = argv[++i];
            else if (!strcmp(a, "-p") && more) o.param_path = argv[++i];
            else if (!strcmp(a, "-pt") && more) o.preset = atoi(argv[++i]), o.param_path = nullptr;
            else if (!strcmp(a, "-prec") && more) o.prec_override = atoi(argv[++i]);
            else if (!strcmp(a, "-gemm") && more) {
                const char *g = argv[++i];
                o.gemm_mode = !strcmp(g, "ozaki") ? SDPG_GEMM_OZAKI : (!strcmp(g, "elem") ? -1 : SDPG_GEMM_PLAIN);
            } else if (!strcmp(a, "-slices") && more) o.slices = atoi(argv[++i]);
            else if (!strcmp(a, "-dev") && more) o.device = atoi(argv[++i]);
            else {
                fprintf(stderr, "unknown argument %s\n", a);
                return usage();
            }
        }
    }
    if (!o.data_path) return usage();
    sdpa_amd_result r;
    int rc = sdpa_amd_run(&o, &r);
    if (rc) {
        fprintf(stderr, "sdpa_amd: error %d: %s\n", rc, sdpg_last_error());
        return 1;
    }
    return 0;
}
