// sdpa_amd — command-line front end (flags follow sdpa_main.cpp:527-636: -ds data -o out -p param)
//   sdpa_amd -ds problem.dat-s [-o out|-] [-p param.sdpa] [-prec bits] [-gemm plain|ozaki] [-slices s] [-dev d]
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sdpa_amd.h"
#include "sdpg.h"

int main(int argc, char **argv) {
    const char *data = nullptr, *out = "-", *par = nullptr;
    int prec = 0, mode = SDPG_GEMM_PLAIN, slices = 0, dev = 0;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-ds") && i + 1 < argc) data = argv[++i];
        else if (!strcmp(argv[i], "-o") && i + 1 < argc) out = argv[++i];
        else if (!strcmp(argv[i], "-p") && i + 1 < argc) par = argv[++i];
        else if (!strcmp(argv[i], "-prec") && i + 1 < argc) prec = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-gemm") && i + 1 < argc) {
            const char *g = argv[++i];
            mode = !strcmp(g, "ozaki") ? SDPG_GEMM_OZAKI : (!strcmp(g, "elem") ? -1 : SDPG_GEMM_PLAIN);
        } else if (!strcmp(argv[i], "-slices") && i + 1 < argc) slices = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-dev") && i + 1 < argc) dev = atoi(argv[++i]);
        else {
            fprintf(stderr, "unknown argument %s\n", argv[i]);
            return 2;
        }
    }
    if (!data) {
        fprintf(stderr, "usage: sdpa_amd -ds problem.dat-s [-o out] [-p param.sdpa] [-prec bits] [-gemm plain|ozaki]\n");
        return 2;
    }
    sdpa_amd_result r;
    int rc = sdpa_amd_solve(data, par, prec, out, mode, slices, dev, &r);
    if (rc) {
        fprintf(stderr, "sdpa_amd: error %d: %s\n", rc, sdpg_last_error());
        return 1;
    }
    return 0;
}
