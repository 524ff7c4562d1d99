/*
 * orc_cli — command-line front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *   orc_cli -ds data.dat-s [-p param.sdpa] [-prec bits] [-o log|-] [-dump dir -iters 0,3,7 -L limbs]
 *   orc_cli -time data.dat-s [-prec bits] [-warm k] [-reps r] [-threads t]
 * Mirrors the reference CLI flags -ds/-p/-o (sdpa_main.cpp:527-636).
 */
#include "orc.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv) {
    const char *data = NULL, *par = NULL, *out = "-", *dump = NULL, *iters = NULL;
    int prec = 0, L = 8, time_mode = 0, warm = 0, reps = 1, threads = 0;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-ds") && i + 1 < argc) data = argv[++i];
        else if (!strcmp(argv[i], "-time") && i + 1 < argc) { data = argv[++i]; time_mode = 1; }
        else if (!strcmp(argv[i], "-p") && i + 1 < argc) par = argv[++i];
        else if (!strcmp(argv[i], "-o") && i + 1 < argc) out = argv[++i];
        else if (!strcmp(argv[i], "-prec") && i + 1 < argc) prec = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-dump") && i + 1 < argc) dump = argv[++i];
        else if (!strcmp(argv[i], "-iters") && i + 1 < argc) iters = argv[++i];
        else if (!strcmp(argv[i], "-L") && i + 1 < argc) L = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-warm") && i + 1 < argc) warm = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-reps") && i + 1 < argc) reps = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-threads") && i + 1 < argc) threads = atoi(argv[++i]);
        else { fprintf(stderr, "unknown argument %s\n", argv[i]); return 2; }
    }
    if (!data) { fprintf(stderr, "usage: orc_cli -ds file.dat-s [...]\n"); return 2; }
    orc_param p;
    orc_param_default(&p);
    if (par && orc_param_read(&p, par) != 0) { fprintf(stderr, "cannot read %s\n", par); return 2; }
    if (time_mode) {
        double tb = 0, tc = 0;
        int rc = orc_time_hotpath(data, &p, prec, warm, reps, threads, &tb, &tc);
        printf("{\"rc\": %d, \"t_bmat\": %.6f, \"t_chol\": %.6f}\n", rc, tb, tc);
        return rc ? 1 : 0;
    }
    int it[256], n = 0;
    if (iters) {
        const char *s = iters;
        while (*s && n < 256) {
            it[n++] = atoi(s);
            while (*s && *s != ',') s++;
            if (*s == ',') s++;
        }
    }
    orc_result r;
    int rc = orc_solve(data, &p, prec, out, dump, it, n, L, &r);
    if (rc < 0) { fprintf(stderr, "read error %d\n", rc); return 2; }
    return 0;
}
