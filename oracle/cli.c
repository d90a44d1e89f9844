/* ORACLE — test infrastructure only. Runs one deck through the oracle
 * pipeline and prints report contract v1 (CLI semantics) or the worker /
 * Python feedpoint record. Usage: fnec_oracle deck.nec [cli|worker|py] [freq_index] */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fnec_oracle.h"

int main(int argc, char **argv)
{
    FILE *f;
    long len;
    char *txt;
    or_result r;
    int sem = OR_SEM_CLI, fi = 0, st;
    if (argc < 2) {
        fprintf(stderr, "usage: %s deck.nec [cli|worker|py] [freq_index]\n", argv[0]);
        return 2;
    }
    if (argc > 2)
        sem = !strcmp(argv[2], "worker") ? OR_SEM_WORKER : !strcmp(argv[2], "py") ? OR_SEM_PY : OR_SEM_CLI;
    if (argc > 3)
        fi = atoi(argv[3]);
    f = fopen(argv[1], "rb");
    if (!f) {
        perror(argv[1]);
        return 2;
    }
    fseek(f, 0, SEEK_END);
    len = ftell(f);
    fseek(f, 0, SEEK_SET);
    txt = (char *)malloc((size_t)len + 1);
    if (fread(txt, 1, (size_t)len, f) != (size_t)len) {
        fclose(f);
        return 2;
    }
    txt[len] = 0;
    fclose(f);
    st = or_run_deck(txt, sem, fi, 0.0, 1, &r);
    if (st) {
        fprintf(stderr, "error: %s\n", r.err);
        or_result_free(&r);
        return 1;
    }
    if (sem == OR_SEM_CLI)
        fputs(r.report, stdout);
    else
        printf("Z %.17g %.17g I %.17g %.17g\n", r.feed_z[0].re, r.feed_z[0].im, r.feed_i[0].re, r.feed_i[0].im);
    fprintf(stderr, "diag: abs_res=%.6e rel_res=%.6e diag_spread=%.6e N=%d S=%d R=%d\n", r.abs_res, r.rel_res, r.diag_spread, r.n, r.s, r.r);
    or_result_free(&r);
    free(txt);
    return 0;
}
