// fnec_hip: command-line front end reproducing the reference CLI's Hallén
// output (apps/nec-cli/src/main.rs:325-420): one report contract v1 block per
// FR frequency, blank-line separated, then a SWEEP_POINTS table when the deck
// sweeps more than one frequency. Usage:
//   fnec_hip [--precision fp64|mixed] deck.nec
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fnec_hip.h"
#include "../../include/fnec_host.h"

int main(int argc, char **argv)
{
    int precision = FNEC_PREC_FP64;
    const char *path = nullptr;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--precision") && i + 1 < argc) {
            ++i;
            precision = !std::strcmp(argv[i], "mixed") ? FNEC_PREC_MIXED : FNEC_PREC_FP64;
        } else
            path = argv[i];
    }
    if (!path) {
        std::fprintf(stderr, "usage: fnec_hip [--precision fp64|mixed] deck.nec\n");
        return 2;
    }
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        std::perror(path);
        return 1;
    }
    std::string text;
    char buf[65536];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0)
        text.append(buf, k);
    std::fclose(f);
    char err[256];
    fnec_deck *d = fnec_deck_parse(text.c_str(), err, sizeof(err));
    if (!d) {
        std::fprintf(stderr, "error: %s\n", err);
        return 1;
    }
    size_t nf = fnec_deck_frequencies(d, FNEC_SEM_CLI, nullptr, 0);
    std::vector<std::string> sweep;
    for (size_t fi = 0; fi < nf; ++fi) {
        size_t need = 0;
        int st = fnec_deck_report(d, fi, precision, nullptr, 0, &need);
        if (st) {
            std::fprintf(stderr, "error: %s (status %d)\n", fnec_hip_last_error(), st);
            fnec_deck_free(d);
            return 1;
        }
        std::vector<char> rep(need);
        fnec_deck_report(d, fi, precision, rep.data(), rep.size(), &need);
        if (fi > 0)
            std::printf("\n");
        std::fputs(rep.data(), stdout);
        // first feedpoint row -> SWEEP_POINTS line (freq tag seg z_re z_im)
        const char *hdr = std::strstr(rep.data(), "TAG SEG V_RE V_IM I_RE I_IM Z_RE Z_IM\n");
        const char *fq = std::strstr(rep.data(), "FREQ_MHZ ");
        if (hdr && fq) {
            const char *row = hdr + std::strlen("TAG SEG V_RE V_IM I_RE I_IM Z_RE Z_IM\n");
            std::string line(row, std::strcspn(row, "\n"));
            std::string freq(fq + 9, std::strcspn(fq + 9, "\n"));
            char t[32], s[32], vr[64], vi[64], ir[64], ii[64], zr[64], zi[64];
            if (std::sscanf(line.c_str(), "%31s %31s %63s %63s %63s %63s %63s %63s", t, s, vr, vi, ir, ii, zr, zi) == 8)
                sweep.push_back(freq + " " + t + " " + s + " " + zr + " " + zi);
        }
    }
    if (sweep.size() > 1) {
        std::printf("\nSWEEP_POINTS\nN_POINTS %zu\nFREQ_MHZ TAG SEG Z_RE Z_IM\n", sweep.size());
        for (auto &r : sweep)
            std::printf("%s\n", r.c_str());
    }
    fnec_deck_free(d);
    return 0;
}
