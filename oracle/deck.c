/*
 * ORACLE — test infrastructure only.
 *
 * NEC deck parser subset, restating crates/nec_parser/src/lib.rs:84-330 and
 * the field helpers at :424-483 (mnemonic = first whitespace token, upper-cased;
 * numeric fields parsed as f64 and cast with Rust's saturating `as u32/i32`;
 * EN stops parsing; unknown cards are skipped). Frequencies follow
 * apps/nec-cli/src/solve_session.rs:984-1012 (first FR card) and
 * bindings/fnec_py/src/lib.rs:27-44 (all FR cards); RP points follow
 * apps/nec-cli/src/main.rs:292-305 with farfield.rs:677-699.
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

int or_parse_f64_strict(const char *s, double *out)
{
    /* Rust's str::parse::<f64>: optional sign, decimal digits with optional
     * fraction and exponent, or inf/infinity/nan. No hex, no whitespace. */
    const char *p = s;
    if (*p == '+' || *p == '-')
        p++;
    if (!strncasecmp(p, "inf", 3) || !strncasecmp(p, "nan", 3)) {
        if (strcasecmp(p, "inf") && strcasecmp(p, "infinity") && strcasecmp(p, "nan"))
            return -1;
    } else {
        int digits = 0;
        while (isdigit((unsigned char)*p)) { p++; digits++; }
        if (*p == '.') {
            p++;
            while (isdigit((unsigned char)*p)) { p++; digits++; }
        }
        if (!digits)
            return -1;
        if (*p == 'e' || *p == 'E') {
            p++;
            if (*p == '+' || *p == '-')
                p++;
            if (!isdigit((unsigned char)*p))
                return -1;
            while (isdigit((unsigned char)*p))
                p++;
        }
        if (*p)
            return -1;
    }
    *out = strtod(s, NULL);
    return 0;
}

uint32_t or_as_u32(double v)
{
    if (isnan(v) || v <= 0.0)
        return 0;
    if (v >= 4294967295.0)
        return 4294967295u;
    return (uint32_t)v;
}

int32_t or_as_i32(double v)
{
    if (isnan(v))
        return 0;
    if (v <= -2147483648.0)
        return INT32_MIN;
    if (v >= 2147483647.0)
        return INT32_MAX;
    return (int32_t)v;
}

/* How many leading fields each card parses (the highest field index the
 * reference reads), and how many it requires. */
static void card_field_rules(const char *mn, int nf, int *need, int *parse)
{
    *need = 0;
    *parse = 0;
    if (!strcmp(mn, "GW")) { *need = 9; *parse = 9; }
    else if (!strcmp(mn, "GE")) { *need = 0; *parse = nf > 0 ? 1 : 0; }
    else if (!strcmp(mn, "GM")) { *need = 9; *parse = 9; }
    else if (!strcmp(mn, "GR")) { *need = 3; *parse = 3; }
    else if (!strcmp(mn, "GN")) { *need = 1; *parse = nf > 5 ? 6 : (nf > 4 ? 5 : 1); }
    else if (!strcmp(mn, "EX")) { *need = 4; *parse = nf > 10 ? 10 : nf; }
    else if (!strcmp(mn, "NE") || !strcmp(mn, "NH")) { *need = 4; *parse = nf > 10 ? 10 : nf; }
    else if (!strcmp(mn, "FR")) { *need = 4; *parse = nf >= 6 ? 6 : 4; }
    else if (!strcmp(mn, "RP")) { *need = 7; *parse = nf >= 8 ? 8 : 7; }
    else if (!strcmp(mn, "LD")) { *need = 4; *parse = nf > 7 ? 7 : nf; }
    else if (!strcmp(mn, "TL")) { *need = 8; *parse = nf > 8 ? 9 : 8; }
}

static int is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }

or_deck *or_deck_parse(const char *text, char *err, int errlen)
{
    or_deck *d = (or_deck *)calloc(1, sizeof(or_deck));
    const char *p = text;
    int lineno = 0;
    if (err && errlen)
        err[0] = 0;
    while (*p) {
        const char *eol = strchr(p, '\n');
        size_t len = eol ? (size_t)(eol - p) : strlen(p);
        char line[4096];
        char *toks[64];
        int nt = 0;
        lineno++;
        if (len >= sizeof(line))
            len = sizeof(line) - 1;
        memcpy(line, p, len);
        line[len] = 0;
        p = eol ? eol + 1 : p + len;
        {
            char *q = line;
            while (*q) {
                while (*q && is_space(*q))
                    q++;
                if (!*q)
                    break;
                if (nt < 64)
                    toks[nt++] = q;
                while (*q && !is_space(*q))
                    q++;
                if (*q)
                    *q++ = 0;
            }
        }
        if (nt == 0)
            continue;
        {
            char mn[8];
            int i, need, nparse, nf = nt - 1;
            or_card c;
            memset(&c, 0, sizeof(c));
            snprintf(mn, sizeof(mn), "%s", toks[0]);
            for (i = 0; mn[i]; i++)
                mn[i] = (char)toupper((unsigned char)mn[i]);
            if (!strcmp(mn, "CM") || !strcmp(mn, "CE"))
                continue;
            if (!strcmp(mn, "EN"))
                break;
            if (strcmp(mn, "GW") && strcmp(mn, "GE") && strcmp(mn, "GM") && strcmp(mn, "GR") &&
                strcmp(mn, "GN") && strcmp(mn, "EX") && strcmp(mn, "NE") && strcmp(mn, "NH") &&
                strcmp(mn, "FR") && strcmp(mn, "RP") && strcmp(mn, "LD") && strcmp(mn, "TL") &&
                strcmp(mn, "NT") && strcmp(mn, "PT"))
                continue; /* unknown card: non-fatal warning in the reference */
            card_field_rules(mn, nf, &need, &nparse);
            if (nf < need) {
                if (err)
                    snprintf(err, errlen, "line %d: %s needs %d fields, got %d", lineno, mn, need, nf);
                or_deck_free(d);
                return NULL;
            }
            memcpy(c.mn, mn, 2);
            c.mn[2] = 0;
            c.nf = nf;
            if (!strcmp(mn, "NT"))
                for (i = 0; i < nf && i < OR_MAX_RAW; i++)
                    snprintf(c.raw[i], sizeof(c.raw[i]), "%s", toks[1 + i]);
            for (i = 0; i < nparse && i < OR_MAX_FIELDS; i++) {
                if (or_parse_f64_strict(toks[1 + i], &c.f[i])) {
                    if (err)
                        snprintf(err, errlen, "line %d: %s field %d: cannot parse '%s'", lineno,
                                 mn, i + 1, toks[1 + i]);
                    or_deck_free(d);
                    return NULL;
                }
            }
            if (d->n == d->cap) {
                d->cap = d->cap ? 2 * d->cap : 16;
                d->cards = (or_card *)realloc(d->cards, (size_t)d->cap * sizeof(or_card));
            }
            d->cards[d->n++] = c;
        }
    }
    return d;
}

void or_deck_free(or_deck *d)
{
    if (!d)
        return;
    free(d->cards);
    free(d);
}

int or_deck_has_card(const or_deck *d, const char *mn)
{
    int i;
    for (i = 0; i < d->n; i++)
        if (!strcmp(d->cards[i].mn, mn))
            return 1;
    return 0;
}

/* compiler-rt __powidf2, the lowering of Rust's f64::powi. */
static double powi_rt(double a, int b)
{
    int recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1)
            r *= a;
        b /= 2;
        if (b == 0)
            break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

static void fr_fields(const or_card *c, uint32_t *type, uint32_t *steps, double *f, double *st)
{
    int fi = c->nf >= 6 ? 4 : 2;
    *type = or_as_u32(c->f[0]);
    *steps = or_as_u32(c->f[1]);
    *f = c->f[fi];
    *st = c->f[fi + 1];
}

/* apps/nec-cli/src/solve_session.rs:984-1012 */
int or_deck_frequencies(const or_deck *d, double *out, int max)
{
    int i, cnt = 0;
    for (i = 0; i < d->n; i++) {
        if (strcmp(d->cards[i].mn, "FR"))
            continue;
        {
            uint32_t type, steps, idx;
            double f, st;
            fr_fields(&d->cards[i], &type, &steps, &f, &st);
            if (steps < 1)
                steps = 1;
            if (type == 0) {
                for (idx = 0; idx < steps; idx++, cnt++)
                    if (cnt < max)
                        out[cnt] = (f + (double)idx * st) * 1e6;
            } else if (type == 1) {
                for (idx = 0; idx < steps; idx++, cnt++)
                    if (cnt < max)
                        out[cnt] = f * powi_rt(st, (int)idx) * 1e6;
            } else {
                if (cnt < max)
                    out[cnt] = f * 1e6;
                cnt++;
            }
        }
        break;
    }
    return cnt;
}

/* bindings/fnec_py/src/lib.rs:27-44 (all FR cards; linear unless type 1). */
int or_deck_frequencies_py(const or_deck *d, double *out, int max)
{
    int i, cnt = 0;
    for (i = 0; i < d->n; i++) {
        uint32_t type, steps, idx;
        double f, st;
        if (strcmp(d->cards[i].mn, "FR"))
            continue;
        fr_fields(&d->cards[i], &type, &steps, &f, &st);
        if (steps < 1)
            steps = 1;
        for (idx = 0; idx < steps; idx++, cnt++) {
            double fm = type == 1 ? f * powi_rt(st, (int)idx) : f + st * (double)idx;
            if (cnt < max)
                out[cnt] = fm * 1e6;
        }
    }
    return cnt;
}

/* RP card grid: φ outer, θ inner (farfield.rs:677-699), all RP cards
 * concatenated in deck order (main.rs:292-305). */
int or_deck_rp_points(const or_deck *d, double *theta, double *phi, int max)
{
    int i, cnt = 0;
    for (i = 0; i < d->n; i++) {
        const or_card *c = &d->cards[i];
        int a;
        uint32_t nt, np, it, ip;
        if (strcmp(c->mn, "RP"))
            continue;
        a = c->nf >= 8 ? 4 : 3;
        nt = or_as_u32(c->f[1]);
        np = or_as_u32(c->f[2]);
        if (nt < 1)
            nt = 1;
        if (np < 1)
            np = 1;
        for (ip = 0; ip < np; ip++) {
            double ph = c->f[a + 1] + (double)ip * c->f[a + 3];
            for (it = 0; it < nt; it++, cnt++) {
                if (cnt < max) {
                    theta[cnt] = c->f[a] + (double)it * c->f[a + 2];
                    phi[cnt] = ph;
                }
            }
        }
    }
    return cnt;
}

/* RP XNDA digits (nec_parser/src/lib.rs:289-322). */
void or_deck_rp_flags(const or_deck *d, int *normalize, int *avg_gain)
{
    int i;
    *normalize = 0;
    *avg_gain = 0;
    for (i = 0; i < d->n; i++) {
        const or_card *c = &d->cards[i];
        if (strcmp(c->mn, "RP") || c->nf < 8)
            continue;
        {
            uint32_t x = or_as_u32(c->f[3]);
            if ((x / 1000) % 10 != 0)
                *normalize = 1;
            if (x % 10 != 0)
                *avg_gain = 1;
        }
    }
}
