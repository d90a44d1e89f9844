/*
 * ORACLE — test infrastructure only. Nothing in the product links this code.
 *
 * The Z-matrix stamps of LD, TL and NT cards, restating
 *   crates/nec_solver/src/loads.rs:34-223   (build_loads, i0_over_i1,
 *                                            wire_internal_impedance,
 *                                            segment_matches)
 *   crates/nec_solver/src/tl.rs:47-244      (build_tl_stamps,
 *                                            find_segment_index,
 *                                            find_center_segment_index)
 *   crates/nec_solver/src/network.rs:55-140 (build_nt_stamps)
 * and their application order (apps/nec-cli/src/solve_session.rs:1285-1309;
 * crates/nec_worker/src/solve.rs:179-185; bindings/fnec_py/src/lib.rs:58-63):
 * the per-segment sum of the LD loads is added to the diagonal, then each TL
 * stamp, then (CLI only) each NT stamp, to the matrix entry it names, in card
 * order. Cards the reference warns about and skips are skipped.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

#define TWO_PI (2.0 * OR_PI)

/* loads.rs:34-59 */
static cx i0_over_i1(cx z)
{
    if (cx_norm(z) > 15.0) {
        const cx inv = cx_div(cx_make(1.0, 0.0), z);
        return cx_add(cx_add(cx_make(1.0, 0.0), cx_scale(inv, 0.5)), cx_scale(cx_mul(inv, inv), 0.375));
    }
    {
        const cx half = cx_scale(z, 0.5);
        const cx half_sq = cx_mul(half, half);
        cx i0 = cx_make(1.0, 0.0), term0 = cx_make(1.0, 0.0), i1 = half, term1 = half;
        int k;
        for (k = 1; k < 300; k++) {
            const double kf = (double)k;
            term0 = cx_mul(term0, cx_divs(half_sq, kf * kf));
            i0 = cx_add(i0, term0);
            term1 = cx_mul(term1, cx_divs(half_sq, kf * (kf + 1.0)));
            i1 = cx_add(i1, term1);
            if (cx_norm(term0) <= 1e-17 * cx_norm(i0) && cx_norm(term1) <= 1e-17 * cx_norm(i1))
                break;
        }
        return cx_div(i0, i1);
    }
}

/* loads.rs:73-80 */
static cx wire_internal_impedance(double sigma, double a, double dl, double omega)
{
    if (omega <= 0.0)
        return cx_make(dl / (sigma * OR_PI * a * a), 0.0);
    {
        const double delta = sqrt(2.0 / (omega * OR_MU0 * sigma));
        const cx gamma = cx_make(1.0 / delta, 1.0 / delta);
        return cx_scale(cx_mul(cx_divs(gamma, TWO_PI * a * sigma), i0_over_i1(cx_scale(gamma, a))), dl);
    }
}

/* loads.rs:209-223 */
static int segment_matches(uint32_t tag, uint32_t first, uint32_t last, const or_seg *s)
{
    if (tag != 0 && s->tag != tag)
        return 0;
    if (first == 0)
        return 1;
    if (last == 0)
        last = first;
    return s->tag_index >= first && s->tag_index <= last;
}

/* loads.rs:103-206: loads[i] for every segment (zero where no card applies) */
void or_build_loads(const or_deck *d, const or_seg *s, int n, double freq_hz, cx *loads)
{
    const double omega = TWO_PI * freq_hz;
    int c, i;
    for (i = 0; i < n; i++)
        loads[i] = cx_make(0.0, 0.0);
    for (c = 0; c < d->n; c++) {
        const or_card *ld = &d->cards[c];
        int32_t type;
        uint32_t tag, first, last;
        double f1, f2, f3;
        if (strcmp(ld->mn, "LD"))
            continue;
        type = or_as_i32(ld->f[0]);
        tag = or_as_u32(ld->f[1]);
        first = or_as_u32(ld->f[2]);
        last = or_as_u32(ld->f[3]);
        f1 = ld->nf > 4 ? ld->f[4] : 0.0;
        f2 = ld->nf > 5 ? ld->f[5] : 0.0;
        f3 = ld->nf > 6 ? ld->f[6] : 0.0;
        for (i = 0; i < n; i++) {
            cx z;
            if (!segment_matches(tag, first, last, &s[i]))
                continue;
            switch (type) {
            case 0: {
                const double x_l = omega * f2, x_c = f3 > 0.0 ? 1.0 / (omega * f3) : 0.0;
                z = cx_make(f1, x_l - x_c);
                break;
            }
            case 1: {
                cx y = cx_make(0.0, 0.0);
                if (f1 > 0.0)
                    y = cx_add(y, cx_make(1.0 / f1, 0.0));
                if (f2 > 0.0)
                    y = cx_add(y, cx_make(0.0, -1.0 / (omega * f2)));
                if (f3 > 0.0)
                    y = cx_add(y, cx_make(0.0, omega * f3));
                z = cx_norm(y) < 1e-30 ? cx_make(0.0, 0.0) : cx_div(cx_make(1.0, 0.0), y);
                break;
            }
            case 2:
                z = cx_make(f1, omega * f2);
                break;
            case 3:
                z = cx_make(f1, -(f3 > 0.0 ? 1.0 / (omega * f3) : 0.0));
                break;
            case 4:
                z = cx_make(f1, f2);
                break;
            case 5:
                if (f1 <= 0.0)
                    continue; /* warned and ignored */
                z = wire_internal_impedance(f1, s[i].radius, s[i].length, omega);
                break;
            default:
                continue; /* unsupported type: warned and ignored */
            }
            loads[i] = cx_add(loads[i], z);
        }
    }
}

/* tl.rs:186-244: index of (tag, segment), segment 0 = the tag's centre */
static int find_segment_index(const or_seg *s, int n, uint32_t tag, uint32_t seg)
{
    int i, cnt = 0, first = -1;
    if (seg != 0) {
        for (i = 0; i < n; i++)
            if (s[i].tag == tag && s[i].tag_index == seg)
                return i;
        return -1;
    }
    for (i = 0; i < n; i++)
        if (s[i].tag == tag) {
            if (first < 0)
                first = i;
            cnt++;
        }
    if (!cnt)
        return -1;
    {
        /* the k-th segment of the tag, in geometry order */
        int pick, k = 0;
        if (cnt % 2 == 1) {
            pick = cnt / 2;
        } else {
            const uint32_t lower = (uint32_t)cnt / 2;
            pick = -1;
            for (i = 0; i < n; i++)
                if (s[i].tag == tag) {
                    if (s[i].tag_index == lower) {
                        pick = k;
                        break;
                    }
                    k++;
                }
            if (pick < 0)
                pick = cnt / 2 >= 1 ? cnt / 2 - 1 : 0;
        }
        k = 0;
        for (i = 0; i < n; i++)
            if (s[i].tag == tag) {
                if (k == pick)
                    return i;
                k++;
            }
    }
    return -1;
}

/* Rust's str::parse::<u32>: an optional '+', then decimal digits, no overflow */
static int parse_u32_str(const char *t, uint32_t *out)
{
    unsigned long long v = 0;
    if (*t == '+')
        t++;
    if (!*t)
        return -1;
    for (; *t; t++) {
        if (*t < '0' || *t > '9')
            return -1;
        v = v * 10 + (unsigned)(*t - '0');
        if (v > 4294967295ull)
            return -1;
    }
    *out = (uint32_t)v;
    return 0;
}

/* The stamps in application order: diagonal loads (segments with a non-zero
 * load), TL stamps, and NT stamps when with_nt. Returns the count (at most
 * max are written). */
int or_build_stamps(const or_deck *d, const or_seg *s, int n, double freq_hz, int with_nt, or_stamp *out, int max)
{
    int cnt = 0, c, i;
    cx *loads = (cx *)malloc(sizeof(cx) * (size_t)(n > 0 ? n : 1));
    const double k = TWO_PI * freq_hz / OR_C0;
#define PUSH(r, cc, z)                                                                                           \
    do {                                                                                                         \
        if (cnt < max) {                                                                                         \
            out[cnt].row = (r);                                                                                  \
            out[cnt].col = (cc);                                                                                 \
            out[cnt].delta = (z);                                                                                \
        }                                                                                                        \
        cnt++;                                                                                                   \
    } while (0)
    or_build_loads(d, s, n, freq_hz, loads);
    for (i = 0; i < n; i++)
        if (loads[i].re != 0.0 || loads[i].im != 0.0)
            PUSH(i, i, loads[i]);
    free(loads);
    for (c = 0; c < d->n; c++) {
        const or_card *tl = &d->cards[c];
        int i1, i2;
        double z0, len, f3;
        uint32_t type;
        if (strcmp(tl->mn, "TL"))
            continue;
        i1 = find_segment_index(s, n, or_as_u32(tl->f[0]), or_as_u32(tl->f[1]));
        i2 = find_segment_index(s, n, or_as_u32(tl->f[2]), or_as_u32(tl->f[3]));
        type = or_as_u32(tl->f[5]);
        z0 = tl->f[6];
        len = tl->f[7];
        f3 = tl->nf > 8 ? tl->f[8] : 1.0;
        if (i1 < 0 || i2 < 0 || i1 == i2 || z0 <= 0.0 || len <= 0.0)
            continue;
        if (type != 0) {
            /* lossy: Z0 coth(gl), Z0 csch(gl), gl = F3 ln10 / 20 + j k len;
             * num-complex sinh / cosh: (sinh a cos b, cosh a sin b), (cosh a cos b, sinh a sin b) */
            const double beta_l = k * len, alpha_l = f3 * 2.302585092994046 / 20.0;
            const cx sh = cx_make(sinh(alpha_l) * cos(beta_l), cosh(alpha_l) * sin(beta_l));
            const cx ch = cx_make(cosh(alpha_l) * cos(beta_l), sinh(alpha_l) * sin(beta_l));
            if (cx_norm(sh) < 1e-9)
                continue;
            {
                const cx z0c = cx_make(z0, 0.0);
                const cx zd = cx_div(cx_mul(z0c, ch), sh), zo = cx_div(z0c, sh);
                PUSH(i1, i1, zd);
                PUSH(i2, i2, zd);
                PUSH(i1, i2, zo);
                PUSH(i2, i1, zo);
            }
            continue;
        }
        {
            const double vf = (f3 >= 0.0 && f3 <= 1.0 && f3 > 0.0) ? f3 : 1.0;
            const double theta = k * len / vf;
            const double sin_t = sin(theta);
            if (fabs(sin_t) < 1e-9)
                continue;
            {
                const double cot = cos(theta) / sin_t, csc = 1.0 / sin_t;
                const cx zd = cx_make(0.0, -z0 * cot), zo = cx_make(0.0, -z0 * csc);
                PUSH(i1, i1, zd);
                PUSH(i2, i2, zd);
                PUSH(i1, i2, zo);
                PUSH(i2, i1, zo);
            }
        }
    }
    if (with_nt)
        for (c = 0; c < d->n; c++) {
            const or_card *nt = &d->cards[c];
            uint32_t t1, s1, t2, s2;
            double y[6];
            int i1, i2, bad = 0;
            if (strcmp(nt->mn, "NT") || nt->nf < 10)
                continue;
            if (parse_u32_str(nt->raw[0], &t1) || parse_u32_str(nt->raw[1], &s1) ||
                parse_u32_str(nt->raw[2], &t2) || parse_u32_str(nt->raw[3], &s2))
                continue;
            for (i = 0; i < 6; i++)
                if (or_parse_f64_strict(nt->raw[4 + i], &y[i]))
                    bad = 1;
            if (bad)
                continue;
            i1 = find_segment_index(s, n, t1, s1);
            i2 = find_segment_index(s, n, t2, s2);
            if (i1 < 0 || i2 < 0 || i1 == i2)
                continue;
            {
                const cx y11 = cx_make(y[0], y[1]), y12 = cx_make(y[2], y[3]), y21 = y12, y22 = cx_make(y[4], y[5]);
                const cx det = cx_sub(cx_mul(y11, y22), cx_mul(y12, y21));
                if (cx_norm(det) < 1e-30)
                    continue;
                PUSH(i1, i1, cx_div(y22, det));
                PUSH(i2, i2, cx_div(y11, det));
                PUSH(i1, i2, cx_div(cx_neg(y12), det));
                PUSH(i2, i1, cx_div(cx_neg(y21), det));
            }
        }
#undef PUSH
    return cnt;
}

/* ZMatrix::add_to_diagonal / add_to_entry (matrix.rs:131-147) on row-major z */
void or_apply_stamps(cx *z, int n, const or_stamp *st, int ns)
{
    int i;
    for (i = 0; i < ns; i++) {
        cx *e = &z[(size_t)st[i].row * (size_t)n + (size_t)st[i].col];
        *e = cx_add(*e, st[i].delta);
    }
}
