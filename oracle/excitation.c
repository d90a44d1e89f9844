/*
 * ORACLE — test infrastructure only.
 *
 * Excitation, restating crates/nec_solver/src/excitation.rs:
 *   build_excitation / apply_ex   :78-90, :450-482
 *   build_hallen_rhs              :119-244
 *   build_hallen_rhs_paths        :264-341
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

int or_build_excitation(const or_deck *d, const or_seg *s, int n, cx *v)
{
    int i;
    for (i = 0; i < n; i++)
        v[i] = cx_make(0.0, 0.0);
    for (i = 0; i < d->n; i++) {
        const or_card *c = &d->cards[i];
        int idx;
        if (strcmp(c->mn, "EX"))
            continue;
        if (ex_is_plane_wave(c) || ex_is_current(c))
            continue;
        if (!ex_is_voltage(c))
            return OR_EXCITATION;
        idx = or_find_seg(s, n, ex_tag(c), ex_seg(c));
        if (idx < 0)
            return OR_EXCITATION;
        v[idx] = cx_add(v[idx], cx_divs(cx_make(ex_vr(c), ex_vi(c)), s[idx].length));
    }
    return OR_OK;
}

int or_build_hallen_rhs(const or_deck *d, const or_seg *s, int n, double freq_hz, cx *rhs,
                        double *cosv)
{
    const or_card *first = NULL;
    int i, feed, nc, m;
    double k, scale;
    uint32_t driven_tag;
    /* Sources keyed by tag, iterated in ascending tag order (BTreeMap). */
    uint32_t *tags;
    int *src_idx;
    cx *src_v;
    int nsrc = 0;
    int *cf, *cl, *comp_of;

    for (i = 0; i < d->n; i++) {
        const or_card *c = &d->cards[i];
        if (strcmp(c->mn, "EX"))
            continue;
        if (ex_is_plane_wave(c) || ex_is_current(c))
            continue;
        if (!ex_is_voltage(c))
            return OR_EXCITATION;
        if (!first)
            first = c;
    }
    for (i = 0; i < n; i++) {
        rhs[i] = cx_make(0.0, 0.0);
        cosv[i] = 0.0;
    }
    if (!first)
        return OR_OK;
    feed = or_find_seg(s, n, ex_tag(first), ex_seg(first));
    if (feed < 0)
        return OR_EXCITATION;
    k = 2.0 * OR_PI * freq_hz / OR_C0;
    scale = 2.0 * OR_PI / OR_ETA0;
    driven_tag = s[feed].tag;

    tags = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(d->n + 1));
    src_idx = (int *)malloc(sizeof(int) * (size_t)(d->n + 1));
    src_v = (cx *)malloc(sizeof(cx) * (size_t)(d->n + 1));
    tags[0] = driven_tag;
    src_idx[0] = feed;
    src_v[0] = cx_make(ex_vr(first), ex_vi(first));
    nsrc = 1;
    for (i = 0; i < d->n; i++) {
        const or_card *c = &d->cards[i];
        int idx, j;
        if (strcmp(c->mn, "EX"))
            continue;
        if (!ex_is_voltage(c) || ex_tag(c) == driven_tag)
            continue;
        idx = or_find_seg(s, n, ex_tag(c), ex_seg(c));
        if (idx < 0)
            continue;
        for (j = 0; j < nsrc; j++)
            if (tags[j] == ex_tag(c))
                break;
        tags[j] = ex_tag(c);
        src_idx[j] = idx;
        src_v[j] = cx_make(ex_vr(c), ex_vi(c));
        if (j == nsrc)
            nsrc++;
    }
    /* sort sources by tag (insertion sort, tags unique) */
    for (i = 1; i < nsrc; i++) {
        int j = i;
        while (j > 0 && tags[j - 1] > tags[j]) {
            uint32_t t = tags[j]; int si = src_idx[j]; cx sv = src_v[j];
            tags[j] = tags[j - 1]; src_idx[j] = src_idx[j - 1]; src_v[j] = src_v[j - 1];
            tags[j - 1] = t; src_idx[j - 1] = si; src_v[j - 1] = sv;
            j--;
        }
    }

    cf = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    cl = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    comp_of = (int *)calloc((size_t)n, sizeof(int));
    nc = or_merge_collinear(s, n, cf, cl, n + 1);
    for (i = 0; i < nc; i++)
        for (m = cf[i]; m <= cl[i] && m < n; m++)
            comp_of[m] = i;

    for (m = 0; m < n; m++) {
        int a = cf[comp_of[m]], b = cl[comp_of[m]];
        const double *wd = s[a].dir;
        double wm[3], dl[3], sl;
        wm[0] = (s[a].mid[0] + s[b].mid[0]) / 2.0;
        wm[1] = (s[a].mid[1] + s[b].mid[1]) / 2.0;
        wm[2] = (s[a].mid[2] + s[b].mid[2]) / 2.0;
        dl[0] = s[m].mid[0] - wm[0];
        dl[1] = s[m].mid[1] - wm[1];
        dl[2] = s[m].mid[2] - wm[2];
        sl = dl[0] * wd[0] + dl[1] * wd[1] + dl[2] * wd[2];
        cosv[m] = cos(k * sl);
    }
    for (i = 0; i < nsrc; i++) {
        int fi = src_idx[i], a = cf[comp_of[fi]], b = cl[comp_of[fi]];
        const double *sm = s[fi].mid, *sd = s[fi].dir;
        for (m = a; m <= b; m++) {
            double dd[3], sv;
            dd[0] = s[m].mid[0] - sm[0];
            dd[1] = s[m].mid[1] - sm[1];
            dd[2] = s[m].mid[2] - sm[2];
            sv = dd[0] * sd[0] + dd[1] * sd[1] + dd[2] * sd[2];
            rhs[m] = cx_add(rhs[m], cx_mul(cx_make(0.0, -scale * sin(k * fabs(sv))), src_v[i]));
        }
    }
    free(tags);
    free(src_idx);
    free(src_v);
    free(cf);
    free(cl);
    free(comp_of);
    return OR_OK;
}

int or_build_hallen_rhs_paths(const or_deck *d, const or_seg *s, int n, double freq_hz,
                              const int *path_of, const double *sign_of, const double *s_of,
                              cx *rhs, double *cosv)
{
    double k = 2.0 * OR_PI * freq_hz / OR_C0;
    double scale = 2.0 * OR_PI / OR_ETA0;
    int i, m, any = 0;
    for (m = 0; m < n; m++) {
        cosv[m] = sign_of[m] * cos(k * s_of[m]);
        rhs[m] = cx_make(0.0, 0.0);
    }
    for (i = 0; i < d->n; i++) {
        const or_card *c = &d->cards[i];
        int fi, sp;
        double ssrc, fs;
        cx vs;
        if (strcmp(c->mn, "EX"))
            continue;
        if (ex_is_plane_wave(c) || ex_is_current(c))
            continue;
        if (!ex_is_voltage(c))
            return OR_EXCITATION;
        fi = or_find_seg(s, n, ex_tag(c), ex_seg(c));
        if (fi < 0)
            return OR_EXCITATION;
        any = 1;
        vs = cx_make(ex_vr(c), ex_vi(c));
        sp = path_of[fi];
        ssrc = s_of[fi];
        fs = sign_of[fi];
        for (m = 0; m < n; m++) {
            double ds;
            if (path_of[m] != sp)
                continue;
            ds = fabs(s_of[m] - ssrc);
            rhs[m] = cx_add(rhs[m], cx_scale(cx_mul(cx_make(0.0, -scale * sin(k * ds)), vs),
                                             sign_of[m] * fs));
        }
    }
    if (!any)
        for (m = 0; m < n; m++)
            rhs[m] = cx_make(0.0, 0.0);
    return OR_OK;
}
