/*
 * ORACLE — test infrastructure only.
 *
 * Whole-deck pipelines and report contract v1:
 *   CLI     apps/nec-cli/src/main.rs:256-305 + solve_session.rs:1186-1775
 *           (Hallén mode: paths route :1338-1378, merged route :1379-1437,
 *            feedpoint rows :846-947, finite-ground efficiency :1664-1684,
 *            average power gain :1701-1719, diag values :118-207)
 *   worker  crates/nec_worker/src/solve.rs:150-289
 *   python  bindings/fnec_py/src/lib.rs:47-115
 *   report  crates/nec_report/src/lib.rs:307-482
 * LD/TL/NT stamps, plane-wave and current-source excitations are outside the
 * Hallén delta-gap path and are reported as unsupported.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

double or_diag_spread(const cx *z, int n)
{
    double mx = 0.0, mn = INFINITY;
    int i;
    if (n == 0)
        return 0.0;
    for (i = 0; i < n; i++) {
        double d = cx_norm(z[(size_t)i * (size_t)n + (size_t)i]);
        mx = fmax(mx, d);
        mn = fmin(mn, d);
    }
    if (!isfinite(mx) || !isfinite(mn))
        return NAN;
    if (mx == 0.0)
        return 0.0;
    return mx / fmax(mn, 1e-30);
}

static double l2(const cx *v, int n)
{
    double s = 0.0;
    int i;
    for (i = 0; i < n; i++)
        s = s + cx_norm_sqr(v[i]);
    return sqrt(s);
}

void or_residual_hallen(const cx *z, int n, const cx *cur, const cx *chom, int nchom,
                        const double *cosv, const cx *rhs, const int *wf, const int *wl, int nw,
                        double *abs_res, double *rel_res)
{
    cx *r = (cx *)calloc((size_t)(n > 0 ? n : 1), sizeof(cx));
    int *row_wire = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    int fb_f = 0, fb_l = n - 1, w = nw, i, row, col;
    const int *ef = wf, *el = wl;
    double res, den;
    if (nw == 0 || n == 0) {
        ef = &fb_f;
        el = &fb_l;
        w = n > 0 ? 1 : 0;
    }
    for (i = 0; i < w; i++)
        for (row = ef[i]; row <= el[i] && row < n; row++)
            row_wire[row] = i;
    for (row = 0; row < n; row++) {
        cx zi = cx_make(0.0, 0.0), c_row;
        for (col = 0; col < n; col++)
            zi = cx_add(zi, cx_mul(z[(size_t)row * (size_t)n + (size_t)col], cur[col]));
        if (row_wire[row] < nchom)
            c_row = chom[row_wire[row]];
        else if (nchom > 0)
            c_row = chom[0];
        else
            c_row = cx_make(0.0, 0.0);
        r[row] = cx_sub(cx_sub(zi, cx_scale(c_row, cosv[row])), rhs[row]);
    }
    res = l2(r, n);
    den = l2(rhs, n);
    *abs_res = res;
    *rel_res = den > 0.0 ? res / den : res;
    free(r);
    free(row_wire);
}

static void residual_paths(const cx *z, int n, const cx *cur, const cx *chom, int nchom,
                           const double *cosv, const cx *rhs, const int *path_of, double *a,
                           double *rr)
{
    cx *r = (cx *)calloc((size_t)(n > 0 ? n : 1), sizeof(cx));
    int row, col;
    double res, den;
    for (row = 0; row < n; row++) {
        cx zi = cx_make(0.0, 0.0), c_row;
        for (col = 0; col < n; col++)
            zi = cx_add(zi, cx_mul(z[(size_t)row * (size_t)n + (size_t)col], cur[col]));
        c_row = path_of[row] < nchom ? chom[path_of[row]] : cx_make(0.0, 0.0);
        r[row] = cx_sub(cx_sub(zi, cx_scale(c_row, cosv[row])), rhs[row]);
    }
    res = l2(r, n);
    den = l2(rhs, n);
    *a = res;
    *rr = den > 0.0 ? res / den : res;
    free(r);
}

/* ---- Rust number formatting (report contract v1) ---- */

/* {:.prec}: fixed, exact decimal rounding (half-even on exact ties), "-0"
 * kept for negative values. glibc printf rounds the exact binary value the
 * same way. Rust prints inf/NaN as "inf"/"NaN". */
void or_fmt_fixed(char *buf, int cap, double v, int prec)
{
    if (isnan(v)) {
        snprintf(buf, cap, "NaN");
        return;
    }
    if (isinf(v)) {
        snprintf(buf, cap, v < 0 ? "-inf" : "inf");
        return;
    }
    snprintf(buf, cap, "%.*f", prec, v);
}

/* {:.prec e}: mantissa with `prec` decimals, exponent without '+' or
 * zero padding ("1.301300e-2", "7.400000e1", "0.000000e0"). */
void or_fmt_exp(char *buf, int cap, double v, int prec)
{
    char tmp[64];
    char *e;
    int ex;
    if (isnan(v) || isinf(v)) {
        or_fmt_fixed(buf, cap, v, prec);
        return;
    }
    snprintf(tmp, sizeof(tmp), "%.*e", prec, v);
    e = strchr(tmp, 'e');
    ex = atoi(e + 1);
    *e = 0;
    snprintf(buf, cap, "%se%d", tmp, ex);
}

typedef struct {
    char *s;
    size_t n, cap;
} sbuf;

static void sb_add(sbuf *b, const char *t)
{
    size_t l = strlen(t);
    if (b->n + l + 1 > b->cap) {
        b->cap = (b->n + l + 1) * 2 + 256;
        b->s = (char *)realloc(b->s, b->cap);
    }
    memcpy(b->s + b->n, t, l + 1);
    b->n += l;
}

static void report_v1(const or_deck *d, or_result *res, const char *mode, const double *avg_gain,
                      int normalize)
{
    sbuf b = {NULL, 0, 0};
    char t[512], f1[64], f2[64], f3[64], f4[64], f5[64], f6[64];
    int i, nsrc = 0, nld = 0;
    sb_add(&b, "FNEC FEEDPOINT REPORT\nFORMAT_VERSION 1\n");
    or_fmt_fixed(f1, 64, res->freq_hz / 1e6, 6);
    snprintf(t, sizeof(t), "FREQ_MHZ %s\nSOLVER_MODE %s\nPULSE_RHS Nec2\n\nFEEDPOINTS\n", f1, mode);
    sb_add(&b, t);
    sb_add(&b, "TAG SEG V_RE V_IM I_RE I_IM Z_RE Z_IM\n");
    for (i = 0; i < res->nfeed; i++) {
        char f7[64], f8[64];
        or_fmt_fixed(f1, 64, res->feed_v[i].re, 6);
        or_fmt_fixed(f2, 64, res->feed_v[i].im, 6);
        or_fmt_fixed(f3, 64, res->feed_i[i].re, 6);
        or_fmt_fixed(f4, 64, res->feed_i[i].im, 6);
        or_fmt_fixed(f7, 64, res->feed_z[i].re, 6);
        or_fmt_fixed(f8, 64, res->feed_z[i].im, 6);
        snprintf(t, sizeof(t), "%d %d %s %s %s %s %s %s\n", res->feed_tag[i], res->feed_seg[i], f1,
                 f2, f3, f4, f7, f8);
        sb_add(&b, t);
    }
    for (i = 0; i < d->n; i++)
        if (!strcmp(d->cards[i].mn, "EX"))
            nsrc++;
        else if (!strcmp(d->cards[i].mn, "LD"))
            nld++;
    if (nsrc) {
        snprintf(t, sizeof(t), "\nSOURCES\nN_SOURCES %d\nTYPE TAG SEG I4 V_RE V_IM\n", nsrc);
        sb_add(&b, t);
        for (i = 0; i < d->n; i++) {
            const or_card *c = &d->cards[i];
            if (strcmp(c->mn, "EX"))
                continue;
            or_fmt_fixed(f1, 64, ex_vr(c), 6);
            or_fmt_fixed(f2, 64, ex_vi(c), 6);
            snprintf(t, sizeof(t), "%u %u %u %u %s %s\n", or_as_u32(c->f[0]), ex_tag(c), ex_seg(c),
                     or_as_u32(c->f[3]), f1, f2);
            sb_add(&b, t);
        }
    }
    /* LD cards as parsed (solve_session.rs:966-982, nec_report lib.rs:335-344) */
    if (nld) {
        snprintf(t, sizeof(t), "\nLOADS\nN_LOADS %d\nTYPE TAG SEG_FIRST SEG_LAST F1 F2 F3\n", nld);
        sb_add(&b, t);
        for (i = 0; i < d->n; i++) {
            const or_card *c = &d->cards[i];
            if (strcmp(c->mn, "LD"))
                continue;
            or_fmt_fixed(f1, 64, c->nf > 4 ? c->f[4] : 0.0, 6);
            or_fmt_fixed(f2, 64, c->nf > 5 ? c->f[5] : 0.0, 6);
            or_fmt_fixed(f3, 64, c->nf > 6 ? c->f[6] : 0.0, 6);
            snprintf(t, sizeof(t), "%d %u %u %u %s %s %s\n", (int)or_as_i32(c->f[0]),
                     or_as_u32(c->f[1]), or_as_u32(c->f[2]), or_as_u32(c->f[3]), f1, f2, f3);
            sb_add(&b, t);
        }
    }
    sb_add(&b, "\nCURRENTS\nTAG SEG I_RE I_IM I_MAG I_PHASE\n");
    for (i = 0; i < res->n; i++) {
        cx c = res->x[i];
        or_fmt_exp(f1, 64, c.re, 6);
        or_fmt_exp(f2, 64, c.im, 6);
        or_fmt_exp(f3, 64, cx_norm(c), 6);
        or_fmt_fixed(f4, 64, rad2deg(cx_arg(c)), 4);
        snprintf(t, sizeof(t), "%u %u %s %s %s %s\n", res->segs[i].tag, res->segs[i].tag_index, f1,
                 f2, f3, f4);
        sb_add(&b, t);
    }
    if (res->np > 0) {
        snprintf(t, sizeof(t), "\nRADIATION_PATTERN\nN_POINTS %d\nTHETA PHI GAIN_DB GAIN_V_DB GAIN_H_DB AXIAL_RATIO\n", res->np);
        sb_add(&b, t);
        for (i = 0; i < res->np; i++) {
            const double *o = res->rp + 4 * (size_t)i;
            or_fmt_fixed(f1, 64, res->rp_theta[i], 4);
            or_fmt_fixed(f2, 64, res->rp_phi[i], 4);
            or_fmt_fixed(f3, 64, o[0], 4);
            or_fmt_fixed(f4, 64, o[1], 4);
            or_fmt_fixed(f5, 64, o[2], 4);
            or_fmt_fixed(f6, 64, o[3], 4);
            snprintf(t, sizeof(t), "%s %s %s %s %s %s\n", f1, f2, f3, f4, f5, f6);
            sb_add(&b, t);
        }
    }
    if (normalize && res->np > 0) {
        double peak = -INFINITY;
        for (i = 0; i < res->np; i++)
            peak = fmax(peak, res->rp[4 * (size_t)i]);
        snprintf(t, sizeof(t), "\nNORMALIZED_PATTERN\nN_POINTS %d\nTHETA PHI GAIN_NORM_DB\n", res->np);
        sb_add(&b, t);
        for (i = 0; i < res->np; i++) {
            double g = res->rp[4 * (size_t)i];
            double nv = g > -900.0 ? g - peak : g;
            or_fmt_fixed(f1, 64, res->rp_theta[i], 4);
            or_fmt_fixed(f2, 64, res->rp_phi[i], 4);
            or_fmt_fixed(f3, 64, nv, 4);
            snprintf(t, sizeof(t), "%s %s %s\n", f1, f2, f3);
            sb_add(&b, t);
        }
    }
    if (avg_gain) {
        or_fmt_fixed(f1, 64, *avg_gain, 4);
        snprintf(t, sizeof(t), "\nAVERAGE_POWER_GAIN %s\n", f1);
        sb_add(&b, t);
    }
    res->report = b.s;
}

static void fail(or_result *res, int st, const char *msg)
{
    res->status = st;
    snprintf(res->err, sizeof(res->err), "%s", msg);
}

int or_run_deck(const char *text, int semantics, int freq_index, double freq_override_hz,
                int with_report, or_result *res)
{
    char err[256];
    or_deck *d;
    int n, i, maxseg = 1 << 20;
    double freqs[4096];
    int nfreq;
    cx *z = NULL, *rhs = NULL, *v = NULL;
    double *cosv = NULL;
    int *wf = NULL, *wl = NULL, nw = 0, *ja = NULL, *jb = NULL, nj = 0;
    double *js = NULL;
    memset(res, 0, sizeof(*res));
    d = or_deck_parse(text, err, sizeof(err));
    if (!d) {
        fail(res, OR_PARSE, err);
        return res->status;
    }
    for (i = 0; i < d->n; i++)
        if (!strcmp(d->cards[i].mn, "EX") && (ex_is_plane_wave(&d->cards[i]) || ex_is_current(&d->cards[i]))) {
            fail(res, OR_UNSUPPORTED, "plane-wave / current-source excitation");
            or_deck_free(d);
            return res->status;
        }
    if (freq_override_hz > 0.0)
        res->freq_hz = freq_override_hz;
    else {
        nfreq = semantics == OR_SEM_PY ? or_deck_frequencies_py(d, freqs, 4096)
                                       : or_deck_frequencies(d, freqs, 4096);
        if (freq_index >= nfreq || freq_index >= 4096) {
            fail(res, OR_PARSE, "no such frequency");
            or_deck_free(d);
            return res->status;
        }
        res->freq_hz = freqs[freq_index];
    }
    res->segs = (or_seg *)malloc(sizeof(or_seg) * (size_t)maxseg);
    n = or_build_geometry(d, res->segs, maxseg, err, sizeof(err));
    if (n < 0) {
        fail(res, -n, err);
        or_deck_free(d);
        return res->status;
    }
    res->n = n;
    res->ground = or_ground_model(d);
    v = (cx *)calloc((size_t)n, sizeof(cx));
    rhs = (cx *)calloc((size_t)n, sizeof(cx));
    cosv = (double *)calloc((size_t)n, sizeof(double));
    z = (cx *)malloc(sizeof(cx) * (size_t)n * (size_t)n);
    wf = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    wl = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    ja = (int *)malloc(sizeof(int) * (size_t)(4 * n + 4));
    jb = (int *)malloc(sizeof(int) * (size_t)(4 * n + 4));
    js = (double *)malloc(sizeof(double) * (size_t)(4 * n + 4));
    if (or_build_excitation(d, res->segs, n, v)) {
        if (semantics != OR_SEM_WORKER) {
            fail(res, OR_EXCITATION, "EX: bad excitation");
            goto out;
        }
    }
    or_fill_z(res->segs, n, res->freq_hz, res->ground, z);
    {
        /* LD, TL (and for the CLI NT) stamps on Z (solve_session.rs:1285-1309,
         * nec_worker solve.rs:179-185, fnec_py lib.rs:58-63) */
        const int ns = or_build_stamps(d, res->segs, n, res->freq_hz, semantics == OR_SEM_CLI, NULL, 0);
        if (ns > 0) {
            or_stamp *st = (or_stamp *)malloc(sizeof(or_stamp) * (size_t)ns);
            or_build_stamps(d, res->segs, n, res->freq_hz, semantics == OR_SEM_CLI, st, ns);
            or_apply_stamps(z, n, st, ns);
            free(st);
        }
    }
    res->diag_spread = or_diag_spread(z, n);

    if (semantics == OR_SEM_CLI) {
        int *path_of = (int *)malloc(sizeof(int) * (size_t)n);
        double *sign_of = (double *)malloc(sizeof(double) * (size_t)n);
        double *s_of = (double *)malloc(sizeof(double) * (size_t)n);
        int *fa = (int *)malloc(sizeof(int) * (size_t)(n + 1));
        int *fbv = (int *)malloc(sizeof(int) * (size_t)(n + 1));
        int nontriv = 0, np;
        np = or_conductor_paths(res->segs, n, path_of, sign_of, s_of, fa, fbv, n + 1, &nontriv);
        if (np > 0 && nontriv) {
            int *fe = (int *)malloc(sizeof(int) * (size_t)(2 * np));
            res->used_paths = 1;
            if (or_build_hallen_rhs_paths(d, res->segs, n, res->freq_hz, path_of, sign_of, s_of, rhs, cosv)) {
                fail(res, OR_EXCITATION, "EX: bad excitation");
            } else {
                for (i = 0; i < np; i++) {
                    fe[2 * i] = fa[i];
                    fe[2 * i + 1] = fbv[i];
                }
                res->s = n + np;
                res->nw = np;
                res->r = n + 2 * np;
                res->x = (cx *)calloc((size_t)res->s, sizeof(cx));
                res->status = or_solve_hallen_paths(z, n, rhs, cosv, path_of, fe, 2 * np, res->x);
                if (!res->status)
                    residual_paths(z, n, res->x, res->x + n, np, cosv, rhs, path_of, &res->abs_res, &res->rel_res);
                else
                    fail(res, res->status, "impedance matrix is singular");
            }
            free(fe);
        } else {
            int *comp = (int *)calloc((size_t)n, sizeof(int));
            int nja = 0, k, *ta, *tb;
            double *ts;
            if (or_build_hallen_rhs(d, res->segs, n, res->freq_hz, rhs, cosv)) {
                fail(res, OR_EXCITATION, "EX: bad excitation");
            } else {
                nw = or_merge_collinear(res->segs, n, wf, wl, n + 1);
                for (i = 0; i < nw; i++)
                    for (k = wf[i]; k <= wl[i]; k++)
                        comp[k] = i;
                nj = or_detect_junctions(res->segs, wf, wl, nw, 1e-6, ja, jb, js, 4 * n + 4);
                ta = (int *)malloc(sizeof(int) * (size_t)(nj + 1));
                tb = (int *)malloc(sizeof(int) * (size_t)(nj + 1));
                ts = (double *)malloc(sizeof(double) * (size_t)(nj + 1));
                for (i = 0; i < nj; i++)
                    if (comp[ja[i]] != comp[jb[i]]) {
                        ta[nja] = ja[i];
                        tb[nja] = jb[i];
                        ts[nja] = js[i];
                        nja++;
                    }
                res->nw = nw > 0 ? nw : 1;
                res->s = n + res->nw;
                res->x = (cx *)calloc((size_t)res->s, sizeof(cx));
                or_hallen_normal_eqs(z, n, cosv, rhs, wf, wl, nw, ta, tb, ts, nja, NULL, NULL, &res->r);
                res->status = or_solve_hallen(z, n, rhs, cosv, wf, wl, nw, ta, tb, ts, nja, res->x);
                if (!res->status)
                    or_residual_hallen(z, n, res->x, res->x + n, res->nw, cosv, rhs, wf, wl, nw,
                                       &res->abs_res, &res->rel_res);
                else
                    fail(res, res->status, "impedance matrix is singular");
                free(ta);
                free(tb);
                free(ts);
            }
            free(comp);
        }
        free(path_of);
        free(sign_of);
        free(s_of);
        free(fa);
        free(fbv);
    } else {
        /* worker / python: raw wires, all junctions */
        if (or_build_hallen_rhs(d, res->segs, n, res->freq_hz, rhs, cosv)) {
            fail(res, OR_EXCITATION, "EX: bad excitation");
            goto out;
        }
        nw = or_wire_endpoints(res->segs, n, wf, wl, n + 1);
        nj = or_detect_junctions(res->segs, wf, wl, nw, 1e-6, ja, jb, js, 4 * n + 4);
        res->nw = nw > 0 ? nw : 1;
        res->s = n + res->nw;
        res->x = (cx *)calloc((size_t)res->s, sizeof(cx));
        or_hallen_normal_eqs(z, n, cosv, rhs, wf, wl, nw, ja, jb, js, nj, NULL, NULL, &res->r);
        res->status = or_solve_hallen(z, n, rhs, cosv, wf, wl, nw, ja, jb, js, nj, res->x);
        if (res->status)
            fail(res, res->status, "impedance matrix is singular");
        else
            or_residual_hallen(z, n, res->x, res->x + n, res->nw, cosv, rhs, wf, wl, nw,
                               &res->abs_res, &res->rel_res);
    }
    if (res->status)
        goto out;

    /* feedpoints */
    res->feed_tag = (int *)malloc(sizeof(int) * (size_t)(d->n + 1));
    res->feed_seg = (int *)malloc(sizeof(int) * (size_t)(d->n + 1));
    res->feed_idx = (int *)malloc(sizeof(int) * (size_t)(d->n + 1));
    res->feed_v = (cx *)malloc(sizeof(cx) * (size_t)(d->n + 1));
    res->feed_i = (cx *)malloc(sizeof(cx) * (size_t)(d->n + 1));
    res->feed_z = (cx *)malloc(sizeof(cx) * (size_t)(d->n + 1));
    for (i = 0; i < d->n; i++) {
        const or_card *c = &d->cards[i];
        int idx;
        cx cur, vs, zin;
        if (strcmp(c->mn, "EX"))
            continue;
        if (ex_is_plane_wave(c))
            continue;
        if (semantics == OR_SEM_WORKER && ex_type(c) != 0)
            continue;
        idx = or_find_seg(res->segs, n, ex_tag(c), ex_seg(c));
        if (idx < 0)
            continue;
        cur = res->x[idx];
        if (semantics == OR_SEM_WORKER)
            vs = cx_make(ex_vr(c), ex_vi(c));
        else
            vs = cx_scale(v[idx], res->segs[idx].length);
        zin = cx_norm(cur) > 1e-60 ? cx_div(vs, cur) : vs;
        res->feed_tag[res->nfeed] = (int)res->segs[idx].tag;
        res->feed_seg[res->nfeed] = (int)res->segs[idx].tag_index;
        res->feed_idx[res->nfeed] = idx;
        res->feed_v[res->nfeed] = vs;
        res->feed_i[res->nfeed] = cur;
        res->feed_z[res->nfeed] = zin;
        res->nfeed++;
        if (semantics != OR_SEM_CLI)
            break;
    }
    if (res->nfeed == 0 && semantics != OR_SEM_CLI) {
        fail(res, OR_NO_FEED, "no EX type-0 card found in deck");
        goto out;
    }

    if (semantics == OR_SEM_CLI) {
        int np = or_deck_rp_points(d, NULL, NULL, 0), normalize, avgf;
        double avg = 0.0;
        int have_avg = 0;
        if (np > 0) {
            res->np = np;
            res->rp_theta = (double *)malloc(sizeof(double) * (size_t)np);
            res->rp_phi = (double *)malloc(sizeof(double) * (size_t)np);
            res->rp = (double *)malloc(sizeof(double) * 4 * (size_t)np);
            or_deck_rp_points(d, res->rp_theta, res->rp_phi, np);
            res->rp_total = or_radiation_pattern(res->segs, n, res->x, res->freq_hz, res->ground,
                                                 res->rp_theta, res->rp_phi, np, res->rp);
            if (res->ground.kind == OR_GROUND_FINITE) {
                double p_in = 0.0;
                for (i = 0; i < res->nfeed; i++)
                    p_in += 0.5 * cx_mul(res->feed_v[i], cx_conj(res->feed_i[i])).re;
                if (p_in > 0.0) {
                    double eta = or_radiation_efficiency(res->segs, n, res->x, res->freq_hz,
                                                         res->ground, p_in);
                    double ddb = 10.0 * log10(eta);
                    for (i = 0; i < np; i++) {
                        int q;
                        for (q = 0; q < 3; q++)
                            if (res->rp[4 * (size_t)i + q] > -900.0)
                                res->rp[4 * (size_t)i + q] += ddb;
                    }
                }
            }
        }
        or_deck_rp_flags(d, &normalize, &avgf);
        if (avgf && np > 0) {
            double num = 0.0, den = 0.0;
            for (i = 0; i < np; i++) {
                double g = res->rp[4 * (size_t)i], w;
                if (g <= -900.0)
                    continue;
                w = sin(deg2rad(res->rp_theta[i]));
                num += pow(10.0, g / 10.0) * w;
                den += w;
            }
            if (den > 0.0) {
                avg = num / den;
                have_avg = 1;
            }
        }
        if (with_report)
            report_v1(d, res, res->used_paths ? "hallen" : "hallen", have_avg ? &avg : NULL, normalize);
    }
out:
    free(z);
    free(rhs);
    free(v);
    free(cosv);
    free(wf);
    free(wl);
    free(ja);
    free(jb);
    free(js);
    or_deck_free(d);
    return res->status;
}

void or_result_free(or_result *res)
{
    free(res->segs);
    free(res->x);
    free(res->feed_tag);
    free(res->feed_seg);
    free(res->feed_idx);
    free(res->feed_v);
    free(res->feed_i);
    free(res->feed_z);
    free(res->rp_theta);
    free(res->rp_phi);
    free(res->rp);
    free(res->report);
    memset(res, 0, sizeof(*res));
}

int or_deck_stamps(const char *text, int semantics, double freq_hz, int *rows, int *cols, double *re,
                   double *im, int max)
{
    char err[256];
    or_deck *d = or_deck_parse(text, err, sizeof(err));
    or_seg *segs;
    or_stamp *st;
    int n, ns, i;
    if (!d)
        return -OR_PARSE;
    segs = (or_seg *)malloc(sizeof(or_seg) * (size_t)(1 << 20));
    n = or_build_geometry(d, segs, 1 << 20, err, sizeof(err));
    if (n < 0) {
        free(segs);
        or_deck_free(d);
        return n;
    }
    ns = or_build_stamps(d, segs, n, freq_hz, semantics == OR_SEM_CLI, NULL, 0);
    st = (or_stamp *)malloc(sizeof(or_stamp) * (size_t)(ns > 0 ? ns : 1));
    or_build_stamps(d, segs, n, freq_hz, semantics == OR_SEM_CLI, st, ns);
    for (i = 0; i < ns && i < max; i++) {
        rows[i] = st[i].row;
        cols[i] = st[i].col;
        re[i] = st[i].delta.re;
        im[i] = st[i].delta.im;
    }
    free(st);
    free(segs);
    or_deck_free(d);
    return ns;
}
