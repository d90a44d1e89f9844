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
    if (res->nrecv > 0) {
        snprintf(t, sizeof(t), "\nRECEIVE_PATTERN\nN_POINTS %d\nTHETA PHI RESPONSE_DB\n", res->nrecv);
        sb_add(&b, t);
        for (i = 0; i < res->nrecv; i++) {
            or_fmt_fixed(f1, 64, res->recv_theta[i], 4);
            or_fmt_fixed(f2, 64, res->recv_phi[i], 4);
            or_fmt_fixed(f3, 64, res->recv_db[i], 4);
            snprintf(t, sizeof(t), "%s %s %s\n", f1, f2, f3);
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

/* solve_plane_wave_hallen (solve_session.rs:262-301): conductor paths when one
 * is not trivial, else the per-wire solve on the raw GW wires. x gets the
 * currents and the homogeneous constants; *s the count. */
static int plane_wave_solve(const or_deck *d, const or_seg *segs, int n, double freq, const cx *z, cx **x,
                            int *s_out, int *nhom)
{
    int *path_of = (int *)malloc(sizeof(int) * (size_t)n), *fa = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *fbv = (int *)malloc(sizeof(int) * (size_t)(n + 1)), *wf = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *wl = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    double *sign_of = (double *)malloc(sizeof(double) * (size_t)n), *s_of = (double *)malloc(sizeof(double) * (size_t)n);
    double *cosv = (double *)malloc(sizeof(double) * (size_t)n), *sinv = (double *)malloc(sizeof(double) * (size_t)n);
    cx *rhs = (cx *)malloc(sizeof(cx) * (size_t)n);
    int nontriv = 0, np, st, i;
    np = or_conductor_paths(segs, n, path_of, sign_of, s_of, fa, fbv, n + 1, &nontriv);
    if (np > 0 && nontriv) {
        int *fe = (int *)malloc(sizeof(int) * (size_t)(2 * np));
        for (i = 0; i < np; i++) {
            fe[2 * i] = fa[i];
            fe[2 * i + 1] = fbv[i];
        }
        st = or_planewave_hallen_paths(d, segs, n, freq, path_of, sign_of, s_of, rhs, cosv, sinv);
        *nhom = 2 * np;
        *x = (cx *)calloc((size_t)(n + *nhom), sizeof(cx));
        if (!st)
            st = or_solve_hallen_planewave_paths(z, n, rhs, cosv, sinv, path_of, fe, 2 * np, *x);
        free(fe);
    } else {
        const int nw = or_wire_endpoints(segs, n, wf, wl, n + 1);
        st = or_planewave_hallen(d, segs, n, freq, rhs, cosv, sinv);
        *nhom = 2 * (nw > 0 ? nw : 1);
        *x = (cx *)calloc((size_t)(n + *nhom), sizeof(cx));
        if (!st)
            st = or_solve_hallen_planewave(z, n, rhs, cosv, sinv, wf, wl, nw, *x);
    }
    *s_out = n + *nhom;
    free(path_of);
    free(fa);
    free(fbv);
    free(wf);
    free(wl);
    free(sign_of);
    free(s_of);
    free(cosv);
    free(sinv);
    free(rhs);
    return st;
}

/* solve_current_source_hallen (solve_session.rs:392-448): the first EX type 4
 * card forces i0 = F1 + j F2 on its segment. */
static int current_source_solve(const or_deck *d, const or_seg *segs, int n, double freq, const cx *z, cx **x,
                                int *s_out, int *nhom, cx *port_v)
{
    const or_card *cs = NULL;
    int *path_of = (int *)malloc(sizeof(int) * (size_t)n), *fa = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *fbv = (int *)malloc(sizeof(int) * (size_t)(n + 1)), *wf = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *wl = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    double *sign_of = (double *)malloc(sizeof(double) * (size_t)n), *s_of = (double *)malloc(sizeof(double) * (size_t)n);
    double *cosv = (double *)malloc(sizeof(double) * (size_t)n);
    cx *shape = (cx *)malloc(sizeof(cx) * (size_t)n), i0;
    int nontriv = 0, np, st = OR_OK, i, src = -1, nw;
    for (i = 0; i < d->n && !cs; i++)
        if (!strcmp(d->cards[i].mn, "EX") && ex_is_current(&d->cards[i]))
            cs = &d->cards[i];
    i0 = cx_make(ex_vr(cs), ex_vi(cs));
    np = or_conductor_paths(segs, n, path_of, sign_of, s_of, fa, fbv, n + 1, &nontriv);
    nw = or_wire_endpoints(segs, n, wf, wl, n + 1);
    *x = NULL;
    if (np > 0 && nontriv) {
        int *fe = (int *)malloc(sizeof(int) * (size_t)(2 * np));
        for (i = 0; i < np; i++) {
            fe[2 * i] = fa[i];
            fe[2 * i + 1] = fbv[i];
        }
        st = or_current_source_shape_paths(d, segs, n, freq, ex_tag(cs), ex_seg(cs), path_of, sign_of, s_of, shape,
                                           cosv, &src);
        *nhom = np;
        *x = (cx *)calloc((size_t)(n + np + 1), sizeof(cx));
        if (!st)
            st = or_solve_hallen_current_source_paths(z, n, shape, cosv, src, i0, path_of, fe, 2 * np, *x);
        free(fe);
    } else {
        if (np <= 0) {
            int *ja = (int *)malloc(sizeof(int) * (size_t)(4 * n + 4)), *jb = (int *)malloc(sizeof(int) * (size_t)(4 * n + 4));
            double *js = (double *)malloc(sizeof(double) * (size_t)(4 * n + 4));
            if (or_detect_junctions(segs, wf, wl, nw, 1e-6, ja, jb, js, 4 * n + 4) > 0)
                st = OR_UNSUPPORTED; /* degree-3+ junctions, closed loops */
            free(ja);
            free(jb);
            free(js);
        }
        *nhom = nw > 0 ? nw : 1;
        *x = (cx *)calloc((size_t)(n + *nhom + 1), sizeof(cx));
        if (!st)
            st = or_current_source_shape(d, segs, n, freq, ex_tag(cs), ex_seg(cs), shape, cosv, &src);
        if (!st)
            st = or_solve_hallen_current_source(z, n, shape, cosv, src, i0, wf, wl, nw, *x);
    }
    *s_out = n + *nhom + 1;
    *port_v = *x ? (*x)[*s_out - 1] : cx_make(0.0, 0.0);
    free(path_of);
    free(fa);
    free(fbv);
    free(wf);
    free(wl);
    free(sign_of);
    free(s_of);
    free(cosv);
    free(shape);
    return st;
}

/* plane_wave_receive_sweep (solve_session.rs:313-371): with NTHETA = max(I2,
 * 1), NPHI = max(I3, 1) and NTHETA NPHI > 1, each incidence (theta0 + it
 * dtheta, phi0 + ip dphi; it outer) is solved as a single-incidence deck and
 * its peak |I| recorded; the response is 20 log10(peak / max peak). */
static int receive_sweep(const or_deck *d, const or_seg *segs, int n, double freq, const cx *z, or_result *res)
{
    const or_card *ex = NULL;
    or_deck dd;
    uint32_t nth, nph, it, ip;
    int i, st = OR_OK, cnt = 0;
    double *peak, max_peak = 0.0;
    for (i = 0; i < d->n && !ex; i++)
        if (!strcmp(d->cards[i].mn, "EX") && ex_is_plane_wave(&d->cards[i]))
            ex = &d->cards[i];
    nth = ex_tag(ex) > 0 ? ex_tag(ex) : 1;
    nph = ex_seg(ex) > 0 ? ex_seg(ex) : 1;
    if ((uint64_t)nth * nph <= 1)
        return OR_OK;
    res->recv_theta = (double *)malloc(sizeof(double) * (size_t)nth * nph);
    res->recv_phi = (double *)malloc(sizeof(double) * (size_t)nth * nph);
    res->recv_db = (double *)malloc(sizeof(double) * (size_t)nth * nph);
    peak = (double *)malloc(sizeof(double) * (size_t)nth * nph);
    dd.n = d->n;
    dd.cap = d->n;
    dd.cards = (or_card *)malloc(sizeof(or_card) * (size_t)(d->n > 0 ? d->n : 1));
    for (it = 0; it < nth && !st; it++)
        for (ip = 0; ip < nph && !st; ip++) {
            const double theta = (ex->nf > 4 ? ex->f[4] : 0.0) + (double)it * (ex->nf > 7 ? ex->f[7] : 0.0);
            const double phi = (ex->nf > 5 ? ex->f[5] : 0.0) + (double)ip * (ex->nf > 8 ? ex->f[8] : 0.0);
            cx *x = NULL;
            int s, nhom, m;
            double pk = 0.0;
            memcpy(dd.cards, d->cards, sizeof(or_card) * (size_t)d->n);
            for (i = 0; i < dd.n; i++) {
                or_card *c = &dd.cards[i];
                if (strcmp(c->mn, "EX") || !ex_is_plane_wave(c))
                    continue;
                while (c->nf < 9)
                    c->f[c->nf++] = 0.0;
                c->f[1] = 1.0;
                c->f[2] = 1.0;
                c->f[4] = theta;
                c->f[5] = phi;
                c->f[7] = 0.0;
                c->f[8] = 0.0;
            }
            st = plane_wave_solve(&dd, segs, n, freq, z, &x, &s, &nhom);
            for (m = 0; m < n && !st; m++)
                pk = fmax(pk, cx_norm(x[m]));
            free(x);
            res->recv_theta[cnt] = theta;
            res->recv_phi[cnt] = phi;
            peak[cnt++] = pk;
        }
    free(dd.cards);
    for (i = 0; i < cnt; i++)
        max_peak = fmax(max_peak, peak[i]);
    for (i = 0; i < cnt; i++)
        res->recv_db[i] = peak[i] > 0.0 && max_peak > 0.0 ? 20.0 * log10(peak[i] / max_peak) : -999.99;
    res->nrecv = st ? 0 : cnt;
    free(peak);
    return st;
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
            /* the CLI solves them on their own Hallén systems; the worker and
             * the Python binding do not */
            if (semantics != OR_SEM_CLI) {
                fail(res, OR_UNSUPPORTED, "plane-wave / current-source excitation");
                or_deck_free(d);
                return res->status;
            }
            if (ex_is_plane_wave(&d->cards[i]))
                res->variant = 1;
            else if (res->variant == 0)
                res->variant = 2;
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

    if (res->variant) {
        /* plane wave first, then current source (solve_session.rs:1327-1337) */
        int nhom = 0;
        res->status = res->variant == 1 ? plane_wave_solve(d, res->segs, n, res->freq_hz, z, &res->x, &res->s, &nhom)
                                        : current_source_solve(d, res->segs, n, res->freq_hz, z, &res->x, &res->s,
                                                               &nhom, &res->port_v);
        res->nw = nhom;
        if (!res->status && res->variant == 1)
            res->status = receive_sweep(d, res->segs, n, res->freq_hz, z, res);
        if (res->status) {
            fail(res, res->status, res->status == OR_SINGULAR ? "impedance matrix is singular"
                                                               : "EX: plane-wave / current-source setup failed");
            goto out;
        }
    } else if (semantics == OR_SEM_CLI) {
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
        else if (res->variant == 2 && ex_is_current(c))
            vs = res->port_v; /* the solved port voltage (solve_session.rs:893-897) */
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
            report_v1(d, res,
                      res->variant == 1 ? "hallen-planewave" : res->variant == 2 ? "hallen-current-source" : "hallen",
                      have_avg ? &avg : NULL, normalize);
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
    free(res->recv_theta);
    free(res->recv_phi);
    free(res->recv_db);
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

/* The inputs the CLI builds for a plane-wave or current-source deck at
 * freq_hz (after the stamps): returns 1 (plane wave) or 2 (current source), 0
 * for a delta-gap deck, or a negative status. route 1 fills path_of and
 * free_ends (*nfree), route 0 the raw wire ranges wf / wl (*nw). Arrays hold
 * n entries (n + 1 for wf, wl, free_ends). */
int or_deck_variant_inputs(const char *text, double freq_hz, int *n_out, int *route, cx *rhs, double *cosv,
                           double *sinv, int *src, cx *i0, int *wf, int *wl, int *nw, int *path_of, int *free_ends,
                           int *nfree)
{
    char err[256];
    or_deck *d = or_deck_parse(text, err, sizeof(err));
    or_seg *segs;
    int n, i, variant = 0, np, nontriv = 0, st;
    int *fa, *fbv;
    double *sign_of, *s_of;
    const or_card *cs = NULL;
    if (!d)
        return -OR_PARSE;
    for (i = 0; i < d->n; i++)
        if (!strcmp(d->cards[i].mn, "EX")) {
            if (ex_is_plane_wave(&d->cards[i]))
                variant = 1;
            else if (ex_is_current(&d->cards[i])) {
                if (!variant)
                    variant = 2;
                if (!cs)
                    cs = &d->cards[i];
            }
        }
    segs = (or_seg *)malloc(sizeof(or_seg) * (size_t)(1 << 20));
    n = or_build_geometry(d, segs, 1 << 20, err, sizeof(err));
    if (n < 0 || !variant) {
        free(segs);
        or_deck_free(d);
        return n < 0 ? n : 0;
    }
    *n_out = n;
    fa = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    fbv = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    sign_of = (double *)malloc(sizeof(double) * (size_t)n);
    s_of = (double *)malloc(sizeof(double) * (size_t)n);
    np = or_conductor_paths(segs, n, path_of, sign_of, s_of, fa, fbv, n + 1, &nontriv);
    *nw = or_wire_endpoints(segs, n, wf, wl, n + 1);
    *route = np > 0 && nontriv;
    *nfree = 0;
    if (*route)
        for (i = 0; i < np; i++) {
            free_ends[2 * i] = fa[i];
            free_ends[2 * i + 1] = fbv[i];
            *nfree += 2;
        }
    if (variant == 1)
        st = *route ? or_planewave_hallen_paths(d, segs, n, freq_hz, path_of, sign_of, s_of, rhs, cosv, sinv)
                    : or_planewave_hallen(d, segs, n, freq_hz, rhs, cosv, sinv);
    else {
        *i0 = cx_make(ex_vr(cs), ex_vi(cs));
        st = *route ? or_current_source_shape_paths(d, segs, n, freq_hz, ex_tag(cs), ex_seg(cs), path_of, sign_of,
                                                    s_of, rhs, cosv, src)
                    : or_current_source_shape(d, segs, n, freq_hz, ex_tag(cs), ex_seg(cs), rhs, cosv, src);
    }
    free(fa);
    free(fbv);
    free(sign_of);
    free(s_of);
    free(segs);
    or_deck_free(d);
    return st ? -st : variant;
}
