/*
 * ORACLE — test infrastructure only. Nothing in the product links this code.
 *
 * The inputs of the plane-wave and current-source Hallén solves, restating
 *   crates/nec_solver/src/planewave.rs:54-119   (IncidentPlaneWave: r_hat, pol_hat)
 *   crates/nec_solver/src/planewave.rs:153-243  (build_planewave_hallen)
 *   crates/nec_solver/src/planewave.rs:271-347  (build_planewave_hallen_paths)
 *   crates/nec_solver/src/excitation.rs:350-447 (build_current_source_shape(_paths))
 * with num-complex 0.4 arithmetic in the reference's operation order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

#define PW_PI 3.14159265358979323846

/* The first plane-wave EX card: theta F1, phi F2, eta F3, axial ratio F6
 * (nec_parser lib.rs:191-245; IncidentPlaneWave::from_ex_card). */
static const or_card *first_plane_wave(const or_deck *d)
{
    int i;
    for (i = 0; i < d->n; i++)
        if (!strcmp(d->cards[i].mn, "EX") && ex_is_plane_wave(&d->cards[i]))
            return &d->cards[i];
    return NULL;
}

static double fld(const or_card *c, int i) { return c->nf > i ? c->f[i] : 0.0; }

static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* r_hat and the complex polarisation vector (planewave.rs:80-118) */
static void wave_vectors(const or_card *ex, double r_hat[3], cx pol[3])
{
    const int type = ex_type(ex);
    const double ar = type == 1 ? 0.0 : fld(ex, 9), sense = type == 2 ? 1.0 : (type == 3 ? -1.0 : 0.0);
    const double t = deg2rad(fld(ex, 4)), p = deg2rad(fld(ex, 5)), eta = deg2rad(fld(ex, 6));
    const double th_hat[3] = {cos(t) * cos(p), cos(t) * sin(p), -sin(t)};
    const double ph_hat[3] = {-sin(p), cos(p), 0.0};
    const double ce = cos(eta), se = sin(eta);
    const cx jminor = cx_make(0.0, sense * ar);
    int i;
    r_hat[0] = sin(t) * cos(p);
    r_hat[1] = sin(t) * sin(p);
    r_hat[2] = cos(t);
    for (i = 0; i < 3; i++) {
        const double major = ce * th_hat[i] + se * ph_hat[i];
        const double minor = -se * th_hat[i] + ce * ph_hat[i];
        pol[i] = cx_add(cx_make(major, 0.0), cx_scale(jminor, minor));
    }
}

/* (ê . d) E0 with E0 = 1: the Sum of pol[c] * d[c] from zero, then * e0 */
static cx coupling(const cx pol[3], const double d[3])
{
    cx acc = cx_make(0.0, 0.0);
    int c;
    for (c = 0; c < 3; c++)
        acc = cx_add(acc, cx_scale(pol[c], d[c]));
    return cx_scale(acc, 1.0);
}

int or_planewave_hallen(const or_deck *d, const or_seg *s, int n, double freq_hz, cx *rhs, double *cosv,
                        double *sinv)
{
    const or_card *ex = first_plane_wave(d);
    int *wf, *wl, nw, i, m, p, nj;
    int *ja, *jb, *wire_of;
    double *js, *s_coord, (*wdir)[3], (*wmid)[3];
    cx *e_tan;
    double r_hat[3], k, scale;
    cx pol[3];
    if (!ex)
        return OR_EXCITATION;
    wf = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    wl = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    nw = or_wire_endpoints(s, n, wf, wl, n + 1);
    ja = (int *)malloc(sizeof(int) * (size_t)(4 * n + 4));
    jb = (int *)malloc(sizeof(int) * (size_t)(4 * n + 4));
    js = (double *)malloc(sizeof(double) * (size_t)(4 * n + 4));
    nj = or_detect_junctions(s, wf, wl, nw, 1e-6, ja, jb, js, 4 * n + 4);
    free(ja);
    free(jb);
    free(js);
    if (nj > 0) {
        free(wf);
        free(wl);
        return OR_UNSUPPORTED; /* PlaneWaveError::JunctionedGeometryNotSupported */
    }
    k = 2.0 * PW_PI * freq_hz / OR_C0;
    wave_vectors(ex, r_hat, pol);
    scale = 2.0 * PW_PI / OR_ETA0;
    wire_of = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    wdir = malloc(sizeof(double[3]) * (size_t)(nw > 0 ? nw : 1));
    wmid = malloc(sizeof(double[3]) * (size_t)(nw > 0 ? nw : 1));
    for (i = 0; i < nw; i++) {
        int c;
        for (m = wf[i]; m <= wl[i]; m++)
            wire_of[m] = i;
        for (c = 0; c < 3; c++) {
            wdir[i][c] = s[wf[i]].dir[c];
            wmid[i][c] = (s[wf[i]].mid[c] + s[wl[i]].mid[c]) / 2.0;
        }
    }
    s_coord = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    e_tan = (cx *)calloc((size_t)(n > 0 ? n : 1), sizeof(cx));
    for (i = 0; i < n; i++) {
        const int w = wire_of[i];
        const double dd[3] = {s[i].mid[0] - wmid[w][0], s[i].mid[1] - wmid[w][1], s[i].mid[2] - wmid[w][2]};
        const double phase = k * dot3(r_hat, s[i].mid);
        s_coord[i] = dot3(dd, wdir[w]);
        e_tan[i] = cx_mul(coupling(pol, wdir[w]), cx_from_polar(1.0, phase));
    }
    for (m = 0; m < n; m++) {
        cx acc = cx_make(0.0, 0.0);
        for (p = 0; p < n; p++) {
            double kern;
            if (wire_of[p] != wire_of[m])
                continue;
            kern = sin(k * fabs(s_coord[m] - s_coord[p]));
            acc = cx_add(acc, cx_scale(cx_scale(e_tan[p], kern), s[p].length));
        }
        rhs[m] = cx_mul(cx_make(0.0, -scale), acc);
        cosv[m] = cos(k * s_coord[m]);
        sinv[m] = sin(k * s_coord[m]);
    }
    free(wf);
    free(wl);
    free(wire_of);
    free(wdir);
    free(wmid);
    free(s_coord);
    free(e_tan);
    return OR_OK;
}

int or_planewave_hallen_paths(const or_deck *d, const or_seg *s, int n, double freq_hz, const int *path_of,
                              const double *sign_of, const double *s_of, cx *rhs, double *cosv, double *sinv)
{
    const or_card *ex = first_plane_wave(d);
    double r_hat[3], k, scale;
    cx pol[3], *e_path;
    int m, p;
    if (!ex)
        return OR_EXCITATION;
    k = 2.0 * PW_PI * freq_hz / OR_C0;
    wave_vectors(ex, r_hat, pol);
    scale = 2.0 * PW_PI / OR_ETA0;
    e_path = (cx *)calloc((size_t)(n > 0 ? n : 1), sizeof(cx));
    for (p = 0; p < n; p++) {
        const double phase = k * dot3(r_hat, s[p].mid);
        e_path[p] = cx_scale(cx_mul(coupling(pol, s[p].dir), cx_from_polar(1.0, phase)), sign_of[p]);
    }
    for (m = 0; m < n; m++) {
        cx acc = cx_make(0.0, 0.0);
        for (p = 0; p < n; p++) {
            double kern;
            if (path_of[p] != path_of[m])
                continue;
            kern = sin(k * fabs(s_of[m] - s_of[p]));
            acc = cx_add(acc, cx_scale(cx_scale(e_path[p], kern), s[p].length));
        }
        rhs[m] = cx_scale(cx_mul(cx_make(0.0, -scale), acc), sign_of[m]);
        cosv[m] = sign_of[m] * cos(k * s_of[m]);
        sinv[m] = sign_of[m] * sin(k * s_of[m]);
    }
    free(e_path);
    return OR_OK;
}

/* The deck with its first EX card replaced by a unit delta-gap at (tag, seg)
 * and every other EX card dropped (excitation.rs:365-392). */
static or_deck *unit_source_deck(const or_deck *d, uint32_t tag, uint32_t seg)
{
    or_deck *o = (or_deck *)calloc(1, sizeof(or_deck));
    int i, placed = 0;
    o->cards = (or_card *)calloc((size_t)(d->n + 1), sizeof(or_card));
    o->cap = d->n + 1;
    for (i = 0; i < d->n; i++) {
        const or_card *c = &d->cards[i];
        if (!strcmp(c->mn, "EX")) {
            if (!placed) {
                or_card *e = &o->cards[o->n++];
                memcpy(e->mn, "EX", 3);
                e->nf = 6;
                e->f[0] = 0.0;
                e->f[1] = (double)tag;
                e->f[2] = (double)seg;
                e->f[3] = 0.0;
                e->f[4] = 1.0;
                e->f[5] = 0.0;
                placed = 1;
            }
            continue;
        }
        o->cards[o->n++] = *c;
    }
    return o;
}

static int current_source_shape(const or_deck *d, const or_seg *s, int n, double freq_hz, uint32_t tag,
                                 uint32_t seg, const int *path_of, const double *sign_of, const double *s_of,
                                 cx *shape, double *cosv, int *src)
{
    or_deck *u;
    int st;
    *src = or_find_seg(s, n, tag, seg);
    if (*src < 0)
        return OR_EXCITATION;
    u = unit_source_deck(d, tag, seg);
    st = path_of ? or_build_hallen_rhs_paths(u, s, n, freq_hz, path_of, sign_of, s_of, shape, cosv)
                 : or_build_hallen_rhs(u, s, n, freq_hz, shape, cosv);
    free(u->cards);
    free(u);
    return st;
}

int or_current_source_shape(const or_deck *d, const or_seg *s, int n, double freq_hz, uint32_t tag,
                            uint32_t seg, cx *shape, double *cosv, int *src)
{
    return current_source_shape(d, s, n, freq_hz, tag, seg, NULL, NULL, NULL, shape, cosv, src);
}

int or_current_source_shape_paths(const or_deck *d, const or_seg *s, int n, double freq_hz, uint32_t tag,
                                  uint32_t seg, const int *path_of, const double *sign_of, const double *s_of,
                                  cx *shape, double *cosv, int *src)
{
    return current_source_shape(d, s, n, freq_hz, tag, seg, path_of, sign_of, s_of, shape, cosv, src);
}
