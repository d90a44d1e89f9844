/*
 * ORACLE — test infrastructure only.
 *
 * Geometry, restating crates/nec_solver/src/geometry.rs:
 *   ground_model_from_deck      :71-111
 *   wire_endpoints_from_segs    :168-185
 *   merge_collinear_wire_endpoints :200-241
 *   detect_wire_junctions       :274-339
 *   build_conductor_paths       :404-584
 *   build_geometry / expand_wire / apply_gm / apply_gr / rotations :676-879
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

int or_find_seg(const or_seg *s, int n, uint32_t tag, uint32_t idx)
{
    int i;
    for (i = 0; i < n; i++)
        if (s[i].tag == tag && s[i].tag_index == idx)
            return i;
    return -1;
}

or_ground or_ground_model(const or_deck *d)
{
    or_ground g = {OR_GROUND_FREE, 0.0, 0.0};
    int i;
    for (i = 0; i < d->n; i++) {
        const or_card *c = &d->cards[i];
        if (strcmp(c->mn, "GN"))
            continue;
        {
            int t = or_as_i32(c->f[0]);
            if (t == -1)
                g.kind = OR_GROUND_FREE;
            else if (t == 1)
                g.kind = OR_GROUND_PEC;
            else if (t == 0 || t == 2) {
                g.kind = OR_GROUND_FINITE;
                g.eps_r = c->nf > 4 ? c->f[4] : 13.0;
                g.sigma = c->nf > 5 ? c->f[5] : 0.005;
            } else
                g.kind = OR_GROUND_DEFERRED;
        }
        return g;
    }
    for (i = 0; i < d->n; i++) {
        const or_card *c = &d->cards[i];
        if (!strcmp(c->mn, "GE") && c->nf > 0 && or_as_i32(c->f[0]) == 1) {
            g.kind = OR_GROUND_PEC;
            return g;
        }
    }
    return g;
}

static void rot_x(double p[3], double c, double s)
{
    double y = c * p[1] - s * p[2], z = s * p[1] + c * p[2];
    p[1] = y;
    p[2] = z;
}
static void rot_y(double p[3], double c, double s)
{
    double x = c * p[0] + s * p[2], z = -s * p[0] + c * p[2];
    p[0] = x;
    p[2] = z;
}
static void rot_z(double p[3], double c, double s)
{
    double x = c * p[0] - s * p[1], y = s * p[0] + c * p[1];
    p[0] = x;
    p[1] = y;
}

static void recompute_dir(or_seg *g)
{
    double dx = g->end[0] - g->start[0], dy = g->end[1] - g->start[1], dz = g->end[2] - g->start[2];
    double len = sqrt(dx * dx + dy * dy + dz * dz);
    if (len == 0.0) {
        g->dir[0] = 0.0;
        g->dir[1] = 0.0;
        g->dir[2] = 1.0;
    } else {
        g->dir[0] = dx / len;
        g->dir[1] = dy / len;
        g->dir[2] = dz / len;
    }
}

static void gm_point(double p[3], const or_card *c)
{
    rot_x(p, cos(deg2rad(c->f[2])), sin(deg2rad(c->f[2])));
    rot_y(p, cos(deg2rad(c->f[3])), sin(deg2rad(c->f[3])));
    rot_z(p, cos(deg2rad(c->f[4])), sin(deg2rad(c->f[4])));
    p[0] = p[0] + c->f[5];
    p[1] = p[1] + c->f[6];
    p[2] = p[2] + c->f[7];
}

static int tag_in_range(uint32_t tag, uint32_t first, uint32_t last)
{
    return (first == 0 || tag >= first) && (last == 0 || tag <= last);
}

typedef struct {
    or_seg *v;
    int n, cap, max;
} segvec;

static int push(segvec *sv, const or_seg *g)
{
    if (sv->n >= sv->max)
        return -1;
    sv->v[sv->n++] = *g;
    return 0;
}

int or_build_geometry(const or_deck *d, or_seg *out, int max, char *err, int errlen)
{
    segvec sv = {out, 0, 0, max};
    int ci;
    for (ci = 0; ci < d->n; ci++) {
        const or_card *c = &d->cards[ci];
        if (!strcmp(c->mn, "GW")) {
            uint32_t tag = or_as_u32(c->f[0]), ns = or_as_u32(c->f[1]), i;
            double x1 = c->f[2], y1 = c->f[3], z1 = c->f[4];
            double wdx = c->f[5] - x1, wdy = c->f[6] - y1, wdz = c->f[7] - z1;
            double wl, nn, seglen, dir[3];
            if (ns == 0) {
                snprintf(err, errlen, "GW tag %u: segment count must be >= 1", tag);
                return -OR_GEOMETRY;
            }
            wl = sqrt(wdx * wdx + wdy * wdy + wdz * wdz);
            if (wl == 0.0) {
                snprintf(err, errlen, "GW tag %u: wire has zero length (start == end)", tag);
                return -OR_GEOMETRY;
            }
            dir[0] = wdx / wl;
            dir[1] = wdy / wl;
            dir[2] = wdz / wl;
            nn = (double)ns;
            seglen = wl / nn;
            for (i = 0; i < ns; i++) {
                or_seg g;
                double t0 = (double)i / nn, t1 = ((double)i + 1.0) / nn, tm = (t0 + t1) * 0.5;
                g.tag = tag;
                g.tag_index = i + 1;
                g.start[0] = x1 + wdx * t0; g.start[1] = y1 + wdy * t0; g.start[2] = z1 + wdz * t0;
                g.end[0] = x1 + wdx * t1;   g.end[1] = y1 + wdy * t1;   g.end[2] = z1 + wdz * t1;
                g.mid[0] = x1 + wdx * tm;   g.mid[1] = y1 + wdy * tm;   g.mid[2] = z1 + wdz * tm;
                memcpy(g.dir, dir, sizeof(dir));
                g.length = seglen;
                g.radius = c->f[8];
                if (push(&sv, &g)) {
                    snprintf(err, errlen, "too many segments");
                    return -OR_GEOMETRY;
                }
            }
        } else if (!strcmp(c->mn, "GM")) {
            uint32_t inc = or_as_u32(c->f[0]), last = or_as_u32(c->f[1]), first = or_as_u32(c->f[8]);
            int i, n0 = sv.n;
            if (inc == 0) {
                for (i = 0; i < n0; i++) {
                    or_seg *g = &sv.v[i];
                    if (!tag_in_range(g->tag, first, last))
                        continue;
                    gm_point(g->start, c);
                    gm_point(g->end, c);
                    gm_point(g->mid, c);
                    recompute_dir(g);
                }
            } else {
                for (i = 0; i < n0; i++) {
                    or_seg g = sv.v[i];
                    if (!tag_in_range(g.tag, first, last))
                        continue;
                    g.tag += inc;
                    gm_point(g.start, c);
                    gm_point(g.end, c);
                    gm_point(g.mid, c);
                    recompute_dir(&g);
                    if (push(&sv, &g)) {
                        snprintf(err, errlen, "too many segments");
                        return -OR_GEOMETRY;
                    }
                }
            }
        } else if (!strcmp(c->mn, "GR")) {
            uint32_t inc = or_as_u32(c->f[0]), count = or_as_u32(c->f[1]), k;
            int i, n0 = sv.n;
            for (k = 1; k <= count; k++) {
                double ang = c->f[2] * (double)k;
                double ca = cos(deg2rad(ang)), sa = sin(deg2rad(ang));
                for (i = 0; i < n0; i++) {
                    or_seg g = sv.v[i];
                    g.tag += inc * k;
                    rot_z(g.start, ca, sa);
                    rot_z(g.end, ca, sa);
                    rot_z(g.mid, ca, sa);
                    recompute_dir(&g);
                    if (push(&sv, &g)) {
                        snprintf(err, errlen, "too many segments");
                        return -OR_GEOMETRY;
                    }
                }
            }
        }
    }
    if (sv.n == 0) {
        snprintf(err, errlen, "deck contains no GW (wire) cards");
        return -OR_GEOMETRY;
    }
    return sv.n;
}

int or_wire_endpoints(const or_seg *s, int n, int *first, int *last, int max)
{
    int i, cnt = 0, f = 0;
    uint32_t cur = 0xFFFFFFFFu;
    for (i = 0; i < n; i++) {
        if (s[i].tag != cur) {
            if (cur != 0xFFFFFFFFu) {
                if (cnt < max) { first[cnt] = f; last[cnt] = i - 1; }
                cnt++;
            }
            cur = s[i].tag;
            f = i;
        }
    }
    if (cur != 0xFFFFFFFFu) {
        if (cnt < max) { first[cnt] = f; last[cnt] = n - 1; }
        cnt++;
    }
    return cnt;
}

static double dist2(const double a[3], const double b[3])
{
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}

static int dir_aligned(const double a[3], const double b[3], double tol)
{
    double cr[3], z[3] = {0.0, 0.0, 0.0};
    double dot;
    cr[0] = a[1] * b[2] - a[2] * b[1];
    cr[1] = a[2] * b[0] - a[0] * b[2];
    cr[2] = a[0] * b[1] - a[1] * b[0];
    dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    return dist2(cr, z) < tol * tol && dot > 0.0;
}

int or_merge_collinear(const or_seg *s, int n, int *first, int *last, int max)
{
    int nb, i, cnt = 0;
    int *bf = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *bl = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int cf, cl;
    const double tol = 1e-6;
    nb = or_wire_endpoints(s, n, bf, bl, n + 1);
    if (nb < 2) {
        for (i = 0; i < nb && i < max; i++) { first[i] = bf[i]; last[i] = bl[i]; }
        free(bf);
        free(bl);
        return nb;
    }
    cf = bf[0];
    cl = bl[0];
    for (i = 1; i < nb; i++) {
        const or_seg *a = &s[cl], *b = &s[bf[i]];
        int contiguous = cl + 1 == bf[i];
        int connects = dist2(a->end, b->start) < tol * tol;
        int collinear = dir_aligned(a->dir, b->dir, tol);
        int same_r = fabs(a->radius - b->radius) < tol;
        if (contiguous && connects && collinear && same_r) {
            cl = bl[i];
        } else {
            if (cnt < max) { first[cnt] = cf; last[cnt] = cl; }
            cnt++;
            cf = bf[i];
            cl = bl[i];
        }
    }
    if (cnt < max) { first[cnt] = cf; last[cnt] = cl; }
    cnt++;
    free(bf);
    free(bl);
    return cnt;
}

int or_detect_junctions(const or_seg *s, const int *first, const int *last, int nw, double tol,
                        int *ja, int *jb, double *jsign, int max)
{
    int i, j, cnt = 0, nn = 2 * nw;
    for (i = 0; i < nn; i++) {
        int si = (i & 1) ? last[i / 2] : first[i / 2];
        int st_i = !(i & 1);
        const double *pi = st_i ? s[si].start : s[si].end;
        for (j = i + 1; j < nn; j++) {
            int sj = (j & 1) ? last[j / 2] : first[j / 2];
            int st_j = !(j & 1);
            const double *pj = st_j ? s[sj].start : s[sj].end;
            double dx, dy, dz;
            if (si == sj)
                continue;
            dx = pi[0] - pj[0];
            dy = pi[1] - pj[1];
            dz = pi[2] - pj[2];
            if (sqrt(dx * dx + dy * dy + dz * dz) > tol)
                continue;
            if (cnt < max) {
                ja[cnt] = si;
                jb[cnt] = sj;
                jsign[cnt] = (st_i == st_j) ? 1.0 : -1.0;
            }
            cnt++;
        }
    }
    return cnt;
}

static int uf_find(int *parent, int x)
{
    int r = x, c = x;
    while (parent[r] != r)
        r = parent[r];
    while (parent[c] != c) {
        int nx = parent[c];
        parent[c] = r;
        c = nx;
    }
    return r;
}

int or_conductor_paths(const or_seg *s, int n, int *path_of, double *sign_of, double *s_of,
                       int *free_a, int *free_b, int max_paths, int *nontrivial)
{
    int w, i, np = 0, result = 0;
    int *wf = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *wl = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    int *node_of, *node_cnt, *node_mem, nnodes = 0, *parent, *visited, *order, *seen_root;
    const double tol = 1e-6;
    *nontrivial = 0;
    w = or_wire_endpoints(s, n, wf, wl, n + 1);
    if (w == 0) {
        free(wf);
        free(wl);
        return 0;
    }
    node_of = (int *)malloc(sizeof(int) * (size_t)(2 * w));
    node_cnt = (int *)calloc((size_t)(2 * w), sizeof(int));
    node_mem = (int *)malloc(sizeof(int) * (size_t)(2 * w) * (size_t)(2 * w));
    parent = (int *)malloc(sizeof(int) * (size_t)w);
    visited = (int *)calloc((size_t)w, sizeof(int));
    order = (int *)malloc(sizeof(int) * (size_t)n);
    seen_root = (int *)calloc((size_t)w, sizeof(int));
    for (i = 0; i < 2 * w; i++)
        node_of[i] = -1;
    /* Cluster the 2W endpoints into coincident nodes (geometry.rs:424-453). */
    for (i = 0; i < 2 * w; i++) {
        int wi = i / 2, end = i & 1, j;
        const double *p;
        if (node_of[i] >= 0)
            continue;
        p = end == 0 ? s[wf[wi]].start : s[wl[wi]].end;
        node_of[i] = nnodes;
        node_mem[nnodes * 2 * w + node_cnt[nnodes]++] = i;
        for (j = 0; j < 2 * w; j++) {
            int wj = j / 2, endj = j & 1;
            const double *q;
            if (node_of[j] >= 0)
                continue;
            q = endj == 0 ? s[wf[wj]].start : s[wl[wj]].end;
            if (dist2(p, q) < tol * tol) {
                node_of[j] = nnodes;
                node_mem[nnodes * 2 * w + node_cnt[nnodes]++] = j;
            }
        }
        nnodes++;
    }
    for (i = 0; i < nnodes; i++)
        if (node_cnt[i] >= 3) {
            result = -1;
            goto done;
        }
    for (i = 0; i < w; i++)
        parent[i] = i;
    for (i = 0; i < nnodes; i++) {
        if (node_cnt[i] == 2) {
            int a = node_mem[i * 2 * w] / 2, b = node_mem[i * 2 * w + 1] / 2;
            int ra, rb;
            if (a == b) {
                result = -1;
                goto done;
            }
            ra = uf_find(parent, a);
            rb = uf_find(parent, b);
            parent[ra] = rb;
        }
    }
    /* Components in ascending root order (BTreeMap iteration). */
    {
        int root;
        for (root = 0; root < w; root++) {
            int fe[4], nfe = 0, wi;
            int cur, enter, nord = 0;
            double arc = 0.0;
            int any = 0;
            for (wi = 0; wi < w; wi++)
                if (uf_find(parent, wi) == root) {
                    int end;
                    any = 1;
                    for (end = 0; end < 2; end++)
                        if (node_cnt[node_of[wi * 2 + end]] == 1) {
                            if (nfe < 4)
                                fe[nfe] = wi * 2 + end;
                            nfe++;
                        }
                }
            if (!any)
                continue;
            if (nfe != 2) {
                result = -1;
                goto done;
            }
            (void)seen_root;
            cur = fe[0] / 2;
            enter = fe[0] & 1;
            for (wi = 0; wi < w; wi++)
                visited[wi] = 0;
            for (;;) {
                int first = wf[cur], last = wl[cur], fwd = enter == 0, m, exit_end, exit_node;
                int nxt_w = -1, nxt_e = -1, k;
                double sg = fwd ? 1.0 : -1.0;
                visited[cur] = 1;
                for (k = 0; k <= last - first; k++) {
                    m = fwd ? first + k : last - k;
                    {
                        double half = s[m].length * 0.5;
                        s_of[m] = arc + half; /* u_mid; re-centred below */
                        arc += s[m].length;
                    }
                    order[nord++] = m;
                    path_of[m] = np;
                    sign_of[m] = sg;
                }
                exit_end = 1 - enter;
                exit_node = node_of[cur * 2 + exit_end];
                if (node_cnt[exit_node] == 1)
                    break;
                for (k = 0; k < node_cnt[exit_node]; k++) {
                    int mem = node_mem[exit_node * 2 * w + k];
                    if (mem / 2 != cur && !visited[mem / 2]) {
                        nxt_w = mem / 2;
                        nxt_e = mem & 1;
                    }
                }
                if (nxt_w < 0)
                    break;
                cur = nxt_w;
                enter = nxt_e;
            }
            for (i = 0; i < nord; i++)
                s_of[order[i]] = s_of[order[i]] - arc * 0.5;
            /* ConductorPath::is_trivial (geometry.rs:377-382). */
            for (i = 0; i < nord; i++) {
                if (sign_of[order[i]] != 1.0)
                    *nontrivial = 1;
                if (i > 0 && order[i] != order[i - 1] + 1)
                    *nontrivial = 1;
            }
            if (np < max_paths) {
                free_a[np] = order[0];
                free_b[np] = order[nord - 1];
            }
            np++;
        }
    }
    result = np;
done:
    free(wf);
    free(wl);
    free(node_of);
    free(node_cnt);
    free(node_mem);
    free(parent);
    free(visited);
    free(order);
    free(seen_root);
    return result;
}
