/*
 * ORACLE — test infrastructure only.
 *
 * Hallén A-matrix fill, restating crates/nec_solver/src/matrix.rs:
 *   constants and Gauss-Legendre tables  :31-72 (literals used verbatim)
 *   assemble_z_matrix_with_ground        :186-218
 *   image_segment                        :265-277
 *   fresnel_reflection_scalar            :285-290
 *   elem                                 :297-331
 *   green_k                              :397-399
 */
#include <math.h>

#include "oracle_internal.h"

static const double GL_N4[4] = {-0.861136311594953, -0.339981043584856, 0.339981043584856,
                                0.861136311594953};
static const double GL_W4[4] = {0.347854845137454, 0.652145154862626, 0.652145154862626,
                                0.347854845137454};
static const double GL_N8[8] = {-0.960289856497536, -0.796666477413627, -0.525532409916329,
                                -0.183434642495650, 0.183434642495650,  0.525532409916329,
                                0.796666477413627,  0.960289856497536};
static const double GL_W8[8] = {0.101228536290376, 0.222381034453374, 0.313706645877887,
                                0.362683783378362, 0.362683783378362, 0.313706645877887,
                                0.222381034453374, 0.101228536290376};

/* exp(-jkR)/R via num-complex: exp((0, -k*r)) = (cos(-k r), sin(-k r)); then / r. */
static cx green_k(double r, double k)
{
    double a = -k * r;
    cx e = cx_make(1.0 * cos(a), 1.0 * sin(a));
    return cx_divs(e, r);
}

cx or_fill_elem(const or_seg *obs, const or_seg *src, double k, int is_self)
{
    double cos_alpha = obs->dir[0] * src->dir[0] + obs->dir[1] * src->dir[1] + obs->dir[2] * src->dir[2];
    double half = src->length * 0.5;
    double a = src->radius;
    cx int_k;
    int m;
    if (is_self) {
        cx smooth = cx_make(0.0, 0.0);
        double r_end, analytic;
        for (m = 0; m < 4; m++) {
            double l = GL_N4[m] * half;
            double r_eff = sqrt(l * l + a * a);
            cx g = green_k(r_eff, k);
            cx dyn = cx_make(g.re - 1.0 / r_eff, g.im);
            smooth = cx_add(smooth, cx_scale(dyn, GL_W4[m]));
        }
        smooth = cx_scale(smooth, half);
        r_end = sqrt(half * half + a * a);
        analytic = 2.0 * log((half + r_end) / a);
        int_k = cx_make(smooth.re + analytic, smooth.im);
    } else {
        cx sum = cx_make(0.0, 0.0);
        for (m = 0; m < 8; m++) {
            double sc = GL_N8[m] * half;
            double rs0 = src->mid[0] + sc * src->dir[0];
            double rs1 = src->mid[1] + sc * src->dir[1];
            double rs2 = src->mid[2] + sc * src->dir[2];
            double p0 = obs->mid[0] - rs0, p1 = obs->mid[1] - rs1, p2 = obs->mid[2] - rs2;
            double r_sq = p0 * p0 + p1 * p1 + p2 * p2;
            double r_eff = sqrt(r_sq + a * a);
            sum = cx_add(sum, cx_scale(green_k(r_eff, k), GL_W8[m]));
        }
        int_k = cx_scale(sum, half);
    }
    return cx_scale(int_k, cos_alpha);
}

static or_seg image_segment(const or_seg *g)
{
    or_seg im = *g;
    im.start[2] = -g->start[2];
    im.end[2] = -g->end[2];
    im.mid[2] = -g->mid[2];
    im.dir[0] = -g->dir[0];
    im.dir[1] = -g->dir[1];
    im.dir[2] = g->dir[2];
    return im;
}

static cx fresnel_reflection_scalar(double freq_hz, double eps_r, double sigma)
{
    double omega = 2.0 * OR_PI * freq_hz;
    double er = eps_r > 1.0e-6 ? eps_r : 1.0e-6;
    double sg = sigma > 0.0 ? sigma : 0.0;
    cx eps_c, sq;
    /* f64::max ignores NaN: max(NaN, x) = x */
    if (isnan(eps_r)) er = 1.0e-6;
    if (isnan(sigma)) sg = 0.0;
    eps_c = cx_make(er, -sg / (omega * OR_EPS0_MATRIX));
    sq = cx_sqrt(eps_c);
    return cx_div(cx_sub(sq, cx_make(1.0, 0.0)), cx_add(sq, cx_make(1.0, 0.0)));
}

void or_fill_z(const or_seg *s, int n, double freq_hz, or_ground g, cx *z)
{
    double k = 2.0 * OR_PI * freq_hz / OR_C0;
    int i;
    cx gamma = cx_make(0.0, 0.0);
    if (g.kind == OR_GROUND_FINITE)
        gamma = fresnel_reflection_scalar(freq_hz, g.eps_r, g.sigma);
#pragma omp parallel for num_threads(or_threads) schedule(dynamic, 4) if (or_threads > 1)
    for (i = 0; i < n; i++) {
        int j;
        for (j = 0; j < n; j++) {
            cx direct = or_fill_elem(&s[i], &s[j], k, i == j);
            cx img = cx_make(0.0, 0.0);
            if (g.kind == OR_GROUND_PEC) {
                or_seg im = image_segment(&s[j]);
                img = or_fill_elem(&s[i], &im, k, 0);
            } else if (g.kind == OR_GROUND_FINITE) {
                or_seg im = image_segment(&s[j]);
                img = cx_mul(gamma, or_fill_elem(&s[i], &im, k, 0));
            }
            z[(size_t)i * (size_t)n + (size_t)j] = cx_add(direct, img);
        }
    }
}
