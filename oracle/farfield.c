/*
 * ORACLE — test infrastructure only.
 *
 * Far field, restating crates/nec_solver/src/farfield.rs:
 *   pec_image_farfield / unit_vectors          :64-94
 *   far_field_components                       :265-291
 *   radiation_efficiency                       :308-322
 *   fresnel_coeffs / far_field_components_fresnel :333-388
 *   integrate_power_for_ground                 :422-464
 *   far_field_components_with_image            :514-546
 *   compute_radiation_pattern                  :563-647
 *   polarization_axial_ratio                   :659-667
 */
#include <math.h>

#include "oracle_internal.h"

static void unit_vectors(double th_deg, double ph_deg, double r[3], double t[3], double p[3])
{
    double th = deg2rad(th_deg), ph = deg2rad(ph_deg);
    double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
    r[0] = st * cp; r[1] = st * sp; r[2] = ct;
    t[0] = ct * cp; t[1] = ct * sp; t[2] = -st;
    p[0] = -sp; p[1] = cp; p[2] = 0.0;
}

static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void fresnel_coeffs(double th_deg, double eps_r, double sigma, double k, cx *gv, cx *gh)
{
    double omega = k * OR_C0;
    cx eps_c = cx_make(eps_r, -sigma / (omega * OR_EPS0_FF));
    double th = deg2rad(th_deg);
    cx cos_i = cx_make(cos(th), 0.0);
    double sin2 = sin(th) * sin(th);
    cx root = cx_sqrt(cx_sub(eps_c, cx_make(sin2, 0.0)));
    cx ec = cx_mul(eps_c, cos_i);
    *gv = cx_div(cx_sub(ec, root), cx_add(ec, root));
    *gh = cx_div(cx_sub(cos_i, root), cx_add(cos_i, root));
}

void or_far_field(const or_seg *s, int n, const cx *cur, double k, or_ground g, double th,
                  double ph, cx *ft_out, cx *fp_out)
{
    double r[3], t[3], p[3];
    cx ft = cx_make(0.0, 0.0), fp = cx_make(0.0, 0.0), gv = cx_make(0, 0), gh = cx_make(0, 0);
    int i;
    unit_vectors(th, ph, r, t, p);
    if (g.kind == OR_GROUND_FINITE)
        fresnel_coeffs(th, g.eps_r, g.sigma, k, &gv, &gh);
    for (i = 0; i < n; i++) {
        cx in = cur[i];
        double pa = k * dot3(s[i].mid, r);
        cx phase = cx_make(cos(pa), sin(pa));
        cx w = cx_mul(in, cx_scale(phase, s[i].length));
        ft = cx_add(ft, cx_scale(w, dot3(s[i].dir, t)));
        fp = cx_add(fp, cx_scale(w, dot3(s[i].dir, p)));
        if (g.kind == OR_GROUND_PEC || g.kind == OR_GROUND_FINITE) {
            double im_mid[3], im_dir[3], ipa;
            cx iphase, iw;
            im_mid[0] = s[i].mid[0]; im_mid[1] = s[i].mid[1]; im_mid[2] = -s[i].mid[2];
            im_dir[0] = -s[i].dir[0]; im_dir[1] = -s[i].dir[1]; im_dir[2] = s[i].dir[2];
            ipa = k * dot3(im_mid, r);
            iphase = cx_make(cos(ipa), sin(ipa));
            iw = cx_mul(in, cx_scale(iphase, s[i].length));
            if (g.kind == OR_GROUND_PEC) {
                ft = cx_add(ft, cx_scale(iw, dot3(im_dir, t)));
                fp = cx_add(fp, cx_scale(iw, dot3(im_dir, p)));
            } else {
                ft = cx_add(ft, cx_mul(gv, cx_scale(iw, dot3(im_dir, t))));
                fp = cx_sub(fp, cx_mul(gh, cx_scale(iw, dot3(im_dir, p))));
            }
        }
    }
    *ft_out = ft;
    *fp_out = fp;
}

double or_power_integral(const or_seg *s, int n, const cx *cur, double k, or_ground g)
{
    int has_ground = g.kind == OR_GROUND_PEC || g.kind == OR_GROUND_FINITE;
    int n_theta = has_ground ? 91 : 181, it, ip;
    double tmax = has_ground ? 90.0 : 180.0;
    double d_theta = deg2rad(tmax) / ((double)n_theta - 1.0);
    double d_phi = 2.0 * OR_PI / 360.0;
    double total = 0.0;
    for (it = 0; it < n_theta; it++) {
        double th = (double)it * tmax / ((double)n_theta - 1.0);
        double st = sin(deg2rad(th));
        double wt;
        if (st == 0.0)
            continue;
        wt = (it == 0 || it == n_theta - 1) ? 0.5 * d_theta : d_theta;
        for (ip = 0; ip < 360; ip++) {
            cx ft, fp;
            or_far_field(s, n, cur, k, g, th, (double)ip, &ft, &fp);
            total += (cx_norm_sqr(ft) + cx_norm_sqr(fp)) * st * wt * d_phi;
        }
    }
    return total;
}

double or_axial_ratio(cx et, cx ep)
{
    double s0 = cx_norm_sqr(et) + cx_norm_sqr(ep), s3, q, chi;
    if (s0 <= 1e-30)
        return 0.0;
    s3 = 2.0 * cx_mul(et, cx_conj(ep)).im;
    q = s3 / s0;
    if (q < -1.0) q = -1.0;
    if (q > 1.0) q = 1.0;
    chi = 0.5 * asin(q);
    return tan(chi);
}

double or_radiation_pattern(const or_seg *s, int n, const cx *cur, double freq_hz, or_ground g,
                            const double *theta, const double *phi, int np, double *out)
{
    double lambda = OR_C0 / freq_hz;
    double k = 2.0 * OR_PI / lambda;
    int has_ground = g.kind == OR_GROUND_PEC || g.kind == OR_GROUND_FINITE;
    double total = or_power_integral(s, n, cur, k, g);
    int i;
    for (i = 0; i < np; i++) {
        double *o = out + 4 * (size_t)i;
        cx ft, fp;
        double ut, uth, uph, norm;
        if (has_ground && theta[i] > 90.0) {
            o[0] = o[1] = o[2] = -999.99;
            o[3] = 0.0;
            continue;
        }
        or_far_field(s, n, cur, k, g, theta[i], phi[i], &ft, &fp);
        ut = cx_norm_sqr(ft) + cx_norm_sqr(fp);
        uth = cx_norm_sqr(ft);
        uph = cx_norm_sqr(fp);
        norm = total > 0.0 ? 4.0 * OR_PI / total : 0.0;
        o[0] = ut * norm > 1e-20 ? 10.0 * log10(ut * norm) : -999.99;
        o[1] = uth * norm > 1e-20 ? 10.0 * log10(uth * norm) : -999.99;
        o[2] = uph * norm > 1e-20 ? 10.0 * log10(uph * norm) : -999.99;
        o[3] = or_axial_ratio(ft, fp);
    }
    return total;
}

double or_radiation_efficiency(const or_seg *s, int n, const cx *cur, double freq_hz,
                               or_ground g, double p_in)
{
    double k, integral, p_rad, e;
    if (p_in <= 0.0)
        return 1.0;
    k = 2.0 * OR_PI * freq_hz / OR_C0;
    integral = or_power_integral(s, n, cur, k, g);
    p_rad = k * k * OR_ETA0 / (32.0 * OR_PI * OR_PI) * integral;
    e = p_rad / p_in;
    if (e < 0.0) e = 0.0;
    if (e > 1.0) e = 1.0;
    return e;
}
