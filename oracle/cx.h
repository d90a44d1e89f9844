/*
 * ORACLE — test infrastructure only. Nothing in the product links this code.
 *
 * Complex-double arithmetic with the exact operation order of the Rust crate
 * num-complex 0.4.6 (the reference's pinned dependency, Cargo.lock:2459-2461;
 * not vendored in the reference tree). Each operation below restates that
 * crate's published definition:
 *   a*b   = (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re)
 *   a/b   = ((a.re*b.re + a.im*b.im)/|b|^2, (a.im*b.re - a.re*b.im)/|b|^2)   (naive)
 *   a*s   = (a.re*s, a.im*s);   a/s = (a.re/s, a.im/s)
 *   norm  = hypot(re, im);      norm_sqr = re*re + im*im
 *   exp   = from_polar(exp(re), im) = (e*cos(im), e*sin(im))
 *   sqrt  = special cases for im==0 / re==0, else from_polar(sqrt(r), theta/2)
 *   arg   = atan2(im, re)
 * The oracle is compiled with -ffp-contract=off so no multiply-add is fused,
 * matching Rust's codegen (it never contracts to FMA).
 */
#ifndef FNEC_ORACLE_CX_H
#define FNEC_ORACLE_CX_H

#include <math.h>

typedef struct {
    double re, im;
} cx;

static inline cx cx_make(double re, double im) { cx z; z.re = re; z.im = im; return z; }
static inline cx cx_add(cx a, cx b) { return cx_make(a.re + b.re, a.im + b.im); }
static inline cx cx_sub(cx a, cx b) { return cx_make(a.re - b.re, a.im - b.im); }
static inline cx cx_mul(cx a, cx b)
{
    return cx_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cx cx_div(cx a, cx b)
{
    double ns = b.re * b.re + b.im * b.im;
    return cx_make((a.re * b.re + a.im * b.im) / ns, (a.im * b.re - a.re * b.im) / ns);
}
static inline cx cx_scale(cx a, double s) { return cx_make(a.re * s, a.im * s); }
static inline cx cx_divs(cx a, double s) { return cx_make(a.re / s, a.im / s); }
static inline cx cx_conj(cx a) { return cx_make(a.re, -a.im); }
static inline cx cx_neg(cx a) { return cx_make(-a.re, -a.im); }
static inline double cx_norm(cx a) { return hypot(a.re, a.im); }
static inline double cx_norm_sqr(cx a) { return a.re * a.re + a.im * a.im; }
static inline double cx_arg(cx a) { return atan2(a.im, a.re); }
static inline cx cx_from_polar(double r, double th) { return cx_make(r * cos(th), r * sin(th)); }

/* exp of a finite complex number (the corner cases for infinite/NaN real
 * parts are not reachable from the Hallén path). */
static inline cx cx_exp(cx a) { return cx_from_polar(exp(a.re), a.im); }

static inline cx cx_sqrt(cx a)
{
    if (a.im == 0.0) {
        if (!signbit(a.re))
            return cx_make(sqrt(a.re), a.im);
        {
            double im = sqrt(-a.re);
            return signbit(a.im) ? cx_make(0.0, -im) : cx_make(0.0, im);
        }
    }
    if (a.re == 0.0) {
        double x = sqrt(fabs(a.im) / 2.0);
        return signbit(a.im) ? cx_make(x, -x) : cx_make(x, x);
    }
    {
        double r = cx_norm(a), th = cx_arg(a);
        return cx_from_polar(sqrt(r), th / 2.0);
    }
}

/* Rust f64::to_radians / to_degrees. */
static inline double deg2rad(double d) { return d * (3.14159265358979323846 / 180.0); }
static inline double rad2deg(double r) { return r * (180.0 / 3.14159265358979323846); }

#endif
