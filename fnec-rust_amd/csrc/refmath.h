// Elementary functions of the plain FP64 path, bitwise equal to the host libm
// the reference runs on (glibc 2.35, x86-64 with FMA), so that the device
// fill, pivot search and elimination round exactly as the reference's CPU
// path does (nec_solver matrix.rs:297-331,397-399 call f64::sin / cos /
// sqrt; linear.rs:1345-1356 calls Complex64::norm = f64::hypot).
//
//   rm_sin, rm_cos,  the IBM Accurate Mathematical Library's table-driven
//                    method (Ziv et al.; glibc dbl-64 s_sin.c): |x| < 0.126
//                    by a degree-11 odd polynomial with a correction term;
//                    |x| < 0.855469 by a table of sin/cos(i/128) as
//                    double-double pairs and short polynomials in the
//                    remainder; up to 2.426265 by the same after subtracting
//                    pi/2 in two parts; up to 105414350 after a Cody-Waite
//   rm_sincos        reduction by pi/2 in four parts, carried as a
//                    double-double. The libm ships this method twice: sin and
//                    cos are dispatched to a build that fuses multiply-adds
//                    (the variant the x86-64 libm selects on an FMA-capable
//                    CPU), sincos runs a build that rounds every product and
//                    sum on its own; the two differ in the last bit on about
//                    0.1% of arguments. rm_sin / rm_cos fuse exactly where
//                    the first does; rm_sincos fuses nowhere. A sin and a cos
//                    of one operand in the same basic block are one sincos
//                    call in the reference's codegen (LLVM turns the pair into
//                    the sincos libcall on GNU targets), so the fill's
//                    exp(-jkR) takes rm_sincos. Beyond 105414350 (never a
//                    phase of this path) the library routine is called.
//   rm_hypot         Borges' corrected hypot ("An improved algorithm for
//                    hypot(a, b)", 2019, the form without FMA), with the
//                    scaling of huge and tiny arguments.
//   rm_div           complex division as num-complex 0.4 defines it (naive).
//
// These functions are written once for the device kernels and for a host
// build that tests them against the host libm on CPU-only machines. Every
// file that includes this header is compiled with -ffp-contract=off.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "sincos_table.h"

#if defined(__HIPCC__)
#define FNEC_HD __host__ __device__ __forceinline__
#else
#define FNEC_HD static inline
#endif

namespace fnec {

FNEC_HD double rm_bits2d(uint64_t u)
{
    double d;
    memcpy(&d, &u, 8);
    return d;
}
FNEC_HD uint64_t rm_d2bits(double d)
{
    uint64_t u;
    memcpy(&u, &d, 8);
    return u;
}

// Polynomials: sin(x) ~ x + x^3 (s1 + x^2 (s2 + ...)) near 0; x + x^3 (sn3 +
// x^2 sn5) and x^2 (cs2 + x^2 (cs4 + x^2 cs6)) on the table's remainder
// interval |r| <= 1/256.
#define RM_S1 (-0x1.5555555555555p-3)
#define RM_S2 0.0083333333333323288
#define RM_S3 (-0.00019841269834414642)
#define RM_S4 2.755729806860771e-06
#define RM_S5 (-2.5022014848318398e-08)
#define RM_SN3 (-1.66666666666664880952546298448555E-01)
#define RM_SN5 8.33333214285722277379541354343671E-03
#define RM_CS2 4.99999999999999999999950396842453E-01
#define RM_CS4 (-4.16666666666664434524222570944589E-02)
#define RM_CS6 1.38888874007937613028114285595617E-03
// 1.5 * 2^45: adding it rounds |x| to a multiple of 1/128 (the table index
// is then in the low word); 1.5 * 2^52 rounds to an integer.
#define RM_BIG 0x1.8p45
#define RM_TOINT 0x1.8p52
// pi/2 = HP0 + HP1 (double-double) and = MP1 + MP2 + PP3 + PP4 (Cody-Waite
// parts: MP1, MP2, PP3 carry few enough bits that n MP1, n MP2 are exact).
#define RM_HP0 0x1.921fb54442d18p+0
#define RM_HP1 0x1.1a62633145c07p-54
#define RM_MP1 0x1.921fb58p+0
#define RM_MP2 (-0x1.dde973cp-27)
#define RM_PP3 (-0x1.cb3b398p-55)
#define RM_PP4 (-0x1.d747f23e32ed7p-83)
#define RM_HPINV 0x1.45f306dc9c883p-1 // 2/pi

// a * b + c: fused in the FMA build of the method, two roundings otherwise.
template <bool Fused> FNEC_HD double rm_mla(double a, double b, double c)
{
    if (Fused)
        return fma(a, b, c);
    return a * b + c;
}

template <bool Fu> FNEC_HD double rm_taylor_sin(double xx, double a, double da)
{
    const double p = rm_mla<Fu>(rm_mla<Fu>(rm_mla<Fu>(rm_mla<Fu>(RM_S5, xx, RM_S4), xx, RM_S3), xx, RM_S2), xx,
                                RM_S1);
    const double t = rm_mla<Fu>(rm_mla<Fu>(p, a, -(0.5 * da)), xx, da);
    return a + t;
}

// cos(x + dx) for |x| < 0.855469 by the table.
template <bool Fu> FNEC_HD double rm_do_cos(double x, double dx, const double *tab)
{
    if (x < 0)
        dx = -dx;
    const double u = RM_BIG + fabs(x);
    x = fabs(x) - (u - RM_BIG) + dx;
    const double xx = x * x;
    const double s = rm_mla<Fu>(x * xx, rm_mla<Fu>(xx, RM_SN5, RM_SN3), x);
    const double c = xx * rm_mla<Fu>(xx, rm_mla<Fu>(xx, RM_CS6, RM_CS4), RM_CS2);
    const int k = (int)(rm_d2bits(u) & 0xffffffffu) << 2;
    const double sn = tab[k], ssn = tab[k + 1], cs = tab[k + 2], ccs = tab[k + 3];
    const double cor = rm_mla<Fu>(-sn, s, rm_mla<Fu>(-cs, c, rm_mla<Fu>(-s, ssn, ccs)));
    return cs + cor;
}

// sin(x + dx) for |x| < 0.855469.
template <bool Fu> FNEC_HD double rm_do_sin(double x, double dx, const double *tab)
{
    const double xold = x;
    if (fabs(x) < 0.126)
        return rm_taylor_sin<Fu>(x * x, x, dx);
    if (x <= 0)
        dx = -dx;
    const double u = RM_BIG + fabs(x);
    x = fabs(x) - (u - RM_BIG);
    const double xx = x * x;
    const double s = x + rm_mla<Fu>(x * xx, rm_mla<Fu>(xx, RM_SN5, RM_SN3), dx);
    const double c = rm_mla<Fu>(x, dx, xx * rm_mla<Fu>(xx, rm_mla<Fu>(xx, RM_CS6, RM_CS4), RM_CS2));
    const int k = (int)(rm_d2bits(u) & 0xffffffffu) << 2;
    const double sn = tab[k], ssn = tab[k + 1], cs = tab[k + 2], ccs = tab[k + 3];
    const double cor = rm_mla<Fu>(cs, s, rm_mla<Fu>(-sn, c, rm_mla<Fu>(s, ccs, ssn)));
    return copysign(sn + cor, xold);
}

// x = n pi/2 + (a + da), |a| <= pi/4 (about); returns n mod 4.
template <bool Fu> FNEC_HD int rm_reduce(double x, double &a, double &da)
{
    const double t = rm_mla<Fu>(x, RM_HPINV, RM_TOINT);
    const double xn = t - RM_TOINT;
    const double y = rm_mla<Fu>(-xn, RM_MP2, rm_mla<Fu>(-xn, RM_MP1, x));
    const int n = (int)(rm_d2bits(t) & 3);
    double t1 = xn * RM_PP3;
    const double t2 = y - t1;
    double db = (y - t2) - t1;
    t1 = xn * RM_PP4;
    const double b = t2 - t1;
    db += (t2 - b) - t1;
    a = b;
    da = db;
    return n;
}

template <bool Fu> FNEC_HD double rm_quadrant(double a, double da, int n, const double *tab)
{
    const double r = (n & 1) ? rm_do_cos<Fu>(a, da, tab) : rm_do_sin<Fu>(a, da, tab);
    return (n & 2) ? -r : r;
}

FNEC_HD double rm_sin(double x, const double *tab)
{
    const uint32_t k = (uint32_t)(rm_d2bits(x) >> 32) & 0x7fffffffu;
    if (k < 0x3e500000u)
        return x;
    if (k < 0x3feb6000u)
        return rm_do_sin<true>(x, 0.0, tab);
    if (k < 0x400368fdu) {
        const double t = RM_HP0 - fabs(x);
        return copysign(rm_do_cos<true>(t, RM_HP1, tab), x);
    }
    if (k < 0x419921fbu) {
        double a, da;
        const int n = rm_reduce<true>(x, a, da);
        return rm_quadrant<true>(a, da, n, tab);
    }
    return sin(x);
}

FNEC_HD double rm_cos(double x, const double *tab)
{
    const uint32_t k = (uint32_t)(rm_d2bits(x) >> 32) & 0x7fffffffu;
    if (k < 0x3e400000u)
        return 1.0;
    if (k < 0x3feb6000u)
        return rm_do_cos<true>(x, 0.0, tab);
    if (k < 0x400368fdu) {
        const double y = RM_HP0 - fabs(x);
        const double a = y + RM_HP1;
        const double da = (y - a) + RM_HP1;
        return rm_do_sin<true>(a, da, tab);
    }
    if (k < 0x419921fbu) {
        double a, da;
        const int n = rm_reduce<true>(x, a, da);
        return rm_quadrant<true>(a, da, n + 1, tab);
    }
    return cos(x);
}

// sin(x) and cos(x) together, as the libm's sincos computes them.
FNEC_HD void rm_sincos(double x, const double *tab, double &sx, double &cx)
{
    const uint32_t k = (uint32_t)(rm_d2bits(x) >> 32) & 0x7fffffffu;
    if (k < 0x400368fdu) {
        if (k < 0x3e400000u) {
            sx = x;
            cx = 1.0;
            return;
        }
        if (k < 0x3feb6000u) {
            sx = rm_do_sin<false>(x, 0.0, tab);
            cx = rm_do_cos<false>(x, 0.0, tab);
            return;
        }
        const double y = RM_HP0 - fabs(x);
        const double a = y + RM_HP1;
        const double da = (y - a) + RM_HP1;
        sx = copysign(rm_do_cos<false>(y, RM_HP1, tab), x);
        cx = rm_do_sin<false>(a, da, tab);
        return;
    }
    if (k < 0x419921fbu) {
        double a, da;
        const int n = rm_reduce<false>(x, a, da);
        sx = rm_quadrant<false>(a, da, n, tab);
        cx = rm_quadrant<false>(a, da, n + 1, tab);
        return;
    }
    sx = sin(x);
    cx = cos(x);
}

FNEC_HD double rm_hypot_kernel(double ax, double ay)
{
    double t1, t2;
    double h = sqrt(ax * ax + ay * ay);
    if (h <= 2.0 * ay) {
        const double delta = h - ay;
        t1 = ax * (2.0 * delta - ax);
        t2 = (delta - 2.0 * (ax - ay)) * delta;
    } else {
        const double delta = h - ax;
        t1 = 2.0 * delta * (ax - 2.0 * ay);
        t2 = (4.0 * delta - ay) * ay + delta * delta;
    }
    h -= (t1 + t2) / (2.0 * h);
    return h;
}

// |x + iy| for finite arguments (the path never forms infinities or NaNs
// from finite data except through overflow, which the callers treat alike).
FNEC_HD double rm_hypot(double x, double y)
{
    if (!isfinite(x) || !isfinite(y))
        return (isinf(x) || isinf(y)) ? INFINITY : x + y;
    x = fabs(x);
    y = fabs(y);
    const double ax = x < y ? y : x, ay = x < y ? x : y;
    if (ax > 0x1p+511) {
        if (ay <= ax * 0x1p-54)
            return ax + ay;
        return rm_hypot_kernel(ax * 0x1p-600, ay * 0x1p-600) / 0x1p-600;
    }
    if (ay < 0x1p-511) {
        if (ax >= ay / 0x1p-54)
            return ax + ay;
        return rm_hypot_kernel(ax / 0x1p-600, ay / 0x1p-600) * 0x1p-600;
    }
    if (ay <= ax * 0x1p-54)
        return ax + ay;
    return rm_hypot_kernel(ax, ay);
}

// num-complex 0.4 a / b: ((ar br + ai bi) / |b|^2, (ai br - ar bi) / |b|^2).
FNEC_HD void rm_div(double ar, double ai, double br, double bi, double &qr, double &qi)
{
    const double ns = br * br + bi * bi;
    qr = (ar * br + ai * bi) / ns;
    qi = (ai * br - ar * bi) / ns;
}

} // namespace fnec
