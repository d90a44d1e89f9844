/*
 * ORACLE — test infrastructure only.
 *
 * The host libm over arrays: f64::sin, f64::cos and f64::hypot, which the
 * reference calls through Rust std (linear.rs:1346 via Complex64::norm), and
 * sincos, which its codegen calls where a sin and a cos of one operand meet
 * (num-complex exp = from_polar, as in green_k, matrix.rs:397-399). Tests
 * compare the plain path's device functions with these, bit for bit.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stddef.h>

#include "oracle_internal.h"

void or_libm_eval(int fn, const double *x, const double *y, size_t n, double *out)
{
    size_t i;
    for (i = 0; i < n; i++) {
        if (fn == 0)
            out[i] = sin(x[i]);
        else if (fn == 1)
            out[i] = cos(x[i]);
        else if (fn == 2)
            out[i] = hypot(x[i], y[i]);
        else {
            double sx, cx;
            sincos(x[i], &sx, &cx);
            out[i] = fn == 5 ? sx : cx;
        }
    }
}
