// Host build of refmath.h, the elementary functions the plain path's kernels
// evaluate, so that CPU-only machines can test them against the host libm.
// Nothing in libfnec_hip.so links this file; the kernels compile the same
// header for the device.
#include <stddef.h>

#include "refmath.h"

namespace {
const double kTab[FNEC_SINCOS_TABLE_LEN] = {FNEC_SINCOS_TABLE_INIT};
}

// fn 0: sin(x), 1: cos(x), 2: hypot(x, y), 5: sin and 6: cos of sincos(x).
// y may be NULL except for fn 2.
extern "C" void fnec_refmath_eval(int fn, const double *x, const double *y, size_t n, double *out)
{
    for (size_t i = 0; i < n; ++i) {
        if (fn == 0)
            out[i] = fnec::rm_sin(x[i], kTab);
        else if (fn == 1)
            out[i] = fnec::rm_cos(x[i], kTab);
        else if (fn == 2)
            out[i] = fnec::rm_hypot(x[i], y[i]);
        else {
            double sx, cx;
            fnec::rm_sincos(x[i], kTab, sx, cx);
            out[i] = fn == 5 ? sx : cx;
        }
    }
}
