// Shared host/device definitions for the fnec_hip engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fnec {

constexpr double kPi = 3.14159265358979323846;
constexpr double kC0 = 299792458.0;
constexpr double kMu0 = 4.0 * kPi * 1e-7;
constexpr double kEps0Matrix = 8.854187817e-12;   // matrix.rs:33
constexpr double kEps0FarField = 8.8541878128e-12; // farfield.rs:293
constexpr double kLambda = 1e-8;                   // linear.rs:782
constexpr double kSingular = 1e-30;                // linear.rs:1354

constexpr int kTile = 64;      // Gram / trailing-update output tile (complex elements)
constexpr int kGramK = 16;     // Gram K chunk
constexpr int kPanel = 32;     // LU panel width
constexpr int kBsThreads = 1024; // k_backsub: one workgroup per problem, many row-update loads in flight
constexpr int kSegStride = 8;  // doubles per segment: mid xyz, dir xyz, length, radius

// One Hallén problem inside a batch. All pointers are device pointers into
// the batch arena. Matrices are column-major with split real / imaginary
// planes so that MFMA operands load as contiguous doubles.
struct Prob {
    int n;      // segments
    int s;      // unknowns (n + homogeneous columns)
    int r;      // rows of the augmented system (n + constraint rows)
    int w;      // homogeneous columns
    int ncons;  // constraint rows (free ends + junctions)
    int nruns;  // runs of equal homogeneous column along the rows
    int ldz;    // leading dimension of Z
    int lda;    // leading dimension of the augmented normal matrix
    int ground; // 0 free, 1 PEC, 2 scalar-Gamma image
    int fill_centred; // k_fill_z: phases taken relative to the source centre's (1) or whole (0)
    double k;        // fill wavenumber 2*pi*f/c0 (matrix.rs:194)
    double gam_re;   // scalar reflection coefficient for ground == 2
    double gam_im;
    const double *seg;     // n * kSegStride
    const double *selfln;  // n: 2 ln((h + R_end) / a) of each segment's self term (host libm)
    const double *stamps;  // nstamps * 4: row, col, re, im of each Z stamp, in order
    int nstamps;
    int pad2;
    double *zr, *zi;       // Z, ldz x n
    double *ar, *ai;       // [G | M^H y], lda x (s + 1)
    const double *cosv;    // n
    const double *rhs;     // n complex, interleaved
    const int *runs;       // nruns * 3: row_begin, row_end (exclusive), column
    const int *cons;       // ncons * 2: a, b (b < 0: single-entry row)
    const double *cons_sign; // ncons
    int *ipiv;             // s
    int *info;             // [0] status (0 ok, 1 singular), [1] swaps
    double *x;             // s complex, interleaved
    double *stats;         // [0] abs_res, [1] rel_res, [2] diag_spread
    double *scratch;       // residual partial sums
    double *pr, *pi;       // panel workspace, lda x kPanel (fast no-swap path)
    unsigned long long *colmax; // kPanel: max |Schur entry| below the diagonal, as bits
    double *diagmag;       // kPanel: |u_jj|
    int *fast_ok;          // [0]: the fast path factored the current panel
    int pad1;
    // ---- plane-wave and current-source systems (plain path only) ----
    // Right-hand sides: nrhs columns s .. s + nrhs - 1 of [G | M^H Y], rhs
    // holds nrhs forcing vectors of n (column k at rhs + 2 n k), x nrhs
    // solutions of s (solution k at x + 2 s k).
    int nrhs;
    int variant;           // 0: delta-gap system; 1: plane wave; 2: current source
    // Entries of M beyond the -cos runs: nxruns records (row begin, row end,
    // column, kind), kind 1: -sinv[r], kind 2: -(shape_re, shape_im)[r].
    const int *xruns;
    int nxruns;
    int pad3;
    const double *sinv;    // n
    const double *shape;   // n complex, interleaved
    // Per row r < n: the extra column holding its -cos, -sin and -shape
    // entries (-1 for none), 3 ints per row.
    const int *rowent;
    // Constraint rows: the value at a (null: 1) and the right-hand side y
    // (null: 0; complex, the same for every right-hand side column).
    const double *cons_aval;
    const double *cons_y;
    // ---- mixed-precision path (null on the plain FP64 path) ----
    float *fr, *fi;        // FP32 copy of Z written by the fill, then its L\U factors
    float *fpr, *fpi;      // FP32 panel workspace, ldz x kPanel
    const int *rowcol;     // n: homogeneous column of each row, -1 for none
    double *fh;            // n x ncons complex, column c = Z^{-H} (constraint row c)^H
    double *lk;            // ncons x ncons complex, Cholesky factor of K = I + F F^H
    double *ym;            // ncons x w complex, Y = Lk^{-1} H with H = -F B
    double *ls;            // w x w complex, Cholesky factor of Y^H Y + lambda I
    double *vr, *vz, *vp, *vq; // PCG vectors, s complex each
    double *vt;            // M p, r complex
    double *vv;            // triangular-solve vector, n complex
    double *part;          // nchunk x n complex partial sums of Z p
    double *colsq;         // n: squared column norms of Z
    double *pcg;           // [0] rho, [1] |h|, [2] |r|, [3] |M|_F^2, [4] |x|
    int *pstate;           // [0] converged, [1] iterations, [2] breakdown
    int *perm;             // n: P Z = L U row permutation of the FP32 factors
    double *tws;           // n x (ncons + 1) complex: triangular-solve workspace per vector
    int *tflags;           // (ncons + 1) x 2 x ceil(n / 64): block completion epochs
    float *dinv;           // ceil(n / 64) x 2 x 64 x 64 complex: L_ii^{-1}, U_ii^{-1}
    unsigned long long *gx; // 2 x ceil(n / 64) x 64 x 4 words: solved blocks a gang publishes
                            // (forward, backward sweep), one data dword + the launch epoch per word
    int nchunk;
    int maxit;
    double tol;
};

// sin and cos of a phase: Cody-Waite reduction by pi/2 in three parts (the
// first two carry 33 significant bits, so n P1 and n P2 are exact for
// |n| < 2^20), then the Taylor polynomials of sin (to r^15) and cos (to r^16)
// on |r| <= pi/4, whose truncation errors stay below 6e-17 and 3e-18. The
// result is within about 2 ulp. Arguments beyond 1e5 (never a phase k R of
// the decks this path accepts) take the library routine.
__device__ __forceinline__ void sincos_phase(double x, double &s, double &c)
{
    if (!(fabs(x) < 1e5)) {
        sincos(x, &s, &c);
        return;
    }
    const double n = rint(x * 0.6366197723675814);      // 2 / pi
    double r = fma(-n, 1.5707963267341256, x);           // P1 = 0x1.921fb544p+0
    r = fma(-n, 6.077100506303966e-11, r);               // P2 = 0x1.0b4611a6p-34
    r = fma(-n, 2.0222662487959506e-21, r);              // P3 = pi/2 - P1 - P2
    const double r2 = r * r;
    double ps = fma(r2, -7.647163731819816e-13, 1.6059043836821613e-10); // 1/15!, 1/13!
    ps = fma(r2, ps, -2.505210838544172e-08);
    ps = fma(r2, ps, 2.7557319223985893e-06);
    ps = fma(r2, ps, -0.0001984126984126984);
    ps = fma(r2, ps, 0.008333333333333333);
    ps = fma(r2, ps, -0.16666666666666666);
    const double sr = fma(r * r2, ps, r);
    double pc = fma(r2, 4.779477332387385e-14, -1.1470745597729725e-11); // 1/16!, 1/14!
    pc = fma(r2, pc, 2.08767569878681e-09);
    pc = fma(r2, pc, -2.755731922398589e-07);
    pc = fma(r2, pc, 2.48015873015873e-05);
    pc = fma(r2, pc, -0.001388888888888889);
    pc = fma(r2, pc, 0.041666666666666664);
    pc = fma(r2, pc, -0.5);
    const double cr = fma(r2, pc, 1.0);
    const int q = (int)n & 3; // quadrant of x
    s = (q & 1) ? cr : sr;
    c = (q & 1) ? sr : cr;
    if (q & 2)
        s = -s;
    if ((q + 1) & 2)
        c = -c;
}

// sin and cos of a small phase by their Taylor polynomials, sin to d^kSinDeg
// and cos to d^(kSinDeg + 1), no range reduction. The fill uses them for
//   kSinDeg  9: |d| <= 0.121, truncation below 1.9e-18 (sin) and 2e-20 (cos)
//   kSinDeg 11: |d| <= 0.250, below 2.4e-18 and 5e-20
//   kSinDeg 13: |d| <= 0.490, below 1.8e-17 and 6e-19.
template <int kSinDeg> __device__ __forceinline__ void sincos_small(double d, double &s, double &c)
{
    static_assert(kSinDeg == 9 || kSinDeg == 11 || kSinDeg == 13, "degree 9, 11 or 13");
    const double d2 = d * d;
    double ps, pc;
    if constexpr (kSinDeg == 13) {
        ps = fma(d2, 1.6059043836821613e-10, -2.505210838544172e-08); // 1/13!, -1/11!
        ps = fma(d2, ps, 2.7557319223985893e-06);
        pc = fma(d2, -1.1470745597729725e-11, 2.08767569878681e-09); // -1/14!, 1/12!
        pc = fma(d2, pc, -2.755731922398589e-07);
    } else if constexpr (kSinDeg == 11) {
        ps = fma(d2, -2.505210838544172e-08, 2.7557319223985893e-06); // -1/11!, 1/9!
        pc = fma(d2, 2.08767569878681e-09, -2.755731922398589e-07);   // 1/12!, -1/10!
    } else {
        ps = 2.7557319223985893e-06; // 1/9!
        pc = -2.755731922398589e-07; // -1/10!
    }
    ps = fma(d2, ps, -0.0001984126984126984); // -1/7!
    ps = fma(d2, ps, 0.008333333333333333);
    ps = fma(d2, ps, -0.16666666666666666);
    s = fma(d * d2, ps, d);
    pc = fma(d2, pc, 2.48015873015873e-05); // 1/8!
    pc = fma(d2, pc, -0.001388888888888889);
    pc = fma(d2, pc, 0.041666666666666664);
    pc = fma(d2, pc, -0.5);
    c = fma(d2, pc, 1.0);
}

constexpr int kMixMaxCons = 256; // constraint rows the mixed path accepts
constexpr int kMixMaxW = 128;    // homogeneous columns the mixed path accepts
constexpr int kMvRows = 256;     // rows per M p work item
constexpr int kMvChunk = 256;    // columns per M p work item
constexpr int kMvCols = 16;      // columns per M^H t work item

// One matrix factored by the batched LU (lu.h): column-major split planes.
template <typename T> struct LuDesc {
    T *ar, *ai;       // lda x ncols, factored in place
    T *pr, *pi;       // panel workspace, lda x kPanel (fast no-swap path)
    int lda;
    int m;            // order: rows and pivot columns
    int ncols;        // columns carried along (m + 1 with an augmented right-hand side)
    int pad0;
    int *ipiv;        // m
    int *info;        // [0] singular, [1] interchanges, [2] panels on the exact path
    unsigned long long *colmax; // kPanel: max |entry| below the diagonal, as double bits
    double *diagmag;  // kPanel: |u_jj|
    int *fast_ok;     // [0]: the fast path factored the current panel
    int pad1;
    T *linv;          // FP32 two-level LU: L11^{-1} of the current 64-column step, 2 x 64 x 64
    void *h16;        // FP32 LU, K = 128 passes: the combining step's L21 and U12 split into
                      // f16 hi / lo planes (lu.h, k_split_f16x2), 2 x 16 x lda x 32 halves
};

// Work lists of one panel step (offsets into the batch's int list).
struct LuStep {
    int p;
    size_t panel_off, panel_cnt, trsm_off, trsm_cnt, upd_off, upd_cnt;
    size_t rows_off, rows_cnt, commit_off, commit_cnt;
};

// Far-field problem (one RP grid + normalisation integral).
struct RpProb {
    int n;
    int npts;
    int ground;       // 0 free, 1 PEC, 2 Fresnel
    int ndir;         // normalisation directions (181*360 or 91*360)
    double k;         // 2*pi/(c0/f) (farfield.rs:570-571)
    double eps_r, sigma;
    const double *seg;     // n * kSegStride
    const double *cur;     // n complex interleaved
    const double *theta;   // npts
    const double *phi;     // npts
    double *field;         // npts * 4: F_theta re/im, F_phi re/im
    double *partial;       // per-block partial sums of the integral
    int npartial;
    int pad;
    double *total;         // [0]
    double *out;           // npts * 4: gains + axial ratio
};

} // namespace fnec
