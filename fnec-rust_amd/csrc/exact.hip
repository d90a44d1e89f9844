// Kernels of the plain FP64 path that reproduce the reference CPU path's
// rounding bit for bit. Compiled with -ffp-contract=off: every product and sum
// rounds on its own, as Rust's codegen does (it never contracts to FMA).
//
//   k_fill_exact    Z fill (matrix.rs:186-218, 297-331, 397-399, 403-418): the
//                   reference's operation order per Gauss point, with the
//                   host libm's sin / cos reproduced on the device (refmath.h).
//                   The self term's 2 ln((h + R_end) / a) is the host's value
//                   (one per segment, computed with the same libm).
//   k_gx_*          A = M^H M + lambda I and M^H y (linear.rs:764-785): every
//                   element is one sequential sum over the rows of M in
//                   ascending order, conj(m_ri) m_rj formed as num-complex
//                   does. Products with a structural zero of M are skipped:
//                   adding +-0 to a sum changes at most the sign of an exactly
//                   zero result. The upper triangle is the conjugate of the
//                   lower, which the reference's sums give bit for bit.
//   k_ge_*          Gaussian elimination with partial pivoting on [A | b]
//                   (linear.rs:1321-1383) in 32-column panels. Every element
//                   receives the reference's updates a -= f u in the
//                   reference's order (f = a_rc / a_cc by naive complex
//                   division), so the factors are the reference's bit for bit:
//                     k_ge_diag    the panel's diagonal block without
//                                  interchanges, one lane per row (backup of
//                                  the panel kept in the workspace);
//                     k_ge_rows    the rows below it, one thread per row; a
//                                  row whose |a_rc| is not strictly below the
//                                  step's |a_cc| would be chosen as the pivot
//                                  (last maximal hypot), so it flags the panel;
//                     k_ge_exact   a flagged panel: restore from the backup,
//                                  then partial pivoting exactly as the
//                                  reference does, column by column;
//                     k_ge_u12     the panel's interchanges on the columns to
//                                  its right, then their U rows;
//                     k_ge_update  the trailing update, one 4x4 element block
//                                  per thread, 32 sequential updates each.
//   k_backsub_exact back substitution in the reference's order (ascending j
//                   in every row sum): a serial chain of S^2/2 subtractions,
//                   selectable for bitwise comparisons; by default the plain
//                   path uses the blocked k_backsub, whose summation order
//                   moves x by about 1e-14 relative.
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.h"
#include "refmath.h"

namespace fnec {

__device__ const double g_sincos_tab[FNEC_SINCOS_TABLE_LEN] = {FNEC_SINCOS_TABLE_INIT};

// Gauss-Legendre nodes and weights, the 15-digit literals of matrix.rs:39-72.
__constant__ double c_xn4[4] = {-0.861136311594953, -0.339981043584856, 0.339981043584856,
                                0.861136311594953};
__constant__ double c_xw4[4] = {0.347854845137454, 0.652145154862626, 0.652145154862626,
                                0.347854845137454};
__constant__ double c_xn8[8] = {-0.960289856497536, -0.796666477413627, -0.525532409916329,
                                -0.183434642495650, 0.183434642495650,  0.525532409916329,
                                0.796666477413627,  0.960289856497536};
__constant__ double c_xw8[8] = {0.101228536290376, 0.222381034453374, 0.313706645877887,
                                0.362683783378362, 0.362683783378362, 0.313706645877887,
                                0.222381034453374, 0.101228536290376};

// ---------------------------------------------------------------------------
// Z fill
// ---------------------------------------------------------------------------

// green_k(r, k) = exp((0, -k r)) / r = (cos(-k r) / r, sin(-k r) / r)
// (num-complex exp = from_polar(exp(0) = 1, -k r); 1 * c = c exactly; the
// sin and cos of one phase are one sincos call in the reference's codegen).
__device__ __forceinline__ void green_exact(double r, double k, double &gr, double &gi)
{
    const double a = -k * r;
    double s, c;
    rm_sincos(a, g_sincos_tab, s, c);
    gr = c / r;
    gi = s / r;
}

// elem() for an observation segment (mid om, dir od) and a source segment
// (mid sm, dir sd, length len, radius rad); selfln = 2 ln((h + R_end)/a).
__device__ __forceinline__ void elem_exact(const double *om, const double *od, double smx,
                                           double smy, double smz, double sdx, double sdy,
                                           double sdz, double len, double rad, double k,
                                           bool self, double selfln, double &re, double &im)
{
    const double cos_alpha = od[0] * sdx + od[1] * sdy + od[2] * sdz;
    const double half = len * 0.5;
    const double a = rad;
    double ir, ii;
    if (self) {
        double sr = 0.0, si = 0.0;
#pragma unroll 1
        for (int m = 0; m < 4; ++m) {
            const double l = c_xn4[m] * half;
            const double r_eff = sqrt(l * l + a * a);
            double gr, gi;
            green_exact(r_eff, k, gr, gi);
            const double dyn_r = gr - 1.0 / r_eff;
            sr = sr + c_xw4[m] * dyn_r;
            si = si + c_xw4[m] * gi;
        }
        sr = sr * half;
        si = si * half;
        ir = sr + selfln;
        ii = si;
    } else {
        double sr = 0.0, si = 0.0;
#pragma unroll 1
        for (int m = 0; m < 8; ++m) {
            const double sc = c_xn8[m] * half;
            const double r0 = smx + sc * sdx, r1 = smy + sc * sdy, r2 = smz + sc * sdz;
            const double p0 = om[0] - r0, p1 = om[1] - r1, p2 = om[2] - r2;
            const double r_sq = p0 * p0 + p1 * p1 + p2 * p2;
            const double r_eff = sqrt(r_sq + a * a);
            double gr, gi;
            green_exact(r_eff, k, gr, gi);
            sr = sr + c_xw8[m] * gr;
            si = si + c_xw8[m] * gi;
        }
        ir = sr * half;
        ii = si * half;
    }
    re = ir * cos_alpha;
    im = ii * cos_alpha;
}

constexpr int kXFillRows = 64;
constexpr int kXFillCols = 16;

// Work item (problem, row tile, column tile): 64 observation rows x 16 source
// columns; each thread owns one row and four columns, and a wavefront writes
// 64 consecutive rows of one column (coalesced column-major stores).
__global__ __launch_bounds__(256) void k_fill_exact(const Prob *__restrict__ probs,
                                                    const int4 *__restrict__ work)
{
    const int4 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    const int i0 = wk.y * kXFillRows, j0 = wk.z * kXFillCols;
    __shared__ double so[kXFillRows][6];
    __shared__ double ss[kXFillCols][9];
    const int t = threadIdx.x;
    for (int e = t; e < kXFillRows * 6; e += 256) {
        const int row = e / 6, c = e % 6, i = i0 + row;
        so[row][c] = i < P.n ? P.seg[(size_t)i * kSegStride + c] : 0.0;
    }
    for (int e = t; e < kXFillCols * 9; e += 256) {
        const int col = e / 9, c = e % 9, j = j0 + col;
        double v = 1.0;
        if (j < P.n)
            v = c < 8 ? P.seg[(size_t)j * kSegStride + c] : P.selfln[j];
        ss[col][c] = v;
    }
    __syncthreads();
    const int row = t & 63;
    const int i = i0 + row;
    if (i >= P.n)
        return;
    const double om[3] = {so[row][0], so[row][1], so[row][2]};
    const double od[3] = {so[row][3], so[row][4], so[row][5]};
    const double k = P.k;
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        const int col = (t >> 6) * 4 + q;
        const int j = j0 + col;
        if (j >= P.n)
            break;
        const double sx = ss[col][0], sy = ss[col][1], sz = ss[col][2];
        const double dx = ss[col][3], dy = ss[col][4], dz = ss[col][5];
        const double len = ss[col][6], rad = ss[col][7];
        double re, im;
        elem_exact(om, od, sx, sy, sz, dx, dy, dz, len, rad, k, i == j, ss[col][8], re, im);
        // direct + image_contrib, the image being (0, 0) in free space
        // (matrix.rs:199-214): the sum turns a -0 of the direct term into +0
        double xr = 0.0, xi = 0.0;
        if (P.ground != 0) {
            // image_segment (matrix.rs:265-277): mid z -> -z, dir (-dx, -dy, dz)
            elem_exact(om, od, sx, sy, -sz, -dx, -dy, dz, len, rad, k, false, 0.0, xr, xi);
            if (P.ground == 2) { // gamma * image, num-complex product
                const double gr = P.gam_re * xr - P.gam_im * xi;
                const double gi = P.gam_re * xi + P.gam_im * xr;
                xr = gr;
                xi = gi;
            }
        }
        re = re + xr;
        im = im + xi;
        const size_t o = (size_t)j * P.ldz + i;
        P.zr[o] = re;
        P.zi[o] = im;
    }
}

// ---------------------------------------------------------------------------
// Normal equations, reference summation order
// ---------------------------------------------------------------------------

constexpr int kXT = 64;  // output tile
constexpr int kXK = 16;  // rows of M per LDS chunk

// Imaginary part of the mirror element G[j][i] = conj(G[i][j]): the
// reference sums G[j][i] on its own, and its products are those of G[i][j]
// with the imaginary parts negated, so the sum is -Im G[i][j] bit for bit,
// except that an exactly zero sum is +0 both ways (round to nearest).
__device__ __forceinline__ double mirror_im(double x) { return x == 0.0 ? 0.0 : -x; }

// conj(a) b as num-complex forms it: (ar br - (-ai) bi, ar bi + (-ai) br),
// which is (ar br + ai bi, ar bi - ai br) bit for bit.
__device__ __forceinline__ void cjmul(double ar, double ai, double br, double bi, double &pr, double &pi)
{
    pr = ar * br + ai * bi;
    pi = ar * bi - ai * br;
}

// Work item (problem, ti, tj), ti >= tj: the 64x64 tile of G[i][j] =
// sum_{r < n} conj(Z_ri) Z_rj, rows r in ascending order. Each thread owns a
// 4x4 block (rows i0 + ty + 16u, columns j0 + tx + 16v): 32 independent
// sequential sums.
__global__ __launch_bounds__(256) void k_gx_tile(const Prob *__restrict__ probs,
                                                 const int4 *__restrict__ work)
{
    const int4 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    const int n = P.n, ldz = P.ldz, lda = P.lda;
    const int i0 = wk.y * kXT, j0 = wk.z * kXT;
    __shared__ double sar[kXK][kXT + 1], sai[kXK][kXT + 1];
    __shared__ double sbr[kXK][kXT + 1], sbi[kXK][kXT + 1];
    const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
    double gr[4][4], gi[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            gr[u][v] = gi[u][v] = 0.0;
    for (int r0 = 0; r0 < n; r0 += kXK) {
        const int cnt = min(kXK, n - r0);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q, c = e >> 4, rr = e & 15;
            double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
            if (rr < cnt) {
                if (i0 + c < n) {
                    ar = P.zr[(size_t)(i0 + c) * ldz + r0 + rr];
                    ai = P.zi[(size_t)(i0 + c) * ldz + r0 + rr];
                }
                if (j0 + c < n) {
                    br = P.zr[(size_t)(j0 + c) * ldz + r0 + rr];
                    bi = P.zi[(size_t)(j0 + c) * ldz + r0 + rr];
                }
            }
            sar[rr][c] = ar;
            sai[rr][c] = ai;
            sbr[rr][c] = br;
            sbi[rr][c] = bi;
        }
        __syncthreads();
        for (int rr = 0; rr < cnt; ++rr) {
            double a_r[4], a_i[4], b_r[4], b_i[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a_r[u] = sar[rr][ty + 16 * u];
                a_i[u] = sai[rr][ty + 16 * u];
                b_r[u] = sbr[rr][tx + 16 * u];
                b_i[u] = sbi[rr][tx + 16 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double pr, pi;
                    cjmul(a_r[u], a_i[u], b_r[v], b_i[v], pr, pi);
                    gr[u][v] = gr[u][v] + pr;
                    gi[u][v] = gi[u][v] + pi;
                }
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + ty + 16 * u, j = j0 + tx + 16 * v;
            if (i >= n || j >= n || i < j)
                continue;
            P.ar[(size_t)j * lda + i] = gr[u][v];
            P.ai[(size_t)j * lda + i] = gi[u][v];
            if (i != j) {
                P.ar[(size_t)i * lda + j] = gr[u][v];
                P.ai[(size_t)i * lda + j] = mirror_im(gi[u][v]);
            }
        }
}

// Entry of M at row r < n in an extra column of the given kind (linear.rs:
// 728-731, 942-945, 1157-1161): -cos_r and -sin_r as (value, 0), and the
// current source's -shape_r (num-complex negation of both parts).
__device__ __forceinline__ void m_extra(const Prob &P, int kind, int r, double &vr, double &vi)
{
    if (kind == 0) {
        vr = -P.cosv[r];
        vi = 0.0;
    } else if (kind == 1) {
        vr = -P.sinv[r];
        vi = 0.0;
    } else {
        vr = -P.shape[2 * r];
        vi = -P.shape[2 * r + 1];
    }
}

// Work item (problem, block of 256 columns): for each column j < n of Z,
//   (M^H y)[j]   = sum_{r < n} conj(Z_rj) rhs_r   (one right-hand side; with
//                  several, k_gx_rhs forms them)
//   G[n+e][j]    = sum_{r in column e} conj(m_{r,n+e}) Z_rj   (and its
//                  conjugate G[j][n+e]) for every extra column e: the rows of
//                  a -cos column are its runs, of a -sin or -shape column its
//                  xruns, in ascending order.
__global__ __launch_bounds__(256) void k_gx_cols(const Prob *__restrict__ probs,
                                                 const int2 *__restrict__ work)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    const int n = P.n, lda = P.lda, s = P.s;
    const int j = wk.y * 256 + threadIdx.x;
    if (j >= n)
        return;
    const double *zr = P.zr + (size_t)j * P.ldz, *zi = P.zi + (size_t)j * P.ldz;
    if (P.nrhs == 1) {
        double yr = 0.0, yi = 0.0;
        for (int r = 0; r < n; ++r) {
            double pr, pi;
            cjmul(zr[r], zi[r], P.rhs[2 * r], P.rhs[2 * r + 1], pr, pi);
            yr = yr + pr;
            yi = yi + pi;
        }
        P.ar[(size_t)s * lda + j] = yr;
        P.ai[(size_t)s * lda + j] = yi;
    }
    for (int w = 0; w < P.w; ++w) {
        double gr = 0.0, gi = 0.0;
        for (int q = 0; q < P.nruns; ++q) {
            const int rb = P.runs[3 * q], re = P.runs[3 * q + 1];
            if (P.runs[3 * q + 2] != w)
                continue;
            for (int r = rb; r < re; ++r) {
                double pr, pi;
                cjmul(-P.cosv[r], 0.0, zr[r], zi[r], pr, pi);
                gr = gr + pr;
                gi = gi + pi;
            }
        }
        for (int q = 0; q < P.nxruns; ++q) {
            const int rb = P.xruns[4 * q], re = P.xruns[4 * q + 1], kind = P.xruns[4 * q + 3];
            if (P.xruns[4 * q + 2] != w)
                continue;
            for (int r = rb; r < re; ++r) {
                double vr, vi, pr, pi;
                m_extra(P, kind, r, vr, vi);
                cjmul(vr, vi, zr[r], zi[r], pr, pi);
                gr = gr + pr;
                gi = gi + pi;
            }
        }
        const int cw = n + w;
        P.ar[(size_t)j * lda + cw] = gr; // G[n+w][j]
        P.ai[(size_t)j * lda + cw] = gi;
        P.ar[(size_t)cw * lda + j] = gr; // G[j][n+w]
        P.ai[(size_t)cw * lda + j] = mirror_im(gi);
    }
}

// Work item (problem, tj, tk): the 64 x 64 tile of (M^H Y)[j][k] =
// sum_{r < n} conj(Z_rj) y_k[r] for columns j < n of Z and right-hand sides
// k (several incidences of a plane wave), each a sequential sum over r in
// ascending order as k_gx_cols forms the single one. Written to column s + k
// of [G | M^H Y].
__global__ __launch_bounds__(256) void k_gx_rhs(const Prob *__restrict__ probs,
                                                const int4 *__restrict__ work)
{
    const int4 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    const int n = P.n, ldz = P.ldz, lda = P.lda, nrhs = P.nrhs;
    const int j0 = wk.y * kXT, k0 = wk.z * kXT;
    __shared__ double sar[kXK][kXT + 1], sai[kXK][kXT + 1];
    __shared__ double sbr[kXK][kXT + 1], sbi[kXK][kXT + 1];
    const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
    double gr[4][4], gi[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            gr[u][v] = gi[u][v] = 0.0;
    for (int r0 = 0; r0 < n; r0 += kXK) {
        const int cnt = min(kXK, n - r0);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q, c = e >> 4, rr = e & 15;
            double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
            if (rr < cnt) {
                if (j0 + c < n) {
                    ar = P.zr[(size_t)(j0 + c) * ldz + r0 + rr];
                    ai = P.zi[(size_t)(j0 + c) * ldz + r0 + rr];
                }
                if (k0 + c < nrhs) {
                    const size_t o = 2 * ((size_t)(k0 + c) * n + r0 + rr);
                    br = P.rhs[o];
                    bi = P.rhs[o + 1];
                }
            }
            sar[rr][c] = ar;
            sai[rr][c] = ai;
            sbr[rr][c] = br;
            sbi[rr][c] = bi;
        }
        __syncthreads();
        for (int rr = 0; rr < cnt; ++rr) {
            double a_r[4], a_i[4], b_r[4], b_i[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a_r[u] = sar[rr][ty + 16 * u];
                a_i[u] = sai[rr][ty + 16 * u];
                b_r[u] = sbr[rr][tx + 16 * u];
                b_i[u] = sbi[rr][tx + 16 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double pr, pi;
                    cjmul(a_r[u], a_i[u], b_r[v], b_i[v], pr, pi);
                    gr[u][v] = gr[u][v] + pr;
                    gi[u][v] = gi[u][v] + pi;
                }
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = j0 + ty + 16 * u, k = k0 + tx + 16 * v;
            if (j >= n || k >= nrhs)
                continue;
            P.ar[(size_t)(P.s + k) * lda + j] = gr[u][v];
            P.ai[(size_t)(P.s + k) * lda + j] = gi[u][v];
        }
}

// Work item (problem, block of 256 tasks) over the extra columns e, f in
// [0, s - n): task p < h (h + 1) / 2 forms G[n+e][n+f], f <= e (and its
// conjugate mirror), as the sum over the rows of column e, ascending, of
// conj(m_{r,n+e}) m_{r,n+f} (rows where column f is zero skipped); the next
// h * nrhs tasks form (M^H y_k)[n+e] over the same rows. Rows of a column are
// its runs (-cos) or xruns (-sin, -shape); rowent gives each row's columns.
__global__ __launch_bounds__(256) void k_gx_extra(const Prob *__restrict__ probs,
                                                  const int2 *__restrict__ work)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    const int n = P.n, s = P.s, lda = P.lda, h = s - n;
    const int npair = h * (h + 1) / 2;
    const int task = wk.y * 256 + threadIdx.x;
    if (task >= npair + h * P.nrhs)
        return;
    int e, f = -1, k = -1;
    if (task < npair) {
        e = (int)((sqrt(8.0 * task + 1.0) - 1.0) * 0.5);
        while ((e + 1) * (e + 2) / 2 <= task)
            ++e;
        while (e * (e + 1) / 2 > task)
            --e;
        f = task - e * (e + 1) / 2;
    } else {
        e = (task - npair) / P.nrhs;
        k = (task - npair) % P.nrhs;
    }
    const double *y = k >= 0 ? P.rhs + 2 * (size_t)k * n : nullptr;
    double ar = 0.0, ai = 0.0;
    if (k < 0 && f != e && !P.rowent) {
        // delta-gap rows have one extra entry: distinct columns share no row
        P.ar[(size_t)(n + f) * lda + n + e] = 0.0;
        P.ai[(size_t)(n + f) * lda + n + e] = 0.0;
        P.ar[(size_t)(n + e) * lda + n + f] = 0.0;
        P.ai[(size_t)(n + e) * lda + n + f] = 0.0;
        return;
    }
    auto row = [&](int kind, int r) {
        double vr, vi, pr, pi;
        m_extra(P, kind, r, vr, vi);
        if (k >= 0) {
            cjmul(vr, vi, y[2 * r], y[2 * r + 1], pr, pi);
        } else {
            double wr = vr, wi = vi;
            if (f != e) {
                // a delta-gap row has one extra entry (no rowent): column f is
                // zero there
                if (!P.rowent)
                    return;
                int fk = -1;
                for (int q = 0; q < 3; ++q)
                    if (P.rowent[3 * r + q] == f)
                        fk = q;
                if (fk < 0)
                    return;
                m_extra(P, fk, r, wr, wi);
            }
            cjmul(vr, vi, wr, wi, pr, pi);
        }
        ar = ar + pr;
        ai = ai + pi;
    };
    for (int q = 0; q < P.nruns; ++q)
        if (P.runs[3 * q + 2] == e)
            for (int r = P.runs[3 * q]; r < P.runs[3 * q + 1]; ++r)
                row(0, r);
    for (int q = 0; q < P.nxruns; ++q)
        if (P.xruns[4 * q + 2] == e)
            for (int r = P.xruns[4 * q]; r < P.xruns[4 * q + 1]; ++r)
                row(P.xruns[4 * q + 3], r);
    if (k >= 0) {
        P.ar[(size_t)(s + k) * lda + n + e] = ar;
        P.ai[(size_t)(s + k) * lda + n + e] = ai;
        return;
    }
    P.ar[(size_t)(n + f) * lda + n + e] = ar; // G[n+e][n+f]
    P.ai[(size_t)(n + f) * lda + n + e] = ai;
    if (f != e) {
        P.ar[(size_t)(n + e) * lda + n + f] = ar;
        P.ai[(size_t)(n + e) * lda + n + f] = mirror_im(ai);
    }
}

// One workgroup per problem, after the kernels above: the constraint rows'
// terms in row order by one lane (each continues its elements' sums of G and,
// where the row's y is not zero, of M^H y), then lambda on the diagonal
// (linear.rs:782-785) by every lane. A row's
// entries are (a, aval) and, for a junction, (b, sign); a zero y adds +-0,
// which leaves every sum as it is, so such rows skip M^H y.
__global__ __launch_bounds__(64) void k_gx_fix(const Prob *__restrict__ probs,
                                               const int *__restrict__ plist)
{
    const Prob &P = probs[plist[blockIdx.x]];
    const int s = P.s, lda = P.lda;
    for (int c = 0; c < P.ncons && threadIdx.x == 0; ++c) {
        int idx[2];
        double val[2];
        int ne = 1;
        idx[0] = P.cons[2 * c];
        val[0] = P.cons_aval ? P.cons_aval[c] : 1.0;
        const int b = P.cons[2 * c + 1];
        if (b >= 0) {
            if (b == idx[0]) {
                val[0] = P.cons_sign[c]; // m[row][b] = sign overwrites m[row][a]
            } else {
                idx[1] = b;
                val[1] = P.cons_sign[c];
                ne = 2;
            }
        }
        for (int e = 0; e < ne; ++e)
            for (int f = 0; f < ne; ++f) {
                const int p = idx[e], q = idx[f];
                if (p < q)
                    continue;
                double pr, pi;
                cjmul(val[e], 0.0, val[f], 0.0, pr, pi);
                const size_t o = (size_t)q * lda + p;
                P.ar[o] = P.ar[o] + pr;
                P.ai[o] = P.ai[o] + pi;
                if (p != q) {
                    P.ar[(size_t)p * lda + q] = P.ar[o];
                    P.ai[(size_t)p * lda + q] = mirror_im(P.ai[o]);
                }
            }
        if (P.cons_y) {
            const double yr = P.cons_y[2 * c], yi = P.cons_y[2 * c + 1];
            if (yr != 0.0 || yi != 0.0)
                for (int k = 0; k < P.nrhs; ++k)
                    for (int e = 0; e < ne; ++e) {
                        double pr, pi;
                        cjmul(val[e], 0.0, yr, yi, pr, pi);
                        const size_t o = (size_t)(s + k) * lda + idx[e];
                        P.ar[o] = P.ar[o] + pr;
                        P.ai[o] = P.ai[o] + pi;
                    }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < s; i += 64) {
        const size_t o = (size_t)i * lda + i;
        P.ar[o] = P.ar[o] + kLambda;
        P.ai[o] = P.ai[o] + 0.0;
    }
}

// ---------------------------------------------------------------------------
// Gaussian elimination, reference order
// ---------------------------------------------------------------------------

constexpr int kXP = 32; // panel width

// One wavefront per problem: rows [p, p + jb) of the panel's columns, one lane
// per row, the block in LDS. Keeps the panel's original rows in the workspace,
// writes the no-interchange result to A, and clears fast_ok when partial
// pivoting would interchange (or the pivot is below the singular threshold).
__global__ __launch_bounds__(64) void k_ge_diag(const Prob *__restrict__ probs,
                                                const int *__restrict__ plist, int p)
{
    const Prob &P = probs[plist[blockIdx.x]];
    const int s = P.s, lda = P.lda;
    if (P.info[0] != 0 || p >= s)
        return;
    const int jb = min(kXP, s - p), lane = threadIdx.x;
    __shared__ double br[kXP][kXP + 1], bi[kXP][kXP + 1]; // [row][column]
    __shared__ int fail;
    if (lane == 0) {
        fail = 0;
        P.fast_ok[0] = 1;
    }
    const int row = p + lane;
    if (lane < jb) {
        P.ipiv[row] = row;
        for (int c = 0; c < jb; ++c) {
            const size_t o = (size_t)(p + c) * lda + row;
            const double vr = P.ar[o], vi = P.ai[o];
            br[lane][c] = vr;
            bi[lane][c] = vi;
            P.pr[(size_t)c * lda + row] = vr;
            P.pi[(size_t)c * lda + row] = vi;
        }
    }
    __syncthreads();
    for (int j = 0; j < jb; ++j) {
        const double dr = br[j][j], di = bi[j][j];
        const double hd = rm_hypot(dr, di);
        if (lane == j) {
            P.diagmag[j] = hd;
            if (!(hd >= kSingular))
                fail = 1;
        }
        if (lane > j && lane < jb) {
            if (!(rm_hypot(br[lane][j], bi[lane][j]) < hd))
                fail = 1;
            double fr, fi;
            rm_div(br[lane][j], bi[lane][j], dr, di, fr, fi);
            br[lane][j] = fr;
            bi[lane][j] = fi;
            for (int c = j + 1; c < jb; ++c) {
                const double ur = br[j][c], ui = bi[j][c];
                const double sr = fr * ur - fi * ui;
                const double si = fr * ui + fi * ur;
                br[lane][c] = br[lane][c] - sr;
                bi[lane][c] = bi[lane][c] - si;
            }
        }
        __syncthreads();
    }
    if (lane < jb)
        for (int c = 0; c < jb; ++c) {
            const size_t o = (size_t)(p + c) * lda + row;
            P.ar[o] = br[lane][c];
            P.ai[o] = bi[lane][c];
        }
    if (lane == 0 && fail)
        P.fast_ok[0] = 0;
}

// Work item (problem, block of 256 rows below the diagonal block): each thread
// eliminates its row's panel columns with the diagonal block's U rows, in the
// reference's order; a row whose |a_rc| at step c is not strictly below
// |a_cc| would be the pivot (ties take the later row), and clears fast_ok.
__global__ __launch_bounds__(256) void k_ge_rows(const Prob *__restrict__ probs,
                                                 const int2 *__restrict__ work, int p)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    const int s = P.s, lda = P.lda;
    if (P.info[0] != 0)
        return;
    const int jb = min(kXP, s - p), c1 = p + jb;
    __shared__ double ur[kXP][kXP + 1], ui[kXP][kXP + 1];
    __shared__ double hd[kXP];
    const int t = threadIdx.x;
    for (int e = t; e < jb * jb; e += 256) {
        const int c = e / jb, r = e % jb; // column c, row r of the block
        ur[r][c] = P.ar[(size_t)(p + c) * lda + p + r];
        ui[r][c] = P.ai[(size_t)(p + c) * lda + p + r];
    }
    if (t < jb)
        hd[t] = P.diagmag[t];
    __syncthreads();
    const int row = c1 + wk.y * 256 + t;
    if (row >= s)
        return;
    double vr[kXP], vi[kXP];
#pragma unroll
    for (int c = 0; c < kXP; ++c) {
        vr[c] = vi[c] = 0.0;
        if (c < jb) {
            const size_t o = (size_t)(p + c) * lda + row;
            vr[c] = P.ar[o];
            vi[c] = P.ai[o];
            P.pr[(size_t)c * lda + row] = vr[c];
            P.pi[(size_t)c * lda + row] = vi[c];
        }
    }
    bool fail = false;
#pragma unroll
    for (int j = 0; j < kXP; ++j) {
        if (j < jb) {
            if (!(rm_hypot(vr[j], vi[j]) < hd[j]))
                fail = true;
            double fr, fi;
            rm_div(vr[j], vi[j], ur[j][j], ui[j][j], fr, fi);
            vr[j] = fr;
            vi[j] = fi;
#pragma unroll
            for (int c = j + 1; c < kXP; ++c) {
                const double xr = ur[j][c], xi = ui[j][c];
                const double sr = fr * xr - fi * xi;
                const double si = fr * xi + fi * xr;
                vr[c] = vr[c] - sr;
                vi[c] = vi[c] - si;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < kXP; ++c)
        if (c < jb) {
            const size_t o = (size_t)(p + c) * lda + row;
            P.ar[o] = vr[c];
            P.ai[o] = vi[c];
        }
    if (fail)
        P.fast_ok[0] = 0;
}

constexpr int kXExThreads = 1024;

// One workgroup per problem, only for a panel the fast path flagged: restore
// the panel from the backup, then eliminate its columns with partial pivoting
// exactly as linear.rs:1343-1370 does (the last row of maximal hypot is the
// pivot; below 1e-30 the matrix is singular).
__global__ __launch_bounds__(kXExThreads) void k_ge_exact(const Prob *__restrict__ probs,
                                                          const int *__restrict__ plist, int p)
{
    const Prob &P = probs[plist[blockIdx.x]];
    const int s = P.s, lda = P.lda;
    if (P.info[0] != 0 || p >= s || P.fast_ok[0])
        return;
    const int jb = min(kXP, s - p), t = threadIdx.x;
    __shared__ double rv[kXExThreads];
    __shared__ int ri[kXExThreads];
    __shared__ double prow_r[kXP], prow_i[kXP];
    for (int c = 0; c < jb; ++c)
        for (int r = p + t; r < s; r += kXExThreads) {
            P.ar[(size_t)(p + c) * lda + r] = P.pr[(size_t)c * lda + r];
            P.ai[(size_t)(p + c) * lda + r] = P.pi[(size_t)c * lda + r];
        }
    if (t == 0)
        P.info[2] += 1;
    __syncthreads();
    for (int j = 0; j < jb; ++j) {
        const int col = p + j;
        const double *cr = P.ar + (size_t)col * lda, *ci = P.ai + (size_t)col * lda;
        double best = -1.0;
        int bidx = col;
        for (int r = col + t; r < s; r += kXExThreads) {
            const double v = rm_hypot(cr[r], ci[r]);
            if (!(v < best)) { // Iterator::max_by keeps the later of equals
                best = v;
                bidx = r;
            }
        }
        rv[t] = best;
        ri[t] = bidx;
        __syncthreads();
        for (int off = kXExThreads / 2; off > 0; off >>= 1) {
            if (t < off) {
                const double v2 = rv[t + off];
                const int i2 = ri[t + off];
                if (v2 > rv[t] || (v2 == rv[t] && i2 > ri[t])) {
                    rv[t] = v2;
                    ri[t] = i2;
                }
            }
            __syncthreads();
        }
        const int piv = ri[0];
        const double pbest = rv[0];
        __syncthreads();
        if (pbest < kSingular) {
            if (t == 0)
                P.info[0] = 1;
            return;
        }
        if (piv != col) {
            for (int c = t; c < jb; c += kXExThreads) {
                const size_t o1 = (size_t)(p + c) * lda + col, o2 = (size_t)(p + c) * lda + piv;
                const double xr = P.ar[o1], xi = P.ai[o1];
                P.ar[o1] = P.ar[o2];
                P.ai[o1] = P.ai[o2];
                P.ar[o2] = xr;
                P.ai[o2] = xi;
            }
            if (t == 0)
                P.info[1] += 1;
        }
        if (t == 0)
            P.ipiv[col] = piv;
        __syncthreads();
        for (int c = j + t; c < jb; c += kXExThreads) {
            prow_r[c] = P.ar[(size_t)(p + c) * lda + col];
            prow_i[c] = P.ai[(size_t)(p + c) * lda + col];
        }
        __syncthreads();
        const double dr = prow_r[j], di = prow_i[j];
        for (int r = col + 1 + t; r < s; r += kXExThreads) {
            double fr, fi;
            rm_div(cr[r], ci[r], dr, di, fr, fi);
            P.ar[(size_t)col * lda + r] = fr;
            P.ai[(size_t)col * lda + r] = fi;
            for (int c = j + 1; c < jb; ++c) {
                const size_t o = (size_t)(p + c) * lda + r;
                const double xr = prow_r[c], xi = prow_i[c];
                const double sr = fr * xr - fi * xi;
                const double si = fr * xi + fi * xr;
                P.ar[o] = P.ar[o] - sr;
                P.ai[o] = P.ai[o] - si;
            }
        }
        __syncthreads();
    }
}

// Work item (problem, block of 64 columns right of the panel, up to and
// including the right-hand side column s): apply the panel's interchanges in
// order, then form the panel rows' U entries: row p + j takes the updates of
// rows p .. p + j - 1 in that order.
__global__ __launch_bounds__(64) void k_ge_u12(const Prob *__restrict__ probs,
                                                const int2 *__restrict__ work, int p)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    const int s = P.s, lda = P.lda;
    if (P.info[0] != 0)
        return;
    const int jb = min(kXP, s - p), c1 = p + jb, t = threadIdx.x;
    __shared__ double lr[kXP][kXP + 1], li[kXP][kXP + 1];
    __shared__ int piv[kXP];
    for (int e = t; e < jb * jb; e += 64) {
        const int c = e / jb, r = e % jb;
        lr[r][c] = P.ar[(size_t)(p + c) * lda + p + r];
        li[r][c] = P.ai[(size_t)(p + c) * lda + p + r];
    }
    if (t < jb)
        piv[t] = P.ipiv[p + t];
    __syncthreads();
    const int col = c1 + wk.y * 64 + t;
    if (col >= s + P.nrhs)
        return;
    double *cr = P.ar + (size_t)col * lda, *ci = P.ai + (size_t)col * lda;
    for (int j = 0; j < jb; ++j) {
        const int q = piv[j];
        if (q != p + j) {
            const double xr = cr[p + j], xi = ci[p + j];
            cr[p + j] = cr[q];
            ci[p + j] = ci[q];
            cr[q] = xr;
            ci[q] = xi;
        }
    }
    double vr[kXP], vi[kXP];
#pragma unroll
    for (int j = 0; j < kXP; ++j) {
        vr[j] = vi[j] = 0.0;
        if (j < jb) {
            vr[j] = cr[p + j];
            vi[j] = ci[p + j];
        }
    }
#pragma unroll
    for (int j = 1; j < kXP; ++j) {
        if (j < jb) {
#pragma unroll
            for (int q = 0; q < j; ++q) {
                const double fr = lr[j][q], fi = li[j][q];
                const double sr = fr * vr[q] - fi * vi[q];
                const double si = fr * vi[q] + fi * vr[q];
                vr[j] = vr[j] - sr;
                vi[j] = vi[j] - si;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kXP; ++j)
        if (j < jb) {
            cr[p + j] = vr[j];
            ci[p + j] = vi[j];
        }
}

// Work item (problem, y, z): the 64x64 tile of rows c1 + 64y.., columns
// c1 + 64z.. (the right-hand side column s included) of the trailing matrix:
// a_rk -= f_rc u_ck for c = p .. p + jb - 1 in order. L21 and U12 staged in
// LDS; each thread owns a 4x4 element block.
__global__ __launch_bounds__(256) void k_ge_update(const Prob *__restrict__ probs,
                                                   const int4 *__restrict__ work, int p)
{
    const int4 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    const int s = P.s, lda = P.lda;
    if (P.info[0] != 0)
        return;
    const int jb = min(kXP, s - p), c1 = p + jb;
    const int r0 = c1 + wk.y * kXT, k0 = c1 + wk.z * kXT;
    __shared__ double lr_[kXP][kXT + 1], li_[kXP][kXT + 1]; // [c][row]
    __shared__ double ur_[kXP][kXT + 1], ui_[kXP][kXT + 1]; // [c][col]
    const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
    for (int e = t; e < jb * kXT; e += 256) {
        const int c = e / kXT, m = e % kXT;
        double a = 0.0, b = 0.0;
        if (r0 + m < s) {
            a = P.ar[(size_t)(p + c) * lda + r0 + m];
            b = P.ai[(size_t)(p + c) * lda + r0 + m];
        }
        lr_[c][m] = a;
        li_[c][m] = b;
    }
    for (int e = t; e < jb * kXT; e += 256) {
        const int m = e / jb, c = e % jb;
        double a = 0.0, b = 0.0;
        if (k0 + m < s + P.nrhs) {
            a = P.ar[(size_t)(k0 + m) * lda + p + c];
            b = P.ai[(size_t)(k0 + m) * lda + p + c];
        }
        ur_[c][m] = a;
        ui_[c][m] = b;
    }
    __syncthreads();
    double cr[4][4], ci[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = r0 + ty + 16 * u, k = k0 + tx + 16 * v;
            cr[u][v] = ci[u][v] = 0.0;
            if (r < s && k < s + P.nrhs) {
                cr[u][v] = P.ar[(size_t)k * lda + r];
                ci[u][v] = P.ai[(size_t)k * lda + r];
            }
        }
    for (int c = 0; c < jb; ++c) {
        double fr[4], fi[4], xr[4], xi[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            fr[u] = lr_[c][ty + 16 * u];
            fi[u] = li_[c][ty + 16 * u];
            xr[u] = ur_[c][tx + 16 * u];
            xi[u] = ui_[c][tx + 16 * u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double sr = fr[u] * xr[v] - fi[u] * xi[v];
                const double si = fr[u] * xi[v] + fi[u] * xr[v];
                cr[u][v] = cr[u][v] - sr;
                ci[u][v] = ci[u][v] - si;
            }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = r0 + ty + 16 * u, k = k0 + tx + 16 * v;
            if (r < s && k < s + P.nrhs) {
                P.ar[(size_t)k * lda + r] = cr[u][v];
                P.ai[(size_t)k * lda + r] = ci[u][v];
            }
        }
}

// ---------------------------------------------------------------------------
// Back substitution in the reference's order (linear.rs:1373-1380)
// ---------------------------------------------------------------------------

constexpr int kXBsThreads = 256;
constexpr int kXBsChunk = 1024;

// One workgroup per (problem, right-hand side). Row i: sum = b_i - a_i,i+1 x_i+1 - ... (ascending
// j, each subtraction rounded), x_i = sum / a_ii. The products of a chunk of
// the row are formed in parallel; one lane runs the subtraction chain.
__global__ __launch_bounds__(kXBsThreads) void k_backsub_exact(const Prob *__restrict__ probs,
                                                               const int2 *__restrict__ work)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    if (P.info[0] != 0)
        return;
    const int s = P.s, lda = P.lda, t = threadIdx.x;
    double *x = P.x + 2 * (size_t)s * wk.y;
    const size_t bcol = (size_t)(s + wk.y) * lda;
    __shared__ double pr[kXBsChunk], pi[kXBsChunk];
    for (int i = s - 1; i >= 0; --i) {
        double sr = P.ar[bcol + i], si = P.ai[bcol + i];
        for (int j0 = i + 1; j0 < s; j0 += kXBsChunk) {
            const int cnt = min(kXBsChunk, s - j0);
            __syncthreads();
            for (int e = t; e < cnt; e += kXBsThreads) {
                const int j = j0 + e;
                const double ur = P.ar[(size_t)j * lda + i], ui = P.ai[(size_t)j * lda + i];
                const double xr = x[2 * j], xi = x[2 * j + 1];
                pr[e] = ur * xr - ui * xi;
                pi[e] = ur * xi + ui * xr;
            }
            __syncthreads();
            if (t == 0)
                for (int e = 0; e < cnt; ++e) {
                    sr = sr - pr[e];
                    si = si - pi[e];
                }
        }
        if (t == 0) {
            double qr, qi;
            rm_div(sr, si, P.ar[(size_t)i * lda + i], P.ai[(size_t)i * lda + i], qr, qi);
            x[2 * i] = qr;
            x[2 * i + 1] = qi;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Probe of the device elementary functions (parity tests): fn 0 sin(x),
// 1 cos(x), 2 hypot(x, y), 3 sqrt(x), 4 x / y, 5 and 6 the sin and cos of
// sincos(x).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_refmath_probe(int fn, const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       double *__restrict__ out, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    double v;
    if (fn == 0)
        v = rm_sin(x[i], g_sincos_tab);
    else if (fn == 1)
        v = rm_cos(x[i], g_sincos_tab);
    else if (fn == 2)
        v = rm_hypot(x[i], y[i]);
    else if (fn == 3)
        v = sqrt(x[i]);
    else if (fn == 4)
        v = x[i] / y[i];
    else {
        double sx, cx;
        rm_sincos(x[i], g_sincos_tab, sx, cx);
        v = fn == 5 ? sx : cx;
    }
    out[i] = v;
}

} // namespace fnec
