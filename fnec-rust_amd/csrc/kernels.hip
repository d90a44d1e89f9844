// HIP kernels for the Hallén path on gfx950 (CDNA4).
//
//   k_fill_z        Z fill (matrix.rs:186-218, 297-331): FP64 Gauss-Legendre
//                   reduced-kernel quadrature + PEC / scalar-Gamma image.
//   k_gram          G = Z^H Z on v_mfma_f64_16x16x4f64, lower-triangle tiles.
//   k_gram_cols     homogeneous columns, M^H y (linear.rs:735-780).
//   k_gram_fix      W x W block, constraint rows, lambda (linear.rs:744-785).
//   (the blocked LU kernels are templates in lu.h, instantiated by engine.hip)
//   k_backsub       U x = b' (blocked, one workgroup per problem and right-hand side).
//   k_residual_*    ||Z I - C cos - rhs|| and diag spread (solve_session.rs:118-207).
//   k_rp_*          far field + normalisation integral (farfield.rs:265-647).
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.h"
#include "phase_table.h"

namespace fnec {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// v_mfma_f64_16x16x4f64 operand layout: lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator item v of lane l is D[(l/16) + 4v][l%16].
__device__ __forceinline__ int acc_row(int lane, int v) { return (lane >> 4) + 4 * v; }

// Gauss-Legendre nodes and weights, the 15-digit literals of matrix.rs:39-72.
__constant__ double c_gl4n[4] = {-0.861136311594953, -0.339981043584856, 0.339981043584856,
                                 0.861136311594953};
__constant__ double c_gl4w[4] = {0.347854845137454, 0.652145154862626, 0.652145154862626,
                                 0.347854845137454};
__constant__ double c_gl8n[8] = {-0.960289856497536, -0.796666477413627, -0.525532409916329,
                                 -0.183434642495650, 0.183434642495650,  0.525532409916329,
                                 0.796666477413627,  0.960289856497536};
__constant__ double c_gl8w[8] = {0.101228536290376, 0.222381034453374, 0.313706645877887,
                                 0.362683783378362, 0.362683783378362, 0.313706645877887,
                                 0.222381034453374, 0.101228536290376};

// ---------------------------------------------------------------------------
// Z fill
// ---------------------------------------------------------------------------

// (sin, cos)(j pi/32), j = 0 .. 63, for sincos_tab
__constant__ double2 c_phase_tab[FNEC_PHASE_TABLE_LEN] = {FNEC_PHASE_TABLE_INIT};

// Copies the table into a kernel's LDS (the caller synchronises).
__device__ __forceinline__ void load_phase_tab(double2 *tab)
{
    for (int e = threadIdx.x; e < FNEC_PHASE_TABLE_LEN; e += blockDim.x)
        tab[e] = c_phase_tab[e];
}

// sin and cos of a phase by Tang's table method: x = n pi/32 + r with
// n = rint(32 x / pi) and r from a three-part Cody-Waite reduction (the parts
// of sincos_phase's pi/2 split divided by 16, so n P1 and n P2 stay exact for
// |n| < 2^20), |r| <= pi/64; the Taylor polynomials of sin r (to r^9) and
// cos r - 1 (to r^8), truncated below 1e-22 and 3e-20; then
//   sin x = S + (S (cos r - 1) + C sin r),  cos x = C + (C (cos r - 1) - S sin r)
// with (S, C) = (sin, cos)(n pi/32) from the table (n mod 64). About 1 ulp.
// kLarge: arguments beyond 1e5 take the library routine; without it the
// caller guarantees |x| < 1e5.
template <bool kLarge>
__device__ __forceinline__ void sincos_tab(double x, const double2 *tab, double &s, double &c)
{
    if (kLarge && !(fabs(x) < 1e5)) {
        sincos(x, &s, &c);
        return;
    }
    const double n = rint(x * 10.185916357881302); // 32 / pi
    double r = fma(-n, 0x1.921fb544p-4, x);
    r = fma(-n, 0x1.0b4611a6p-38, r);
    r = fma(-n, 0x1.3198a2e037073p-73, r);
    const double r2 = r * r;
    double ps = fma(r2, 2.7557319223985893e-06, -0.0001984126984126984); // 1/9!, -1/7!
    ps = fma(r2, ps, 0.008333333333333333);
    ps = fma(r2, ps, -0.16666666666666666);
    const double sr = fma(r * r2, ps, r);
    double pc = fma(r2, 2.48015873015873e-05, -0.001388888888888889); // 1/8!, -1/6!
    pc = fma(r2, pc, 0.041666666666666664);
    pc = fma(r2, pc, -0.5);
    const double cm1 = r2 * pc;
    const double2 t = tab[(int)n & 63]; // two's complement: n mod 64 also for n < 0
    s = fma(t.y, sr, fma(t.x, cm1, t.x));
    c = fma(-t.x, sr, fma(t.y, cm1, t.y));
}

// 1/sqrt(r2) for r2 >= a^2 > 0 (finite and normal, so rsqrt()'s special-case
// handling is not needed): the hardware estimate y and one third-order
// correction y += y e (1/2 + 3/8 e), e = 1 - r2 y^2, which takes the
// estimate's relative error d to about (5/16) d^3, below double rounding.
__device__ __forceinline__ double rsqrt_fill(double r2)
{
    double ir = __builtin_amdgcn_rsq(r2);
    const double e = fma(-r2 * ir, ir, 1.0);
    return fma(ir * e, fma(0.375, e, 0.5), ir);
}

// Self element: 4-point GL on G - 1/R plus 2 ln((h + R_end)/a) (matrix.rs:302-316).
template <bool kLarge>
__device__ __forceinline__ void self_term(double h, double a, double k, const double2 *tab, double &re,
                                          double &im)
{
    double sr = 0.0, si = 0.0;
#pragma unroll 1
    for (int m = 0; m < 4; ++m) {
        double l = c_gl4n[m] * h;
        double r = sqrt(l * l + a * a);
        double s, c;
        sincos_tab<kLarge>(k * r, tab, s, c);
        double inv = 1.0 / r;
        sr += c_gl4w[m] * (c * inv - inv);
        si -= c_gl4w[m] * (s * inv);
    }
    double rend = sqrt(h * h + a * a);
    re = sr * h + 2.0 * log((h + rend) / a);
    im = si * h;
}

// Off-diagonal element integral  half * sum_m w_m exp(-jkR_m)/R_m  over the
// source segment (src mid s, dir d, half length h, radius a) seen from obs;
// sc[m] = node_m h.
// kSinDeg 0: every point's phase k R_m reduced (sincos_tab).
// kSinDeg 9, 11, 13, the centred form: exp(-jkR_m) = exp(-jkR_0) exp(-jd_m)
// with R_0 the distance to the source centre and d_m = k R_m - k R_0. Since
// |R_m - R_0| <= |sc_m| <= 0.9603 h, |d_m| <= 0.9603 k h needs no range
// reduction (sincos_small of that degree), and the full phase's sin and cos
// are taken once per element.
template <int kSinDeg, bool kLarge>
__device__ __forceinline__ void quad8(double ox, double oy, double oz, double sx, double sy,
                                      double sz, double dx, double dy, double dz, double h,
                                      double a2, double k, const double *sc8, const double2 *tab,
                                      double &re, double &im)
{
    // R^2 at the point sc along the unit direction: |d - sc dir|^2 + a^2 =
    // (|d|^2 + a^2) + sc (sc - 2 d.dir), d = observation - source centre
    const double ex = ox - sx, ey = oy - sy, ez = oz - sz;
    const double A = ex * ex + ey * ey + ez * ez + a2;
    const double B = -2.0 * (ex * dx + ey * dy + ez * dz);
    double sr = 0.0, si = 0.0;
    if constexpr (kSinDeg > 0) {
        const double ph0 = k * (A * rsqrt_fill(A));
#pragma unroll 4
        for (int m = 0; m < 8; ++m) {
            const double sc = sc8[m];
            const double r2 = fma(sc, sc + B, A);
            const double ir = rsqrt_fill(r2);
            double s, c;
            sincos_small<kSinDeg>(fma(k, r2 * ir, -ph0), s, c);
            const double wr = c_gl8w[m] * ir;
            sr = fma(wr, c, sr);
            si = fma(-wr, s, si);
        }
        double s0, c0;
        sincos_tab<kLarge>(ph0, tab, s0, c0);
        re = fma(sr, c0, si * s0) * h;
        im = fma(si, c0, -(sr * s0)) * h;
    } else {
        // two points per iteration: two independent dependency chains per lane
#pragma unroll 2
        for (int m = 0; m < 8; ++m) {
            const double sc = sc8[m];
            const double r2 = fma(sc, sc + B, A);
            const double ir = rsqrt_fill(r2);
            double s, c;
            sincos_tab<kLarge>(k * (r2 * ir), tab, s, c);
            const double wr = c_gl8w[m] * ir;
            sr += wr * c;
            si -= wr * s;
        }
        re = sr * h;
        im = si * h;
    }
}

// The direct term of an element (self: the self term) and, with ground, the
// image term, each scaled by its cos alpha; the image of a finite ground by
// the scalar Gamma.
template <int kSinDeg, bool kLarge>
__device__ __forceinline__ void element(const Prob &P, double ox, double oy, double oz, double odx,
                                        double ody, double odz, const double *sg, const double *sc8,
                                        const double2 *tab, bool self, double &re, double &im)
{
    const double sx = sg[0], sy = sg[1], sz = sg[2], dx = sg[3], dy = sg[4], dz = sg[5];
    const double h = sg[6] * 0.5, a2 = sg[7] * sg[7], k = P.k;
    if (self)
        self_term<kLarge>(h, sg[7], k, tab, re, im);
    else
        quad8<kSinDeg, kLarge>(ox, oy, oz, sx, sy, sz, dx, dy, dz, h, a2, k, sc8, tab, re, im);
    const double ca = odx * dx + ody * dy + odz * dz;
    re *= ca;
    im *= ca;
    if (P.ground != 0) {
        // image: z -> -z, direction (-dx, -dy, +dz) (matrix.rs:265-277)
        double ir, ii;
        quad8<kSinDeg, kLarge>(ox, oy, oz, sx, sy, -sz, -dx, -dy, dz, h, a2, k, sc8, tab, ir, ii);
        const double ica = odx * (-dx) + ody * (-dy) + odz * dz;
        ir *= ica;
        ii *= ica;
        if (P.ground == 2) {
            const double gr = P.gam_re * ir - P.gam_im * ii;
            const double gi = P.gam_re * ii + P.gam_im * ir;
            ir = gr;
            ii = gi;
        }
        re += ir;
        im += ii;
    }
}


// Regular work lists from segment records (first item, problem, ny,
// nz | gy << 24): items (problem, y, z, 0). gy = 0: y-major order. gy > 0:
// groups of gy rows, each group in z-major order (z outer, y inner), so that
// workgroups running together share their column operands as well as their
// row operands in L2. One workgroup per segment.
__global__ __launch_bounds__(256) void k_gen_items(const int4 *__restrict__ segs, int4 *__restrict__ out)
{
    const int4 sg = segs[blockIdx.x];
    const int ny = sg.z, nz = sg.w & 0xffffff, gy = sg.w >> 24;
    const int total = ny * nz;
    for (int e = threadIdx.x; e < total; e += 256) {
        int y = e / nz, z = e % nz;
        if (gy > 0) {
            const int g = e / (gy * nz), i = e - g * gy * nz;
            const int rows = min(gy, ny - g * gy); // rows in this group
            y = g * gy + i % rows;
            z = i / rows;
        }
        out[(size_t)sg.x + e] = make_int4(sg.y, y, z, 0);
    }
}

constexpr int kFillRows = 64;
constexpr int kFillCols = 16;

// Work item: (problem, row tile, column tile). Each thread owns one
// observation row and four source columns; a wavefront writes 64 consecutive
// rows of one column, so the column-major stores are fully coalesced. The
// column's k h (uniform in a wavefront) picks the element form: the centred
// form with the shortest polynomial its phase offsets allow (k h <= 0.126,
// 0.26, 0.51), else every point's phase reduced. kLarge keeps the library
// sin / cos for phases beyond 1e5 (the host bounds every phase by
// k (2 max|p| + max a) and launches the kLarge form only when that reaches
// 9e4).
template <bool kLarge> __device__ __forceinline__ void fill_z_body(const Prob *__restrict__ probs,
                                                                  const int4 *__restrict__ work)
{
    const int4 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    const int i0 = wk.y * kFillRows, j0 = wk.z * kFillCols;
    __shared__ double so[kFillRows][6];
    __shared__ double ss[kFillCols][8];
    __shared__ double sn[kFillCols][8]; // node_m h of each column
    __shared__ double2 tab[FNEC_PHASE_TABLE_LEN];
    const int t = threadIdx.x;
    load_phase_tab(tab);
    for (int e = t; e < kFillRows * 6; e += 256) {
        int row = e / 6, c = e % 6;
        int i = i0 + row;
        so[row][c] = i < P.n ? P.seg[(size_t)i * kSegStride + c] : 0.0;
    }
    for (int e = t; e < kFillCols * 8; e += 256) {
        int col = e / 8, c = e % 8;
        int j = j0 + col;
        ss[col][c] = j < P.n ? P.seg[(size_t)j * kSegStride + c] : 1.0;
        sn[col][c] = c_gl8n[c] * (j < P.n ? P.seg[(size_t)j * kSegStride + 6] * 0.5 : 0.5);
    }
    __syncthreads();
    const int row = t & 63;
    const int i = i0 + row;
    if (i >= P.n) {
        // mixed path: the FP32 copy's padding rows [n, ldz) are 0, so that the
        // factorisation's kernels read and write whole float4 row quads
        if (P.fr)
            for (int q = 0; q < 4; ++q) {
                const int j = j0 + (t >> 6) * 4 + q;
                if (j >= P.n)
                    break;
                P.fr[(size_t)j * P.ldz + i] = 0.f;
                P.fi[(size_t)j * P.ldz + i] = 0.f;
            }
        return;
    }
    const double ox = so[row][0], oy = so[row][1], oz = so[row][2];
    const double odx = so[row][3], ody = so[row][4], odz = so[row][5];
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
        const int col = (t >> 6) * 4 + q;
        const int j = j0 + col;
        if (j >= P.n)
            break;
        const double *sg = ss[col];
        const double kh = P.k * sg[6] * 0.5;
        const bool self = i == j;
        double re, im;
        if (!P.fill_centred || !(kh <= 0.51))
            element<0, kLarge>(P, ox, oy, oz, odx, ody, odz, sg, sn[col], tab, self, re, im);
        else if (kh <= 0.126)
            element<9, kLarge>(P, ox, oy, oz, odx, ody, odz, sg, sn[col], tab, self, re, im);
        else if (kh <= 0.26)
            element<11, kLarge>(P, ox, oy, oz, odx, ody, odz, sg, sn[col], tab, self, re, im);
        else
            element<13, kLarge>(P, ox, oy, oz, odx, ody, odz, sg, sn[col], tab, self, re, im);
        const size_t o = (size_t)j * P.ldz + i;
        P.zr[o] = re;
        P.zi[o] = im;
        if (P.fr) { // mixed path: FP32 copy for the factorisation
            P.fr[o] = (float)re;
            P.fi[o] = (float)im;
        }
    }
}

__global__ __launch_bounds__(256) void k_fill_z(const Prob *__restrict__ probs, const int4 *__restrict__ work)
{
    fill_z_body<false>(probs, work);
}

__global__ __launch_bounds__(256) void k_fill_z_large(const Prob *__restrict__ probs,
                                                      const int4 *__restrict__ work)
{
    fill_z_body<true>(probs, work);
}

// The LD / TL / NT stamps of one problem, Z[row][col] += delta
// (ZMatrix::add_to_diagonal / add_to_entry, matrix.rs:131-147; applied after
// assembly, solve_session.rs:1285-1309). The mixed path's FP32 copy of an
// entry is rounded again from the stamped FP64 value, as the fill rounds it.
__global__ __launch_bounds__(256) void k_stamp_z(const Prob *__restrict__ probs,
                                                 const int *__restrict__ plist)
{
    // The stamps arrive grouped by matrix entry (their order kept within an
    // entry): the thread at the head of a group adds that group's deltas in
    // order, so each entry sees the sequence of additions of the CPU path.
    const Prob &P = probs[plist[blockIdx.x]];
    for (int q = threadIdx.x; q < P.nstamps; q += blockDim.x) {
        const double *st = P.stamps + 4 * (size_t)q;
        if (q > 0 && st[-4] == st[0] && st[-3] == st[1])
            continue;
        const size_t o = (size_t)st[1] * P.ldz + (size_t)st[0]; // column-major: col * ldz + row
        double re = P.zr[o], im = P.zi[o];
        for (int k = q; k < P.nstamps; ++k) {
            const double *sk = P.stamps + 4 * (size_t)k;
            if (sk[0] != st[0] || sk[1] != st[1])
                break;
            re += sk[2];
            im += sk[3];
        }
        P.zr[o] = re;
        P.zi[o] = im;
        if (P.fr) {
            P.fr[o] = (float)re;
            P.fi[o] = (float)im;
        }
    }
}

// ---------------------------------------------------------------------------
// Gram G = Z^H Z (complex HERK on FP64 MFMA)
// ---------------------------------------------------------------------------

// Work item: (problem, bi, bj) with bi >= bj, 64x64 output tile. Four waves
// in a 2x2 arrangement each own a 32x32 block = 2x2 MFMA 16x16 tiles, with
// separate real and imaginary accumulators:
//   Re += Zr_i Zr_j + Zi_i Zi_j ;  Im += Zr_i Zi_j - Zi_i Zr_j
constexpr int kLdsPad = 4;

__global__ __launch_bounds__(256) void k_gram(const Prob *__restrict__ probs,
                                              const int4 *__restrict__ work)
{
    const int4 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    const int i0 = wk.y * kTile, j0 = wk.z * kTile;
    const int n = P.n;
    __shared__ double sa_r[kGramK][kTile + kLdsPad], sa_i[kGramK][kTile + kLdsPad];
    __shared__ double sb_r[kGramK][kTile + kLdsPad], sb_i[kGramK][kTile + kLdsPad];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    d4 accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accr[a][b] = d4{0.0, 0.0, 0.0, 0.0};
            acci[a][b] = d4{0.0, 0.0, 0.0, 0.0};
        }
    const int lrow = lane & 15, lk = lane >> 4;
    for (int k0 = 0; k0 < n; k0 += kGramK) {
        // 64 columns x 16 rows per plane and operand; 4 elements per thread.
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            const int c = e >> 4, kk = e & 15;
            const int kr = k0 + kk;
            const int ci = i0 + c, cj = j0 + c;
            double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
            if (kr < n) {
                if (ci < n) {
                    ar = P.zr[(size_t)ci * P.ldz + kr];
                    ai = P.zi[(size_t)ci * P.ldz + kr];
                }
                if (cj < n) {
                    br = P.zr[(size_t)cj * P.ldz + kr];
                    bi = P.zi[(size_t)cj * P.ldz + kr];
                }
            }
            sa_r[kk][c] = ar;
            sa_i[kk][c] = ai;
            sb_r[kk][c] = br;
            sb_i[kk][c] = bi;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kGramK; kk += 4) {
            double xr[2], xi[2], yr[2], yi[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                xr[a] = sa_r[kk + lk][wr * 32 + a * 16 + lrow];
                xi[a] = sa_i[kk + lk][wr * 32 + a * 16 + lrow];
                yr[a] = sb_r[kk + lk][wc * 32 + a * 16 + lrow];
                yi[a] = sb_i[kk + lk][wc * 32 + a * 16 + lrow];
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    accr[a][b] = mfma(xr[a], yr[b], accr[a][b]);
                    accr[a][b] = mfma(xi[a], yi[b], accr[a][b]);
                    acci[a][b] = mfma(xr[a], yi[b], acci[a][b]);
                    acci[a][b] = mfma(-xi[a], yr[b], acci[a][b]);
                }
        }
        __syncthreads();
    }
    const bool diag = wk.y == wk.z;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = i0 + wr * 32 + a * 16 + acc_row(lane, v);
                const int gj = j0 + wc * 32 + b * 16 + lrow;
                if (gi >= n || gj >= n)
                    continue;
                if (diag && gi < gj)
                    continue;
                const double re = accr[a][b][v];
                // the diagonal of Z^H Z is real: every term re*im - im*re cancels exactly
                const double im = gi == gj ? 0.0 : acci[a][b][v];
                P.ar[(size_t)gj * P.lda + gi] = re;
                P.ai[(size_t)gj * P.lda + gi] = im;
                if (gi != gj) {
                    P.ar[(size_t)gi * P.lda + gj] = re;
                    P.ai[(size_t)gi * P.lda + gj] = -im;
                }
            }
}

__device__ __forceinline__ double wave_sum(double v)
{
    // Fixed butterfly order: deterministic for a given lane assignment.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// One wavefront per Z column i: G[i][n+w] = -sum_{r in w} conj(Z_ri) cos_r,
// its Hermitian mirror, and (M^H y)_i = sum_r conj(Z_ri) rhs_r.
__global__ __launch_bounds__(256) void k_gram_cols(const Prob *__restrict__ probs,
                                                   const int2 *__restrict__ work)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    const int lane = threadIdx.x & 63;
    const int i = wk.y * 4 + (threadIdx.x >> 6);
    if (i >= P.n)
        return;
    const double *zr = P.zr + (size_t)i * P.ldz, *zi = P.zi + (size_t)i * P.ldz;
    const int s = P.s;
    // M^H y over the n segment rows (constraint rows carry y = 0)
    double yr = 0.0, yi = 0.0;
    for (int r = lane; r < P.n; r += 64) {
        const double a = zr[r], b = zi[r];
        const double c = P.rhs[2 * r], d = P.rhs[2 * r + 1];
        yr += a * c + b * d;
        yi += a * d - b * c;
    }
    yr = wave_sum(yr);
    yi = wave_sum(yi);
    if (lane == 0) {
        P.ar[(size_t)s * P.lda + i] = yr;
        P.ai[(size_t)s * P.lda + i] = yi;
    }
    // homogeneous columns, accumulated run by run in row order
    for (int w = 0; w < P.w; ++w) {
        double gr = 0.0, gi = 0.0;
        for (int q = 0; q < P.nruns; ++q) {
            const int rb = P.runs[3 * q], re = P.runs[3 * q + 1], col = P.runs[3 * q + 2];
            if (col != w)
                continue;
            for (int r = rb + lane; r < re; r += 64) {
                const double c = P.cosv[r];
                gr -= zr[r] * c;
                gi += zi[r] * c;
            }
        }
        gr = wave_sum(gr);
        gi = wave_sum(gi);
        if (lane == 0) {
            const int cw = P.n + w;
            P.ar[(size_t)cw * P.lda + i] = gr; // G[i][n+w]
            P.ai[(size_t)cw * P.lda + i] = gi;
            P.ar[(size_t)i * P.lda + cw] = gr; // G[n+w][i] = conj
            P.ai[(size_t)i * P.lda + cw] = -gi;
        }
    }
}

// One workgroup per problem: the W x W block, (M^H y) for the homogeneous
// columns, the sparse constraint rows (K^H K) and lambda on the diagonal.
__global__ __launch_bounds__(256) void k_gram_fix(const Prob *__restrict__ probs,
                                                  const int *__restrict__ plist)
{
    const Prob &P = probs[plist[blockIdx.x]];
    const int t = threadIdx.x;
    const int n = P.n, s = P.s;
    __shared__ double red[256];
    for (int w = 0; w < P.w; ++w) {
        double c2 = 0.0, yr = 0.0, yi = 0.0;
        for (int q = 0; q < P.nruns; ++q) {
            const int rb = P.runs[3 * q], re = P.runs[3 * q + 1], col = P.runs[3 * q + 2];
            if (col != w)
                continue;
            for (int r = rb + t; r < re; r += 256) {
                const double c = P.cosv[r];
                c2 += c * c;
                yr -= c * P.rhs[2 * r];
                yi -= c * P.rhs[2 * r + 1];
            }
        }
        double vals[3] = {c2, yr, yi};
        for (int v = 0; v < 3; ++v) {
            red[t] = vals[v];
            __syncthreads();
            for (int off = 128; off > 0; off >>= 1) {
                if (t < off)
                    red[t] += red[t + off];
                __syncthreads();
            }
            vals[v] = red[0];
            __syncthreads();
        }
        if (t == 0) {
            const int cw = n + w;
            for (int v = 0; v < P.w; ++v) {
                P.ar[(size_t)(n + v) * P.lda + cw] = v == w ? vals[0] : 0.0;
                P.ai[(size_t)(n + v) * P.lda + cw] = 0.0;
            }
            P.ar[(size_t)s * P.lda + cw] = vals[1];
            P.ai[(size_t)s * P.lda + cw] = vals[2];
        }
    }
    __syncthreads();
    if (t == 0) {
        for (int c = 0; c < P.ncons; ++c) {
            const int a = P.cons[2 * c], b = P.cons[2 * c + 1];
            const double sg = P.cons_sign[c];
            if (b < 0 || b == a) {
                const double v = b < 0 ? 1.0 : sg * sg;
                P.ar[(size_t)a * P.lda + a] += v;
            } else {
                P.ar[(size_t)a * P.lda + a] += 1.0;
                P.ar[(size_t)b * P.lda + a] += sg; // G[a][b]
                P.ar[(size_t)a * P.lda + b] += sg; // G[b][a]
                P.ar[(size_t)b * P.lda + b] += sg * sg;
            }
        }
    }
    __syncthreads();
    for (int i = t; i < s; i += 256)
        P.ar[(size_t)i * P.lda + i] += kLambda;
}

// Back substitution U x = b' with b' a transformed right-hand side column.
// Blocks of 64 rows from the bottom; one workgroup per (problem, right-hand
// side).
constexpr int kBsBlock = 64;

__global__ __launch_bounds__(kBsThreads) void k_backsub(const Prob *__restrict__ probs,
                                                 const int2 *__restrict__ work)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    if (P.info[0] != 0)
        return;
    const int s = P.s, lda = P.lda, t = threadIdx.x;
    double *x = P.x + 2 * (size_t)s * wk.y;
    const size_t bcol = (size_t)(s + wk.y) * lda;
    for (int i = t; i < s; i += kBsThreads) {
        x[2 * i] = P.ar[bcol + i];
        x[2 * i + 1] = P.ai[bcol + i];
    }
    __syncthreads();
    __shared__ double xb_r[kBsBlock], xb_i[kBsBlock];
    // the block's triangle, [column i - r0][row m - r0]
    __shared__ double ut_r[kBsBlock][kBsBlock + 1], ut_i[kBsBlock][kBsBlock + 1];
    for (int r1 = s; r1 > 0; r1 -= kBsBlock) {
        const int r0 = max(0, r1 - kBsBlock), nb = r1 - r0;
        // stage the block's columns (rows r0 .. r1) and its part of x in LDS
        for (int e = t; e < nb * nb; e += kBsThreads) {
            const int c = e / nb, m = e % nb;
            ut_r[c][m] = P.ar[(size_t)(r0 + c) * lda + r0 + m];
            ut_i[c][m] = P.ai[(size_t)(r0 + c) * lda + r0 + m];
        }
        for (int m = t; m < nb; m += kBsThreads) {
            xb_r[m] = x[2 * (r0 + m)];
            xb_i[m] = x[2 * (r0 + m) + 1];
        }
        __syncthreads();
        // triangle [r0, r1): serial in i, parallel over the rows above it
        for (int i = nb - 1; i >= 0; --i) {
            if (t == 0) {
                const double dr = ut_r[i][i], di = ut_i[i][i];
                const double nsq = dr * dr + di * di;
                const double br = xb_r[i], bi = xb_i[i];
                xb_r[i] = (br * dr + bi * di) / nsq;
                xb_i[i] = (bi * dr - br * di) / nsq;
            }
            __syncthreads();
            const double vr = xb_r[i], vi = xb_i[i];
            if (t < i) {
                const double ur = ut_r[i][t], ui = ut_i[i][t];
                xb_r[t] -= ur * vr - ui * vi;
                xb_i[t] -= ur * vi + ui * vr;
            }
            __syncthreads();
        }
        for (int m = t; m < nb; m += kBsThreads) {
            x[2 * (r0 + m)] = xb_r[m];
            x[2 * (r0 + m) + 1] = xb_i[m];
        }
        // rows above the block
        for (int m = t; m < r0; m += kBsThreads) {
            double sr = 0.0, si = 0.0;
#pragma unroll 16
            for (int i = r0; i < r1; ++i) {
                const double ur = P.ar[(size_t)i * lda + m], ui = P.ai[(size_t)i * lda + m];
                const double vr = xb_r[i - r0], vi = xb_i[i - r0];
                sr += ur * vr - ui * vi;
                si += ur * vi + ui * vr;
            }
            x[2 * m] -= sr;
            x[2 * m + 1] -= si;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Residual ||Z I - C cos - rhs||, ||rhs|| and the diagonal spread
// ---------------------------------------------------------------------------

constexpr int kResRows = 256;

// Work item: (problem, row block). Thread per row; partial sums per block.
__global__ __launch_bounds__(256) void k_residual_rows(const Prob *__restrict__ probs,
                                                       const int2 *__restrict__ work)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    if (P.info[0] != 0)
        return;
    const int t = threadIdx.x;
    const int row = wk.y * kResRows + t;
    __shared__ double red[3][256];
    double r2 = 0.0, b2 = 0.0, dmag = 0.0;
    // the residual is that of the delta-gap system; plane-wave and
    // current-source solves report only the diagonal spread, as the
    // reference does (solve_session.rs:1327-1337)
    if (row < P.n && P.variant != 0) {
        dmag = hypot(P.zr[(size_t)row * P.ldz + row], P.zi[(size_t)row * P.ldz + row]);
    } else if (row < P.n) {
        double sr = 0.0, si = 0.0;
        for (int c = 0; c < P.n; ++c) {
            const double zr = P.zr[(size_t)c * P.ldz + row], zi = P.zi[(size_t)c * P.ldz + row];
            const double xr = P.x[2 * c], xi = P.x[2 * c + 1];
            sr += zr * xr - zi * xi;
            si += zr * xi + zi * xr;
        }
        int col = 0;
        for (int q = 0; q < P.nruns; ++q)
            if (row >= P.runs[3 * q] && row < P.runs[3 * q + 1])
                col = P.runs[3 * q + 2];
        const double cr = P.x[2 * (P.n + col)], ci = P.x[2 * (P.n + col) + 1];
        const double cv = P.cosv[row];
        const double er = sr - cr * cv - P.rhs[2 * row];
        const double ei = si - ci * cv - P.rhs[2 * row + 1];
        r2 = er * er + ei * ei;
        b2 = P.rhs[2 * row] * P.rhs[2 * row] + P.rhs[2 * row + 1] * P.rhs[2 * row + 1];
        dmag = hypot(P.zr[(size_t)row * P.ldz + row], P.zi[(size_t)row * P.ldz + row]);
    }
    red[0][t] = r2;
    red[1][t] = b2;
    red[2][t] = row < P.n ? dmag : 0.0;
    __syncthreads();
    // sums for the norms; max for the diagonal, min kept separately below
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
            red[0][t] += red[0][t + off];
            red[1][t] += red[1][t + off];
            red[2][t] = fmax(red[2][t], red[2][t + off]);
        }
        __syncthreads();
    }
    __shared__ double mn[256];
    mn[t] = row < P.n ? dmag : INFINITY;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off)
            mn[t] = fmin(mn[t], mn[t + off]);
        __syncthreads();
    }
    if (t == 0) {
        double *o = P.scratch + 4 * (size_t)wk.y;
        o[0] = red[0][0];
        o[1] = red[1][0];
        o[2] = red[2][0];
        o[3] = mn[0];
    }
}

__global__ __launch_bounds__(64) void k_residual_finish(const Prob *__restrict__ probs,
                                                        const int *__restrict__ plist)
{
    const Prob &P = probs[plist[blockIdx.x]];
    if (threadIdx.x != 0)
        return;
    const int nb = (P.n + kResRows - 1) / kResRows;
    double r2 = 0.0, b2 = 0.0, mx = 0.0, mnv = INFINITY;
    for (int b = 0; b < nb; ++b) {
        const double *o = P.scratch + 4 * (size_t)b;
        r2 += o[0];
        b2 += o[1];
        mx = fmax(mx, o[2]);
        mnv = fmin(mnv, o[3]);
    }
    const double res = sqrt(r2), den = sqrt(b2);
    P.stats[0] = res;
    P.stats[1] = den > 0.0 ? res / den : res;
    double spread;
    if (!isfinite(mx) || !isfinite(mnv))
        spread = NAN;
    else if (mx == 0.0)
        spread = 0.0;
    else
        spread = mx / fmax(mnv, 1e-30);
    P.stats[2] = spread;
}

// ---------------------------------------------------------------------------
// Radiation pattern
// ---------------------------------------------------------------------------

constexpr int kRpThreads = 256;

__device__ __forceinline__ void cdiv(double ar, double ai, double br, double bi, double &cr,
                                     double &ci)
{
    const double ns = br * br + bi * bi;
    cr = (ar * br + ai * bi) / ns;
    ci = (ai * br - ar * bi) / ns;
}

__device__ __forceinline__ void csqrt_polar(double re, double im, double &sr, double &si)
{
    if (im == 0.0) {
        if (!signbit(re)) {
            sr = sqrt(re);
            si = im;
        } else {
            sr = 0.0;
            si = signbit(im) ? -sqrt(-re) : sqrt(-re);
        }
        return;
    }
    if (re == 0.0) {
        const double x = sqrt(fabs(im) / 2.0);
        sr = x;
        si = signbit(im) ? -x : x;
        return;
    }
    const double r = hypot(re, im), th = atan2(im, re);
    const double rs = sqrt(r);
    double s, c;
    sincos(th / 2.0, &s, &c);
    sr = rs * c;
    si = rs * s;
}

// Fresnel coefficients (farfield.rs:333-344).
__device__ void fresnel(double th_deg, double eps_r, double sigma, double k, double &gvr,
                        double &gvi, double &ghr, double &ghi)
{
    const double omega = k * kC0;
    const double er = eps_r, ei = -sigma / (omega * kEps0FarField);
    const double th = th_deg * (kPi / 180.0);
    double st, ct;
    sincos(th, &st, &ct);
    const double sin2 = st * st;
    double rr, ri;
    csqrt_polar(er - sin2, ei, rr, ri);
    const double ecr = er * ct, eci = ei * ct;
    cdiv(ecr - rr, eci - ri, ecr + rr, eci + ri, gvr, gvi);
    cdiv(ct - rr, -ri, ct + rr, ri, ghr, ghi);
}

// Segments staged in LDS as {mid x,y,z, dir x,y,z, length, I re, I im}.
constexpr int kRpStage = 10; // mid xyz, dir xyz, length, run flag, current x length (re, im)
constexpr int kRpChunk = 256;

struct Dir {
    double rx, ry, rz, tx, ty, tz, px, py;
};

__device__ __forceinline__ Dir unit_vectors(double th_deg, double ph_deg)
{
    const double th = th_deg * (kPi / 180.0), ph = ph_deg * (kPi / 180.0);
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    Dir d;
    d.rx = st * cp; d.ry = st * sp; d.rz = ct;
    d.tx = ct * cp; d.ty = ct * sp; d.tz = -st;
    d.px = -sp; d.py = cp;
    return d;
}

// Accumulate one LDS chunk of segments into (F_theta, F_phi) for direction u,
// in segment index order (farfield.rs:277-288, :526-543, :368-385).
// Far-field sum over one staged chunk of segments. The phase factor
// e^{j k u.mid} of a segment that continues a uniform run (flag in slot 7:
// midpoint = previous midpoint + length * dir) is the previous one times the
// run's rotation e^{j k length u.dir}; it is evaluated directly at the start
// of a run, at the start of a chain and every kRpAnchor segments of it, so a
// rotated factor is at most kRpAnchor - 1 products from a direct one. Along a
// run the direction and length are the same, so the projections of dir on
// the theta and phi unit vectors are computed at those evaluation points only.
// The chunk is walked as two interleaved chains (its first and second half),
// so that each lane has two independent rotation recurrences in flight.
constexpr int kRpAnchor = 16;

struct FfChain {
    double c = 1.0, s = 0.0, rc = 1.0, rs = 0.0;     // phase factor, rotation
    double c2 = 1.0, s2 = 0.0, rc2 = 1.0, rs2 = 0.0; // the same for the image
    double qt = 0.0, qp = 0.0, it = 0.0, ip = 0.0;   // dir . theta, dir . phi (and the image's)
};

__device__ __forceinline__ void ff_term(const double *__restrict__ g, bool anchor, double k, int ground,
                                        const Dir &u, double gvr, double gvi, double ghr, double ghi,
                                        FfChain &q, double &ftr, double &fti, double &fpr, double &fpi)
{
    const double ir = g[8], ii = g[9]; // current x length
    const bool rotate = g[7] != 0.0 && !anchor;
    if (rotate) {
        const double nc = q.c * q.rc - q.s * q.rs, ns = q.s * q.rc + q.c * q.rs;
        q.c = nc;
        q.s = ns;
    } else {
        const double mx = g[0], my = g[1], mz = g[2], dx = g[3], dy = g[4], dz = g[5], len = g[6];
        sincos_phase(k * (mx * u.rx + my * u.ry + mz * u.rz), q.s, q.c);
        sincos_phase(k * len * (dx * u.rx + dy * u.ry + dz * u.rz), q.rs, q.rc);
        q.qt = dx * u.tx + dy * u.ty + dz * u.tz;
        q.qp = dx * u.px + dy * u.py;
        if (ground != 0) {
            // image: position (x, y, -z), direction (-dx, -dy, dz) (farfield.rs:64-82)
            sincos_phase(k * (mx * u.rx + my * u.ry - mz * u.rz), q.s2, q.c2);
            sincos_phase(k * len * (dx * u.rx + dy * u.ry - dz * u.rz), q.rs2, q.rc2);
            q.it = -dx * u.tx - dy * u.ty + dz * u.tz;
            q.ip = -dx * u.px - dy * u.py;
        }
    }
    const double wr = ir * q.c - ii * q.s, wi = ir * q.s + ii * q.c;
    ftr += wr * q.qt;
    fti += wi * q.qt;
    fpr += wr * q.qp;
    fpi += wi * q.qp;
    if (ground != 0) {
        if (rotate) {
            const double nc = q.c2 * q.rc2 - q.s2 * q.rs2, ns = q.s2 * q.rc2 + q.c2 * q.rs2;
            q.c2 = nc;
            q.s2 = ns;
        }
        const double vr = ir * q.c2 - ii * q.s2, vi = ir * q.s2 + ii * q.c2;
        const double atr = vr * q.it, ati = vi * q.it, apr = vr * q.ip, api = vi * q.ip;
        if (ground == 1) {
            ftr += atr;
            fti += ati;
            fpr += apr;
            fpi += api;
        } else {
            ftr += gvr * atr - gvi * ati;
            fti += gvr * ati + gvi * atr;
            fpr -= ghr * apr - ghi * api;
            fpi -= ghr * api + ghi * apr;
        }
    }
}

__device__ __forceinline__ void ff_chunk(const double *__restrict__ sg, int cnt, double k,
                                         int ground, const Dir &u, double gvr, double gvi,
                                         double ghr, double ghi, double &ftr, double &fti,
                                         double &fpr, double &fpi)
{
    const int half = (cnt + 1) / 2; // chain A: [0, half), chain B: [half, cnt)
    FfChain qa, qb;
    for (int mm = 0; mm < half; ++mm) {
        const bool anchor = (mm % kRpAnchor) == 0;
        ff_term(sg + mm * kRpStage, anchor, k, ground, u, gvr, gvi, ghr, ghi, qa, ftr, fti, fpr, fpi);
        if (half + mm < cnt)
            ff_term(sg + (half + mm) * kRpStage, anchor, k, ground, u, gvr, gvi, ghr, ghi, qb, ftr, fti, fpr,
                    fpi);
    }
}

__device__ __forceinline__ void stage_chunk(const RpProb &R, double *sg, int base, int t)
{
    const int cnt = min(kRpChunk, R.n - base);
    for (int e = t; e < cnt * kRpStage; e += kRpThreads) {
        const int m = e / kRpStage, c = e % kRpStage;
        const int gm = base + m;
        double v;
        if (c < 8)
            v = R.seg[(size_t)gm * kSegStride + c];
        else
            v = R.cur[2 * (size_t)gm + (c - 8)] * R.seg[(size_t)gm * kSegStride + 6];
        sg[e] = v;
    }
}

// Work item: (rp problem, block of 256 points). Writes F_theta, F_phi.
__global__ __launch_bounds__(kRpThreads) void k_rp_points(const RpProb *__restrict__ probs,
                                                          const int2 *__restrict__ work)
{
    const int2 wk = work[blockIdx.x];
    const RpProb &R = probs[wk.x];
    __shared__ double sg[kRpChunk * kRpStage];
    const int t = threadIdx.x;
    const int pt = wk.y * kRpThreads + t;
    const bool active = pt < R.npts && !(R.ground != 0 && R.theta[pt] > 90.0);
    double th = 0.0, ph = 0.0;
    if (pt < R.npts) {
        th = R.theta[pt];
        ph = R.phi[pt];
    }
    const Dir u = unit_vectors(th, ph);
    double gvr = 1.0, gvi = 0.0, ghr = -1.0, ghi = 0.0;
    if (active && R.ground == 2)
        fresnel(th, R.eps_r, R.sigma, R.k, gvr, gvi, ghr, ghi);
    double ftr = 0.0, fti = 0.0, fpr = 0.0, fpi = 0.0;
    for (int base = 0; base < R.n; base += kRpChunk) {
        __syncthreads();
        stage_chunk(R, sg, base, t);
        __syncthreads();
        if (active)
            ff_chunk(sg, min(kRpChunk, R.n - base), R.k, R.ground, u, gvr, gvi, ghr, ghi, ftr,
                     fti, fpr, fpi);
    }
    if (pt < R.npts) {
        double *o = R.field + 4 * (size_t)pt;
        o[0] = ftr;
        o[1] = fti;
        o[2] = fpr;
        o[3] = fpi;
    }
}

// Work item: (rp problem, block of 256 normalisation directions). The grid is
// theta_i = i*tmax/(nt-1) (nt = 181 free / 91 ground), phi = 0..359 degrees,
// trapezoid half weights at both theta ends, sin(theta) == 0 skipped
// (farfield.rs:422-464). One partial sum per block, fixed reduction order.
__global__ __launch_bounds__(kRpThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_rp_norm(const RpProb *__restrict__ probs,
                                                        const int2 *__restrict__ work)
{
    const int2 wk = work[blockIdx.x];
    const RpProb &R = probs[wk.x];
    __shared__ double sg[kRpChunk * kRpStage];
    __shared__ double red[kRpThreads];
    const int t = threadIdx.x;
    const int d = wk.y * kRpThreads + t;
    const bool has_ground = R.ground != 0;
    const int nt = has_ground ? 91 : 181;
    const double tmax = has_ground ? 90.0 : 180.0;
    const int it = d / 360, ip = d % 360;
    const double th_deg = (double)it * tmax / ((double)nt - 1.0);
    const double st = sin(th_deg * (kPi / 180.0));
    const bool active = d < R.ndir && st != 0.0;
    const double dth = tmax * (kPi / 180.0) / ((double)nt - 1.0);
    const double dph = 2.0 * kPi / 360.0;
    const double wt = (it == 0 || it == nt - 1) ? 0.5 * dth : dth;
    const Dir u = unit_vectors(th_deg, (double)ip);
    double gvr = 1.0, gvi = 0.0, ghr = -1.0, ghi = 0.0;
    if (active && R.ground == 2)
        fresnel(th_deg, R.eps_r, R.sigma, R.k, gvr, gvi, ghr, ghi);
    double ftr = 0.0, fti = 0.0, fpr = 0.0, fpi = 0.0;
    for (int base = 0; base < R.n; base += kRpChunk) {
        __syncthreads();
        stage_chunk(R, sg, base, t);
        __syncthreads();
        if (active)
            ff_chunk(sg, min(kRpChunk, R.n - base), R.k, R.ground, u, gvr, gvi, ghr, ghi, ftr,
                     fti, fpr, fpi);
    }
    const double term =
        active ? (ftr * ftr + fti * fti + fpr * fpr + fpi * fpi) * st * wt * dph : 0.0;
    red[t] = term;
    __syncthreads();
    for (int off = kRpThreads / 2; off > 0; off >>= 1) {
        if (t < off)
            red[t] += red[t + off];
        __syncthreads();
    }
    if (t == 0)
        R.partial[wk.y] = red[0];
}

// One workgroup per rp problem: total = ordered sum of partials, then gains
// (farfield.rs:594-646) and the Stokes axial ratio (:659-667).
__global__ __launch_bounds__(kRpThreads) void k_rp_finish(const RpProb *__restrict__ probs,
                                                          const int *__restrict__ plist)
{
    const RpProb &R = probs[plist[blockIdx.x]];
    __shared__ double s_total;
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int b = 0; b < R.npartial; ++b)
            tot += R.partial[b];
        s_total = tot;
        R.total[0] = tot;
    }
    __syncthreads();
    const double total = s_total;
    const double norm = total > 0.0 ? 4.0 * kPi / total : 0.0;
    for (int pt = threadIdx.x; pt < R.npts; pt += kRpThreads) {
        double *o = R.out + 4 * (size_t)pt;
        if (R.ground != 0 && R.theta[pt] > 90.0) {
            o[0] = o[1] = o[2] = -999.99;
            o[3] = 0.0;
            continue;
        }
        const double *f = R.field + 4 * (size_t)pt;
        const double ut = f[0] * f[0] + f[1] * f[1], up = f[2] * f[2] + f[3] * f[3];
        const double uu = ut + up;
        o[0] = uu * norm > 1e-20 ? 10.0 * log10(uu * norm) : -999.99;
        o[1] = ut * norm > 1e-20 ? 10.0 * log10(ut * norm) : -999.99;
        o[2] = up * norm > 1e-20 ? 10.0 * log10(up * norm) : -999.99;
        double ar = 0.0;
        if (uu > 1e-30) {
            // S3 = 2 Im(E_theta conj(E_phi))
            const double s3 = 2.0 * (f[1] * f[2] - f[0] * f[3]);
            double q = s3 / uu;
            q = q < -1.0 ? -1.0 : (q > 1.0 ? 1.0 : q);
            ar = tan(0.5 * asin(q));
        }
        o[3] = ar;
    }
}

// Fill a buffer with zeros (graph-capturable replacement for a memset).
__global__ void k_zero(double *__restrict__ p, size_t count)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// MFMA operand-layout probe: D = A B for A (16x4) and B (4x16) given in
// row-major; lane l loads A[l%16][l/16] and B[l/16][l%16]. Writes the raw
// per-lane accumulator so the host can check the output mapping.
__global__ __launch_bounds__(64) void k_mfma_probe(const double *__restrict__ a,
                                                   const double *__restrict__ b,
                                                   double *__restrict__ out)
{
    const int l = threadIdx.x;
    d4 acc = d4{0.0, 0.0, 0.0, 0.0};
    acc = mfma(a[(l % 16) * 4 + l / 16], b[(l / 16) * 16 + l % 16], acc);
    for (int v = 0; v < 4; ++v)
        out[l * 4 + v] = acc[v];
}

} // namespace fnec
