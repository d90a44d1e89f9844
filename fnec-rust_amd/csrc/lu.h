// Blocked complex LU with partial pivoting, templated on the scalar type.
//
// double: the plain path factors the augmented normal matrix [G | M^H y]
//         (linear.rs:1321-1383), ncols = m + 1.
// float:  the mixed path factors the Z matrix itself (ncols = m); the factors
//         precondition an FP64 iteration.
//
// Matrices are column-major split planes (re, im) with leading dimension lda.
// Row interchanges are applied to the panel and to the columns right of it;
// with swap_left (the mixed path) also to the multipliers of earlier panels,
// so that the factors read P Z = L U with one row permutation P. Without it
// (the plain path, which needs U and the transformed right-hand side only)
// earlier multipliers keep their original row order.
//
// Kernels per panel step p (jb = min(kPanel, m - p) columns):
//   k_panel_diag / k_panel_rows / k_panel_commit  no-interchange fast path:
//       factor without pivoting, then prove that partial pivoting (last row
//       of maximal modulus, linear.rs:1345-1352) would have kept every
//       diagonal row; commit only then.
//   k_lu_panel   exact path with interchanges when the fast path declined.
//   k_lu_trsm    interchanges + U12 = L11^{-1} A12 (one thread per column).
//   k_lu_update  A22 -= L21 U12 on MFMA (16x16x4, four real products per
//                complex product, 64x64 output tiles, K = jb staged in LDS).
//   k_lu_update_f32  the FP32 variant: 128x64 tiles, accumulators seeded
//                with C, K = 32 nk columns of L21 per launch.
//   k_l11_inv    FP32 outer step: L11^{-1} of the 64 x 64 diagonal block.
//   k_lu_trsm64_f32  FP32 outer step: interchanges of two 32-column panels
//                and U12 = L11^{-1} A12 on MFMA for the columns right of both
//                (together, or as two launches).
//   k_split_f16x2  the K = 128 passes' L21 and U12 split into hi + lo f16
//                pairs, once per combining step.
//   k_lu_update_f16x2s / k_lu_update_f16x2d  the K = 128 trailing update of
//                two outer steps on the f16 matrix cores (operands streamed
//                through an LDS ring by DMA / loaded into registers).
// The FP32 factorisation (mixed path) is two-level: each 64-column step
// factors two 32-column panels (the second after a narrow K = 32 update of
// its columns only), then updates the trailing matrix once with K = 64, so
// the trailing matrix is read and written once per 64 columns. With K = 128
// passes, every other step updates only the next step's 64 columns, and the
// step after it first brings its own U12 rows up to date (a thin K = 64
// update between its interchanges and its L11^{-1} product), then updates
// the rest with both steps' factors at once: once per 128 columns.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "common.h"

namespace fnec {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;

// v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and B[l/16][l%16];
// accumulator item v of lane l is D[(l/16) + 4v][l%16].
template <> struct Mfma<double> {
    typedef d4 acc;
    static __device__ __forceinline__ acc zero() { return d4{0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ acc fma(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

// v_mfma_f32_16x16x4f32: same operand layout; accumulator item v of lane l
// is D[4(l/16) + v][l%16].
template <> struct Mfma<float> {
    typedef f4 acc;
    static __device__ __forceinline__ acc zero() { return f4{0.f, 0.f, 0.f, 0.f}; }
    static __device__ __forceinline__ acc fma(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// |z|: hypot for FP64 (the reference's norm(), linear.rs:1346); for the FP32
// factorisation, which only ranks pivot candidates, a plain sqrt keeps the
// call inline (hypotf is an out-of-line call that forces register spills).
__device__ __forceinline__ double cmod(double re, double im) { return hypot(re, im); }
__device__ __forceinline__ float cmod(float re, float im) { return __builtin_sqrtf(re * re + im * im); }

constexpr int kPanelThreads = 512;
constexpr int kLuPad = 4;

// One workgroup per problem factors columns [p, p + jb) of rows [p, m).
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void k_lu_panel(const LuDesc<T> *__restrict__ descs,
                                                            const int *__restrict__ plist, int p)
{
    const LuDesc<T> &D = descs[plist[blockIdx.x]];
    const int s = D.m;
    if (D.info[0] != 0 || p >= s || D.fast_ok[0])
        return;
    const int jb = min(kPanel, s - p);
    const int t = threadIdx.x;
    const int lda = D.lda;
    T *ar = D.ar, *ai = D.ai;
    __shared__ T rv[kPanelThreads];
    __shared__ int ri[kPanelThreads];
    __shared__ T prow_r[kPanel], prow_i[kPanel];
    __shared__ int s_piv;
    __shared__ T s_best;
    if (t == 0)
        D.info[2] += 1; // panels factored on the exact interchange path
    for (int j = 0; j < jb; ++j) {
        const int col = p + j;
        const T *cr = ar + (size_t)col * lda, *ci = ai + (size_t)col * lda;
        T best = -1;
        int bidx = col;
        for (int r = col + t; r < s; r += kPanelThreads) {
            const T v = cmod(cr[r], ci[r]);
            if (!(v < best)) { // later rows win ties (Iterator::max_by)
                best = v;
                bidx = r;
            }
        }
        rv[t] = best;
        ri[t] = bidx;
        __syncthreads();
        for (int off = kPanelThreads / 2; off > 0; off >>= 1) {
            if (t < off) {
                const T v2 = rv[t + off];
                const int i2 = ri[t + off];
                if (v2 > rv[t] || (v2 == rv[t] && i2 > ri[t])) {
                    rv[t] = v2;
                    ri[t] = i2;
                }
            }
            __syncthreads();
        }
        if (t == 0) {
            s_piv = ri[0];
            s_best = rv[0];
        }
        __syncthreads();
        const int piv = s_piv;
        if (s_best < (T)kSingular) { // NaN passes, as in linear.rs:1354
            if (t == 0)
                D.info[0] = 1;
            return;
        }
        if (piv != col) {
            for (int c = p + t; c < p + jb; c += kPanelThreads) {
                const size_t o1 = (size_t)c * lda + col, o2 = (size_t)c * lda + piv;
                const T xr = ar[o1], xi = ai[o1];
                ar[o1] = ar[o2];
                ai[o1] = ai[o2];
                ar[o2] = xr;
                ai[o2] = xi;
            }
            if (t == 0)
                D.info[1] += 1;
        }
        if (t == 0)
            D.ipiv[col] = piv;
        __syncthreads();
        for (int c = col + 1 + t; c < p + jb; c += kPanelThreads) {
            prow_r[c - col - 1] = ar[(size_t)c * lda + col];
            prow_i[c - col - 1] = ai[(size_t)c * lda + col];
        }
        __syncthreads();
        const T dr = cr[col], di = ci[col];
        const T nsq = dr * dr + di * di;
        const T ir = dr / nsq, ii = -di / nsq; // 1/d
        const int ncols = p + jb - col - 1;
        for (int r = col + 1 + t; r < s; r += kPanelThreads) {
            const T xr = cr[r], xi = ci[r];
            const T lr = xr * ir - xi * ii, li = xr * ii + xi * ir;
            ar[(size_t)col * lda + r] = lr;
            ai[(size_t)col * lda + r] = li;
            for (int c = 0; c < ncols; ++c) {
                const size_t o = (size_t)(col + 1 + c) * lda + r;
                const T ur = prow_r[c], ui = prow_i[c];
                ar[o] -= lr * ur - li * ui;
                ai[o] -= lr * ui + li * ur;
            }
        }
        __syncthreads();
    }
}

// ---- fast panel path: factor without interchanges, then prove none was due ----
//
// The fast path records, for every panel column j, the largest |a_rj| below
// the diagonal at elimination step j. Partial pivoting keeps row j exactly
// when every such entry is strictly smaller than |u_jj| (a tie selects the
// later row), so when that holds for all columns the no-interchange
// factorisation is the partial-pivoting one and is committed; otherwise the
// panel is left untouched for k_lu_panel.

// One wavefront per problem: LU of the jb x jb diagonal block, one lane per
// row. A lane keeps its row in registers (the loops over the block's columns
// are unrolled, so every register index is static); at step j the pivot row
// is broadcast through LDS and each lane below it eliminates in registers.
template <typename T>
__global__ __launch_bounds__(64) void k_panel_diag(const LuDesc<T> *__restrict__ descs,
                                                   const int *__restrict__ plist, int p)
{
    const LuDesc<T> &D = descs[plist[blockIdx.x]];
    const int s = D.m;
    if (D.info[0] != 0 || p >= s)
        return;
    const int jb = min(kPanel, s - p);
    const int lda = D.lda, lane = threadIdx.x;
    const bool mine = lane < jb;
    __shared__ T rowr[kPanel], rowi[kPanel];
    T ar[kPanel], ai[kPanel];
#pragma unroll
    for (int c = 0; c < kPanel; ++c) {
        ar[c] = ai[c] = (T)0;
        if (mine && c < jb) {
            ar[c] = D.ar[(size_t)(p + c) * lda + p + lane];
            ai[c] = D.ai[(size_t)(p + c) * lda + p + lane];
        }
    }
    T cmax_mine = 0, dmag = 0; // lane j: column j's maximum below the diagonal, |u_jj|
#pragma unroll
    for (int j = 0; j < kPanel; ++j) {
        if (j < jb) { // uniform
            if (lane == j) {
#pragma unroll
                for (int c = j; c < kPanel; ++c) {
                    rowr[c] = ar[c];
                    rowi[c] = ai[c];
                }
            }
            __syncthreads();
            const T ur = rowr[j], ui = rowi[j];
            const T nsq = ur * ur + ui * ui;
            const T ir = ur / nsq, ii = -ui / nsq;
            // |a_rj| below the diagonal at step j (max is order independent)
            T v = (lane > j && mine) ? cmod(ar[j], ai[j]) : (T)0;
            for (int off = 32; off > 0; off >>= 1)
                v = fmax(v, __shfl_xor(v, off, 64));
            if (lane == j) {
                cmax_mine = v;
                dmag = cmod(ur, ui);
            }
            if (lane > j && mine) {
                const T xr = ar[j], xi = ai[j];
                const T lr = xr * ir - xi * ii, li = xr * ii + xi * ir;
                ar[j] = lr;
                ai[j] = li;
#pragma unroll
                for (int c = j + 1; c < kPanel; ++c) {
                    const T vr = rowr[c], vi = rowi[c];
                    ar[c] -= lr * vr - li * vi;
                    ai[c] -= lr * vi + li * vr;
                }
            }
            __syncthreads(); // row j is consumed before row j + 1 is published
        }
    }
    if (mine) {
#pragma unroll
        for (int c = 0; c < kPanel; ++c)
            if (c < jb) {
                D.pr[(size_t)c * lda + p + lane] = ar[c];
                D.pi[(size_t)c * lda + p + lane] = ai[c];
            }
        D.diagmag[lane] = (double)dmag;
        D.colmax[lane] = (unsigned long long)__double_as_longlong((double)cmax_mine);
    }
}

// Rows below the diagonal block: l_r U_D = a_r by forward substitution, one
// thread per row; the intermediate t_rj is a_rj at elimination step j. The
// row's panel entries are loaded up front so the loads overlap. Arithmetic is
// FP64 for both factor types (the FP32 instantiation of this fully unrolled
// recurrence spills registers; FP64 does not, and is at least as accurate).
template <typename T>
__global__ __launch_bounds__(256) void k_panel_rows(const LuDesc<T> *__restrict__ descs,
                                                    const int2 *__restrict__ work, int p)
{
    const int2 wk = work[blockIdx.x];
    const LuDesc<T> &D = descs[wk.x];
    const int s = D.m;
    if (D.info[0] != 0 || p >= s)
        return;
    const int jb = min(kPanel, s - p);
    const int lda = D.lda, t = threadIdx.x;
    __shared__ double ur_[kPanel][kPanel + 1], ui_[kPanel][kPanel + 1];
    __shared__ double cm[kPanel][8];
    for (int e = t; e < jb * jb; e += 256) {
        const int r = e % jb, c = e / jb;
        ur_[r][c] = (double)D.pr[(size_t)c * lda + p + r];
        ui_[r][c] = (double)D.pi[(size_t)c * lda + p + r];
    }
    const int row = p + jb + wk.y * 256 + t;
    const bool active = row < s;
    double lr[kPanel], li[kPanel];
#pragma unroll
    for (int j = 0; j < kPanel; ++j) {
        lr[j] = 0.0;
        li[j] = 0.0;
        if (active && j < jb) {
            lr[j] = (double)D.ar[(size_t)(p + j) * lda + row];
            li[j] = (double)D.ai[(size_t)(p + j) * lda + row];
        }
    }
    __syncthreads();
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int j = 0; j < kPanel; ++j) {
        if (j < jb) {
            double vr = lr[j], vi = li[j];
            if (active) {
#pragma unroll
                for (int m = 0; m < j; ++m) {
                    vr -= lr[m] * ur_[m][j] - li[m] * ui_[m][j];
                    vi -= lr[m] * ui_[m][j] + li[m] * ur_[m][j];
                }
            }
            // |a_rj| at step j, reduced over the wave (inactive lanes give 0)
            double v = active ? hypot(vr, vi) : 0.0;
            for (int off = 32; off > 0; off >>= 1)
                v = fmax(v, __shfl_xor(v, off, 64));
            if (lane == 0)
                cm[j][wave] = v;
            if (active) {
                const double dr = ur_[j][j], di = ui_[j][j];
                const double nsq = dr * dr + di * di;
                const double ir = dr / nsq, ii = -di / nsq;
                lr[j] = vr * ir - vi * ii;
                li[j] = vr * ii + vi * ir;
            } else {
                lr[j] = 0.0;
                li[j] = 0.0;
            }
        }
    }
    if (active)
#pragma unroll
        for (int j = 0; j < kPanel; ++j)
            if (j < jb) {
                D.pr[(size_t)j * lda + row] = (T)lr[j];
                D.pi[(size_t)j * lda + row] = (T)li[j];
            }
    __syncthreads();
    if (t < jb) {
        const double v = fmax(fmax(cm[t][0], cm[t][1]), fmax(cm[t][2], cm[t][3]));
        atomicMax(&D.colmax[t], (unsigned long long)__double_as_longlong(v));
    }
}

// ---- FP32 fast panel (mixed path) ----
// The multipliers satisfy |t_rj| = |l_rj| |u_jj|, so "no interchange was due"
// reads max_r |l_rj| < 1 (strict: a tie selects the later row). The diagonal
// block is factored in LDS and U_D^{-1} formed; the rows below are then
// l_r = a_r U_D^{-1}, a (rows x 32) x (32 x 32) complex product on MFMA.

// One wavefront per problem: LU of the diagonal block (one lane per row, the
// block padded to 32 x 32 with the identity), max |l| per column, then
// U_D^{-1} right-looking across the wave into D.linv + 2 x 64 x 64 ([k][c]).
__global__ __launch_bounds__(64) void k_panel_diag_f32(const LuDesc<float> *__restrict__ descs,
                                                       const int *__restrict__ plist, int p)
{
    constexpr int NB = kPanel;
    const LuDesc<float> &D = descs[plist[blockIdx.x]];
    const int s = D.m;
    if (D.info[0] != 0 || p >= s)
        return;
    const int jb = min(NB, s - p);
    const int lda = D.lda, lane = threadIdx.x;
    const bool mine = lane < NB;
    __shared__ float rowr[NB], rowi[NB];
    __shared__ float ur_s[NB][NB + 1], ui_s[NB][NB + 1]; // U, for U_D^{-1}
    // lane r keeps row r in registers; a partial block is completed with the
    // identity so that the factorisation and the inverse are defined on NB x NB
    float ar[NB], ai[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const bool in = lane < jb && c < jb;
        ar[c] = in ? D.ar[(size_t)(p + c) * lda + p + lane] : (lane == c ? 1.f : 0.f);
        ai[c] = in ? D.ai[(size_t)(p + c) * lda + p + lane] : 0.f;
    }
    float lmax = 0.f; // lane j: max |l_ij|^2 of column j within the block
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (lane == j) {
#pragma unroll
            for (int c = j; c < NB; ++c) {
                rowr[c] = ar[c];
                rowi[c] = ai[c];
            }
        }
        __syncthreads();
        const float ur = rowr[j], ui = rowi[j];
        const float nsq = ur * ur + ui * ui;
        const float ir = ur / nsq, ii = -ui / nsq;
        float m = 0.f;
        if (lane > j && mine) {
            const float x = ar[j], y = ai[j];
            const float lr = x * ir - y * ii, li = x * ii + y * ir;
            ar[j] = lr;
            ai[j] = li;
            m = lr * lr + li * li;
#pragma unroll
            for (int c = j + 1; c < NB; ++c) {
                const float vr = rowr[c], vi = rowi[c];
                ar[c] -= lr * vr - li * vi;
                ai[c] -= lr * vi + li * vr;
            }
        }
        for (int off = 32; off > 0; off >>= 1)
            m = fmaxf(m, __shfl_xor(m, off, 64));
        if (lane == j)
            lmax = m;
        __syncthreads(); // row j is consumed before row j + 1 is published
    }
    if (mine) {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            ur_s[lane][c] = ar[c];
            ui_s[lane][c] = ai[c];
        }
    }
    __syncthreads();
    // U_D^{-1} in LDS: row i of X is final once scaled by 1/u_ii; then every
    // row k < i takes -= u_ki X[i][:]
    __shared__ float xr[NB][NB + 1], xi[NB][NB + 1];
    for (int e = lane; e < NB * NB; e += 64) {
        xr[e / NB][e % NB] = (e / NB == e % NB) ? 1.f : 0.f;
        xi[e / NB][e % NB] = 0.f;
    }
    __syncthreads();
    for (int i = NB - 1; i >= 0; --i) {
        const float dr = ur_s[i][i], di = ui_s[i][i];
        const float nsq = dr * dr + di * di;
        const float ir = dr / nsq, ii = -di / nsq;
        if (lane >= i && lane < NB) {
            const float x = xr[i][lane], y = xi[i][lane];
            xr[i][lane] = x * ir - y * ii;
            xi[i][lane] = x * ii + y * ir;
        }
        __syncthreads();
        const int w = NB - i;
        for (int e = lane; e < i * w; e += 64) {
            const int k = e / w, c = i + e % w;
            const float a = ur_s[k][i], b = ui_s[k][i], x = xr[i][c], y = xi[i][c];
            xr[k][c] -= a * x - b * y;
            xi[k][c] -= a * y + b * x;
        }
        __syncthreads();
    }
    if (mine) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < jb && lane < jb) {
                D.pr[(size_t)c * lda + p + lane] = ar[c];
                D.pi[(size_t)c * lda + p + lane] = ai[c];
            }
    }
    float *ui_r = D.linv + 2 * 64 * 64, *ui_i = ui_r + NB * NB;
    for (int e = lane; e < NB * NB; e += 64) {
        ui_r[e] = xr[e / NB][e % NB];
        ui_i[e] = xi[e / NB][e % NB];
    }
    if (lane < jb) {
        const float dr = ur_s[lane][lane], di = ui_s[lane][lane];
        D.diagmag[lane] = (double)__builtin_sqrtf(dr * dr + di * di);
        D.colmax[lane] = (unsigned long long)__double_as_longlong((double)__builtin_sqrtf(lmax));
    }
}

// Rows [p + 32 + 256 y, + 256) below a full diagonal block: l_r = a_r U_D^{-1}
// on MFMA (four waves of 64 rows x 32 columns, K = 32) into the panel
// workspace, and max |l| per column into colmax. Row quads are whole (the
// padding rows past m hold 0).
__global__ __launch_bounds__(256) void k_panel_rows_f32(const LuDesc<float> *__restrict__ descs,
                                                        const int2 *__restrict__ work, int p)
{
    typedef Mfma<float> MF;
    constexpr int NB = kPanel, TM = 256, LA = TM + 16, LB = NB + 4;
    const int2 wk = work[blockIdx.x];
    const LuDesc<float> &D = descs[wk.x];
    const int s = D.m;
    if (D.info[0] != 0 || p + NB >= s)
        return;
    const int lda = D.lda, t = threadIdx.x;
    const int r0 = p + NB + wk.y * TM;
    __shared__ float sa_r[NB][LA], sa_i[NB][LA]; // a: [k][row]
    __shared__ float sb_r[NB][LB], sb_i[NB][LB]; // U_D^{-1}: [k][c]
    __shared__ unsigned cmx[NB];
    if (t < NB)
        cmx[t] = 0u;
#pragma unroll
    for (int i = 0; i < (NB * TM / 4) / 256; ++i) {
        const int e = t + 256 * i;
        const int k = e / (TM / 4), r4 = (e % (TM / 4)) * 4;
        const int gr = r0 + r4;
        float4 vr = make_float4(0.f, 0.f, 0.f, 0.f), vi = vr;
        if (gr < s) {
            const size_t o = (size_t)(p + k) * lda + gr;
            vr = *(const float4 *)(D.ar + o);
            vi = *(const float4 *)(D.ai + o);
        }
        *(float4 *)&sa_r[k][r4] = vr;
        *(float4 *)&sa_i[k][r4] = vi;
    }
    {
        const float *ui_r = D.linv + 2 * 64 * 64, *ui_i = ui_r + NB * NB;
        for (int e = t; e < NB * NB; e += 256) {
            sb_r[e / NB][e % NB] = ui_r[e];
            sb_i[e / NB][e % NB] = ui_i[e];
        }
    }
    __syncthreads();
    const int lane = t & 63, wave = t >> 6, lrow = lane & 15, lk = lane >> 4;
    typename MF::acc accr[4][2], acci[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accr[a][b] = MF::zero();
            acci[a][b] = MF::zero();
        }
#pragma unroll 2
    for (int kk = 0; kk < NB; kk += 4) {
        float xr[4], xi[4], yr[2], yi[2];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            xr[a] = sa_r[kk + lk][wave * 64 + a * 16 + lrow];
            xi[a] = sa_i[kk + lk][wave * 64 + a * 16 + lrow];
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            yr[b] = sb_r[kk + lk][b * 16 + lrow];
            yi[b] = sb_i[kk + lk][b * 16 + lrow];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                accr[a][b] = MF::fma(xr[a], yr[b], accr[a][b]);
                accr[a][b] = MF::fma(-xi[a], yi[b], accr[a][b]);
                acci[a][b] = MF::fma(xr[a], yi[b], acci[a][b]);
                acci[a][b] = MF::fma(xi[a], yr[b], acci[a][b]);
            }
    }
    float m[2] = {0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int r = r0 + wave * 64 + a * 16 + 4 * lk; // rows r .. r + 3
            const int c = b * 16 + lrow;
            if (r >= s)
                continue;
            const size_t o = (size_t)c * lda + r;
            *(float4 *)(D.pr + o) = make_float4(accr[a][b][0], accr[a][b][1], accr[a][b][2], accr[a][b][3]);
            *(float4 *)(D.pi + o) = make_float4(acci[a][b][0], acci[a][b][1], acci[a][b][2], acci[a][b][3]);
#pragma unroll
            for (int v = 0; v < 4; ++v)
                m[b] = fmaxf(m[b], accr[a][b][v] * accr[a][b][v] + acci[a][b][v] * acci[a][b][v]);
        }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        m[b] = fmaxf(m[b], __shfl_xor(m[b], 16, 64));
        m[b] = fmaxf(m[b], __shfl_xor(m[b], 32, 64));
        if (lk == 0)
            atomicMax(&cmx[b * 16 + lrow], __float_as_uint(m[b])); // non-negative: bits order as values
    }
    __syncthreads();
    if (t < NB)
        atomicMax(&D.colmax[t], (unsigned long long)__double_as_longlong(
                                    (double)__builtin_sqrtf(__uint_as_float(cmx[t]))));
}

// Decide and commit: every workgroup reads the same per-column maxima, so all
// agree; on success the panel rows of the workspace replace the matrix panel.
template <typename T, bool kUnitL = false>
__global__ __launch_bounds__(256) void k_panel_commit(const LuDesc<T> *__restrict__ descs,
                                                      const int2 *__restrict__ work, int p)
{
    const int2 wk = work[blockIdx.x];
    const LuDesc<T> &D = descs[wk.x];
    const int s = D.m;
    if (D.info[0] != 0 || p >= s)
        return;
    const int jb = min(kPanel, s - p);
    const int lda = D.lda, t = threadIdx.x;
    int bad = 0;
    if (t < jb) {
        const double m = __longlong_as_double((long long)D.colmax[t]);
        const double d = D.diagmag[t];
        // colmax holds max |t_rj| (compared with |u_jj|) or, for the FP32
        // panel, max |l_rj| = max |t_rj| / |u_jj| (compared with 1)
        bad = (!(m < (kUnitL ? 1.0 : d)) || d < kSingular) ? 1 : 0;
    }
    const int ok = !__syncthreads_or(bad);
    if (t == 0 && wk.y == 0)
        D.fast_ok[0] = ok;
    if (!ok)
        return;
    const int row = p + wk.y * 256 + t;
    if (row < s) {
        T vr[kPanel], vi[kPanel];
#pragma unroll
        for (int j = 0; j < kPanel; ++j)
            if (j < jb) {
                vr[j] = D.pr[(size_t)j * lda + row];
                vi[j] = D.pi[(size_t)j * lda + row];
            }
#pragma unroll
        for (int j = 0; j < kPanel; ++j)
            if (j < jb) {
                D.ar[(size_t)(p + j) * lda + row] = vr[j];
                D.ai[(size_t)(p + j) * lda + row] = vi[j];
            }
    }
    if (wk.y == 0 && t < jb)
        D.ipiv[p + t] = p + t;
}

// Row interchanges of the panel and U12 = L11^{-1} A12 for columns
// [p + jb, ncols). One thread per column. Work items with a negative column
// block (-1 - b) apply the panel's interchanges to columns [64 b, 64 b + 64)
// left of the panel (requested by lu_plan when swap_left is set).
template <typename T>
__global__ __launch_bounds__(64) void k_lu_trsm(const LuDesc<T> *__restrict__ descs,
                                                const int2 *__restrict__ work, int p, int cend)
{
    const int2 wk = work[blockIdx.x];
    const LuDesc<T> &D = descs[wk.x];
    const int s = D.m;
    if (D.info[0] != 0 || p >= s)
        return;
    const int jb = min(kPanel, s - p);
    const int lda = D.lda;
    __shared__ T lr[kPanel][kPanel + 1], li[kPanel][kPanel + 1];
    __shared__ int piv_s[kPanel];
    const int t = threadIdx.x;
    if (t < jb)
        piv_s[t] = D.ipiv[p + t];
    for (int e = t; e < jb * jb; e += 64) {
        const int rr = e % jb, cc = e / jb;
        lr[rr][cc] = D.ar[(size_t)(p + cc) * lda + p + rr];
        li[rr][cc] = D.ai[(size_t)(p + cc) * lda + p + rr];
    }
    __syncthreads();
    const bool any_swap = __syncthreads_or(t < jb && piv_s[t] != p + t);
    if (wk.y < 0) {
        // columns left of the panel (L of earlier panels): interchanges only,
        // so that the factors read P Z = L U with one row permutation P
        const int cl = (-wk.y - 1) * 64 + t;
        if (!any_swap || cl >= p)
            return;
        T *cr = D.ar + (size_t)cl * lda, *ci = D.ai + (size_t)cl * lda;
        for (int j = 0; j < jb; ++j) {
            const int piv = piv_s[j];
            if (piv != p + j) {
                const T xr = cr[p + j], xi = ci[p + j];
                cr[p + j] = cr[piv];
                ci[p + j] = ci[piv];
                cr[piv] = xr;
                ci[piv] = xi;
            }
        }
        return;
    }
    const int c = p + jb + wk.y * 64 + t;
    if (c >= D.ncols || c >= cend)
        return;
    T *cr = D.ar + (size_t)c * lda, *ci = D.ai + (size_t)c * lda;
    if (any_swap)
        for (int j = 0; j < jb; ++j) {
            const int piv = piv_s[j];
            if (piv != p + j) {
                const T xr = cr[p + j], xi = ci[p + j];
                cr[p + j] = cr[piv];
                ci[p + j] = ci[piv];
                cr[piv] = xr;
                ci[piv] = xi;
            }
        }
    T ur[kPanel], ui[kPanel];
#pragma unroll
    for (int j = 0; j < kPanel; ++j)
        if (j < jb) {
            ur[j] = cr[p + j];
            ui[j] = ci[p + j];
        }
#pragma unroll
    for (int j = 0; j < kPanel; ++j) {
        if (j < jb) {
            T vr = ur[j], vi = ui[j];
            for (int m = 0; m < j; ++m) {
                vr -= lr[j][m] * ur[m] - li[j][m] * ui[m];
                vi -= lr[j][m] * ui[m] + li[j][m] * ur[m];
            }
            ur[j] = vr;
            ui[j] = vi;
            cr[p + j] = vr;
            ci[p + j] = vi;
        }
    }
}

// Trailing update A22 -= L21 U12 over rows [p+jb, m) x columns [p+jb, ncols).
// Work item: (problem, row tile, column tile). Four waves in a 2x2
// arrangement each own a 32x32 block = 2x2 MFMA 16x16 tiles with separate
// real and imaginary accumulators; K = jb <= 32 is staged whole in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_lu_update(const LuDesc<T> *__restrict__ descs,
                                                   const int4 *__restrict__ work, int p)
{
    typedef Mfma<T> MF;
    const int4 wk = work[blockIdx.x];
    const LuDesc<T> &D = descs[wk.x];
    const int s = D.m, nc = D.ncols;
    if (D.info[0] != 0 || p >= s)
        return;
    const int jb = min(kPanel, s - p);
    const int lda = D.lda;
    const int r0 = p + jb + wk.y * kTile, c0 = p + jb + wk.z * kTile;
    __shared__ T sa_r[kPanel][kTile + kLuPad], sa_i[kPanel][kTile + kLuPad];
    __shared__ T sb_r[kPanel][kTile + kLuPad], sb_i[kPanel][kTile + kLuPad];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    // A tile: rows r0.., k columns p..p+jb (column-major: contiguous rows)
    for (int e = t; e < kPanel * kTile; e += 256) {
        const int kk = e / kTile, rr = e % kTile;
        const int gr = r0 + rr;
        T vr = 0, vi = 0;
        if (kk < jb && gr < s) {
            vr = D.ar[(size_t)(p + kk) * lda + gr];
            vi = D.ai[(size_t)(p + kk) * lda + gr];
        }
        sa_r[kk][rr] = vr;
        sa_i[kk][rr] = vi;
    }
    // B tile: k rows p..p+jb, columns c0.. (contiguous k within a column)
    for (int e = t; e < kPanel * kTile; e += 256) {
        const int cc = e / kPanel, kk = e % kPanel;
        const int gc = c0 + cc;
        T vr = 0, vi = 0;
        if (kk < jb && gc < nc) {
            vr = D.ar[(size_t)gc * lda + p + kk];
            vi = D.ai[(size_t)gc * lda + p + kk];
        }
        sb_r[kk][cc] = vr;
        sb_i[kk][cc] = vi;
    }
    __syncthreads();
    typename MF::acc accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accr[a][b] = MF::zero();
            acci[a][b] = MF::zero();
        }
    const int lrow = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < kPanel; kk += 4) {
        T xr[2], xi[2], yr[2], yi[2];
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
                accr[a][b] = MF::fma(xr[a], yr[b], accr[a][b]);
                accr[a][b] = MF::fma(-xi[a], yi[b], accr[a][b]);
                acci[a][b] = MF::fma(xr[a], yi[b], acci[a][b]);
                acci[a][b] = MF::fma(xi[a], yr[b], acci[a][b]);
            }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = r0 + wr * 32 + a * 16 + MF::row(lane, v);
                const int gj = c0 + wc * 32 + b * 16 + lrow;
                if (gi >= s || gj >= nc)
                    continue;
                const size_t o = (size_t)gj * lda + gi;
                D.ar[o] -= accr[a][b][v];
                D.ai[o] -= acci[a][b][v];
            }
}

// FP32 trailing update for the mixed path: A22 -= L21 U12 with TM x TN
// output tiles. Four waves in a 2 x 2 arrangement each own a TM/2 x TN/2
// block of MFMA 16x16x4 tiles; the A operand enters negated, so the MFMA
// chain produces -L21 U12, and C (float4 loads in the accumulator's own row
// layout: 4 consecutive rows per lane) is added in the epilogue; C is read and
// written once per launch whatever K. The loads of C and of the next K
// chunk's operands are issued before the current chunk's products, so their
// latency overlaps the MFMA work. The update covers rows
// [rbeg, m) x columns [cbeg, min(ncols, cend)) with L21 = columns
// [kbeg, kbeg + 32 nk); those panels are full (a partial panel is always the
// last one and has no rows below it). Row quads are read and written whole: a
// quad that starts below m ends inside the column (lda is a multiple of 64),
// and the padding rows past m hold 0 (the fill writes them, and here their C
// and L21 entries are 0, so they stay 0).
constexpr int kTile32 = 128;   // rows per FP32 update tile
constexpr int kTile32N = 64;   // columns per FP32 update tile

constexpr int kXcds = 8;

// kGauss: three real products per complex multiply-add (Gauss):
// p1 = Σ ar·br, p2 = Σ ai·bi, p3 = Σ (ar + ai)(br + bi), then
// C -= (p1 - p2) + i (p3 - p1 - p2). Three MFMA chains instead of four; the
// FP32 rounding differs from the four-product form at the 1e-6 relative level,
// which the factor (a CG preconditioner) absorbs.
template <int TM, int TN, bool kGauss = false>
__global__ __launch_bounds__(256) void k_lu_update_f32(const LuDesc<float> *__restrict__ descs,
                                                       const int4 *__restrict__ work, int cnt, int kbeg, int nk,
                                                       int rbeg, int cbeg, int cend)
{
    typedef Mfma<float> MF;
    constexpr int WM = TM / 2, WN = TN / 2;       // per-wave block (2 x 2 waves)
    constexpr int AM = WM / 16, AN = WN / 16;     // MFMA tiles per wave
    // A rows padded by 16 floats: the 4 k-rows a wave reads hit disjoint banks
    constexpr int LA = TM + 16, LB = TN + 4;
    constexpr int NA = (kPanel * TM / 4) / 256;   // float4 of A per thread and chunk
    constexpr int NB = (kPanel * TN / 4) / 256;   // float4 of B per thread and chunk
    // Workgroups are dealt to the XCDs round-robin by index; XCD x takes the
    // contiguous item range [x per, (x + 1) per), so the tiles that share
    // panels of one problem share that XCD's L2.
    const int per = (cnt + kXcds - 1) / kXcds;
    const int item = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (item >= cnt)
        return;
    const int4 wk = work[item];
    const LuDesc<float> &D = descs[wk.x];
    const int s = D.m, nc = min(D.ncols, cend);
    if (D.info[0] != 0 || rbeg >= s || kbeg + kPanel * nk > s)
        return;
    const int lda = D.lda;
    const int r0 = rbeg + wk.y * TM, c0 = cbeg + wk.z * TN;
    __shared__ float sa_r[kPanel][LA], sa_i[kPanel][LA];
    __shared__ float sb_r[kPanel][LB], sb_i[kPanel][LB];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const float *ar = D.ar, *ai = D.ai;
    const int lrow = lane & 15, lk = lane >> 4;
    // operands of one K chunk in registers: A column k0 + k, rows r0 .. (TM / 4
    // float4 per column); B column c0 + c, rows k0 .. k0 + 31 (8 float4 per column)
    float4 va_r[NA], va_i[NA], vb_r[NB], vb_i[NB];
    auto load_ab = [&](int k0) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int e = t + 256 * i;
            const int k = e / (TM / 4), r4 = (e % (TM / 4)) * 4;
            const int gr = r0 + r4;
            const size_t o = (size_t)(k0 + k) * lda + gr;
            va_r[i] = va_i[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gr < s) { // whole quad: padding rows hold 0
                va_r[i] = *(const float4 *)(ar + o);
                va_i[i] = *(const float4 *)(ai + o);
            }
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int e = t + 256 * i;
            const int c = e >> 3, k4 = (e & 7) * 4;
            const int gc = c0 + c;
            vb_r[i] = vb_i[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gc < nc) {
                const size_t o = (size_t)gc * lda + k0 + k4;
                vb_r[i] = *(const float4 *)(ar + o);
                vb_i[i] = *(const float4 *)(ai + o);
            }
        }
    };
    auto store_ab = [&]() {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int e = t + 256 * i;
            const int k = e / (TM / 4), r4 = (e % (TM / 4)) * 4;
            *(float4 *)&sa_r[k][r4] = va_r[i];
            *(float4 *)&sa_i[k][r4] = va_i[i];
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int e = t + 256 * i;
            const int c = e >> 3, k4 = (e & 7) * 4;
            sb_r[k4][c] = vb_r[i].x;
            sb_r[k4 + 1][c] = vb_r[i].y;
            sb_r[k4 + 2][c] = vb_r[i].z;
            sb_r[k4 + 3][c] = vb_r[i].w;
            sb_i[k4][c] = vb_i[i].x;
            sb_i[k4 + 1][c] = vb_i[i].y;
            sb_i[k4 + 2][c] = vb_i[i].z;
            sb_i[k4 + 3][c] = vb_i[i].w;
        }
    };
    load_ab(kbeg);
    // C, read after the first operands so that its loads stay in flight during
    // the first chunk's products (4 consecutive rows per lane: the
    // accumulator's row layout); it is added in the epilogue
    f4 cr[AM][AN], ci[AM][AN];
#pragma unroll
    for (int a = 0; a < AM; ++a)
#pragma unroll
        for (int b = 0; b < AN; ++b) {
            const int gi = r0 + wr * WM + a * 16 + 4 * lk; // rows gi .. gi + 3
            const int gj = c0 + wc * WN + b * 16 + lrow;
            cr[a][b] = ci[a][b] = f4{0.f, 0.f, 0.f, 0.f};
            if (gj < nc && gi < s) {
                const size_t o = (size_t)gj * lda + gi;
                const float4 x = *(const float4 *)(ar + o), y = *(const float4 *)(ai + o);
                cr[a][b] = f4{x.x, x.y, x.z, x.w};
                ci[a][b] = f4{y.x, y.y, y.z, y.w};
            }
        }
    // four-product form: accr = -Re, acci = -Im of the product sum; Gauss
    // form: accr = p1, acci = p2, accs = p3
    typename MF::acc accr[AM][AN], acci[AM][AN], accs[kGauss ? AM : 1][kGauss ? AN : 1];
#pragma unroll
    for (int a = 0; a < AM; ++a)
#pragma unroll
        for (int b = 0; b < AN; ++b) {
            accr[a][b] = MF::zero();
            acci[a][b] = MF::zero();
            if constexpr (kGauss)
                accs[a][b] = MF::zero();
        }
    store_ab();
    __syncthreads();
    for (int chunk = 0; chunk < nk; ++chunk) {
        if (chunk + 1 < nk)
            load_ab(kbeg + kPanel * (chunk + 1)); // in flight during this chunk's products
#pragma unroll
        for (int kk = 0; kk < kPanel; kk += 4) {
            float xr[AM], xi[AM], yr[AN], yi[AN];
#pragma unroll
            for (int a = 0; a < AM; ++a) {
                xr[a] = sa_r[kk + lk][wr * WM + a * 16 + lrow];
                xi[a] = sa_i[kk + lk][wr * WM + a * 16 + lrow];
            }
#pragma unroll
            for (int b = 0; b < AN; ++b) {
                yr[b] = sb_r[kk + lk][wc * WN + b * 16 + lrow];
                yi[b] = sb_i[kk + lk][wc * WN + b * 16 + lrow];
            }
            if constexpr (kGauss) {
                float xs[AM], ys[AN];
#pragma unroll
                for (int a = 0; a < AM; ++a)
                    xs[a] = xr[a] + xi[a];
#pragma unroll
                for (int b = 0; b < AN; ++b)
                    ys[b] = yr[b] + yi[b];
#pragma unroll
                for (int a = 0; a < AM; ++a)
#pragma unroll
                    for (int b = 0; b < AN; ++b) {
                        accr[a][b] = MF::fma(xr[a], yr[b], accr[a][b]);
                        acci[a][b] = MF::fma(xi[a], yi[b], acci[a][b]);
                        accs[a][b] = MF::fma(xs[a], ys[b], accs[a][b]);
                    }
            } else {
#pragma unroll
                for (int a = 0; a < AM; ++a)
#pragma unroll
                    for (int b = 0; b < AN; ++b) {
                        // -(xr + i xi)(yr + i yi)
                        accr[a][b] = MF::fma(-xr[a], yr[b], accr[a][b]);
                        accr[a][b] = MF::fma(xi[a], yi[b], accr[a][b]);
                        acci[a][b] = MF::fma(-xr[a], yi[b], acci[a][b]);
                        acci[a][b] = MF::fma(-xi[a], yr[b], acci[a][b]);
                    }
            }
        }
        if (chunk + 1 < nk) {
            __syncthreads(); // this chunk's operands consumed
            store_ab();
            __syncthreads();
        }
    }
#pragma unroll
    for (int a = 0; a < AM; ++a)
#pragma unroll
        for (int b = 0; b < AN; ++b) {
            const int gi = r0 + wr * WM + a * 16 + 4 * lk;
            const int gj = c0 + wc * WN + b * 16 + lrow;
            if (gj >= nc || gi >= s)
                continue;
            const size_t o = (size_t)gj * lda + gi;
            f4 vr, vi;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if constexpr (kGauss) {
                    vr[q] = cr[a][b][q] - accr[a][b][q] + acci[a][b][q];
                    vi[q] = ci[a][b][q] - accs[a][b][q] + accr[a][b][q] + acci[a][b][q];
                } else {
                    vr[q] = cr[a][b][q] + accr[a][b][q];
                    vi[q] = ci[a][b][q] + acci[a][b][q];
                }
            }
            // rows past m (at most 3, inside the padded column) keep 0:
            // their C and L21 entries are 0
            *(float4 *)(D.ar + o) = make_float4(vr[0], vr[1], vr[2], vr[3]);
            *(float4 *)(D.ai + o) = make_float4(vi[0], vi[1], vi[2], vi[3]);
        }
}

typedef _Float16 h8 __attribute__((ext_vector_type(8)));

// x = hi + lo 2^-11 with hi and lo in f16: hi is x rounded to f16 (0 below
// the f16 normal range, so that no hi is subnormal), lo the remainder scaled
// into the f16 range. x - hi and the scaling are exact in FP32, so
// |x - hi - lo 2^-11| <= 2^-22 |x| + 2^-26 for |x| < 65504.
__device__ __forceinline__ void split_f16x2(float x, _Float16 &hi, _Float16 &lo)
{
    const bool normal = fabsf(x) >= 0x1p-14f;
    const _Float16 h = normal ? (_Float16)x : (_Float16)0.f;
    hi = h;
    lo = (_Float16)((x - (float)h) * 2048.f);
}

// The K = 128 passes' operands, split once per combining step: L21 (rows
// [rbeg, ldz) of the columns [kbeg, kbeg + 128)) and U12 (rows [kbeg, kbeg +
// 128) of the columns [rbeg, ldz)) into f16 hi / lo planes of the real and
// imaginary parts. Halves h16[kind][chunk * 4 + plane][g][index][8] (kind 0:
// L21 by row; 1: U12 by column; chunk: 32 k; g: 8 k of the chunk), so that
// the 8 k one MFMA lane supplies are 16 contiguous bytes and 16 consecutive
// rows of one k group are 256. Rows and columns at or past the order are
// written as 0. Work item (problem, block of 64 rows / columns, chunk * 2 +
// kind); 8 k of one index per thread: for L21 thread t takes row (t & 63) and
// group (t >> 6) (a load instruction reads 64 consecutive rows of a column),
// for U12 column (t >> 2) and group (t & 3) (four lanes read a column's 32
// contiguous k).
__global__ __launch_bounds__(256) void k_split_f16x2(const LuDesc<float> *__restrict__ descs,
                                                     const int4 *__restrict__ work, int kbeg, int rbeg)
{
    const int4 wk = work[blockIdx.x];
    const LuDesc<float> &D = descs[wk.x];
    if (D.info[0] != 0)
        return;
    const int s = D.m, lda = D.lda;
    const int c = wk.z >> 1, kind = wk.z & 1, t = threadIdx.x;
    const int li = kind == 0 ? (t & 63) : (t >> 2), g = kind == 0 ? (t >> 6) : (t & 3);
    const int idx = rbeg + wk.y * 64 + li, k0 = kbeg + 32 * c + 8 * g;
    const float *ar = D.ar, *ai = D.ai;
    float xr[8], xi[8];
    if (kind == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            xr[q] = xi[q] = 0.f;
            if (idx < s) {
                const size_t o = (size_t)(k0 + q) * lda + idx;
                xr[q] = ar[o];
                xi[q] = ai[o];
            }
        }
    } else {
        float4 b0r = make_float4(0.f, 0.f, 0.f, 0.f), b1r = b0r, b0i = b0r, b1i = b0r;
        if (idx < min(D.ncols, s)) {
            const size_t o = (size_t)idx * lda + k0;
            b0r = *(const float4 *)(ar + o);
            b1r = *(const float4 *)(ar + o + 4);
            b0i = *(const float4 *)(ai + o);
            b1i = *(const float4 *)(ai + o + 4);
        }
        xr[0] = b0r.x, xr[1] = b0r.y, xr[2] = b0r.z, xr[3] = b0r.w;
        xr[4] = b1r.x, xr[5] = b1r.y, xr[6] = b1r.z, xr[7] = b1r.w;
        xi[0] = b0i.x, xi[1] = b0i.y, xi[2] = b0i.z, xi[3] = b0i.w;
        xi[4] = b1i.x, xi[5] = b1i.y, xi[6] = b1i.z, xi[7] = b1i.w;
    }
    h8 pl[4];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        _Float16 a, b;
        split_f16x2(xr[q], a, b);
        pl[0][q] = a, pl[1][q] = b;
        split_f16x2(xi[q], a, b);
        pl[2][q] = a, pl[3][q] = b;
    }
    _Float16 *h = (_Float16 *)D.h16 + (size_t)kind * 16 * lda * 32;
#pragma unroll
    for (int p = 0; p < 4; ++p)
        *(h8 *)(h + (((size_t)(c * 4 + p) * 4 + g) * lda + idx) * 8) = pl[p];
}

// FP32 trailing update on the f16 matrix cores, for K = 128 passes: A22 -=
// L21 U12 over 64 x 64 tiles, with the work list, bounds and launch shape of
// k_lu_update_f32<64, 64>. Every FP32 operand is split by split_f16x2 (once
// per combining step, k_split_f16x2), and a real product a b is taken as
// hi_a hi_b + 2^-11 (hi_a lo_b + lo_a hi_b) with FP32 accumulation: relative
// error about 2^-21 per product (the dropped lo lo term and the rounding of
// lo), where FP32 products have 2^-24. The factor is a CG preconditioner, so
// this changes the iteration's rate, not the solution it converges to.
// Operands must stay below 65504 in modulus (the f16 range of hi): L21 is
// bounded by 1 through partial pivoting and U12 by the pivot growth times
// max |Z|, whose Hallén entries are O(10). A larger operand would turn the
// factor non-finite, which the solve reports as a failure status (a singular
// pivot or a CG breakdown), never as a wrong result. Per 16 x 16 tile and
// 32-deep K chunk: 12 v_mfma_f32_16x16x32_f16 (hi hi and two cross terms for
// each of the four real products of the complex product), where the FP32
// forms issue 24 or 32 v_mfma_f32_16x16x4f32 of the same issue cost each.
//
// Two forms. k_lu_update_f16x2d: each wavefront loads its MFMA operands
// straight from the split layout into registers: lane l of a 16-row operand
// reads the 16 bytes of row l % 16 in k group l / 16, so one load
// instruction reads four runs of 256 contiguous bytes. A wavefront owns a
// 32 x 32 block of the tile (2 x 2 MFMA tiles); chunk c + 1's operands are
// loaded while chunk c's products issue (double buffer in registers), and C
// is loaded during the last chunk. k_lu_update_f16x2s streams the operands
// through LDS (below). Both issue the same products in the same
// order per accumulator, so they give the same bits.

// A pointer into global memory (address space 1), for kernels whose
// pointers arrive through descriptors and would otherwise be generic (flat).
template <typename T> using gptr = __attribute__((address_space(1))) T *;

__device__ __forceinline__ h8 neg_h8(h8 x)
{
    typedef unsigned u4v __attribute__((ext_vector_type(4)));
    u4v u = __builtin_bit_cast(u4v, x);
    u ^= 0x80008000u; // the sign bits of the eight halves
    return __builtin_bit_cast(h8, u);
}

__global__ __launch_bounds__(256, 2) void k_lu_update_f16x2d(const LuDesc<float> *__restrict__ descs,
                                                             const int4 *__restrict__ work, int cnt, int kbeg,
                                                             int rbeg, int cbeg, int cend)
{
    constexpr int TM = 64, TN = 64, NK = 4;
    const int per = (cnt + kXcds - 1) / kXcds;
    const int item = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (item >= cnt)
        return;
    const int4 wk = work[item];
    const LuDesc<float> &D = descs[wk.x];
    const int s = D.m, nc = min(D.ncols, cend);
    if (D.info[0] != 0 || rbeg >= s || kbeg + kPanel * NK > s)
        return;
    const int lda = D.lda;
    const int r0 = rbeg + wk.y * TM, c0 = cbeg + wk.z * TN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int lrow = lane & 15, lk = lane >> 4;
    const size_t ps = (size_t)lda * 32; // halves per plane of a chunk
    // global-address-space views: loads then count on vmcnt alone, so a wait
    // for one buffer leaves the other's loads in flight
    const gptr<_Float16> h = (gptr<_Float16>)D.h16;
    const gptr<_Float16> pa = h + ((size_t)lk * lda + r0 + wr * 32 + lrow) * 8;
    const gptr<_Float16> pb = h + 16 * ps + ((size_t)lk * lda + c0 + wc * 32 + lrow) * 8;
    h8 A[2][2][4], B[2][2][4]; // [buffer][16-row block][plane]
    auto load = [&](int buf, int chunk) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                A[buf][a][p] = *(gptr<h8>)(pa + (size_t)(chunk * 4 + p) * ps + a * 16 * 8);
                B[buf][a][p] = *(gptr<h8>)(pb + (size_t)(chunk * 4 + p) * ps + a * 16 * 8);
            }
    };
    f4 re_m[2][2], re_x[2][2], im_m[2][2], im_x[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
            re_m[a][b] = re_x[a][b] = im_m[a][b] = im_x[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    const gptr<float> ar = (gptr<float>)D.ar, ai = (gptr<float>)D.ai;
    f4 cr[2][2], ci[2][2];
    load(0, 0);
#pragma unroll
    for (int chunk = 0; chunk < NK; ++chunk) {
        const int cur = chunk & 1;
        if (chunk + 1 < NK) {
            load(cur ^ 1, chunk + 1);
        } else {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    // unconditional (a branch here would make the compiler
                    // wait for every load in flight): rows below the order
                    // stay inside the padded column, and a column past the
                    // last carried one reads the last, its result unused
                    const int gi = r0 + wr * 32 + a * 16 + 4 * lk;
                    const int gj = min(c0 + wc * 32 + b * 16 + lrow, nc - 1);
                    const size_t o = (size_t)gj * lda + gi;
                    cr[a][b] = *(gptr<f4>)(ar + o);
                    ci[a][b] = *(gptr<f4>)(ai + o);
                }
        }
        // keep this chunk's products after the next chunk's loads, so that
        // the waits for this chunk's operands leave those loads in flight
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const h8 *fa = A[cur][a];
            const h8 na0 = neg_h8(fa[2]), na1 = neg_h8(fa[3]);
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const h8 *fb = B[cur][b];
                re_m[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], fb[0], re_m[a][b], 0, 0, 0);
                re_m[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(na0, fb[2], re_m[a][b], 0, 0, 0);
                re_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], fb[1], re_x[a][b], 0, 0, 0);
                re_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[1], fb[0], re_x[a][b], 0, 0, 0);
                re_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(na0, fb[3], re_x[a][b], 0, 0, 0);
                re_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(na1, fb[2], re_x[a][b], 0, 0, 0);
                im_m[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], fb[2], im_m[a][b], 0, 0, 0);
                im_m[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[2], fb[0], im_m[a][b], 0, 0, 0);
                im_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], fb[3], im_x[a][b], 0, 0, 0);
                im_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[1], fb[2], im_x[a][b], 0, 0, 0);
                im_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[2], fb[1], im_x[a][b], 0, 0, 0);
                im_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[3], fb[0], im_x[a][b], 0, 0, 0);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    constexpr float kLo = 0x1p-11f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int gi = r0 + wr * 32 + a * 16 + 4 * lk; // rows gi .. gi + 3
            const int gj = c0 + wc * 32 + b * 16 + lrow;
            if (gj >= nc || gi >= s)
                continue;
            const size_t o = (size_t)gj * lda + gi;
            f4 vr, vi;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                vr[q] = cr[a][b][q] - fmaf(re_x[a][b][q], kLo, re_m[a][b][q]);
                vi[q] = ci[a][b][q] - fmaf(im_x[a][b][q], kLo, im_m[a][b][q]);
            }
            *(gptr<f4>)(ar + o) = vr;
            *(gptr<f4>)(ai + o) = vi;
        }
}

// The update streamed through a five-stage LDS ring by the gfx950 LDS DMA
// loads (global_load_lds_dwordx4): a persistent grid, one workgroup per CU,
// each taking every kUpdRingSlots-th tile of its XCD's contiguous share of
// the list. The 32 KiB stage of a chunk holds 32 MFMA fragments (L21 row
// blocks and U12 column blocks of 16, four planes each), each written by one
// DMA instruction in lane order (lane l: row l % 16, k group l / 16), so a
// wavefront reads a fragment with one conflict-free ds_read_b128 at lane x
// 16. Chunks stream across tiles: while chunk q is multiplied, chunks q + 1
// .. q + 4 (the next tile's during this one) are in flight, and C is loaded
// into registers a tile ahead, so the start-up latency of a tile is hidden
// behind the products of the one before. Before stage q is read, each
// wavefront waits for its own DMA of chunk q (the loads issued later stay in
// flight: loads complete in order) and the workgroup meets at a barrier,
// which also frees stage q - 1 for chunk q + 4. The products and their order
// per accumulator are those of k_lu_update_f16x2d.
constexpr int kUpdRingSlots = 32;  // workgroups per XCD: one on each of its CUs
// 5 x 32 KiB: the whole LDS of a CU. The DMA runs kUpdRingStages - 1 = 4
// chunks (one tile) ahead, which the loop's chunk and count bookkeeping
// assumes: chunk c of the next tile is issued at chunk c of this one.
constexpr int kUpdRingStages = 5;

// The ring's stage hand-over: this wavefront's DMA of the chunk about to be
// read has landed (at most N of its vector memory operations outstanding,
// which complete in order), its LDS reads of the previous chunk have
// returned, then the workgroup's barrier. Nothing is scheduled across it.
template <int N> __device__ __forceinline__ void ring_handover()
{
    __builtin_amdgcn_sched_barrier(0);
    // s_waitcnt vmcnt(N) lgkmcnt(0), expcnt left alone (gfx9 encoding)
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4));
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_sched_barrier(0);
}

__global__ __launch_bounds__(256, 1) void k_lu_update_f16x2s(const LuDesc<float> *__restrict__ descs,
                                                             const int4 *__restrict__ work, int cnt, int kbeg,
                                                             int rbeg, int cbeg, int cend)
{
    constexpr int TM = 64, TN = 64;
    __shared__ __align__(16) h8 ring[kUpdRingStages][32][64];
    const int per = (cnt + kXcds - 1) / kXcds;
    const int lo = (blockIdx.x % kXcds) * per, hi = min(cnt, lo + per);
    const int slot = blockIdx.x / kXcds, nslot = gridDim.x / kXcds;
    // the wavefront index as a scalar: the DMA's fragment choice and LDS
    // address then stay in scalar registers
    const int t = threadIdx.x, lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wr = wave >> 1, wc = wave & 1;
    const int lrow = lane & 15, lk = lane >> 4;
    typedef const __attribute__((address_space(4))) LuDesc<float> CDesc;
    typedef const __attribute__((address_space(4))) int CInt;
    struct Tile {
        bool ok;
        int it, s, nc, lda, r0, c0;
        float *ar, *ai;
        const _Float16 *h;
    };
    // the first tile the update computes among items it, it + nslot, ...
    // (a problem that is singular or too small for this pass is skipped);
    // ok = false when there is none. Scalar loads through the constant
    // address space, which leave the vector loads' counter alone.
    auto next_tile = [&](int it) {
        Tile T;
        T.ok = false;
        for (; it < hi; it += nslot) {
            CInt *wk = (CInt *)work + 4 * (size_t)it;
            CDesc &D = ((CDesc *)descs)[wk[0]];
            T.s = D.m;
            if (*(CInt *)D.info != 0 || rbeg >= T.s || kbeg + 4 * kPanel > T.s)
                continue;
            T.ok = true;
            T.it = it;
            T.nc = min(D.ncols, cend);
            T.lda = D.lda;
            T.r0 = rbeg + wk[1] * TM;
            T.c0 = cbeg + wk[2] * TN;
            T.ar = D.ar;
            T.ai = D.ai;
            T.h = (const _Float16 *)D.h16;
            break;
        }
        return T;
    };
    // this wavefront's 8 of the 32 fragments of chunk c of tile X into stage S
    auto dma = [&](const Tile &X, int chunk, h8 (*S)[64]) {
        const unsigned ps = (unsigned)X.lda * 64; // bytes per (chunk, plane)
        const gptr<const char> hb = (gptr<const char>)X.h;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int f = wave * 8 + j, kind = f >> 4, blk = (f >> 2) & 3, p = f & 3;
            const int idx = (kind ? X.c0 : X.r0) + blk * 16 + lrow;
            const unsigned off =
                kind * 16 * ps + (unsigned)(chunk * 4 + p) * ps + ((unsigned)(lk * X.lda + idx)) * 16;
            __builtin_amdgcn_global_load_lds((gptr<const void>)(hb + off),
                                             (__attribute__((address_space(3))) void *)&S[f][0], 16, 0, 0);
        }
    };
    f4 re_m[2][2], re_x[2][2], im_m[2][2], im_x[2][2];
    // A chunk's products from stage S. The fragments are read by inline
    // ds_read_b128 blocks that wait for their own results: the compiler
    // then has no LDS read to order after the DMA, which it would otherwise
    // guard with a wait for every vector memory operation in flight (it
    // cannot tell the DMA of this stage from the later ones once stores are
    // outstanding); the hand-over already covers this stage's DMA.
    auto products = [&](h8 (*S)[64]) {
        const unsigned base = (unsigned)(size_t)&S[0][lane]; // LDS byte address
        const unsigned va = base + (wr * 2) * 4 * 1024, vb = base + (16 + wc * 2 * 4) * 1024;
        h8 fa[4], fb[2][4];
        asm volatile("ds_read_b128 %0, %12\n\t"
                     "ds_read_b128 %1, %12 offset:1024\n\t"
                     "ds_read_b128 %2, %12 offset:2048\n\t"
                     "ds_read_b128 %3, %12 offset:3072\n\t"
                     "ds_read_b128 %4, %13\n\t"
                     "ds_read_b128 %5, %13 offset:1024\n\t"
                     "ds_read_b128 %6, %13 offset:2048\n\t"
                     "ds_read_b128 %7, %13 offset:3072\n\t"
                     "ds_read_b128 %8, %13 offset:4096\n\t"
                     "ds_read_b128 %9, %13 offset:5120\n\t"
                     "ds_read_b128 %10, %13 offset:6144\n\t"
                     "ds_read_b128 %11, %13 offset:7168\n\t"
                     "s_waitcnt lgkmcnt(0)"
                     : "=&v"(fa[0]), "=&v"(fa[1]), "=&v"(fa[2]), "=&v"(fa[3]), "=&v"(fb[0][0]), "=&v"(fb[0][1]),
                       "=&v"(fb[0][2]), "=&v"(fb[0][3]), "=&v"(fb[1][0]), "=&v"(fb[1][1]), "=&v"(fb[1][2]),
                       "=&v"(fb[1][3])
                     : "v"(va), "v"(vb));
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (a == 1)
                asm volatile("ds_read_b128 %0, %4 offset:4096\n\t"
                             "ds_read_b128 %1, %4 offset:5120\n\t"
                             "ds_read_b128 %2, %4 offset:6144\n\t"
                             "ds_read_b128 %3, %4 offset:7168\n\t"
                             "s_waitcnt lgkmcnt(0)"
                             : "=&v"(fa[0]), "=&v"(fa[1]), "=&v"(fa[2]), "=&v"(fa[3])
                             : "v"(va));
            const h8 na0 = neg_h8(fa[2]), na1 = neg_h8(fa[3]);
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const h8 *g = fb[b];
                re_m[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], g[0], re_m[a][b], 0, 0, 0);
                re_m[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(na0, g[2], re_m[a][b], 0, 0, 0);
                re_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], g[1], re_x[a][b], 0, 0, 0);
                re_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[1], g[0], re_x[a][b], 0, 0, 0);
                re_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(na0, g[3], re_x[a][b], 0, 0, 0);
                re_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(na1, g[2], re_x[a][b], 0, 0, 0);
                im_m[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], g[2], im_m[a][b], 0, 0, 0);
                im_m[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[2], g[0], im_m[a][b], 0, 0, 0);
                im_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[0], g[3], im_x[a][b], 0, 0, 0);
                im_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[1], g[2], im_x[a][b], 0, 0, 0);
                im_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[2], g[1], im_x[a][b], 0, 0, 0);
                im_x[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fa[3], g[0], im_x[a][b], 0, 0, 0);
            }
        }
    };
    // C of tile X (unconditional: rows below the order stay inside the
    // padded column, and a column past the last carried one reads the last,
    // its result unused)
    auto load_c = [&](const Tile &X, f4 (&r)[2][2], f4 (&m)[2][2]) {
        // uniform plane bases plus 32-bit byte offsets (a plane is at most
        // ldz x n FP32): one VGPR per load and no 64-bit temporaries
        const gptr<const char> br = (gptr<const char>)X.ar, bi = (gptr<const char>)X.ai;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int gi = X.r0 + wr * 32 + a * 16 + 4 * lk;
                const int gj = min(X.c0 + wc * 32 + b * 16 + lrow, X.nc - 1);
                const unsigned o = ((unsigned)gj * X.lda + gi) * 4;
                r[a][b] = *(gptr<const f4>)(br + o);
                m[a][b] = *(gptr<const f4>)(bi + o);
            }
    };
    // one tile: its C is in (cr, ci), loaded a tile ahead; the next tile's C
    // goes to (cn, cin). Returns the next tile.
    // one tile: its C is in (cr, ci), loaded a tile ahead; the next tile's C
    // goes to (cn, cin). Chunk q of the stream (four per tile) sits in stage
    // q mod kUpdRingStages; sq is the current chunk's stage. Returns the
    // next tile.
    int sq = 0;
    auto run_tile = [&](const Tile &T, f4 (&cr)[2][2], f4 (&ci)[2][2], f4 (&cn)[2][2], f4 (&cin)[2][2]) {
        const Tile N = next_tile(T.it + nslot);
        const Tile Nd = N.ok ? N : T; // with no next tile, a harmless reload keeps the counts fixed
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                re_m[a][b] = re_x[a][b] = im_m[a][b] = im_x[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            // Issued after this chunk's DMA: the next three chunks and one
            // tile's C (the DMA runs four chunks ahead and C is issued once
            // per tile): 32 loads of this wavefront. Outstanding stores
            // only add to the count.
            ring_handover<32>();
            const int sd = sq == 0 ? kUpdRingStages - 1 : sq - 1; // stage of chunk q + 4, freed by the barrier
            dma(Nd, c, ring[sd]);
            if (c == 0)
                load_c(Nd, cn, cin); // the next tile's C, a tile ahead
            __builtin_amdgcn_sched_barrier(0); // the loads issue before the products
            products(ring[sq]);
            sq = sq + 1 == kUpdRingStages ? 0 : sq + 1;
        }
        constexpr float kLo = 0x1p-11f;
        const gptr<float> ar = (gptr<float>)T.ar, ai = (gptr<float>)T.ai;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int gi = T.r0 + wr * 32 + a * 16 + 4 * lk; // rows gi .. gi + 3
                const int gj = T.c0 + wc * 32 + b * 16 + lrow;
                if (gj >= T.nc || gi >= T.s)
                    continue;
                const size_t o = (size_t)gj * T.lda + gi;
                f4 vr, vi;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    vr[q] = cr[a][b][q] - fmaf(re_x[a][b][q], kLo, re_m[a][b][q]);
                    vi[q] = ci[a][b][q] - fmaf(im_x[a][b][q], kLo, im_m[a][b][q]);
                }
                *(gptr<f4>)(ar + o) = vr;
                *(gptr<f4>)(ai + o) = vi;
            }
        return N;
    };
    // two register sets for C, alternating between tiles (a copy from one
    // to the other would wait for the loads in flight)
    f4 c0r[2][2], c0i[2][2], c1r[2][2], c1i[2][2];
    Tile T = next_tile(lo + slot);
    if (!T.ok)
        return;
#pragma unroll
    for (int c = 0; c < 4; ++c)
        dma(T, c, ring[c]);
    load_c(T, c0r, c0i);
    while (true) {
        T = run_tile(T, c0r, c0i, c1r, c1i);
        if (!T.ok)
            break;
        T = run_tile(T, c1r, c1i, c0r, c0i);
        if (!T.ok)
            break;
    }
    // every DMA landed before the workgroup ends (the last tile's reloads)
    __builtin_amdgcn_s_waitcnt(0);
}

// L11^{-1} of a 64-column outer step of the FP32 LU (L11 unit lower at
// (P, P), after both panels' interchanges) into D.linv, column-major, real
// then imaginary plane. Right-looking on X = I in LDS: at step j every
// X[i][c] with i > j, c <= j takes -= L[i][j] X[j][c], spread over the
// workgroup, one barrier per step.
__global__ __launch_bounds__(256) void k_l11_inv(const LuDesc<float> *__restrict__ descs,
                                                 const int *__restrict__ plist, int P)
{
    constexpr int NB = 2 * kPanel;
    const LuDesc<float> &D = descs[plist[blockIdx.x]];
    if (D.info[0] != 0 || P + NB >= D.m)
        return;
    __shared__ float lr[NB][NB + 1], li[NB][NB + 1]; // L11 [row][col]
    __shared__ float xr[NB][NB + 1], xi[NB][NB + 1]; // L11^{-1} [row][col]
    const int t = threadIdx.x, lda = D.lda;
    for (int e = t; e < NB * NB; e += 256) {
        const int rr = e % NB, cc = e / NB;
        lr[rr][cc] = D.ar[(size_t)(P + cc) * lda + P + rr];
        li[rr][cc] = D.ai[(size_t)(P + cc) * lda + P + rr];
        xr[rr][cc] = rr == cc ? 1.f : 0.f;
        xi[rr][cc] = 0.f;
    }
    __syncthreads();
    for (int j = 0; j < NB - 1; ++j) {
        const int w = j + 1, cnt = (NB - 1 - j) * w;
        for (int e = t; e < cnt; e += 256) {
            const int i = j + 1 + e / w, c = e % w;
            const float a = lr[i][j], b = li[i][j], x = xr[j][c], y = xi[j][c];
            xr[i][c] -= a * x - b * y;
            xi[i][c] -= a * y + b * x;
        }
        __syncthreads();
    }
    for (int e = t; e < NB * NB; e += 256) {
        const int rr = e % NB, cc = e / NB;
        D.linv[e] = xr[rr][cc];
        D.linv[NB * NB + e] = xi[rr][cc];
    }
}

// Outer step of the two-level FP32 LU. Right items (problem, g >= 0): for the
// 64 columns [P + 64 + 64 g, + 64) apply the interchanges of both 32-column
// panels in order, then U12 = L11^{-1} A12 in place as a 64 x 64 x 64 complex
// product on MFMA (four waves, 2 x 2 blocks of 32 x 32), with L11^{-1} from
// k_l11_inv. Left items (problem, -1 - g): the same interchanges on the
// columns [64 g, 64 g + 64) left of P (the multipliers of earlier steps), so
// the factors read P Z = L U.
constexpr int kTrsm64Cols = 64;

// k_lu_trsm64_f32 modes: the interchanges (left items, and the columns of
// right items), the L11^{-1} product (right items), or both in one launch.
constexpr int kTrsmSwap = 1, kTrsmSolve = 2;

__global__ __launch_bounds__(256) void k_lu_trsm64_f32(const LuDesc<float> *__restrict__ descs,
                                                       const int2 *__restrict__ work, int P, int mode)
{
    typedef Mfma<float> MF;
    // K in two chunks of 32 staged in turn: 37 KB of LDS, four workgroups per CU
    constexpr int NB = 2 * kPanel, KC = NB / 2, LPA = NB + 16, LPB = NB + 4;
    const int2 wk = work[blockIdx.x];
    const LuDesc<float> &D = descs[wk.x];
    const int s = D.m;
    if (D.info[0] != 0 || P >= s)
        return;
    const bool left = wk.y < 0;
    if (!left && P + NB >= s)
        return;
    if (left && !(mode & kTrsmSwap))
        return;
    const int lda = D.lda, t = threadIdx.x;
    const int jbt = min(NB, s - P); // interchanges of this step
    __shared__ float sa_r[KC][LPA], sa_i[KC][LPA]; // L11^{-1}, one K chunk: [k][row]
    __shared__ float sb_r[KC][LPB], sb_i[KC][LPB]; // A12 block, one K chunk: [k][column]
    __shared__ int piv_s[NB];
    if (t < NB)
        piv_s[t] = t < jbt ? D.ipiv[P + t] : P + t;
    __syncthreads();
    const bool any_swap = __syncthreads_or(t < NB && piv_s[t] != P + t);
    auto swap_rows = [&](float *cr, float *ci) {
        for (int j = 0; j < jbt; ++j) {
            const int piv = piv_s[j];
            if (piv != P + j) {
                const float xr = cr[P + j], xi = ci[P + j];
                cr[P + j] = cr[piv];
                ci[P + j] = ci[piv];
                cr[piv] = xr;
                ci[piv] = xi;
            }
        }
    };
    if (left) {
        const int cl = (-wk.y - 1) * 64 + t;
        if (t < 64 && any_swap && cl < P)
            swap_rows(D.ar + (size_t)cl * lda, D.ai + (size_t)cl * lda);
        return;
    }
    const int c0 = P + NB + wk.y * kTrsm64Cols;
    const int nc = min(kTrsm64Cols, D.ncols - c0);
    if ((mode & kTrsmSwap) && any_swap && t < nc)
        swap_rows(D.ar + (size_t)(c0 + t) * lda, D.ai + (size_t)(c0 + t) * lda);
    if (!(mode & kTrsmSolve))
        return;
    __syncthreads(); // the interchanged rows are visible to the workgroup
    // stage K chunk h of L11^{-1} ([k][row] from its column-major copy) and
    // of the A12 block (rows P + 32 h .. + 31 of the columns c0 ..), one
    // float4 of 4 rows per access
    auto stage = [&](int h) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int e = t + 256 * i;       // 512 float4 per plane and chunk
            const int k = e >> 4, r4 = (e & 15) * 4;
            const int kg = KC * h + k;
            *(float4 *)&sa_r[k][r4] = *(const float4 *)(D.linv + kg * NB + r4);
            *(float4 *)&sa_i[k][r4] = *(const float4 *)(D.linv + NB * NB + kg * NB + r4);
            const int c = e >> 3, q4 = (e & 7) * 4; // column of the block, chunk rows q4 .. q4 + 3
            float4 br = make_float4(0.f, 0.f, 0.f, 0.f), bi = br;
            if (c < nc) {
                const size_t o = (size_t)(c0 + c) * lda + P + KC * h + q4;
                br = *(const float4 *)(D.ar + o);
                bi = *(const float4 *)(D.ai + o);
            }
            sb_r[q4][c] = br.x;
            sb_r[q4 + 1][c] = br.y;
            sb_r[q4 + 2][c] = br.z;
            sb_r[q4 + 3][c] = br.w;
            sb_i[q4][c] = bi.x;
            sb_i[q4 + 1][c] = bi.y;
            sb_i[q4 + 2][c] = bi.z;
            sb_i[q4 + 3][c] = bi.w;
        }
    };
    const int lane = t & 63, wave = t >> 6, wr = wave >> 1, wc = wave & 1;
    const int lrow = lane & 15, lk = lane >> 4;
    // three-product (Gauss) form, as in k_lu_update_f32: p1 = Σ xr yr,
    // p2 = Σ xi yi, p3 = Σ (xr + xi)(yr + yi); product (p1 - p2) + i (p3 - p1 - p2)
    typename MF::acc p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            p1[a][b] = MF::zero();
            p2[a][b] = MF::zero();
            p3[a][b] = MF::zero();
        }
    for (int h = 0; h < 2; ++h) {
    if (h)
        __syncthreads(); // chunk 0 consumed
    stage(h);
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < KC; kk += 4) {
        float xr[2], xi[2], xs[2], yr[2], yi[2], ys[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            xr[a] = sa_r[kk + lk][wr * 32 + a * 16 + lrow];
            xi[a] = sa_i[kk + lk][wr * 32 + a * 16 + lrow];
            yr[a] = sb_r[kk + lk][wc * 32 + a * 16 + lrow];
            yi[a] = sb_i[kk + lk][wc * 32 + a * 16 + lrow];
            xs[a] = xr[a] + xi[a];
            ys[a] = yr[a] + yi[a];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                p1[a][b] = MF::fma(xr[a], yr[b], p1[a][b]);
                p2[a][b] = MF::fma(xi[a], yi[b], p2[a][b]);
                p3[a][b] = MF::fma(xs[a], ys[b], p3[a][b]);
            }
    }
    }
    // both chunks of the block were read before any write: the product goes
    // back in place
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int r = wr * 32 + a * 16 + 4 * lk; // rows r .. r + 3 of the block
            const int c = wc * 32 + b * 16 + lrow;
            if (c >= nc)
                continue;
            const size_t o = (size_t)(c0 + c) * lda + P + r;
            f4 vr, vi;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                vr[q] = p1[a][b][q] - p2[a][b][q];
                vi[q] = p3[a][b][q] - p1[a][b][q] - p2[a][b][q];
            }
            *(float4 *)(D.ar + o) = make_float4(vr[0], vr[1], vr[2], vr[3]);
            *(float4 *)(D.ai + o) = make_float4(vi[0], vi[1], vi[2], vi[3]);
        }
}

// tm: row tile (the work list's row-tile count follows it): 64 with the
// three-product (Gauss) form, or 128 with the four-product form, each the
// faster form at its tile size.
inline void lu_launch_update_f32(const LuDesc<float> *d, size_t off, size_t cnt, const int *ints,
                                 int kbeg, int nk, int rbeg, int cbeg, int cend, int tm, hipStream_t st)
{
    const unsigned grid = (unsigned)((cnt + kXcds - 1) / kXcds * kXcds);
    if (tm == 64)
        hipLaunchKernelGGL((k_lu_update_f32<64, kTile32N, true>), dim3(grid), dim3(256), 0, st, d,
                           (const int4 *)(ints + off), (int)cnt, kbeg, nk, rbeg, cbeg, cend);
    else
        hipLaunchKernelGGL((k_lu_update_f32<kTile32, kTile32N>), dim3(grid), dim3(256), 0, st, d,
                           (const int4 *)(ints + off), (int)cnt, kbeg, nk, rbeg, cbeg, cend);
}

// The 32-column update inside a 64-column outer step (rows [P + 32, m) x
// columns [P + 32, P + 64), K = 32): 128 x 32 tiles, one column tile.
inline void lu_launch_update_narrow_f32(const LuDesc<float> *d, size_t off, size_t cnt, const int *ints,
                                        int P, hipStream_t st)
{
    const unsigned grid = (unsigned)((cnt + kXcds - 1) / kXcds * kXcds);
    hipLaunchKernelGGL((k_lu_update_f32<kTile32, kPanel>), dim3(grid), dim3(256), 0, st, d,
                       (const int4 *)(ints + off), (int)cnt, P, 1, P + kPanel, P + kPanel, P + 2 * kPanel);
}

inline void lu_launch_trsm_f32(const LuDesc<float> *d, size_t off, size_t cnt, const int *ints, int p,
                               int cend, hipStream_t st)
{
    hipLaunchKernelGGL(k_lu_trsm<float>, dim3((unsigned)cnt), dim3(64), 0, st, d,
                       (const int2 *)(ints + off), p, cend);
}

inline void lu_launch_trsm64_f32(const LuDesc<float> *d, size_t off, size_t cnt, const int *ints, int P,
                                 int mode, hipStream_t st)
{
    hipLaunchKernelGGL(k_lu_trsm64_f32, dim3((unsigned)cnt), dim3(256), 0, st, d,
                       (const int2 *)(ints + off), P, mode);
}

// K = 128 trailing update (two outer steps' L21 and U12, columns / rows
// [kbeg, kbeg + 128)) on the f16 matrix cores, 64 x 64 tiles.
// K = 128 trailing update (two outer steps' L21 and U12, columns / rows
// [kbeg, kbeg + 128), split by k_split_f16x2) on the f16 matrix cores, 64 x
// 64 tiles: the LDS ring (persistent, one workgroup per CU) or the
// register-direct form (one workgroup per tile).
inline void lu_launch_update_f16x2(const LuDesc<float> *d, size_t off, size_t cnt, const int *ints, int kbeg,
                                   int rbeg, int cbeg, int cend, bool ring, hipStream_t st)
{
    const unsigned grid = (unsigned)((cnt + kXcds - 1) / kXcds * kXcds);
    if (ring)
        hipLaunchKernelGGL(k_lu_update_f16x2s, dim3(std::min<unsigned>(grid, kXcds * kUpdRingSlots)), dim3(256), 0,
                           st, d, (const int4 *)(ints + off), (int)cnt, kbeg, rbeg, cbeg, cend);
    else
        hipLaunchKernelGGL(k_lu_update_f16x2d, dim3(grid), dim3(256), 0, st, d, (const int4 *)(ints + off),
                           (int)cnt, kbeg, rbeg, cbeg, cend);
}

inline void lu_launch_split_f16x2(const LuDesc<float> *d, size_t off, size_t cnt, const int *ints, int kbeg,
                                  int rbeg, hipStream_t st)
{
    hipLaunchKernelGGL(k_split_f16x2, dim3((unsigned)cnt), dim3(256), 0, st, d, (const int4 *)(ints + off), kbeg,
                       rbeg);
}

inline void lu_launch_l11_inv(const LuDesc<float> *d, size_t off, size_t cnt, const int *ints, int P,
                              hipStream_t st)
{
    hipLaunchKernelGGL(k_l11_inv, dim3((unsigned)cnt), dim3(256), 0, st, d, ints + off, P);
}

// Host side: one panel step of the batched factorisation on stream st.
template <typename T>
void lu_launch_panel(const LuDesc<T> *d, const LuStep &s, const int *ints, hipStream_t st)
{
    hipLaunchKernelGGL(k_panel_diag<T>, dim3((unsigned)s.panel_cnt), dim3(64), 0, st, d,
                       ints + s.panel_off, s.p);
    if (s.rows_cnt)
        hipLaunchKernelGGL(k_panel_rows<T>, dim3((unsigned)s.rows_cnt), dim3(256), 0, st, d,
                           (const int2 *)(ints + s.rows_off), s.p);
    hipLaunchKernelGGL(k_panel_commit<T>, dim3((unsigned)s.commit_cnt), dim3(256), 0, st, d,
                       (const int2 *)(ints + s.commit_off), s.p);
    hipLaunchKernelGGL(k_lu_panel<T>, dim3((unsigned)s.panel_cnt), dim3(kPanelThreads), 0, st, d,
                       ints + s.panel_off, s.p);
}

// FP32 panel step of the mixed path (U_D^{-1} on the diagonal block, rows on MFMA).
inline void lu_launch_panel_f32(const LuDesc<float> *d, const LuStep &s, const int *ints, hipStream_t st)
{
    hipLaunchKernelGGL(k_panel_diag_f32, dim3((unsigned)s.panel_cnt), dim3(64), 0, st, d, ints + s.panel_off,
                       s.p);
    if (s.rows_cnt)
        hipLaunchKernelGGL(k_panel_rows_f32, dim3((unsigned)s.rows_cnt), dim3(256), 0, st, d,
                           (const int2 *)(ints + s.rows_off), s.p);
    hipLaunchKernelGGL((k_panel_commit<float, true>), dim3((unsigned)s.commit_cnt), dim3(256), 0, st, d,
                       (const int2 *)(ints + s.commit_off), s.p);
    hipLaunchKernelGGL(k_lu_panel<float>, dim3((unsigned)s.panel_cnt), dim3(kPanelThreads), 0, st, d,
                       ints + s.panel_off, s.p);
}

template <typename T>
void lu_launch_trsm(const LuDesc<T> *d, const LuStep &s, const int *ints, hipStream_t st)
{
    hipLaunchKernelGGL(k_lu_trsm<T>, dim3((unsigned)s.trsm_cnt), dim3(64), 0, st, d,
                       (const int2 *)(ints + s.trsm_off), s.p, 0x7fffffff);
}

template <typename T>
void lu_launch_update(const LuDesc<T> *d, const LuStep &s, const int *ints, hipStream_t st)
{
    hipLaunchKernelGGL(k_lu_update<T>, dim3((unsigned)s.upd_cnt), dim3(256), 0, st, d,
                       (const int4 *)(ints + s.upd_off), s.p);
}

} // namespace fnec
