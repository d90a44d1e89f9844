// Mixed-precision Hallén solve: FP32 LU of Z + FP64 preconditioned CG on the
// regularised normal equations of linear.rs:672-797.
//
// The reference solves (M^H M + lambda I) x = M^H y with
//     M = [ Z  B ]   (n rows: B[r][w] = -cos[r] on the rows of column w)
//         [ E  0 ]   (constraint rows: free ends, KCL junctions)
// and y = [rhs; 0]. Z is well conditioned (kappa ~ 2 on thin-wire decks), the
// ill-conditioning of M^H M comes from the few homogeneous columns. With
// T = [Z B; 0 I], M T^{-1} = [I 0; F H], F = E Z^{-1}, H = -F B, so
//     P = M^H M + lambda diag(0, I) = T^H [I + F^H F, F^H H; H^H F, H^H H + lambda] T
// is inverted exactly with one Z^{-H} and one Z^{-1} solve plus dense algebra
// of order ncons and w (Woodbury with K = I + F F^H, Schur complement
// Sc = H^H K^{-1} H + lambda). P differs from the true normal matrix only by
// lambda on the current block, so CG preconditioned by P (applied with the
// FP32 factors of Z) converges in a few iterations to FP64 accuracy. The
// residual h - (M^H M + lambda) x and every vector update are FP64.
//
// Kernels:
//   k_trsv0/1    Z v = b or Z^H v = b with the FP32 factors, FP64 vectors;
//                one workgroup per 64-row block, blocks chained by flags.
//   k_zsolve     the same solves in batches: one workgroup per problem, or a
//                gang of workgroups sharing a large problem.
//   k_fsetup     Z^H solves of the constraint rows by gangs, 16 right-hand sides.
//   k_perm       the row permutation of the factors.
//   k_diag_inv   inverses of the 64 x 64 diagonal blocks of L and U.
//   k_pre_setup  K, H, their Cholesky / Schur factors (one workgroup per problem).
//   k_prec_mid   the dense middle of P^{-1} between the two Z solves.
//   k_mv_z, k_mv_fin, k_mv_zh   t = M p and q = M^H t + lambda p (FP64 Z).
//   k_mix_*      CG control: init, direction, step.
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.h"

namespace fnec {

constexpr int kCtlThreads = 1024;

__device__ __forceinline__ double2 cm(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmc(double2 a, double2 b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cdivd(double2 a, double2 b)
{
    const double ns = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / ns, (a.y * b.x - a.x * b.y) / ns);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

__device__ __forceinline__ double2 wave_sum2(double2 v)
{
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
    }
    return v;
}

__device__ __forceinline__ double wave_sum1(double v)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum in a fixed order (deterministic); red holds blockDim doubles.
__device__ double block_sum(double v, double *red)
{
    const int t = threadIdx.x, nt = blockDim.x;
    v = wave_sum1(v);
    __syncthreads();
    if ((t & 63) == 0)
        red[t >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < nt / 64; ++w)
        s += red[w];
    __syncthreads();
    return s;
}

__device__ double2 block_sum2(double2 v, double *red)
{
    const int t = threadIdx.x, nt = blockDim.x;
    v = wave_sum2(v);
    __syncthreads();
    if ((t & 63) == 0) {
        red[2 * (t >> 6)] = v.x;
        red[2 * (t >> 6) + 1] = v.y;
    }
    __syncthreads();
    double2 s = make_double2(0.0, 0.0);
    for (int w = 0; w < nt / 64; ++w) {
        s.x += red[2 * w];
        s.y += red[2 * w + 1];
    }
    __syncthreads();
    return s;
}

// Coefficient of unknown j in constraint row c (linear.rs:744-762: a free-end
// row is e_a; a junction row sets entry a to 1, then entry b to sign).
__device__ __forceinline__ double cons_coef(const Prob &P, int c, int j)
{
    const int a = P.cons[2 * c], b = P.cons[2 * c + 1];
    const double sg = P.cons_sign[c];
    if (b < 0)
        return a == j ? 1.0 : 0.0;
    if (a == b)
        return a == j ? sg : 0.0;
    return (a == j ? 1.0 : 0.0) + (b == j ? sg : 0.0);
}

// ---------------------------------------------------------------------------
// Multi-workgroup triangular solves (one launch per Z or Z^H solve)
// ---------------------------------------------------------------------------
//
// The factors read P Z = L U (the LU applies every panel's interchanges to
// the earlier multipliers too), perm[i] = row of Z that became row i.
//   Z v = b  : phase 0 y = L^{-1} (P b) into ws, phase 1 v = U^{-1} y into vec.
//   Z^H v = b: phase 0 w = U^{-H} b into ws, phase 1 u = L^{-H} w into ws,
//              v = P^T u scattered into vec.
// Work item (problem, vector, phase, block); each workgroup solves one
// 64-row block. Blocks publish completion by writing the launch epoch into a
// flag; a block waits for the flags of the blocks it reads. Workgroups take
// their item from an atomic ticket in list order, and every item's
// dependencies precede it in the list, so a workgroup only ever waits on
// items already taken by running workgroups. Every wait is bounded: past
// kSpinLimit polls the problem is marked failed and the workgroup moves on.

constexpr int kTb = 64;              // rows per block
constexpr long kSpinLimit = 1L << 20; // polls (>= 0.1 s); a legitimate wait is a few ms

__device__ __forceinline__ int *trsv_flag(const Prob &P, int vec, int phase, int blk)
{
    const int nb = (P.n + kTb - 1) / kTb;
    return P.tflags + ((size_t)(vec + 1) * 2 + phase) * nb + blk;
}

// Thread 0 polls the flag, then the workgroup synchronises; returns false on
// timeout (the problem is then marked failed).
__device__ bool wait_flag(const Prob &P, const int *flag, int epoch, int *s_ok)
{
    if (threadIdx.x == 0) {
        long spins = 0;
        while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
            __builtin_amdgcn_s_sleep(2);
            ++spins;
            // another block of this problem gave up: its flags will never come
            if ((spins & 255) == 0 &&
                __hip_atomic_load(&P.pstate[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                *s_ok = 0;
                break;
            }
            if (spins > kSpinLimit) {
                *s_ok = 0;
                __hip_atomic_store(&P.pstate[2], 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return *s_ok != 0;
}

template <int mode>
__device__ __forceinline__ void trsv_body(const Prob *__restrict__ probs, const int4 *__restrict__ work,
                                          int epoch, unsigned long long *__restrict__ ticket)
{
    __shared__ int s_item, s_ok;
    __shared__ float dr[kTb][kTb + 1], di[kTb][kTb + 1];
    __shared__ double2 xs[kTb];
    __shared__ double2 part[4][kTb];
    const int t = threadIdx.x, lane = t & 63, q = t >> 6;
    if (t == 0) {
        s_item = (int)atomicAdd(ticket, 1ULL);
        s_ok = 1;
    }
    __syncthreads();
    const int4 wk = work[s_item];
    const Prob &P = probs[wk.x];
    const int vecid = wk.y, phase = wk.z, blk = wk.w;
    if (P.info[0] != 0 || (vecid < 0 && P.pstate[0] != 0))
        return;
    const int n = P.n, ld = P.ldz, nb = (n + kTb - 1) / kTb;
    const int i0 = blk * kTb, jb = min(kTb, n - i0);
    const float *Fr = P.fr, *Fi = P.fi;
    double2 *vec = (double2 *)(vecid < 0 ? P.vv : P.fh + 2 * (size_t)n * vecid);
    double2 *ws = (double2 *)P.tws + (size_t)n * (vecid + 1);
    // right-hand side of this block
    double2 acc = make_double2(0.0, 0.0);
    if (phase == 0) {
        if (t < jb) {
            const int r = i0 + t;
            if (mode == 0)
                acc = vec[P.perm[r]];
            else
                acc = vecid >= 0 ? make_double2(cons_coef(P, vecid, r), 0.0) : vec[r];
        }
    } else {
        if (!wait_flag(P, trsv_flag(P, vecid, 0, nb - 1), epoch, &s_ok))
            return;
        if (t < jb)
            acc = ws[i0 + t];
    }
    // stage the inverse of the diagonal block: L_ii^{-1} (mode 0 phase 0, mode 1
    // phase 1) or U_ii^{-1}; dr[r][c] = inverse[r][c]
    {
        const int which = (mode == 0) == (phase == 0) ? 0 : 1;
        const float *src = P.dinv + ((size_t)blk * 2 + which) * 2 * kTb * kTb;
        for (int e = t; e < kTb * kTb; e += 256) {
            const int r = e % kTb, c = e / kTb;
            dr[r][c] = src[2 * e];
            di[r][c] = src[2 * e + 1];
        }
    }
    // off-diagonal blocks, in the order their flags become ready
    const bool fwd = phase == 0;
    const int jfirst = fwd ? 0 : nb - 1, jlast = fwd ? blk : blk, jstep = fwd ? 1 : -1;
    double2 pacc = make_double2(0.0, 0.0);   // mode 0: row lane, columns 16q..16q+15
    double2 racc[16];                        // mode 1: rows 16q..16q+15, column lane
#pragma unroll
    for (int k = 0; k < 16; ++k)
        racc[k] = make_double2(0.0, 0.0);
    for (int j = jfirst; j != jlast; j += jstep) {
        if (!wait_flag(P, trsv_flag(P, vecid, phase, j), epoch, &s_ok))
            return;
        const int j0 = j * kTb, jn = min(kTb, n - j0);
        // solved values of block j: phase 0 in ws; phase 1 in vec (mode 0) or ws (mode 1)
        const double2 *src = (phase == 1 && mode == 0) ? vec : ws;
        if (mode == 0) {
            if (t < jn)
                xs[t] = src[j0 + t];
            __syncthreads();
            if (lane < jb) {
                const int r = i0 + lane;
                for (int c = 16 * q; c < min(16 * q + 16, jn); ++c) {
                    const size_t o = (size_t)(j0 + c) * ld + r;
                    const double ar = Fr[o], ai = Fi[o];
                    const double2 x = xs[c];
                    pacc.x += ar * x.x - ai * x.y;
                    pacc.y += ar * x.y + ai * x.x;
                }
            }
            __syncthreads();
        } else {
            // rows r of this block, column c = lane of block j:
            // U^H[r][c] = conj(U[c][r]) (phase 0), L^H[r][c] = conj(L[c][r]) (phase 1)
            const double2 x = lane < jn ? src[j0 + lane] : make_double2(0.0, 0.0);
            if (lane < jn)
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const int r = 16 * q + k;
                    if (r < jb) {
                        const size_t o = (size_t)(i0 + r) * ld + j0 + lane;
                        const double ar = Fr[o], ai = Fi[o];
                        racc[k].x += ar * x.x + ai * x.y;
                        racc[k].y += ar * x.y - ai * x.x;
                    }
                }
        }
    }
    // combine partial sums into acc (rows on lanes of wave 0)
    if (mode == 0) {
        part[q][lane] = pacc;
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const double2 v = wave_sum2(racc[k]);
            if (lane == 0)
                part[0][16 * q + k] = v;
        }
    }
    __syncthreads();
    if (q == 0) {
        double2 sub;
        if (mode == 0)
            sub = cadd(cadd(part[0][lane], part[1][lane]), cadd(part[2][lane], part[3][lane]));
        else
            sub = part[0][lane];
        xs[lane] = lane < jb ? csub(acc, sub) : make_double2(0.0, 0.0);
    }
    __syncthreads();
    // diagonal block: x = T_ii^{-1} acc with T_ii^{-1} = inverse (mode 0) or its
    // conjugate transpose (mode 1); rows on lanes, columns split over waves
    {
        double2 sdiag = make_double2(0.0, 0.0);
        for (int c = 16 * q; c < 16 * q + 16; ++c) {
            const double2 tv = mode == 0 ? make_double2(dr[lane][c], di[lane][c])
                                         : make_double2(dr[c][lane], -di[c][lane]);
            sdiag = cadd(sdiag, cm(tv, xs[c]));
        }
        part[q][lane] = sdiag;
    }
    __syncthreads();
    if (q == 0) {
        const double2 x = cadd(cadd(part[0][lane], part[1][lane]), cadd(part[2][lane], part[3][lane]));
        if (lane < jb) {
            const int r = i0 + lane;
            if (phase == 0 || mode == 1)
                ws[r] = x;
            if (phase == 1)
                vec[mode == 0 ? r : P.perm[r]] = x;
        }
        // publish: this wave wrote every output of the block
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (lane == 0)
            __hip_atomic_store(trsv_flag(P, vecid, phase, blk), epoch, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One kernel per mode, so each instance keeps only its own accumulators.
__global__ __launch_bounds__(256) void k_trsv0(const Prob *__restrict__ probs, const int4 *__restrict__ work,
                                               int epoch, unsigned long long *__restrict__ ticket)
{
    trsv_body<0>(probs, work, epoch, ticket);
}

__global__ __launch_bounds__(256) void k_trsv1(const Prob *__restrict__ probs, const int4 *__restrict__ work,
                                               int epoch, unsigned long long *__restrict__ ticket)
{
    trsv_body<1>(probs, work, epoch, ticket);
}

// Inverses of the 64 x 64 diagonal blocks of the FP32 factors: L_ii^{-1}
// (unit lower) and U_ii^{-1} (upper), computed in FP32 one column per lane,
// stored column-major per block with zero padding beyond the matrix order.
// Work item: (problem, block).
__global__ __launch_bounds__(64) void k_diag_inv(const Prob *__restrict__ probs,
                                                 const int2 *__restrict__ work)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    if (P.info[0] != 0)
        return;
    const int n = P.n, ld = P.ldz, blk = wk.y, c = threadIdx.x;
    const int i0 = blk * kTb, jb = min(kTb, n - i0);
    __shared__ float ar[kTb][kTb + 1], ai[kTb][kTb + 1];
    __shared__ float xr[kTb][kTb + 1], xi[kTb][kTb + 1];
    for (int e = c; e < jb * jb; e += 64) {
        const int r = e % jb, k = e / jb;
        ar[r][k] = P.fr[(size_t)(i0 + k) * ld + i0 + r];
        ai[r][k] = P.fi[(size_t)(i0 + k) * ld + i0 + r];
    }
    __syncthreads();
    for (int which = 0; which < 2; ++which) {
        // column c of the inverse, held in xr/xi[.][c] (this lane's own column)
        if (which == 0) {
            for (int r = 0; r < kTb; ++r) {
                float sr = 0.f, si = 0.f;
                if (c < jb && r < jb && r > c) {
                    for (int k = c; k < r; ++k) {
                        const float lr = ar[r][k], li = ai[r][k];
                        const float vr = xr[k][c], vi = xi[k][c];
                        sr -= lr * vr - li * vi;
                        si -= lr * vi + li * vr;
                    }
                } else if (c < jb && r == c) {
                    sr = 1.f;
                }
                xr[r][c] = sr;
                xi[r][c] = si;
            }
        } else {
            for (int r = kTb - 1; r >= 0; --r) {
                float sr = 0.f, si = 0.f;
                if (c < jb && r < jb && r <= c) {
                    sr = r == c ? 1.f : 0.f;
                    for (int k = r + 1; k <= c; ++k) {
                        const float ur = ar[r][k], ui = ai[r][k];
                        const float vr = xr[k][c], vi = xi[k][c];
                        sr -= ur * vr - ui * vi;
                        si -= ur * vi + ui * vr;
                    }
                    const float dr = ar[r][r], di = ai[r][r];
                    const float ns = dr * dr + di * di;
                    const float qr = (sr * dr + si * di) / ns, qi = (si * dr - sr * di) / ns;
                    sr = qr;
                    si = qi;
                }
                xr[r][c] = sr;
                xi[r][c] = si;
            }
        }
        __syncthreads();
        float *dst = P.dinv + ((size_t)blk * 2 + which) * 2 * kTb * kTb;
        for (int e = c; e < kTb * kTb; e += 64) {
            const int r = e % kTb, k = e / kTb;
            dst[2 * e] = xr[r][k];
            dst[2 * e + 1] = xi[r][k];
        }
        __syncthreads();
    }
}

// Single-workgroup Z or Z^H solve of one problem (the batched form: one
// workgroup per problem, no inter-workgroup waiting). The whole vector lives
// in LDS (16 n bytes; the host sends problems with n <= 8192); 64-row blocks use the diagonal-block
// inverses. Z v = b: L forward and U backward, right-looking (column-major
// factors are read down columns, coalesced); Z^H v = b: U^H forward and L^H
// backward, left-looking (a row of U^H or L^H is a column of U or L).
// Work item: plist[blockIdx.x] = problem.
constexpr int kSolve1Threads = 1024;

__device__ __forceinline__ double2 lds2(const double2 *v, int i) { return v[i]; }

__device__ __forceinline__ void zsolve1_body(const Prob &P, int mode, double2 *__restrict__ vec,
                                             double2 (*__restrict__ part)[kTb], double2 *__restrict__ xb)
{
    const int n = P.n, ld = P.ldz, nb = (n + kTb - 1) / kTb;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6; // 16 waves
    const float *Fr = P.fr, *Fi = P.fi;
    double2 *io = (double2 *)P.vv;
    for (int i = t; i < n; i += kSolve1Threads)
        vec[i] = mode == 0 ? io[P.perm[i]] : io[i];
    __syncthreads();
    // x_blk = op(inverse) * xb, op = identity (mode 0) or conjugate transpose (mode 1)
    auto apply_inv = [&](int blk, int which) {
        const float *src = P.dinv + ((size_t)blk * 2 + which) * 2 * kTb * kTb;
        const int r = t & 63, g = t >> 6; // 16 column groups of 4
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = 4 * g + q;
            const size_t e = mode == 0 ? (size_t)c * kTb + r : (size_t)r * kTb + c; // [r][c] or [c][r]
            const double ar = src[2 * e], ai = mode == 0 ? src[2 * e + 1] : -src[2 * e + 1];
            const double2 x = xb[c];
            acc.x += ar * x.x - ai * x.y;
            acc.y += ar * x.y + ai * x.x;
        }
        part[g][r] = acc;
        __syncthreads();
        if (t < kTb) {
            double2 sum = make_double2(0.0, 0.0);
            for (int q = 0; q < 16; ++q)
                sum = cadd(sum, part[q][t]);
            const int i = blk * kTb + t;
            if (i < n)
                vec[i] = sum;
        }
        __syncthreads();
    };
    // vec[r] -= F[r, j0 .. j0 + jn) xb for rows [ra, rb): four consecutive rows
    // per thread (float4 loads down each column; ld and the row offsets used
    // here are multiples of 4 except at the matrix end, handled one by one)
    auto rows_update = [&](int j0, int jn, int ra, int rb) {
        const int a4 = (ra + 3) & ~3; // first 4-aligned row
        for (int r = ra + t; r < min(a4, rb); r += kSolve1Threads) {
            double sr = 0.0, si = 0.0;
            for (int c = 0; c < jn; ++c) {
                const size_t o = (size_t)(j0 + c) * ld + r;
                const double2 x = xb[c];
                sr += (double)Fr[o] * x.x - (double)Fi[o] * x.y;
                si += (double)Fr[o] * x.y + (double)Fi[o] * x.x;
            }
            vec[r] = make_double2(vec[r].x - sr, vec[r].y - si);
        }
        for (int r4 = a4 + 4 * t; r4 < rb; r4 += 4 * kSolve1Threads) {
            double sr[4] = {0.0, 0.0, 0.0, 0.0}, si[4] = {0.0, 0.0, 0.0, 0.0};
            if (r4 + 3 < rb) {
#pragma unroll 4
                for (int c = 0; c < jn; ++c) {
                    const size_t o = (size_t)(j0 + c) * ld + r4;
                    const float4 fr = *(const float4 *)(Fr + o), fi = *(const float4 *)(Fi + o);
                    const double2 x = xb[c];
                    sr[0] += (double)fr.x * x.x - (double)fi.x * x.y;
                    si[0] += (double)fr.x * x.y + (double)fi.x * x.x;
                    sr[1] += (double)fr.y * x.x - (double)fi.y * x.y;
                    si[1] += (double)fr.y * x.y + (double)fi.y * x.x;
                    sr[2] += (double)fr.z * x.x - (double)fi.z * x.y;
                    si[2] += (double)fr.z * x.y + (double)fi.z * x.x;
                    sr[3] += (double)fr.w * x.x - (double)fi.w * x.y;
                    si[3] += (double)fr.w * x.y + (double)fi.w * x.x;
                }
                for (int q = 0; q < 4; ++q)
                    vec[r4 + q] = make_double2(vec[r4 + q].x - sr[q], vec[r4 + q].y - si[q]);
            } else {
                for (int r = r4; r < rb; ++r) {
                    double ur = 0.0, ui = 0.0;
                    for (int c = 0; c < jn; ++c) {
                        const size_t o = (size_t)(j0 + c) * ld + r;
                        const double2 x = xb[c];
                        ur += (double)Fr[o] * x.x - (double)Fi[o] * x.y;
                        ui += (double)Fr[o] * x.y + (double)Fi[o] * x.x;
                    }
                    vec[r] = make_double2(vec[r].x - ur, vec[r].y - ui);
                }
            }
        }
    };
    // rows of U^H (mode 1 phase 0) or L^H (phase 1) against solved entries
    // [ca, cb): a row is a contiguous column of the factor; four entries per lane
    auto row_dot = [&](size_t col, int ca, int cb, double &sr, double &si) {
        sr = 0.0;
        si = 0.0;
        const int a4 = min((ca + 3) & ~3, cb);
        for (int c = ca + lane; c < a4; c += 64) {
            const double ar = Fr[col + c], ai = Fi[col + c];
            const double2 x = vec[c];
            sr += ar * x.x + ai * x.y;
            si += ar * x.y - ai * x.x;
        }
        const int b4 = a4 + ((cb - a4) & ~3);
        for (int c = a4 + 4 * lane; c < b4; c += 256) {
            const float4 fr = *(const float4 *)(Fr + col + c), fi = *(const float4 *)(Fi + col + c);
            const double2 x0 = vec[c], x1 = vec[c + 1], x2 = vec[c + 2], x3 = vec[c + 3];
            sr += (double)fr.x * x0.x + (double)fi.x * x0.y + (double)fr.y * x1.x + (double)fi.y * x1.y +
                  (double)fr.z * x2.x + (double)fi.z * x2.y + (double)fr.w * x3.x + (double)fi.w * x3.y;
            si += (double)fr.x * x0.y - (double)fi.x * x0.x + (double)fr.y * x1.y - (double)fi.y * x1.x +
                  (double)fr.z * x2.y - (double)fi.z * x2.x + (double)fr.w * x3.y - (double)fi.w * x3.x;
        }
        for (int c = b4 + lane; c < cb; c += 64) {
            const double ar = Fr[col + c], ai = Fi[col + c];
            const double2 x = vec[c];
            sr += ar * x.x + ai * x.y;
            si += ar * x.y - ai * x.x;
        }
    };
    if (mode == 0) {
        // forward L (unit lower): solve block j, then update every row below it
        for (int j = 0; j < nb; ++j) {
            const int j0 = j * kTb, jn = min(kTb, n - j0);
            if (t < kTb)
                xb[t] = t < jn ? vec[j0 + t] : make_double2(0.0, 0.0);
            __syncthreads();
            apply_inv(j, 0);
            if (t < kTb)
                xb[t] = t < jn ? vec[j0 + t] : make_double2(0.0, 0.0);
            __syncthreads();
            rows_update(j0, jn, j0 + jn, n);
            __syncthreads();
        }
        // backward U: solve block j (from the bottom), then update every row above it
        for (int j = nb - 1; j >= 0; --j) {
            const int j0 = j * kTb, jn = min(kTb, n - j0);
            if (t < kTb)
                xb[t] = t < jn ? vec[j0 + t] : make_double2(0.0, 0.0);
            __syncthreads();
            apply_inv(j, 1);
            if (t < kTb)
                xb[t] = t < jn ? vec[j0 + t] : make_double2(0.0, 0.0);
            __syncthreads();
            rows_update(j0, jn, 0, j0);
            __syncthreads();
        }
        for (int i = t; i < n; i += kSolve1Threads)
            io[i] = vec[i];
    } else {
        // forward U^H: rows of block i against every solved entry above it
        for (int i = 0; i < nb; ++i) {
            const int i0 = i * kTb, in = min(kTb, n - i0);
            for (int rr = wave; rr < in; rr += 16) {
                double sr, si; // conj(U[c][r]) x_c over the solved entries c < i0
                row_dot((size_t)(i0 + rr) * ld, 0, i0, sr, si);
                sr = wave_sum1(sr);
                si = wave_sum1(si);
                if (lane == 0)
                    xb[rr] = make_double2(vec[i0 + rr].x - sr, vec[i0 + rr].y - si);
            }
            if (t < kTb && t >= in)
                xb[t] = make_double2(0.0, 0.0);
            __syncthreads();
            apply_inv(i, 1); // (U_ii^{-1})^H
        }
        // backward L^H: rows of block i against every solved entry below it
        for (int i = nb - 1; i >= 0; --i) {
            const int i0 = i * kTb, in = min(kTb, n - i0);
            for (int rr = wave; rr < in; rr += 16) {
                double sr, si; // conj(L[c][r]) x_c over the solved entries below the block
                row_dot((size_t)(i0 + rr) * ld, i0 + in, n, sr, si);
                sr = wave_sum1(sr);
                si = wave_sum1(si);
                if (lane == 0)
                    xb[rr] = make_double2(vec[i0 + rr].x - sr, vec[i0 + rr].y - si);
            }
            if (t < kTb && t >= in)
                xb[t] = make_double2(0.0, 0.0);
            __syncthreads();
            apply_inv(i, 0); // (L_ii^{-1})^H
        }
        for (int i = t; i < n; i += kSolve1Threads)
            io[P.perm[i]] = vec[i];
    }
}

// ---------------------------------------------------------------------------
// Gang-scheduled Z / Z^H solves: G workgroups share one problem
// ---------------------------------------------------------------------------
//
// Member m of a gang owns the 64-row blocks m, m + G, m + 2G, ... and keeps
// only those entries of the vector in LDS. Both triangular sweeps are
// right-looking: when block j is solved, its owner publishes x_j in P.gx
// (epoch-tagged words, below), and every member subtracts B(b, j) x_j from
// its owned blocks b beyond j, where
//   B(b, j) = F[b rows, j columns]            (Z v = b: L forward, U backward)
//   B(b, j) = conj(F[j rows, b columns])^T    (Z^H v = b: U^H forward, L^H backward).
// The owner of block j + 1 first applies x_j to that block alone, solves and
// publishes it, and only then updates its other blocks, so the chain of
// solved blocks does not wait for the bulk of the updates; the operands of
// that step are loaded before x_j is waited for. Reductions run in a fixed
// order, so a result depends on (n, G) only.

constexpr int kGangWaves = kSolve1Threads / 64;

// own[k] -= B(b_k, j) x for the owned local blocks k in [ka, kb), b_k = m + k G.
// Z (mode 0): a lane holds 4 consecutive rows (float4 down a column) and one
// column of every 4; the 4 column lanes of a quad are summed with shuffles.
// When fewer than 16 blocks remain, each block's columns are split over
// several waves whose partial sums meet in red[][] in a fixed order.
// Z^H (mode 1): a row of B is a column segment of F; a lane reads 4
// consecutive entries of it (float4) at 4 column offsets and the quad sums.
__device__ void gang_sweep(const Prob &P, int mode, int m, int G, int j, int ka, int kb,
                           const double2 *__restrict__ x, double2 *__restrict__ own,
                           double2 (*__restrict__ red)[kTb])
{
    const int nS = kb - ka;
    if (nS <= 0)
        return;
    const int n = P.n, ld = P.ldz;
    const int j0 = j * kTb, jn = min(kTb, n - j0);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int U = 1; // units per block: nS * U <= 16 when nS < 16
    while (U < 16 && nS * U * 2 <= kGangWaves)
        U *= 2;
    const int nunits = nS * U;
    const float *__restrict__ Fr = P.fr;
    const float *__restrict__ Fi = P.fi;
    if (mode == 0) {
        const int rq = lane >> 2, cs = lane & 3;
        const int ncol = kTb / U, iters = ncol / 4;
        for (int u = wave; u < nunits; u += kGangWaves) {
            const int k = ka + u / U, p = u % U;
            const int r = (m + k * G) * kTb + 4 * rq; // first row of this lane's quad
            double sr0 = 0.0, si0 = 0.0, sr1 = 0.0, si1 = 0.0, sr2 = 0.0, si2 = 0.0, sr3 = 0.0, si3 = 0.0;
            const int cbeg = p * ncol;
#pragma unroll 4
            for (int i = 0; i < iters; ++i) {
                const int c = cbeg + 4 * i + cs;
                if (c < jn) {
                    const size_t o = (size_t)(j0 + c) * ld + r;
                    float4 fr = *(const float4 *)(Fr + o), fi = *(const float4 *)(Fi + o);
                    if (r + 3 >= n) { // rows past n: column padding, never written
                        if (r >= n) fr.x = fi.x = 0.f;
                        if (r + 1 >= n) fr.y = fi.y = 0.f;
                        if (r + 2 >= n) fr.z = fi.z = 0.f;
                        fr.w = fi.w = 0.f;
                    }
                    const double2 xv = x[c];
                    sr0 += (double)fr.x * xv.x - (double)fi.x * xv.y;
                    si0 += (double)fr.x * xv.y + (double)fi.x * xv.x;
                    sr1 += (double)fr.y * xv.x - (double)fi.y * xv.y;
                    si1 += (double)fr.y * xv.y + (double)fi.y * xv.x;
                    sr2 += (double)fr.z * xv.x - (double)fi.z * xv.y;
                    si2 += (double)fr.z * xv.y + (double)fi.z * xv.x;
                    sr3 += (double)fr.w * xv.x - (double)fi.w * xv.y;
                    si3 += (double)fr.w * xv.y + (double)fi.w * xv.x;
                }
            }
#pragma unroll
            for (int d = 1; d <= 2; d <<= 1) {
                sr0 += __shfl_xor(sr0, d);
                si0 += __shfl_xor(si0, d);
                sr1 += __shfl_xor(sr1, d);
                si1 += __shfl_xor(si1, d);
                sr2 += __shfl_xor(sr2, d);
                si2 += __shfl_xor(si2, d);
                sr3 += __shfl_xor(sr3, d);
                si3 += __shfl_xor(si3, d);
            }
            // lane cs of the quad takes row 4 rq + cs
            const double vr = cs == 0 ? sr0 : cs == 1 ? sr1 : cs == 2 ? sr2 : sr3;
            const double vi = cs == 0 ? si0 : cs == 1 ? si1 : cs == 2 ? si2 : si3;
            if (U == 1) {
                double2 &o = own[k * kTb + 4 * rq + cs];
                o = make_double2(o.x - vr, o.y - vi);
            } else {
                red[u][4 * rq + cs] = make_double2(vr, vi);
            }
        }
        if (U > 1) {
            __syncthreads();
            for (int e = t; e < nS * kTb; e += kSolve1Threads) {
                const int kk = e / kTb, rr = e % kTb;
                double2 sum = red[kk * U][rr];
                for (int p = 1; p < U; ++p)
                    sum = cadd(sum, red[kk * U + p][rr]);
                double2 &o = own[(ka + kk) * kTb + rr];
                o = make_double2(o.x - sum.x, o.y - sum.y);
            }
        }
    } else {
        const int cq = lane & 3, rs = lane >> 2;
        const int nrow = kTb / U; // rows per unit
        for (int u = wave; u < nunits; u += kGangWaves) {
            const int k = ka + u / U, p = u % U;
            const int b0 = (m + k * G) * kTb;
            for (int i = 0; i * 16 < nrow; ++i) {
                const int rr = p * nrow + i * 16 + rs; // row of B within the block
                const bool live = i * 16 + rs < nrow && b0 + rr < n;
                double sr = 0.0, si = 0.0;
                if (live) {
                    const size_t col = (size_t)(b0 + rr) * ld + j0;
#pragma unroll 2
                    for (int h = 0; h < 4; ++h) {
                        const int c = 16 * h + 4 * cq;
                        if (c < jn) {
                            float4 fr = *(const float4 *)(Fr + col + c), fi = *(const float4 *)(Fi + col + c);
                            if (c + 3 >= jn) { // entries past n: column padding
                                if (c + 1 >= jn) fr.y = fi.y = 0.f;
                                if (c + 2 >= jn) fr.z = fi.z = 0.f;
                                fr.w = fi.w = 0.f;
                            }
                            const double2 x0 = x[c], x1 = x[c + 1], x2 = x[c + 2], x3 = x[c + 3];
                            sr += (double)fr.x * x0.x + (double)fi.x * x0.y + (double)fr.y * x1.x +
                                  (double)fi.y * x1.y + (double)fr.z * x2.x + (double)fi.z * x2.y +
                                  (double)fr.w * x3.x + (double)fi.w * x3.y;
                            si += (double)fr.x * x0.y - (double)fi.x * x0.x + (double)fr.y * x1.y -
                                  (double)fi.y * x1.x + (double)fr.z * x2.y - (double)fi.z * x2.x +
                                  (double)fr.w * x3.y - (double)fi.w * x3.x;
                        }
                    }
                }
                sr += __shfl_xor(sr, 1);
                si += __shfl_xor(si, 1);
                sr += __shfl_xor(sr, 2);
                si += __shfl_xor(si, 2);
                if (live && cq == 0) {
                    double2 &o = own[k * kTb + rr];
                    o = make_double2(o.x - sr, o.y - si);
                }
            }
        }
    }
}

// Publication of solved blocks inside a gang. Each 32-bit half of an entry
// travels in one 64-bit word together with the launch epoch, written and read
// with relaxed agent-scope atomics: a word is single-copy atomic and coherent
// across the device, so a reader that sees the epoch in every word of an
// entry has its value, with no fence and no separate flag. The words are
// cleared before a batch and every launch of a batch has its own epoch.
__device__ __forceinline__ void ll_publish(unsigned long long *w, double2 v, unsigned epoch)
{
    const unsigned long long a = (unsigned long long)__double_as_longlong(v.x);
    const unsigned long long b = (unsigned long long)__double_as_longlong(v.y);
    const unsigned long long e = (unsigned long long)epoch << 32;
    __hip_atomic_store(w + 0, e | (a & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 1, e | (a >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 2, e | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 3, e | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-0 lanes t < valid poll the words of entry t until all carry epoch;
// the entry goes to x[t] (0 for t >= valid). Bounded like wait_flag: past
// kSpinLimit polls, or when another workgroup of the problem gave up, the
// problem is marked failed. Ends with a workgroup barrier; returns false on failure.
__device__ bool ll_fetch(const Prob &P, const unsigned long long *w, int valid, unsigned epoch, double2 *x,
                         int *s_ok)
{
    const int t = threadIdx.x;
    if (t < kTb) {
        double2 v = make_double2(0.0, 0.0);
        if (t < valid) {
            const unsigned long long *p = w + 4 * t;
            long spins = 0;
            unsigned long long q0, q1, q2, q3;
            for (;;) {
                q0 = __hip_atomic_load(p + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                q1 = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                q2 = __hip_atomic_load(p + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                q3 = __hip_atomic_load(p + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((unsigned)(q0 >> 32) == epoch && (unsigned)(q1 >> 32) == epoch &&
                    (unsigned)(q2 >> 32) == epoch && (unsigned)(q3 >> 32) == epoch)
                    break;
                ++spins;
                if ((spins & 255) == 0 &&
                    __hip_atomic_load(&P.pstate[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    *s_ok = 0;
                    break;
                }
                if (spins > kSpinLimit) {
                    *s_ok = 0;
                    __hip_atomic_store(&P.pstate[2], 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            v.x = __longlong_as_double((long long)((q0 & 0xffffffffull) | (q1 << 32)));
            v.y = __longlong_as_double((long long)((q2 & 0xffffffffull) | (q3 << 32)));
        }
        x[t] = v;
    }
    __syncthreads();
    return *s_ok != 0;
}

// Operands of one look-ahead step (block jn updated with x_j, then solved),
// loaded into registers before x_j is waited for, so that the chain of solved
// blocks pays only the flag and x_j latencies.
//   b: B(jn, j), one float4 per plane per thread. Z: thread (column c = t/16,
//      rows 4 (t%16) .. +3). Z^H: thread (row r = t/16, columns 4 (t%16) .. +3).
//   d: the inverse diagonal block of jn, thread (row t%64, columns 4 (t/64) .. +3).
struct GangAhead {
    float4 br, bi;
    float dr[4], di[4];
};

__device__ __forceinline__ void gang_prefetch(const Prob &P, int mode, int which, int jn, int j, GangAhead &a)
{
    const int n = P.n, ld = P.ldz, t = threadIdx.x;
    const int j0 = j * kTb, jc = min(kTb, n - j0), b0 = jn * kTb;
    a.br = a.bi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (mode == 0) {
        const int c = t >> 4, r = b0 + 4 * (t & 15);
        if (c < jc) {
            const size_t o = (size_t)(j0 + c) * ld + r;
            a.br = *(const float4 *)(P.fr + o);
            a.bi = *(const float4 *)(P.fi + o);
            if (r + 3 >= n) { // rows past n: column padding, never written
                if (r >= n) a.br.x = a.bi.x = 0.f;
                if (r + 1 >= n) a.br.y = a.bi.y = 0.f;
                if (r + 2 >= n) a.br.z = a.bi.z = 0.f;
                a.br.w = a.bi.w = 0.f;
            }
        }
    } else {
        const int r = t >> 4, c = 4 * (t & 15);
        if (b0 + r < n && c < jc) {
            const size_t o = (size_t)(b0 + r) * ld + j0 + c;
            a.br = *(const float4 *)(P.fr + o);
            a.bi = *(const float4 *)(P.fi + o);
            if (c + 3 >= jc) {
                if (c + 1 >= jc) a.br.y = a.bi.y = 0.f;
                if (c + 2 >= jc) a.br.z = a.bi.z = 0.f;
                a.br.w = a.bi.w = 0.f;
            }
        }
    }
    const int jnc = min(kTb, n - b0);
    const float *src = P.dinv + ((size_t)jn * 2 + which) * 2 * kTb * kTb;
    const int r = t & 63, g = t >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = 4 * g + q;
        a.dr[q] = a.di[q] = 0.f;
        if (c < jnc) {
            const size_t e = mode == 0 ? (size_t)c * kTb + r : (size_t)r * kTb + c;
            a.dr[q] = src[2 * e];
            a.di[q] = mode == 0 ? src[2 * e + 1] : -src[2 * e + 1];
        }
    }
}

// x = op(D) v for the block whose inverse sits in a.d (v: its entries in LDS,
// overwritten with x); x also goes to dst (LDS) and is published in pub.
__device__ void gang_solve_block(const Prob &P, int j, const GangAhead &a, double2 *__restrict__ v,
                                 double2 *__restrict__ dst, double2 (*__restrict__ red)[kTb],
                                 unsigned long long *pub, unsigned epoch)
{
    const int n = P.n, j0 = j * kTb, jn = min(kTb, n - j0);
    const int t = threadIdx.x, r = t & 63, g = t >> 6; // 16 column groups of 4
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double2 xv = v[4 * g + q];
        const double ar = a.dr[q], ai = a.di[q];
        acc.x += ar * xv.x - ai * xv.y;
        acc.y += ar * xv.y + ai * xv.x;
    }
    red[g][r] = acc;
    __syncthreads();
    if (t < kTb) {
        double2 sum = red[0][t];
        for (int q = 1; q < kGangWaves; ++q)
            sum = cadd(sum, red[q][t]);
        if (t >= jn)
            sum = make_double2(0.0, 0.0);
        v[t] = sum;
        dst[t] = sum;
        if (t < jn)
            ll_publish(pub + (size_t)j0 * 4 + 4 * t, sum, epoch);
    }
    __syncthreads();
}

// v -= B(jn, j) x with B from a.b (one block, the look-ahead update).
__device__ void gang_ahead_update(int mode, const GangAhead &a, const double2 *__restrict__ x,
                                  double2 *__restrict__ v, double2 (*__restrict__ red)[kTb])
{
    const int t = threadIdx.x;
    if (mode == 0) {
        const int c = t >> 4, rq = t & 15, w = t >> 6, cl = c & 3;
        const double2 xv = x[c];
        double sr[4], si[4];
        const float fr[4] = {a.br.x, a.br.y, a.br.z, a.br.w}, fi[4] = {a.bi.x, a.bi.y, a.bi.z, a.bi.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            sr[q] = (double)fr[q] * xv.x - (double)fi[q] * xv.y;
            si[q] = (double)fr[q] * xv.y + (double)fi[q] * xv.x;
        }
#pragma unroll
        for (int d = 16; d <= 32; d <<= 1)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                sr[q] += __shfl_xor(sr[q], d);
                si[q] += __shfl_xor(si[q], d);
            }
        // the 4 column lanes of a row quad hold equal sums; lane cl keeps row 4 rq + cl
        const double vr = cl == 0 ? sr[0] : cl == 1 ? sr[1] : cl == 2 ? sr[2] : sr[3];
        const double vi = cl == 0 ? si[0] : cl == 1 ? si[1] : cl == 2 ? si[2] : si[3];
        red[w][4 * rq + cl] = make_double2(vr, vi);
        __syncthreads();
        if (t < kTb) {
            double2 sum = red[0][t];
            for (int q = 1; q < kGangWaves; ++q)
                sum = cadd(sum, red[q][t]);
            v[t] = make_double2(v[t].x - sum.x, v[t].y - sum.y);
        }
    } else {
        const int r = t >> 4, c = 4 * (t & 15);
        const double2 x0 = x[c], x1 = x[c + 1], x2 = x[c + 2], x3 = x[c + 3];
        double sr = (double)a.br.x * x0.x + (double)a.bi.x * x0.y + (double)a.br.y * x1.x +
                    (double)a.bi.y * x1.y + (double)a.br.z * x2.x + (double)a.bi.z * x2.y +
                    (double)a.br.w * x3.x + (double)a.bi.w * x3.y;
        double si = (double)a.br.x * x0.y - (double)a.bi.x * x0.x + (double)a.br.y * x1.y -
                    (double)a.bi.y * x1.x + (double)a.br.z * x2.y - (double)a.bi.z * x2.x +
                    (double)a.br.w * x3.y - (double)a.bi.w * x3.x;
#pragma unroll
        for (int d = 1; d <= 8; d <<= 1) {
            sr += __shfl_xor(sr, d);
            si += __shfl_xor(si, d);
        }
        if ((t & 15) == 0)
            v[r] = make_double2(v[r].x - sr, v[r].y - si);
    }
    __syncthreads();
}

__device__ void zsolve_gang_body(const Prob &P, int mode, int m, int G, int epoch, double2 *__restrict__ own,
                                 double2 (*__restrict__ red)[kTb], double2 *__restrict__ xa,
                                 double2 *__restrict__ xn, int *s_ok)
{
    const int n = P.n, nb = (n + kTb - 1) / kTb;
    const int t = threadIdx.x;
    double2 *io = (double2 *)P.vv;
    const int nown = (nb - m + G - 1) / G;
    for (int e = t; e < nown * kTb; e += kSolve1Threads) {
        const int i = (m + (e / kTb) * G) * kTb + e % kTb;
        own[e] = i < n ? (mode == 0 ? io[P.perm[i]] : io[i]) : make_double2(0.0, 0.0);
    }
    __syncthreads();
    GangAhead ahd;
    for (int phase = 0; phase < 2; ++phase) {
        const bool fwd = phase == 0;
        const int which = (mode == 0) == fwd ? 0 : 1;
        unsigned long long *pub = P.gx + (size_t)phase * nb * kTb * 4;
        // owned local blocks beyond block j in the sweep direction: [ka, kb)
        auto range = [&](int j, int &ka, int &kb) {
            if (fwd) {
                ka = j < m ? 0 : (j - m) / G + 1;
                kb = nown;
            } else {
                ka = 0;
                kb = j > m ? (j - m + G - 1) / G : 0;
            }
        };
        auto next = [&](int j) { return fwd ? j + 1 : j - 1; };
        auto mine = [&](int j) { return j >= 0 && j < nb && j % G == m; };
        auto fetch = [&](int j) { // x_j published by another member -> xa
            return ll_fetch(P, pub + (size_t)j * kTb * 4, min(kTb, n - j * kTb), (unsigned)epoch, xa, s_ok);
        };
        int j = fwd ? 0 : nb - 1;
        bool have = false; // ahd holds the operands of the next look-ahead step
        if (mine(j)) {
            gang_prefetch(P, mode, which, j, j, ahd); // only the inverse is used
            gang_solve_block(P, j, ahd, own + ((j - m) / G) * kTb, xa, red, pub, (unsigned)epoch);
        } else {
            if (mine(next(j))) {
                gang_prefetch(P, mode, which, next(j), j, ahd);
                have = true;
            }
            if (!fetch(j))
                return;
        }
        for (int s = 0; s < nb; ++s) {
            j = fwd ? s : nb - 1 - s;
            int ka, kb;
            range(j, ka, kb);
            if (ka >= kb)
                break; // nothing of this member lies beyond block j
            const int jn = next(j); // in range: an owned block lies beyond j
            if (mine(jn)) {
                const int kn = (jn - m) / G; // the first owned block beyond j
                if (!have)
                    gang_prefetch(P, mode, which, jn, j, ahd);
                have = false;
                gang_ahead_update(mode, ahd, xa, own + kn * kTb, red);
                gang_solve_block(P, jn, ahd, own + kn * kTb, xn, red, pub, (unsigned)epoch);
                if (fwd)
                    gang_sweep(P, mode, m, G, j, kn + 1, kb, xa, own, red);
                else
                    gang_sweep(P, mode, m, G, j, ka, kn, xa, own, red);
                __syncthreads();
                if (t < kTb)
                    xa[t] = xn[t];
                __syncthreads();
            } else {
                gang_sweep(P, mode, m, G, j, ka, kb, xa, own, red);
                __syncthreads();
                int na, nbk;
                range(jn, na, nbk);
                if (na >= nbk)
                    break;
                if (mine(next(jn))) {
                    gang_prefetch(P, mode, which, next(jn), jn, ahd);
                    have = true;
                }
                if (!fetch(jn))
                    return;
            }
        }
        __syncthreads();
    }
    for (int e = t; e < nown * kTb; e += kSolve1Threads) {
        const int i = (m + (e / kTb) * G) * kTb + e % kTb;
        if (i < n)
            io[mode == 0 ? i : P.perm[i]] = own[e];
    }
}

// ---------------------------------------------------------------------------
// Preconditioner set-up: Z^H solves of the constraint rows by gangs
// ---------------------------------------------------------------------------
//
// F^H = Z^{-H} E^H for up to kFsRhs constraint rows per item: the Z^H form of
// the gang solve above, with kFsRhs vectors per row. A thread owns one
// (row, vector) pair; the factor row segment it reads is the same for the 16
// vector lanes of a row. The published words carry kFsRhs entries per row.
// Work item (problem, member, G, group): constraint rows [16 group, + 16).
constexpr int kFsRhs = 16;

__device__ bool ll_fetch_multi(const Prob &P, const unsigned long long *w, int valid_rows, unsigned epoch,
                               double2 *x, int *s_ok)
{
    const int t = threadIdx.x, r = t >> 4; // entry t = (row r, vector t % 16)
    double2 v = make_double2(0.0, 0.0);
    if (r < valid_rows) {
        const unsigned long long *p = w + 4 * t;
        long spins = 0;
        unsigned long long q0, q1, q2, q3;
        for (;;) {
            q0 = __hip_atomic_load(p + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            q1 = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            q2 = __hip_atomic_load(p + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            q3 = __hip_atomic_load(p + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((unsigned)(q0 >> 32) == epoch && (unsigned)(q1 >> 32) == epoch &&
                (unsigned)(q2 >> 32) == epoch && (unsigned)(q3 >> 32) == epoch)
                break;
            ++spins;
            if ((spins & 255) == 0 &&
                __hip_atomic_load(&P.pstate[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                *s_ok = 0;
                break;
            }
            if (spins > kSpinLimit) {
                *s_ok = 0;
                __hip_atomic_store(&P.pstate[2], 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        v.x = __longlong_as_double((long long)((q0 & 0xffffffffull) | (q1 << 32)));
        v.y = __longlong_as_double((long long)((q2 & 0xffffffffull) | (q3 << 32)));
    }
    x[t] = v;
    __syncthreads();
    return *s_ok != 0;
}

__global__ __launch_bounds__(kSolve1Threads) void k_fsetup(const Prob *__restrict__ probs,
                                                           const int4 *__restrict__ items, int epoch,
                                                           unsigned long long *__restrict__ ticket)
{
    extern __shared__ double2 own[]; // owned blocks: [block][row][vector]
    __shared__ double2 xa[kTb * kFsRhs];
    __shared__ int s_item, s_ok;
    const int t = threadIdx.x, r = t >> 4, v = t & 15;
    if (t == 0) {
        s_item = (int)atomicAdd(ticket, 1ULL);
        s_ok = 1;
    }
    __syncthreads();
    const int4 it = items[s_item];
    const Prob &P = probs[it.x];
    if (P.info[0] != 0)
        return;
    const int m = it.y, G = it.z, v0 = it.w * kFsRhs;
    const int n = P.n, ld = P.ldz, nb = (n + kTb - 1) / kTb;
    const int nv = min(kFsRhs, P.ncons - v0);
    const int nown = (nb - m + G - 1) / G;
    const float *__restrict__ Fr = P.fr;
    const float *__restrict__ Fi = P.fi;
    constexpr int BS = kTb * kFsRhs; // entries per block
    for (int k = 0; k < nown; ++k) {
        const int i = (m + k * G) * kTb + r;
        own[k * BS + t] = (i < n && v < nv) ? make_double2(cons_coef(P, v0 + v, i), 0.0) : make_double2(0.0, 0.0);
    }
    __syncthreads();
    // own_b[r][v] -= sum_c conj(F[b0 + r][j0 + c]) x[c][v]: a row of U^H / L^H
    // is a column segment of the factor
    auto update = [&](int j, int k, const double2 *x) {
        const int b0 = (m + k * G) * kTb, j0 = j * kTb, jc = min(kTb, n - j0);
        if (b0 + r >= n)
            return;
        const size_t col = (size_t)(b0 + r) * ld + j0;
        double sr = 0.0, si = 0.0;
#pragma unroll 4
        for (int c4 = 0; c4 < kTb; c4 += 4) {
            if (c4 >= jc)
                break;
            float4 fr = *(const float4 *)(Fr + col + c4), fi = *(const float4 *)(Fi + col + c4);
            if (c4 + 3 >= jc) { // entries past n: column padding
                if (c4 + 1 >= jc) fr.y = fi.y = 0.f;
                if (c4 + 2 >= jc) fr.z = fi.z = 0.f;
                fr.w = fi.w = 0.f;
            }
            const double2 x0 = x[(c4 + 0) * kFsRhs + v], x1 = x[(c4 + 1) * kFsRhs + v];
            const double2 x2 = x[(c4 + 2) * kFsRhs + v], x3 = x[(c4 + 3) * kFsRhs + v];
            sr += (double)fr.x * x0.x + (double)fi.x * x0.y + (double)fr.y * x1.x + (double)fi.y * x1.y +
                  (double)fr.z * x2.x + (double)fi.z * x2.y + (double)fr.w * x3.x + (double)fi.w * x3.y;
            si += (double)fr.x * x0.y - (double)fi.x * x0.x + (double)fr.y * x1.y - (double)fi.y * x1.x +
                  (double)fr.z * x2.y - (double)fi.z * x2.x + (double)fr.w * x3.y - (double)fi.w * x3.x;
        }
        double2 &o = own[k * BS + t];
        o = make_double2(o.x - sr, o.y - si);
    };
    for (int phase = 0; phase < 2; ++phase) {
        const bool fwd = phase == 0;
        const int which = fwd ? 1 : 0; // (U_jj^{-1})^H, then (L_jj^{-1})^H
        unsigned long long *pub = P.gx + (size_t)phase * nb * BS * 4;
        auto range = [&](int j, int &ka, int &kb) {
            if (fwd) {
                ka = j < m ? 0 : (j - m) / G + 1;
                kb = nown;
            } else {
                ka = 0;
                kb = j > m ? (j - m + G - 1) / G : 0;
            }
        };
        // x_j = (D_j^{-1})^H own_j for an owned block (in place), published
        auto solve = [&](int j) {
            const int k = (j - m) / G, j0 = j * kTb, jc = min(kTb, n - j0);
            const float *src = P.dinv + ((size_t)j * 2 + which) * 2 * kTb * kTb;
            double sr = 0.0, si = 0.0;
            for (int c = 0; c < jc; ++c) {
                const size_t e = (size_t)r * kTb + c; // conj(D[c][r])
                const double ar = src[2 * e], ai = -src[2 * e + 1];
                const double2 x = own[k * BS + c * kFsRhs + v];
                sr += ar * x.x - ai * x.y;
                si += ar * x.y + ai * x.x;
            }
            __syncthreads(); // every thread has read the block
            const double2 xv = r < jc ? make_double2(sr, si) : make_double2(0.0, 0.0);
            own[k * BS + t] = xv;
            if (r < jc)
                ll_publish(pub + ((size_t)j * BS + t) * 4, xv, (unsigned)epoch);
            __syncthreads();
        };
        int j = fwd ? 0 : nb - 1;
        const double2 *xj;
        if (j % G == m) {
            solve(j);
            xj = own + ((j - m) / G) * BS;
        } else {
            if (!ll_fetch_multi(P, pub + (size_t)j * BS * 4, min(kTb, n - j * kTb), (unsigned)epoch, xa, &s_ok))
                return;
            xj = xa;
        }
        for (int s = 0; s < nb; ++s) {
            j = fwd ? s : nb - 1 - s;
            int ka, kb;
            range(j, ka, kb);
            if (ka >= kb)
                break;
            const int jn = fwd ? j + 1 : j - 1;
            const bool ahead = jn % G == m;
            if (ahead) {
                const int kn = (jn - m) / G;
                update(j, kn, xj);
                __syncthreads();
                solve(jn);
                for (int k = (fwd ? kn + 1 : ka); k < (fwd ? kb : kn); ++k)
                    update(j, k, xj);
                __syncthreads();
                xj = own + kn * BS;
            } else {
                for (int k = ka; k < kb; ++k)
                    update(j, k, xj);
                __syncthreads();
                int na, nbk;
                range(jn, na, nbk);
                if (na >= nbk)
                    break;
                if (!ll_fetch_multi(P, pub + (size_t)jn * BS * 4, min(kTb, n - jn * kTb), (unsigned)epoch, xa,
                                    &s_ok))
                    return;
                xj = xa;
            }
        }
        __syncthreads();
    }
    // F^H columns: fh[(v0 + v) n + perm[i]]
    for (int k = 0; k < nown; ++k) {
        const int i = (m + k * G) * kTb + r;
        if (i < n && v < nv)
            ((double2 *)P.fh)[(size_t)n * (v0 + v) + P.perm[i]] = own[k * BS + t];
    }
}

// Batched Z or Z^H solves. Work item (problem, member, G): G = 1 runs the
// one-workgroup solve with the whole vector in LDS; G > 1 is one member of a
// gang. Workgroups take items from an atomic ticket in list order; a gang's
// members are consecutive, so every member a workgroup waits on has taken its
// ticket earlier or will take the next ones while the waiting workgroups hold
// fewer slots than the device has, and every wait is bounded.
__global__ __launch_bounds__(kSolve1Threads) void k_zsolve(const Prob *__restrict__ probs,
                                                           const int4 *__restrict__ items, int mode, int epoch,
                                                           unsigned long long *__restrict__ ticket)
{
    extern __shared__ double2 vec[]; // G = 1: n entries; gang member: its owned blocks
    __shared__ double2 part[kGangWaves][kTb];
    __shared__ double2 xa[kTb], xn[kTb];
    __shared__ int s_item, s_ok;
    if (threadIdx.x == 0) {
        s_item = (int)atomicAdd(ticket, 1ULL);
        s_ok = 1;
    }
    __syncthreads();
    const int4 it = items[s_item];
    const Prob &P = probs[it.x];
    if (P.info[0] != 0 || P.pstate[0] != 0)
        return;
    if (it.z == 1)
        zsolve1_body(P, mode, vec, part, xa);
    else
        zsolve_gang_body(P, mode, it.y, it.z, epoch, vec, part, xa, xn, &s_ok);
}

// perm[i]: row of Z that the LU moved to row i (ipiv applied in order).
// One workgroup per problem; interchanges are rare, so one wave finds them
// with a ballot and lane 0 applies them in order.
__global__ __launch_bounds__(256) void k_perm(const Prob *__restrict__ probs,
                                              const int *__restrict__ plist)
{
    const Prob &P = probs[plist[blockIdx.x]];
    if (P.info[0] != 0)
        return;
    const int n = P.n, t = threadIdx.x;
    for (int i = t; i < n; i += 256)
        P.perm[i] = i;
    __syncthreads();
    if (t < 64) {
        for (int j0 = 0; j0 < n; j0 += 64) {
            const int j = j0 + t;
            const bool sw = j < n && P.ipiv[j] != j;
            unsigned long long m = __ballot(sw);
            while (m) {
                const int k = __ffsll((long long)m) - 1;
                m &= m - 1;
                if (t == 0) {
                    const int a = j0 + k, b = P.ipiv[j0 + k];
                    const int x = P.perm[a];
                    P.perm[a] = P.perm[b];
                    P.perm[b] = x;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Preconditioner set-up: K = I + F F^H, H = -F B, Cholesky factors
// ---------------------------------------------------------------------------

// In-place Cholesky of the Hermitian m x m matrix a (row-major complex, lower
// triangle used); returns false on a non-positive pivot. Whole workgroup.
__device__ bool block_cholesky(double2 *a, int m, double *red)
{
    const int t = threadIdx.x, nt = blockDim.x;
    __shared__ int ok;
    if (t == 0)
        ok = 1;
    __syncthreads();
    for (int j = 0; j < m; ++j) {
        double part = 0.0;
        for (int k = t; k < j; k += nt) {
            const double2 l = a[(size_t)j * m + k];
            part += l.x * l.x + l.y * l.y;
        }
        const double ss = block_sum(part, red);
        const double d = a[(size_t)j * m + j].x - ss;
        if (!(d > 0.0)) {
            if (t == 0)
                ok = 0;
            __syncthreads();
            return false;
        }
        const double ljj = sqrt(d);
        __syncthreads();
        for (int i = j + 1 + t; i < m; i += nt) {
            double2 s = a[(size_t)i * m + j];
            for (int k = 0; k < j; ++k) // L[i][k] conj(L[j][k])
                s = csub(s, cm(a[(size_t)i * m + k], make_double2(a[(size_t)j * m + k].x, -a[(size_t)j * m + k].y)));
            a[(size_t)i * m + j] = make_double2(s.x / ljj, s.y / ljj);
        }
        if (t == 0)
            a[(size_t)j * m + j] = make_double2(ljj, 0.0);
        __syncthreads();
    }
    return ok != 0;
}

// One workgroup per problem, after fh holds Z^{-H} E^H.
__global__ __launch_bounds__(kCtlThreads) void k_pre_setup(const Prob *__restrict__ probs,
                                                           const int *__restrict__ plist)
{
    const Prob &P = probs[plist[blockIdx.x]];
    if (P.info[0] != 0)
        return;
    const int n = P.n, nc = P.ncons, w = P.w, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int nw = kCtlThreads / 64;
    __shared__ double red[64];
    const double2 *fh = (const double2 *)P.fh;
    double2 *lk = (double2 *)P.lk, *ym = (double2 *)P.ym, *ls = (double2 *)P.ls;
    // K (lower triangle) and H, one wave per entry
    const int npair = nc * (nc + 1) / 2;
    for (int e = wave; e < npair + nc * w; e += nw) {
        double2 s = make_double2(0.0, 0.0);
        if (e < npair) {
            int c = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
            while (c * (c + 1) / 2 > e)
                --c;
            while ((c + 1) * (c + 2) / 2 <= e)
                ++c;
            const int d = e - c * (c + 1) / 2;
            const double2 *fc = fh + (size_t)c * n, *fd = fh + (size_t)d * n;
            // K[c][d] = delta + sum_r F[c][r] conj(F[d][r]) with F[c][r] = conj(fh_c[r])
            for (int r = lane; r < n; r += 64)
                s = cadd(s, cmc(fc[r], fd[r]));
            s = wave_sum2(s);
            if (lane == 0) {
                const double2 kv = make_double2(s.x + (c == d ? 1.0 : 0.0), s.y);
                lk[(size_t)c * nc + d] = kv;
                if (c != d)
                    lk[(size_t)d * nc + c] = make_double2(kv.x, -kv.y);
            }
        } else {
            const int c = (e - npair) / w, col = (e - npair) % w;
            const double2 *fc = fh + (size_t)c * n;
            for (int r = lane; r < n; r += 64)
                if (P.rowcol[r] == col) { // H[c][w] = -sum F[c][r] B[r][w] = sum conj(fh) cos
                    const double cv = P.cosv[r];
                    s = cadd(s, make_double2(fc[r].x * cv, -fc[r].y * cv));
                }
            s = wave_sum2(s);
            if (lane == 0)
                ym[(size_t)c * w + col] = s;
        }
    }
    __syncthreads();
    if (nc > 0 && !block_cholesky(lk, nc, red)) {
        if (t == 0)
            P.pstate[2] = 1;
        return;
    }
    // Y = Lk^{-1} H, one thread per column
    for (int col = t; col < w; col += kCtlThreads)
        for (int c = 0; c < nc; ++c) {
            double2 s = ym[(size_t)c * w + col];
            for (int k = 0; k < c; ++k)
                s = csub(s, cm(lk[(size_t)c * nc + k], ym[(size_t)k * w + col]));
            ym[(size_t)c * w + col] = cdivd(s, lk[(size_t)c * nc + c]);
        }
    __syncthreads();
    // Sc = Y^H Y + lambda I (lower triangle and its mirror)
    for (int e = wave; e < w * w; e += nw) {
        const int a = e / w, b = e % w;
        if (b > a)
            continue;
        double2 s = make_double2(0.0, 0.0);
        for (int c = lane; c < nc; c += 64)
            s = cadd(s, cmc(ym[(size_t)c * w + a], ym[(size_t)c * w + b]));
        s = wave_sum2(s);
        if (lane == 0) {
            if (a == b)
                s = make_double2(s.x + kLambda, 0.0);
            ls[(size_t)a * w + b] = s;
            if (a != b)
                ls[(size_t)b * w + a] = make_double2(s.x, -s.y);
        }
    }
    __syncthreads();
    if (!block_cholesky(ls, w, red)) {
        if (t == 0)
            P.pstate[2] = 1;
    }
}

// ---------------------------------------------------------------------------
// Dense middle of the preconditioner
// ---------------------------------------------------------------------------

// In: vv = Z^{-H} r_I, r_c = vr[n..s). Out: vv = right-hand side of the Z
// solve, vz[n..s) = c.
__global__ __launch_bounds__(kCtlThreads) void k_prec_mid(const Prob *__restrict__ probs,
                                                          const int *__restrict__ plist)
{
    const Prob &P = probs[plist[blockIdx.x]];
    if (P.info[0] != 0 || P.pstate[0] != 0)
        return;
    const int n = P.n, nc = P.ncons, w = P.w, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int nw = kCtlThreads / 64;
    __shared__ double2 f[kMixMaxCons], a[kMixMaxCons], sc[kMixMaxW], cc[kMixMaxW];
    const double2 *fh = (const double2 *)P.fh, *lk = (const double2 *)P.lk;
    const double2 *ym = (const double2 *)P.ym, *ls = (const double2 *)P.ls;
    double2 *vv = (double2 *)P.vv;
    const double2 *vr = (const double2 *)P.vr;
    // f = F s_u (F[c][r] = conj(fh[c][r])) and s_c = r_c - B^H s_u
    for (int e = wave; e < nc + w; e += nw) {
        double2 s = make_double2(0.0, 0.0);
        if (e < nc) {
            const double2 *fc = fh + (size_t)e * n;
            for (int r = lane; r < n; r += 64)
                s = cadd(s, cmc(fc[r], vv[r]));
        } else {
            const int col = e - nc;
            for (int r = lane; r < n; r += 64)
                if (P.rowcol[r] == col) {
                    const double cv = P.cosv[r];
                    s = cadd(s, make_double2(cv * vv[r].x, cv * vv[r].y));
                }
        }
        s = wave_sum2(s);
        if (lane == 0) {
            if (e < nc)
                f[e] = s;
            else
                sc[e - nc] = cadd(vr[n + e - nc], s);
        }
    }
    __syncthreads();
    // a = Lk^{-1} f (wave 0, one row at a time)
    if (wave == 0)
        for (int i = 0; i < nc; ++i) {
            double2 s = make_double2(0.0, 0.0);
            for (int k = lane; k < i; k += 64)
                s = cadd(s, cm(lk[(size_t)i * nc + k], a[k]));
            s = wave_sum2(s);
            if (lane == 0)
                a[i] = cdivd(csub(f[i], s), lk[(size_t)i * nc + i]);
            __builtin_amdgcn_wave_barrier();
        }
    __syncthreads();
    // rhs_c = s_c - Y^H a
    for (int col = wave; col < w; col += nw) {
        double2 s = make_double2(0.0, 0.0);
        for (int c = lane; c < nc; c += 64)
            s = cadd(s, cmc(ym[(size_t)c * w + col], a[c]));
        s = wave_sum2(s);
        if (lane == 0)
            sc[col] = csub(sc[col], s);
    }
    __syncthreads();
    // c = Ls^{-H} Ls^{-1} rhs_c (thread 0; w is small)
    if (t == 0) {
        for (int i = 0; i < w; ++i) {
            double2 s = sc[i];
            for (int k = 0; k < i; ++k)
                s = csub(s, cm(ls[(size_t)i * w + k], cc[k]));
            cc[i] = cdivd(s, ls[(size_t)i * w + i]);
        }
        for (int i = w - 1; i >= 0; --i) {
            double2 s = cc[i];
            for (int k = i + 1; k < w; ++k) // Ls^H[i][k] = conj(Ls[k][i])
                s = csub(s, cmc(ls[(size_t)k * w + i], cc[k]));
            cc[i] = cdivd(s, make_double2(ls[(size_t)i * w + i].x, -ls[(size_t)i * w + i].y));
        }
    }
    __syncthreads();
    // b = a + Y c, then e = Lk^{-H} b (in a), wave 0
    for (int c = t; c < nc; c += kCtlThreads) {
        double2 s = a[c];
        for (int k = 0; k < w; ++k)
            s = cadd(s, cm(ym[(size_t)c * w + k], cc[k]));
        f[c] = s;
    }
    __syncthreads();
    if (wave == 0)
        for (int i = nc - 1; i >= 0; --i) {
            double2 s = make_double2(0.0, 0.0);
            for (int k = i + 1 + lane; k < nc; k += 64) // Lk^H[i][k] = conj(Lk[k][i])
                s = cadd(s, cmc(lk[(size_t)k * nc + i], a[k]));
            s = wave_sum2(s);
            if (lane == 0) {
                const double2 d = lk[(size_t)i * nc + i];
                a[i] = cdivd(csub(f[i], s), make_double2(d.x, -d.y));
            }
            __builtin_amdgcn_wave_barrier();
        }
    __syncthreads();
    // u = s_u - F^H e ; rhs of the Z solve = u - B c = u + cos * c
    for (int r = t; r < n; r += kCtlThreads) {
        double2 s = vv[r];
        for (int c = 0; c < nc; ++c)
            s = csub(s, cm(fh[(size_t)c * n + r], a[c]));
        const int col = P.rowcol[r];
        if (col >= 0) {
            const double cv = P.cosv[r];
            s = cadd(s, make_double2(cv * cc[col].x, cv * cc[col].y));
        }
        vv[r] = s;
    }
    double2 *vz = (double2 *)P.vz;
    for (int col = t; col < w; col += kCtlThreads)
        vz[n + col] = cc[col];
}

// ---------------------------------------------------------------------------
// t = M p and q = M^H t + lambda p in FP64
// ---------------------------------------------------------------------------

// Work item: (problem, row block, column chunk). part[chunk][row] = Z[row, chunk] p.
__global__ __launch_bounds__(kMvRows) void k_mv_z(const Prob *__restrict__ probs,
                                                  const int4 *__restrict__ work)
{
    const int4 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    if (P.info[0] != 0 || P.pstate[0] != 0)
        return;
    const int n = P.n, t = threadIdx.x;
    const int j0 = wk.z * kMvChunk, j1 = min(n, j0 + kMvChunk);
    __shared__ double2 ps[kMvChunk];
    const double2 *p = (const double2 *)P.vp;
    if (j0 + t < j1)
        ps[t] = p[j0 + t];
    __syncthreads();
    const int row = wk.y * kMvRows + t;
    if (row >= n)
        return;
    double sr = 0.0, si = 0.0;
    for (int j = j0; j < j1; ++j) {
        const size_t o = (size_t)j * P.ldz + row;
        const double zr = P.zr[o], zi = P.zi[o];
        const double2 pj = ps[j - j0];
        sr += zr * pj.x - zi * pj.y;
        si += zr * pj.y + zi * pj.x;
    }
    ((double2 *)P.part)[(size_t)wk.z * n + row] = make_double2(sr, si);
}

// Work item: (problem, row block). t[r] = sum of partials + B p_c; block 0
// also writes the constraint rows.
__global__ __launch_bounds__(kMvRows) void k_mv_fin(const Prob *__restrict__ probs,
                                                    const int2 *__restrict__ work)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    if (P.info[0] != 0 || P.pstate[0] != 0)
        return;
    const int n = P.n, t = threadIdx.x;
    const double2 *p = (const double2 *)P.vp, *part = (const double2 *)P.part;
    double2 *vt = (double2 *)P.vt;
    const int row = wk.y * kMvRows + t;
    if (row < n) {
        double2 s = make_double2(0.0, 0.0);
        for (int c = 0; c < P.nchunk; ++c)
            s = cadd(s, part[(size_t)c * n + row]);
        const int col = P.rowcol[row];
        if (col >= 0) {
            const double cv = P.cosv[row];
            const double2 pc = p[n + col];
            s = csub(s, make_double2(cv * pc.x, cv * pc.y));
        }
        vt[row] = s;
    }
    if (wk.y == 0)
        for (int c = t; c < P.ncons; c += kMvRows) {
            const int a = P.cons[2 * c], b = P.cons[2 * c + 1];
            const double sg = P.cons_sign[c];
            double2 s;
            if (b < 0)
                s = p[a];
            else if (a == b)
                s = make_double2(sg * p[a].x, sg * p[a].y);
            else
                s = make_double2(p[a].x + sg * p[b].x, p[a].y + sg * p[b].y);
            vt[n + c] = s;
        }
}

// Work item: (problem, column block of kMvCols). One wave per column:
// q_j = sum_r conj(Z[r][j]) t_r + sum_c E[c][j] t_{n+c} + lambda p_j.
// With colsq set, the first call also records |Z[:, j]|^2.
__global__ __launch_bounds__(256) void k_mv_zh(const Prob *__restrict__ probs,
                                               const int2 *__restrict__ work, int with_colsq)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    if (P.info[0] != 0 || P.pstate[0] != 0)
        return;
    const int n = P.n, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double2 *vt = (const double2 *)P.vt, *p = (const double2 *)P.vp;
    double2 *q = (double2 *)P.vq;
    for (int jj = wave; jj < kMvCols; jj += 4) {
        const int j = wk.y * kMvCols + jj;
        if (j >= n)
            break;
        const double *zr = P.zr + (size_t)j * P.ldz, *zi = P.zi + (size_t)j * P.ldz;
        double sr = 0.0, si = 0.0, z2 = 0.0;
        for (int r = lane; r < n; r += 64) {
            const double a = zr[r], b = zi[r];
            const double2 tv = vt[r];
            sr += a * tv.x + b * tv.y;
            si += a * tv.y - b * tv.x;
            z2 += a * a + b * b;
        }
        sr = wave_sum1(sr);
        si = wave_sum1(si);
        if (with_colsq)
            z2 = wave_sum1(z2);
        if (lane == 0) {
            for (int c = 0; c < P.ncons; ++c) {
                const double e = cons_coef(P, c, j);
                if (e != 0.0) {
                    sr += e * vt[n + c].x;
                    si += e * vt[n + c].y;
                }
            }
            q[j] = make_double2(sr + kLambda * p[j].x, si + kLambda * p[j].y);
            if (with_colsq)
                P.colsq[j] = z2;
        }
    }
}

// ---------------------------------------------------------------------------
// CG control (one workgroup per problem)
// ---------------------------------------------------------------------------

// t = [rhs; 0], p = 0, x = 0, state cleared: the next k_mv_zh forms h = M^H y.
__global__ __launch_bounds__(kCtlThreads) void k_mix_init(const Prob *__restrict__ probs,
                                                          const int *__restrict__ plist)
{
    const Prob &P = probs[plist[blockIdx.x]];
    const int n = P.n, s = P.s, t = threadIdx.x;
    double2 *vt = (double2 *)P.vt, *vp = (double2 *)P.vp, *x = (double2 *)P.x;
    const double2 *rhs = (const double2 *)P.rhs;
    for (int r = t; r < n + P.ncons; r += kCtlThreads)
        vt[r] = r < n ? rhs[r] : make_double2(0.0, 0.0);
    for (int r = t; r < s; r += kCtlThreads) {
        vp[r] = make_double2(0.0, 0.0);
        x[r] = make_double2(0.0, 0.0);
    }
    if (t < 3)
        P.pstate[t] = 0;
    if (t < 4)
        P.pcg[t] = 0.0;
}

// Homogeneous block of q: q_c[w] = -sum_{rows of w} cos[r] t[r] + lambda p_c[w].
__device__ void q_homog(const Prob &P, double *red)
{
    const int n = P.n, t = threadIdx.x;
    const double2 *vt = (const double2 *)P.vt, *p = (const double2 *)P.vp;
    double2 *q = (double2 *)P.vq;
    for (int col = 0; col < P.w; ++col) {
        double2 s = make_double2(0.0, 0.0);
        for (int r = t; r < n; r += kCtlThreads)
            if (P.rowcol[r] == col) {
                const double cv = P.cosv[r];
                s = csub(s, make_double2(cv * vt[r].x, cv * vt[r].y));
            }
        s = block_sum2(s, red);
        if (t == 0)
            q[n + col] = make_double2(s.x + kLambda * p[n + col].x, s.y + kLambda * p[n + col].y);
    }
    __syncthreads();
}

// After k_mv_zh on t = [rhs; 0]: h = q; r = h; |h|, |M|_F^2; vv = r_I.
__global__ __launch_bounds__(kCtlThreads) void k_mix_start(const Prob *__restrict__ probs,
                                                           const int *__restrict__ plist)
{
    const Prob &P = probs[plist[blockIdx.x]];
    if (P.info[0] != 0)
        return;
    __shared__ double red[64];
    q_homog(P, red);
    const int n = P.n, s = P.s, t = threadIdx.x;
    const double2 *q = (const double2 *)P.vq;
    double2 *vr = (double2 *)P.vr, *vv = (double2 *)P.vv;
    double h2 = 0.0, m2 = 0.0;
    for (int r = t; r < s; r += kCtlThreads) {
        const double2 v = q[r];
        vr[r] = v;
        if (r < n) {
            vv[r] = v;
            m2 += P.colsq[r] + P.cosv[r] * P.cosv[r];
        }
        h2 += v.x * v.x + v.y * v.y;
    }
    for (int c = t; c < P.ncons; c += kCtlThreads) {
        const int a = P.cons[2 * c], b = P.cons[2 * c + 1];
        const double sg = P.cons_sign[c];
        m2 += b < 0 ? 1.0 : (a == b ? sg * sg : 1.0 + sg * sg);
    }
    h2 = block_sum(h2, red);
    m2 = block_sum(m2, red);
    if (t == 0) {
        P.pcg[1] = sqrt(h2);
        P.pcg[2] = sqrt(h2);
        P.pcg[3] = m2;
        if (!(h2 > 0.0))
            P.pstate[0] = 1; // zero excitation: x = 0
        if (P.pstate[2] != 0)
            P.pstate[0] = 1;
    }
}

// z = [vv; vz_c]; rho' = <r, z>; p = z (first) or z + (rho'/rho) p.
__global__ __launch_bounds__(kCtlThreads) void k_mix_dir(const Prob *__restrict__ probs,
                                                         const int *__restrict__ plist)
{
    const Prob &P = probs[plist[blockIdx.x]];
    if (P.info[0] != 0 || P.pstate[0] != 0)
        return;
    __shared__ double red[64];
    const int n = P.n, s = P.s, t = threadIdx.x;
    const double2 *vr = (const double2 *)P.vr, *vv = (const double2 *)P.vv, *vz = (const double2 *)P.vz;
    double2 *vp = (double2 *)P.vp;
    double rz = 0.0;
    for (int r = t; r < s; r += kCtlThreads) {
        const double2 z = r < n ? vv[r] : vz[r];
        rz += vr[r].x * z.x + vr[r].y * z.y; // Re <r, z>
    }
    rz = block_sum(rz, red);
    const bool first = P.pstate[1] == 0;
    const double beta = first ? 0.0 : rz / P.pcg[0];
    for (int r = t; r < s; r += kCtlThreads) {
        const double2 z = r < n ? vv[r] : vz[r];
        const double2 pv = first ? make_double2(0.0, 0.0) : vp[r];
        vp[r] = make_double2(z.x + beta * pv.x, z.y + beta * pv.y);
    }
    __syncthreads();
    if (t == 0) {
        P.pcg[0] = rz;
        if (!(rz > 0.0) || !isfinite(rz)) {
            P.pstate[2] = 2; // preconditioner lost positive definiteness
            P.pstate[0] = 1;
        }
    }
}

// After q = G p: alpha = rho / <p, q>; x += alpha p; r -= alpha q; test.
__global__ __launch_bounds__(kCtlThreads) void k_mix_step(const Prob *__restrict__ probs,
                                                          const int *__restrict__ plist)
{
    const Prob &P = probs[plist[blockIdx.x]];
    if (P.info[0] != 0 || P.pstate[0] != 0)
        return;
    __shared__ double red[64];
    q_homog(P, red);
    const int n = P.n, s = P.s, t = threadIdx.x;
    const double2 *vp = (const double2 *)P.vp, *q = (const double2 *)P.vq;
    double2 *vr = (double2 *)P.vr, *x = (double2 *)P.x, *vv = (double2 *)P.vv;
    double pq = 0.0;
    for (int r = t; r < s; r += kCtlThreads)
        pq += vp[r].x * q[r].x + vp[r].y * q[r].y;
    pq = block_sum(pq, red);
    const double rho = P.pcg[0];
    if (!(pq > 0.0) || !isfinite(pq)) {
        if (t == 0) {
            P.pstate[2] = 3;
            P.pstate[0] = 1;
        }
        return;
    }
    const double alpha = rho / pq;
    double r2 = 0.0, x2 = 0.0;
    for (int r = t; r < s; r += kCtlThreads) {
        x[r] = make_double2(x[r].x + alpha * vp[r].x, x[r].y + alpha * vp[r].y);
        x2 += x[r].x * x[r].x + x[r].y * x[r].y;
        const double2 nr = make_double2(vr[r].x - alpha * q[r].x, vr[r].y - alpha * q[r].y);
        vr[r] = nr;
        if (r < n)
            vv[r] = nr;
        r2 += nr.x * nr.x + nr.y * nr.y;
    }
    r2 = block_sum(r2, red);
    x2 = block_sum(x2, red);
    if (t == 0) {
        const double rn = sqrt(r2);
        P.pcg[4] = sqrt(x2);
        const int it = P.pstate[1] + 1;
        P.pstate[1] = it;
        P.pcg[2] = rn;
        if (rn <= P.tol * P.pcg[1] || it >= P.maxit)
            P.pstate[0] = 1;
    }
}

// Convergence flags of a batch, gathered into one array for the host.
__global__ __launch_bounds__(256) void k_mix_flags(const Prob *__restrict__ probs, unsigned np,
                                                   int *__restrict__ flags)
{
    const unsigned b = blockIdx.x * 256 + threadIdx.x;
    if (b < np)
        flags[b] = (probs[b].info[0] != 0 || probs[b].pstate[0] != 0) ? 1 : 0;
}

// FP32 MFMA operand-layout probe: D = A B for A (16x4), B (4x16) row-major;
// writes the raw per-lane accumulator.
__global__ __launch_bounds__(64) void k_mfma_probe32(const float *__restrict__ a,
                                                     const float *__restrict__ b,
                                                     float *__restrict__ out)
{
    typedef float f4v __attribute__((ext_vector_type(4)));
    const int l = threadIdx.x;
    f4v acc = f4v{0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[(l % 16) * 4 + l / 16], b[(l / 16) * 16 + l % 16],
                                               acc, 0, 0, 0);
    for (int v = 0; v < 4; ++v)
        out[l * 4 + v] = acc[v];
}

} // namespace fnec
