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
//   k_zsolve     Z v = b or Z^H v = b with the FP32 factors, FP64 vector;
//                one workgroup per (problem, vector), panel-interleaved
//                interchanges as recorded by the LU (lu.h).
//   k_pre_setup  K, H, their Cholesky / Schur factors (one workgroup per problem).
//   k_prec_mid   the dense middle of P^{-1} between the two Z solves.
//   k_mv_z, k_mv_fin, k_mv_zh   t = M p and q = M^H t + lambda p (FP64 Z).
//   k_mix_*      CG control: init, direction, step.
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.h"

namespace fnec {

constexpr int kSolveThreads = 1024;
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
// Triangular solves with the FP32 factors
// ---------------------------------------------------------------------------

// Work item: (problem, vector). vector < 0: the CG vector vv (skipped once the
// problem has converged); vector c >= 0: column c of fh, initialised here to
// the conjugate transpose of constraint row c and solved with Z^H.
// mode 0: Z v = b  = forward L with interchanges, then backward U.
// mode 1: Z^H v = b = forward U^H, then backward L^H with interchanges undone.
__global__ __launch_bounds__(kSolveThreads) void k_zsolve(const Prob *__restrict__ probs,
                                                          const int2 *__restrict__ work, int mode)
{
    const int2 wk = work[blockIdx.x];
    const Prob &P = probs[wk.x];
    if (P.info[0] != 0)
        return;
    if (wk.y < 0 && P.pstate[0] != 0)
        return;
    const int n = P.n, ld = P.ldz, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int nt = kSolveThreads, nw = kSolveThreads / 64;
    const float *Fr = P.fr, *Fi = P.fi;
    double2 *vec = (double2 *)(wk.y < 0 ? P.vv : P.fh + 2 * (size_t)n * wk.y);
    __shared__ float dr[kPanel][kPanel + 1], di[kPanel][kPanel + 1];
    __shared__ double2 xs[kPanel];
    __shared__ double2 accs[kPanel];
    if (wk.y >= 0) {
        for (int r = t; r < n; r += nt)
            vec[r] = make_double2(cons_coef(P, wk.y, r), 0.0);
        __syncthreads();
    }
    auto stage = [&](int r0, int jb) {
        for (int e = t; e < jb * jb; e += nt) {
            const int r = e % jb, c = e / jb;
            dr[r][c] = Fr[(size_t)(r0 + c) * ld + r0 + r];
            di[r][c] = Fi[(size_t)(r0 + c) * ld + r0 + r];
        }
    };
    if (mode == 0) {
        // forward: per panel, interchanges then unit-lower elimination
        for (int p = 0; p < n; p += kPanel) {
            const int jb = min(kPanel, n - p);
            stage(p, jb);
            if (t == 0)
                for (int j = 0; j < jb; ++j) {
                    const int piv = P.ipiv[p + j];
                    if (piv != p + j) {
                        const double2 a = vec[p + j];
                        vec[p + j] = vec[piv];
                        vec[piv] = a;
                    }
                }
            __syncthreads();
            if (wave == 0) {
                double2 x = lane < jb ? vec[p + lane] : make_double2(0.0, 0.0);
                for (int k = 0; k < jb; ++k) {
                    const double2 xk = make_double2(__shfl(x.x, k, 64), __shfl(x.y, k, 64));
                    if (lane > k && lane < jb)
                        x = csub(x, cm(make_double2(dr[lane][k], di[lane][k]), xk));
                }
                if (lane < jb) {
                    vec[p + lane] = x;
                    xs[lane] = x;
                }
            }
            __syncthreads();
            for (int r = p + jb + t; r < n; r += nt) {
                double2 acc = vec[r];
                for (int k = 0; k < jb; ++k) {
                    const size_t o = (size_t)(p + k) * ld + r;
                    acc = csub(acc, cm(make_double2(Fr[o], Fi[o]), xs[k]));
                }
                vec[r] = acc;
            }
            __syncthreads();
        }
        // backward: U, blocks from the bottom
        for (int r1 = n; r1 > 0; r1 -= kPanel) {
            const int r0 = max(0, r1 - kPanel), jb = r1 - r0;
            stage(r0, jb);
            __syncthreads();
            if (wave == 0) {
                double2 x = lane < jb ? vec[r0 + lane] : make_double2(0.0, 0.0);
                for (int k = jb - 1; k >= 0; --k) {
                    if (lane == k)
                        x = cdivd(x, make_double2(dr[k][k], di[k][k]));
                    const double2 xk = make_double2(__shfl(x.x, k, 64), __shfl(x.y, k, 64));
                    if (lane < k)
                        x = csub(x, cm(make_double2(dr[lane][k], di[lane][k]), xk));
                }
                if (lane < jb) {
                    vec[r0 + lane] = x;
                    xs[lane] = x;
                }
            }
            __syncthreads();
            for (int m = t; m < r0; m += nt) {
                double2 acc = vec[m];
                for (int k = 0; k < jb; ++k) {
                    const size_t o = (size_t)(r0 + k) * ld + m;
                    acc = csub(acc, cm(make_double2(Fr[o], Fi[o]), xs[k]));
                }
                vec[m] = acc;
            }
            __syncthreads();
        }
    } else {
        // forward: U^H, row i of U^H is the conjugate of column i of U
        for (int r0 = 0; r0 < n; r0 += kPanel) {
            const int jb = min(kPanel, n - r0);
            stage(r0, jb);
            for (int i = wave; i < jb; i += nw) {
                const size_t col = (size_t)(r0 + i) * ld;
                double2 s = make_double2(0.0, 0.0);
                for (int c = lane; c < r0; c += 64)
                    s = cadd(s, cmc(make_double2(Fr[col + c], Fi[col + c]), vec[c]));
                s = wave_sum2(s);
                if (lane == 0)
                    accs[i] = csub(vec[r0 + i], s);
            }
            __syncthreads();
            if (wave == 0) {
                double2 x = lane < jb ? accs[lane] : make_double2(0.0, 0.0);
                for (int k = 0; k < jb; ++k) {
                    if (lane == k)
                        x = cdivd(x, make_double2(dr[k][k], -di[k][k]));
                    const double2 xk = make_double2(__shfl(x.x, k, 64), __shfl(x.y, k, 64));
                    if (lane > k && lane < jb) // U^H[lane][k] = conj(U[k][lane])
                        x = csub(x, cmc(make_double2(dr[k][lane], di[k][lane]), xk));
                }
                if (lane < jb)
                    vec[r0 + lane] = x;
            }
            __syncthreads();
        }
        // backward: L^H per panel (last first), then the panel's interchanges
        for (int p = ((n - 1) / kPanel) * kPanel; p >= 0; p -= kPanel) {
            const int jb = min(kPanel, n - p), pe = p + jb;
            stage(p, jb);
            for (int i = wave; i < jb; i += nw) {
                const size_t col = (size_t)(p + i) * ld;
                double2 s = make_double2(0.0, 0.0);
                for (int c = pe + lane; c < n; c += 64)
                    s = cadd(s, cmc(make_double2(Fr[col + c], Fi[col + c]), vec[c]));
                s = wave_sum2(s);
                if (lane == 0)
                    accs[i] = csub(vec[p + i], s);
            }
            __syncthreads();
            if (wave == 0) {
                double2 x = lane < jb ? accs[lane] : make_double2(0.0, 0.0);
                for (int k = jb - 1; k >= 0; --k) {
                    const double2 xk = make_double2(__shfl(x.x, k, 64), __shfl(x.y, k, 64));
                    if (lane < k) // L^H[lane][k] = conj(L[k][lane])
                        x = csub(x, cmc(make_double2(dr[k][lane], di[k][lane]), xk));
                }
                if (lane < jb)
                    vec[p + lane] = x;
            }
            __syncthreads();
            if (t == 0)
                for (int j = jb - 1; j >= 0; --j) {
                    const int piv = P.ipiv[p + j];
                    if (piv != p + j) {
                        const double2 a = vec[p + j];
                        vec[p + j] = vec[piv];
                        vec[piv] = a;
                    }
                }
            __syncthreads();
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
