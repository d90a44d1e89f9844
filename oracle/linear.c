/*
 * ORACLE — test infrastructure only.
 *
 * Dense complex solvers, restating crates/nec_solver/src/linear.rs:
 *   solve_hallen            :672-797
 *   solve_hallen_paths      :823-894
 *   solve_square_in_place   :1321-1383
 * Data layout follows the reference: M, the Gram matrix and the augmented
 * elimination matrix are arrays of separately allocated row vectors
 * (Vec<Vec<Complex64>>), the Gram loop walks down a column of M across rows,
 * and the pivot is the last row of maximal hypot (Iterator::max_by).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

/* Threads for the row loops below (fixture generation only; 1 by default,
 * as the reference solve runs on one thread). Each parallel iteration owns
 * whole elements or rows and keeps their operation order, so results do not
 * depend on the thread count. */
int or_threads = 1;
void or_set_threads(int n) { or_threads = n > 0 ? n : 1; }

/* Cache-blocked loop nests for the Gram matrix and the elimination (0 by
 * default: the reference's loop order, which the CPU baseline times). Every
 * element still receives the same operations in the same order, so the
 * results are bit for bit those of the reference-order loops; only the order
 * in which different elements are visited changes. For fixtures of large
 * systems, whose reference-order column walks over M and full sweeps of the
 * trailing matrix per column do not fit in any cache. */
int or_blocked = 0;
void or_set_blocked(int on) { or_blocked = on != 0; }

#define OR_RB 128 /* rows of M per block (Gram) */
#define OR_TB 256 /* Gram output tile */
#define OR_PB 64  /* elimination panel width */

static cx **rows_alloc(int r, int c)
{
    cx **m = (cx **)malloc(sizeof(cx *) * (size_t)(r > 0 ? r : 1));
    int i;
    for (i = 0; i < r; i++)
        m[i] = (cx *)calloc((size_t)(c > 0 ? c : 1), sizeof(cx));
    return m;
}

static void rows_free(cx **m, int r)
{
    int i;
    for (i = 0; i < r; i++)
        free(m[i]);
    free(m);
}

/* Elimination on [A | b] held as row vectors; consumes `aug`. piv (may be
 * NULL) receives the pivot row chosen at each column. */
static int solve_aug_piv(cx **aug, int n, cx *x, int *pivots)
{
    int col, row, i, j;
    for (col = 0; col < n; col++) {
        int piv = col;
        double best = cx_norm(aug[col][col]);
        cx diag;
        for (row = col + 1; row < n; row++) {
            double v = cx_norm(aug[row][col]);
            /* max_by keeps the later element when it compares >= ; a NaN
             * comparison counts as Equal, which also keeps the later one. */
            if (!(v < best)) {
                best = v;
                piv = row;
            }
        }
        if (pivots)
            pivots[col] = piv;
        if (cx_norm(aug[piv][col]) < 1e-30)
            return OR_SINGULAR;
        if (piv != col) {
            cx *t = aug[col];
            aug[col] = aug[piv];
            aug[piv] = t;
        }
        diag = aug[col][col];
#pragma omp parallel for num_threads(or_threads) schedule(static) if (or_threads > 1 && n - col > 256)
        for (row = col + 1; row < n; row++) {
            cx factor = cx_div(aug[row][col], diag);
            cx *ar = aug[row], *ac = aug[col];
            int kk;
            for (kk = col; kk <= n; kk++) {
                cx sub = cx_mul(factor, ac[kk]);
                ar[kk] = cx_sub(ar[kk], sub);
            }
        }
    }
    for (i = n - 1; i >= 0; i--) {
        cx sum = aug[i][n];
        for (j = i + 1; j < n; j++)
            sum = cx_sub(sum, cx_mul(aug[i][j], x[j]));
        x[i] = cx_div(sum, aug[i][i]);
    }
    return OR_OK;
}

/* solve_aug_piv with the elimination in panels of OR_PB columns: inside a
 * panel, column col is searched and its rows below are eliminated on the
 * panel's columns only, with the factors kept; then every row below the
 * panel's first row takes, on the columns right of the panel (the right-hand
 * side included), the panel's updates a -= f u in column order. Row pointers
 * are swapped whole, as the reference swaps its row vectors, so each element
 * sees the reference's sequence of values. */
static int solve_aug_piv_blocked(cx **aug, int n, cx *x, int *pivots)
{
    int p, i, j;
    cx *fac = (cx *)malloc(sizeof(cx) * (size_t)(n > 0 ? n : 1) * OR_PB);
    for (p = 0; p < n; p += OR_PB) {
        const int pe = p + OR_PB < n ? p + OR_PB : n;
        int col, row;
        for (col = p; col < pe; col++) {
            int piv = col;
            double best = cx_norm(aug[col][col]);
            cx diag;
            for (row = col + 1; row < n; row++) {
                double v = cx_norm(aug[row][col]);
                if (!(v < best)) {
                    best = v;
                    piv = row;
                }
            }
            if (pivots)
                pivots[col] = piv;
            if (cx_norm(aug[piv][col]) < 1e-30) {
                free(fac);
                return OR_SINGULAR;
            }
            if (piv != col) {
                cx *t = aug[col];
                aug[col] = aug[piv];
                aug[piv] = t;
                /* the factors of earlier panel columns move with their rows */
                for (j = 0; j < col - p; j++) {
                    cx f = fac[(size_t)col * OR_PB + j];
                    fac[(size_t)col * OR_PB + j] = fac[(size_t)piv * OR_PB + j];
                    fac[(size_t)piv * OR_PB + j] = f;
                }
            }
            diag = aug[col][col];
#pragma omp parallel for num_threads(or_threads) schedule(static) if (or_threads > 1 && n - col > 256)
            for (row = col + 1; row < n; row++) {
                cx factor = cx_div(aug[row][col], diag);
                cx *ar = aug[row], *ac = aug[col];
                int kk;
                fac[(size_t)row * OR_PB + (col - p)] = factor;
                for (kk = col; kk < pe; kk++) {
                    cx sub = cx_mul(factor, ac[kk]);
                    ar[kk] = cx_sub(ar[kk], sub);
                }
            }
        }
        /* columns right of the panel: row r takes the updates of the panel
         * columns below it, in order; a pivot row is final before any row
         * below it reads it (rows in ascending order within the panel) */
        for (row = p + 1; row < pe; row++) {
            const int ce = row < pe ? row : pe;
            for (col = p; col < ce; col++) {
                const cx factor = fac[(size_t)row * OR_PB + (col - p)];
                cx *ar = aug[row], *ac = aug[col];
                int kk;
                for (kk = pe; kk <= n; kk++) {
                    cx sub = cx_mul(factor, ac[kk]);
                    ar[kk] = cx_sub(ar[kk], sub);
                }
            }
        }
#pragma omp parallel for num_threads(or_threads) schedule(dynamic, 16) if (or_threads > 1)
        for (row = pe; row < n; row++) {
            cx *ar = aug[row];
            int c2, kk;
            for (c2 = p; c2 < pe; c2++) {
                const cx factor = fac[(size_t)row * OR_PB + (c2 - p)];
                const cx *ac = aug[c2];
                for (kk = pe; kk <= n; kk++) {
                    cx sub = cx_mul(factor, ac[kk]);
                    ar[kk] = cx_sub(ar[kk], sub);
                }
            }
        }
    }
    free(fac);
    for (i = n - 1; i >= 0; i--) {
        cx sum = aug[i][n];
        for (j = i + 1; j < n; j++)
            sum = cx_sub(sum, cx_mul(aug[i][j], x[j]));
        x[i] = cx_div(sum, aug[i][i]);
    }
    return OR_OK;
}

static int solve_aug(cx **aug, int n, cx *x)
{
    return or_blocked ? solve_aug_piv_blocked(aug, n, x, NULL) : solve_aug_piv(aug, n, x, NULL);
}

/* Iterator::max_by with partial_cmp(...).unwrap_or(Equal): the fold keeps
 * the later element unless the earlier one is strictly greater. The loop
 * above implements "replace when !(v < best)". For NaN, partial_cmp returns
 * None -> Equal -> later element kept, which `!(v < best)` also does. */

int or_solve_square_piv(const cx *a, const cx *b, int n, cx *x, int *piv)
{
    cx **aug = rows_alloc(n, n + 1);
    int i, st;
    for (i = 0; i < n; i++) {
        memcpy(aug[i], a + (size_t)i * (size_t)n, sizeof(cx) * (size_t)n);
        aug[i][n] = b[i];
    }
    st = or_blocked ? solve_aug_piv_blocked(aug, n, x, piv) : solve_aug_piv(aug, n, x, piv);
    rows_free(aug, n);
    return st;
}

int or_solve_square(const cx *a, const cx *b, int n, cx *x) { return or_solve_square_piv(a, b, n, x, NULL); }

/* Build M and y of the augmented Hallén system (linear.rs:688-762). */
static int build_m(const cx *z, int n, const double *cosv, const cx *rhs, const int *wf,
                   const int *wl, int nw, const int *ja, const int *jb, const double *js, int nj,
                   cx ***m_out, cx **y_out, int *rows_out, int *cols_out)
{
    int fb_f = 0, fb_l = n - 1;
    const int *ef = wf, *el = wl;
    int w = nw, i, r, c, free_cnt = 0, rows, cols, crow;
    int *row_wire;
    unsigned char *in_junc;
    cx **m, *y;
    if (nw == 0 || n == 0) {
        ef = &fb_f;
        el = &fb_l;
        w = n > 0 ? 1 : 0;
    }
    in_junc = (unsigned char *)calloc((size_t)(n > 0 ? n : 1), 1);
    for (i = 0; i < nj; i++) {
        in_junc[ja[i]] = 1;
        in_junc[jb[i]] = 1;
    }
    for (i = 0; i < w; i++) {
        if (!in_junc[ef[i]])
            free_cnt++;
        if (!in_junc[el[i]])
            free_cnt++;
    }
    rows = n + free_cnt + nj;
    cols = n + w;
    m = rows_alloc(rows, cols);
    y = (cx *)calloc((size_t)(rows > 0 ? rows : 1), sizeof(cx));
    row_wire = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    for (i = 0; i < w; i++)
        for (r = ef[i]; r <= el[i] && r < n; r++)
            row_wire[r] = i;
    for (r = 0; r < n; r++) {
        for (c = 0; c < n; c++)
            m[r][c] = z[(size_t)r * (size_t)n + (size_t)c];
        m[r][n + row_wire[r]] = cx_make(-cosv[r], 0.0);
        y[r] = rhs[r];
    }
    crow = n;
    for (i = 0; i < w; i++) {
        if (!in_junc[ef[i]])
            m[crow++][ef[i]] = cx_make(1.0, 0.0);
        if (!in_junc[el[i]])
            m[crow++][el[i]] = cx_make(1.0, 0.0);
    }
    for (i = 0; i < nj; i++) {
        m[crow][ja[i]] = cx_make(1.0, 0.0);
        m[crow][jb[i]] = cx_make(js[i], 0.0);
        crow++;
    }
    free(in_junc);
    free(row_wire);
    *m_out = m;
    *y_out = y;
    *rows_out = rows;
    *cols_out = cols;
    return OR_OK;
}

/* The Gram part of normal_eqs below, cache-blocked: output tiles of OR_TB x
 * OR_TB accumulate the rows of M in blocks of OR_RB, blocks in ascending
 * order, so every ata[i][j] is the same sequential sum over ascending r. */
static void gram_blocked(cx **m, int rows, int cols, cx **ata)
{
    const int nt = (cols + OR_TB - 1) / OR_TB;
    int r0, i;
    for (i = 0; i < cols; i++) {
        int j;
        for (j = 0; j < cols; j++)
            ata[i][j] = cx_make(0.0, 0.0);
    }
    for (r0 = 0; r0 < rows; r0 += OR_RB) {
        const int r1 = r0 + OR_RB < rows ? r0 + OR_RB : rows;
        long t;
#pragma omp parallel for num_threads(or_threads) schedule(dynamic, 1) if (or_threads > 1)
        for (t = 0; t < (long)nt * nt; t++) {
            const int i0 = (int)(t / nt) * OR_TB, j0 = (int)(t % nt) * OR_TB;
            const int i1 = i0 + OR_TB < cols ? i0 + OR_TB : cols, j1 = j0 + OR_TB < cols ? j0 + OR_TB : cols;
            int r, ii, jj;
            for (r = r0; r < r1; r++) {
                const cx *mr = m[r];
                for (ii = i0; ii < i1; ii++) {
                    const cx a = cx_conj(mr[ii]);
                    cx *out = ata[ii];
                    for (jj = j0; jj < j1; jj++)
                        out[jj] = cx_add(out[jj], cx_mul(a, mr[jj]));
                }
            }
        }
    }
}

/* Regularised normal equations (linear.rs:764-785). Returns aug rows (S x S+1). */
static cx **normal_eqs(cx **m, const cx *y, int rows, int cols)
{
    cx **ata = rows_alloc(cols, cols + 1);
    int i;
    if (or_blocked) {
        gram_blocked(m, rows, cols, ata);
#pragma omp parallel for num_threads(or_threads) schedule(dynamic, 16) if (or_threads > 1)
        for (i = 0; i < cols; i++) {
            cx sum = cx_make(0.0, 0.0);
            int r;
            for (r = 0; r < rows; r++)
                sum = cx_add(sum, cx_mul(cx_conj(m[r][i]), y[r]));
            ata[i][cols] = sum;
        }
        for (i = 0; i < cols; i++)
            ata[i][i] = cx_add(ata[i][i], cx_make(1e-8, 0.0));
        return ata;
    }
#pragma omp parallel for num_threads(or_threads) schedule(dynamic, 4) if (or_threads > 1)
    for (i = 0; i < cols; i++) {
        cx sum;
        int j, r;
        for (j = 0; j < cols; j++) {
            sum = cx_make(0.0, 0.0);
            for (r = 0; r < rows; r++)
                sum = cx_add(sum, cx_mul(cx_conj(m[r][i]), m[r][j]));
            ata[i][j] = sum;
        }
        sum = cx_make(0.0, 0.0);
        for (r = 0; r < rows; r++)
            sum = cx_add(sum, cx_mul(cx_conj(m[r][i]), y[r]));
        ata[i][cols] = sum; /* aty, appended as the augmented column */
    }
    for (i = 0; i < cols; i++)
        ata[i][i] = cx_add(ata[i][i], cx_make(1e-8, 0.0));
    return ata;
}

int or_hallen_normal_eqs(const cx *z, int n, const double *cosv, const cx *rhs, const int *wf,
                         const int *wl, int nw, const int *ja, const int *jb, const double *js,
                         int nj, cx *ata_out, cx *aty_out, int *rows_out)
{
    cx **m, *y, **ata;
    int rows, cols, i;
    build_m(z, n, cosv, rhs, wf, wl, nw, ja, jb, js, nj, &m, &y, &rows, &cols);
    ata = normal_eqs(m, y, rows, cols);
    for (i = 0; i < cols; i++) {
        if (ata_out)
            memcpy(ata_out + (size_t)i * (size_t)cols, ata[i], sizeof(cx) * (size_t)cols);
        if (aty_out)
            aty_out[i] = ata[i][cols];
    }
    if (rows_out)
        *rows_out = rows;
    rows_free(m, rows);
    rows_free(ata, cols);
    free(y);
    return cols;
}

int or_solve_hallen(const cx *z, int n, const cx *rhs, const double *cosv, const int *wf,
                    const int *wl, int nw, const int *ja, const int *jb, const double *js, int nj,
                    cx *x)
{
    cx **m, *y, **ata;
    int rows, cols, st;
    build_m(z, n, cosv, rhs, wf, wl, nw, ja, jb, js, nj, &m, &y, &rows, &cols);
    ata = normal_eqs(m, y, rows, cols);
    rows_free(m, rows);
    free(y);
    st = solve_aug(ata, cols, x);
    rows_free(ata, cols);
    return st;
}

int or_solve_hallen_paths(const cx *z, int n, const cx *rhs, const double *cosv,
                          const int *path_of, const int *free_ends, int nfree, cx *x)
{
    int np = 0, i, r, c, rows, cols, st;
    cx **m, *y, **ata;
    for (i = 0; i < n; i++)
        if (path_of[i] + 1 > np)
            np = path_of[i] + 1;
    rows = n + nfree;
    cols = n + np;
    m = rows_alloc(rows, cols);
    y = (cx *)calloc((size_t)(rows > 0 ? rows : 1), sizeof(cx));
    for (r = 0; r < n; r++) {
        for (c = 0; c < n; c++)
            m[r][c] = z[(size_t)r * (size_t)n + (size_t)c];
        m[r][n + path_of[r]] = cx_make(-cosv[r], 0.0);
        y[r] = rhs[r];
    }
    for (i = 0; i < nfree; i++)
        m[n + i][free_ends[i]] = cx_make(1.0, 0.0);
    ata = normal_eqs(m, y, rows, cols);
    rows_free(m, rows);
    free(y);
    st = solve_aug(ata, cols, x);
    rows_free(ata, cols);
    return st;
}

/* Shared tail of the variant solvers below: the regularised normal equations
 * of the dense M, y (linear.rs:964-984 and its copies) and the elimination. */
static int solve_dense(cx **m, cx *y, int rows, int cols, cx *x)
{
    cx **ata = normal_eqs(m, y, rows, cols);
    int st;
    rows_free(m, rows);
    free(y);
    st = solve_aug(ata, cols, x);
    rows_free(ata, cols);
    return st;
}

/* solve_hallen_planewave (linear.rs:910-985): two homogeneous columns (cos,
 * sin) per wire, I = 0 at both ends of each wire. x: n + 2 w (w = max(nw, 1)). */
int or_solve_hallen_planewave(const cx *z, int n, const cx *rhs, const double *cosv, const double *sinv,
                              const int *wf, const int *wl, int nw, cx *x)
{
    int fb0 = 0, lb0 = n - 1, w, rows, cols, r, c, wi;
    const int *f = wf, *l = wl;
    int *row_wire;
    cx **m, *y;
    if (nw == 0 || n == 0) {
        f = &fb0;
        l = &lb0;
        nw = n > 0 ? 1 : 0;
    }
    w = nw;
    rows = n + 2 * w;
    cols = n + 2 * w;
    m = rows_alloc(rows, cols);
    y = (cx *)calloc((size_t)(rows > 0 ? rows : 1), sizeof(cx));
    row_wire = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    for (wi = 0; wi < w; wi++)
        for (r = f[wi]; r <= l[wi]; r++)
            row_wire[r] = wi;
    for (r = 0; r < n; r++) {
        for (c = 0; c < n; c++)
            m[r][c] = z[(size_t)r * (size_t)n + (size_t)c];
        m[r][n + 2 * row_wire[r]] = cx_make(-cosv[r], 0.0);
        m[r][n + 2 * row_wire[r] + 1] = cx_make(-sinv[r], 0.0);
        y[r] = rhs[r];
    }
    for (wi = 0; wi < w; wi++) {
        m[n + 2 * wi][f[wi]] = cx_make(1.0, 0.0);
        m[n + 2 * wi + 1][l[wi]] = cx_make(1.0, 0.0);
    }
    free(row_wire);
    return solve_dense(m, y, rows, cols, x);
}

/* solve_hallen_planewave_paths (linear.rs:1017-1091): two homogeneous columns
 * per conductor path, I = 0 at the listed free ends. x: n + 2 P. */
int or_solve_hallen_planewave_paths(const cx *z, int n, const cx *rhs, const double *cosv,
                                    const double *sinv, const int *path_of, const int *free_ends,
                                    int nfree, cx *x)
{
    int np = 0, i, r, c, rows, cols;
    cx **m, *y;
    for (i = 0; i < n; i++)
        if (path_of[i] + 1 > np)
            np = path_of[i] + 1;
    cols = n + 2 * np;
    rows = n + nfree;
    m = rows_alloc(rows, cols);
    y = (cx *)calloc((size_t)(rows > 0 ? rows : 1), sizeof(cx));
    for (r = 0; r < n; r++) {
        for (c = 0; c < n; c++)
            m[r][c] = z[(size_t)r * (size_t)n + (size_t)c];
        m[r][n + 2 * path_of[r]] = cx_make(-cosv[r], 0.0);
        m[r][n + 2 * path_of[r] + 1] = cx_make(-sinv[r], 0.0);
        y[r] = rhs[r];
    }
    for (i = 0; i < nfree; i++)
        m[n + i][free_ends[i]] = cx_make(1.0, 0.0);
    return solve_dense(m, y, rows, cols, x);
}

/* solve_hallen_current_source (linear.rs:1124-1210): one homogeneous column
 * per wire plus the port voltage V (column n + w), I = 0 at both ends of each
 * wire and I[src] = i0, those rows weighted by 1e6. x: n + w + 1. */
int or_solve_hallen_current_source(const cx *z, int n, const cx *shape, const double *cosv, int src,
                                   cx i0, const int *wf, const int *wl, int nw, cx *x)
{
    int fb0 = 0, lb0 = n - 1, w, rows, cols, vcol, r, c, wi, crow;
    const int *f = wf, *l = wl;
    int *row_wire;
    const cx cw = cx_make(1.0e6, 0.0);
    cx **m, *y;
    if (nw == 0 || n == 0) {
        f = &fb0;
        l = &lb0;
        nw = n > 0 ? 1 : 0;
    }
    w = nw;
    vcol = n + w;
    cols = n + w + 1;
    rows = n + 2 * w + 1;
    m = rows_alloc(rows, cols);
    y = (cx *)calloc((size_t)rows, sizeof(cx));
    row_wire = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    for (wi = 0; wi < w; wi++)
        for (r = f[wi]; r <= l[wi]; r++)
            row_wire[r] = wi;
    for (r = 0; r < n; r++) {
        for (c = 0; c < n; c++)
            m[r][c] = z[(size_t)r * (size_t)n + (size_t)c];
        m[r][n + row_wire[r]] = cx_make(-cosv[r], 0.0);
        m[r][vcol] = cx_neg(shape[r]);
    }
    crow = n;
    for (wi = 0; wi < w; wi++) {
        m[crow++][f[wi]] = cw;
        m[crow++][l[wi]] = cw;
    }
    m[crow][src] = cw;
    y[crow] = cx_mul(cw, i0);
    free(row_wire);
    return solve_dense(m, y, rows, cols, x);
}

/* solve_hallen_current_source_paths (linear.rs:1241-1319): one homogeneous
 * column per path plus V (column n + P), weighted I = 0 rows at the free ends
 * and the weighted source row. x: n + P + 1. */
int or_solve_hallen_current_source_paths(const cx *z, int n, const cx *shape, const double *cosv,
                                         int src, cx i0, const int *path_of, const int *free_ends,
                                         int nfree, cx *x)
{
    int np = 0, i, r, c, rows, cols, vcol, crow;
    const cx cw = cx_make(1.0e6, 0.0);
    cx **m, *y;
    for (i = 0; i < n; i++)
        if (path_of[i] + 1 > np)
            np = path_of[i] + 1;
    vcol = n + np;
    cols = n + np + 1;
    rows = n + nfree + 1;
    m = rows_alloc(rows, cols);
    y = (cx *)calloc((size_t)rows, sizeof(cx));
    for (r = 0; r < n; r++) {
        for (c = 0; c < n; c++)
            m[r][c] = z[(size_t)r * (size_t)n + (size_t)c];
        m[r][n + path_of[r]] = cx_make(-cosv[r], 0.0);
        m[r][vcol] = cx_neg(shape[r]);
    }
    crow = n;
    for (i = 0; i < nfree; i++)
        m[crow++][free_ends[i]] = cw;
    m[crow][src] = cw;
    y[crow] = cx_mul(cw, i0);
    return solve_dense(m, y, rows, cols, x);
}
