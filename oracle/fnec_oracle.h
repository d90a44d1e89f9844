/*
 * ORACLE — test infrastructure only.
 *
 * A CPU restatement, in C, of the reference's f64 Hallén path
 * (crates/nec_solver: geometry.rs, excitation.rs, matrix.rs, linear.rs,
 * farfield.rs; apps/nec-cli/src/solve_session.rs; crates/nec_worker/src/solve.rs;
 * crates/nec_parser/src/lib.rs; crates/nec_report/src/lib.rs). It follows the
 * reference's operation order, loop order and data layout (row vectors for M,
 * the Gram matrix and the augmented elimination matrix) so that it serves both
 * as the parity checker for the HIP library and as the timed CPU baseline.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product library never links, loads or calls anything here.
 *
 * Pinning (tests/test_oracle.py): the reference's golden corpus and README
 * excerpt to the printed digits, its unit-test closed forms, mpmath
 * evaluations of the fill formula and of the regularised normal equations,
 * an independently assembled Gram matrix and KCL system, the Hertzian-dipole
 * and PEC-image closed forms, the split-dipole invariant of the conductor-path
 * route, and independent evaluations of the residual / diagonal spread.
 * Parity unpinned: or_radiation_efficiency (the GN 0/2 gain correction of
 * farfield.rs:308-322); no reference fixture or closed form covers it.
 */
#ifndef FNEC_ORACLE_H
#define FNEC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "cx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Segment (geometry.rs:115-134). */
typedef struct {
    uint32_t tag;
    uint32_t tag_index;
    double start[3];
    double end[3];
    double mid[3];
    double dir[3];
    double length;
    double radius;
} or_seg;

/* Ground model (geometry.rs:24-54). */
enum { OR_GROUND_FREE = 0, OR_GROUND_PEC = 1, OR_GROUND_FINITE = 2, OR_GROUND_DEFERRED = 3 };
typedef struct {
    int kind;
    double eps_r;
    double sigma;
} or_ground;

/* Status codes shared with the product ABI. */
enum {
    OR_OK = 0,
    OR_SINGULAR = 1,
    OR_DIM_MISMATCH = 2,
    OR_UNSUPPORTED = 3,
    OR_PARSE = 6,
    OR_GEOMETRY = 7,
    OR_EXCITATION = 8,
    OR_NO_FEED = 9,
};

/* ---- deck (nec_parser/src/lib.rs:84-330) ---- */
typedef struct or_deck or_deck;
or_deck *or_deck_parse(const char *text, char *err, int errlen);
void or_deck_free(or_deck *d);
int or_deck_frequencies(const or_deck *d, double *out, int max);
int or_deck_frequencies_py(const or_deck *d, double *out, int max);
int or_deck_rp_points(const or_deck *d, double *theta, double *phi, int max);
int or_deck_has_card(const or_deck *d, const char *mn);

/* ---- geometry (geometry.rs) ---- */
int or_build_geometry(const or_deck *d, or_seg *out, int max, char *err, int errlen);
or_ground or_ground_model(const or_deck *d);
int or_wire_endpoints(const or_seg *s, int n, int *first, int *last, int max);
int or_merge_collinear(const or_seg *s, int n, int *first, int *last, int max);
int or_detect_junctions(const or_seg *s, const int *first, const int *last, int nw, double tol,
                        int *ja, int *jb, double *jsign, int max);
/* Conductor paths: returns the number of paths, or -1 when the topology is out
 * of scope (geometry.rs:404-584 returns None). path_of/sign_of/s_of are per
 * segment; free_a/free_b per path; nontrivial is set when any path is not a
 * forward contiguous chain (ConductorPath::is_trivial). */
int or_conductor_paths(const or_seg *s, int n, int *path_of, double *sign_of, double *s_of,
                       int *free_a, int *free_b, int max_paths, int *nontrivial);

/* ---- excitation (excitation.rs) ---- */
int or_build_excitation(const or_deck *d, const or_seg *s, int n, cx *v);
int or_build_hallen_rhs(const or_deck *d, const or_seg *s, int n, double freq_hz, cx *rhs,
                        double *cosv);
int or_build_hallen_rhs_paths(const or_deck *d, const or_seg *s, int n, double freq_hz,
                              const int *path_of, const double *sign_of, const double *s_of,
                              cx *rhs, double *cosv);

/* ---- plane-wave and current-source inputs (planewave.rs, excitation.rs:350-447) ---- */
int or_planewave_hallen(const or_deck *d, const or_seg *s, int n, double freq_hz, cx *rhs, double *cosv,
                        double *sinv);
int or_planewave_hallen_paths(const or_deck *d, const or_seg *s, int n, double freq_hz, const int *path_of,
                              const double *sign_of, const double *s_of, cx *rhs, double *cosv, double *sinv);
int or_current_source_shape(const or_deck *d, const or_seg *s, int n, double freq_hz, uint32_t tag,
                            uint32_t seg, cx *shape, double *cosv, int *src);
int or_current_source_shape_paths(const or_deck *d, const or_seg *s, int n, double freq_hz, uint32_t tag,
                                  uint32_t seg, const int *path_of, const double *sign_of, const double *s_of,
                                  cx *shape, double *cosv, int *src);

/* ---- matrix (matrix.rs:186-218, 297-331) ---- */
void or_fill_z(const or_seg *s, int n, double freq_hz, or_ground g, cx *z /* n*n row-major */);
cx or_fill_elem(const or_seg *obs, const or_seg *src, double k, int is_self);

/* ---- LD / TL / NT stamps (loads.rs:103-223, tl.rs:47-244, network.rs:55-140) ---- */
typedef struct {
    int row, col;
    cx delta;
} or_stamp;
void or_build_loads(const or_deck *d, const or_seg *s, int n, double freq_hz, cx *loads);
int or_build_stamps(const or_deck *d, const or_seg *s, int n, double freq_hz, int with_nt, or_stamp *out,
                    int max);
void or_apply_stamps(cx *z /* n*n row-major */, int n, const or_stamp *st, int ns);
/* The stamps a deck's text gets at freq_hz under the caller semantics (CLI:
 * loads, TL, NT; worker and Python: loads, TL). Returns the count, or a
 * negative status when the deck does not parse or build. */
int or_deck_stamps(const char *text, int semantics, double freq_hz, int *rows, int *cols, double *re,
                   double *im, int max);

/* ---- linear (linear.rs) ---- */
/* Threads for the fill, Gram and elimination row loops (default 1, as the
 * reference runs; results do not depend on it). Fixture generation only. */
void or_set_threads(int n);
/* Cache-blocked Gram and elimination loops (default 0): the same operations
 * per element in the same order, bit for bit; for fixtures of large systems. */
void or_set_blocked(int on);
int or_solve_square(const cx *a, const cx *b, int n, cx *x);
/* the same, recording the pivot row chosen at each elimination column */
int or_solve_square_piv(const cx *a, const cx *b, int n, cx *x, int *piv);
int or_solve_hallen(const cx *z, int n, const cx *rhs, const double *cosv, const int *wf,
                    const int *wl, int nw, const int *ja, const int *jb, const double *js, int nj,
                    cx *x /* n + max(nw,1) */);
int or_solve_hallen_paths(const cx *z, int n, const cx *rhs, const double *cosv,
                          const int *path_of, const int *free_ends, int nfree, cx *x);
/* linear.rs:910-1319: x holds the currents then the homogeneous constants
 * (plane wave: cos, sin per wire or path) and, for a current source, V last. */
int or_solve_hallen_planewave(const cx *z, int n, const cx *rhs, const double *cosv, const double *sinv,
                              const int *wf, const int *wl, int nw, cx *x /* n + 2 max(nw,1) */);
int or_solve_hallen_planewave_paths(const cx *z, int n, const cx *rhs, const double *cosv,
                                    const double *sinv, const int *path_of, const int *free_ends,
                                    int nfree, cx *x /* n + 2 P */);
int or_solve_hallen_current_source(const cx *z, int n, const cx *shape, const double *cosv, int src,
                                   cx i0, const int *wf, const int *wl, int nw, cx *x /* n + w + 1 */);
int or_solve_hallen_current_source_paths(const cx *z, int n, const cx *shape, const double *cosv,
                                         int src, cx i0, const int *path_of, const int *free_ends,
                                         int nfree, cx *x /* n + P + 1 */);
/* The augmented system pieces, for kernel-level parity tests. Returns S. */
int or_hallen_normal_eqs(const cx *z, int n, const double *cosv, const cx *rhs, const int *wf,
                         const int *wl, int nw, const int *ja, const int *jb, const double *js,
                         int nj, cx *ata /* S*S row-major, or NULL */, cx *aty /* S, or NULL */,
                         int *rows_out);

/* ---- far field (farfield.rs) ---- */
/* out: per point {gain_total, gain_theta, gain_phi, axial_ratio}. Returns the
 * normalisation integral. */
double or_radiation_pattern(const or_seg *s, int n, const cx *cur, double freq_hz, or_ground g,
                            const double *theta, const double *phi, int np, double *out);
double or_power_integral(const or_seg *s, int n, const cx *cur, double k, or_ground g);
void or_far_field(const or_seg *s, int n, const cx *cur, double k, or_ground g, double th,
                  double ph, cx *ft, cx *fp);
double or_axial_ratio(cx et, cx ep);
double or_radiation_efficiency(const or_seg *s, int n, const cx *cur, double freq_hz,
                               or_ground g, double p_in);

/* ---- solve_session.rs helpers ---- */
double or_diag_spread(const cx *z, int n);
/* host libm over arrays: fn 0 sin(x), 1 cos(x), 2 hypot(x, y), 5 and 6 the sin
 * and cos of sincos(x) (libm.c) */
void or_libm_eval(int fn, const double *x, const double *y, size_t n, double *out);
void or_residual_hallen(const cx *z, int n, const cx *cur, const cx *chom, int nchom,
                        const double *cosv, const cx *rhs, const int *wf, const int *wl, int nw,
                        double *abs_res, double *rel_res);

/* ---- whole-deck pipelines ---- */
enum { OR_SEM_CLI = 0, OR_SEM_WORKER = 1, OR_SEM_PY = 2 };
typedef struct {
    int status;
    char err[256];
    double freq_hz;
    int n;         /* segments */
    int s;         /* unknowns */
    int r;         /* rows of M */
    int nw;        /* homogeneous columns */
    int used_paths;
    or_ground ground;
    or_seg *segs;  /* n */
    cx *x;         /* s: currents then homogeneous constants */
    int nfeed;
    int *feed_tag, *feed_seg, *feed_idx;
    cx *feed_v, *feed_i, *feed_z;
    int np;
    double *rp_theta, *rp_phi, *rp; /* rp: 4 per point */
    double rp_total;
    double abs_res, rel_res, diag_spread;
    char *report;  /* report contract v1 text (CLI semantics) */
    /* CLI: 0 delta gap, 1 plane wave (receive), 2 current source */
    int variant;
    cx port_v;     /* current source: the solved port voltage V */
    int nrecv;     /* plane-wave receive sweep (NTHETA NPHI > 1) */
    double *recv_theta, *recv_phi, *recv_db;
} or_result;

int or_run_deck(const char *text, int semantics, int freq_index, double freq_override_hz,
                int with_report, or_result *res);
void or_result_free(or_result *res);
int or_deck_variant_inputs(const char *text, double freq_hz, int *n_out, int *route, cx *rhs, double *cosv,
                           double *sinv, int *src, cx *i0, int *wf, int *wl, int *nw, int *path_of, int *free_ends,
                           int *nfree);

/* Report contract v1 number formats (nec_report/src/lib.rs:437-482). */
void or_fmt_fixed(char *buf, int cap, double v, int prec);
void or_fmt_exp(char *buf, int cap, double v, int prec);

#ifdef __cplusplus
}
#endif
#endif
