/*
 * fnec_host — the host side above the fnec_hip C ABI: a NEC deck parser subset
 * and the geometry / excitation / topology builders the reference's callers
 * run before they reach nec_accel, plus the CLI pipeline and report contract
 * v1. Mirrors (file:line in the reference):
 *   nec_parser::parse                       crates/nec_parser/src/lib.rs:84-330
 *   build_geometry (GW/GM/GR)               crates/nec_solver/src/geometry.rs:676-879
 *   ground_model_from_deck                  geometry.rs:71-111
 *   wire_endpoints_from_segs / merge        geometry.rs:168-241
 *   detect_wire_junctions                   geometry.rs:274-339
 *   build_conductor_paths                   geometry.rs:404-584
 *   build_excitation / build_hallen_rhs(_paths) crates/nec_solver/src/excitation.rs:78-341
 *   solve_frequency_point (Hallén route)    apps/nec-cli/src/solve_session.rs:1186-1775
 *   worker solve                            crates/nec_worker/src/solve.rs:150-289
 *   render_text_report                      crates/nec_report/src/lib.rs:307-482
 *   LD / TL / NT impedance stamps           crates/nec_solver/src/loads.rs,
 *                                           tl.rs, network.rs
 * Plane-wave and current-source EX (solve_session.rs:262-448) are built and
 * solved for the CLI; for the worker and the Python binding, which do not solve
 * them, fnec_deck_build returns FNEC_UNAVAILABLE.
 */
#ifndef FNEC_HOST_H
#define FNEC_HOST_H

#include "fnec_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { FNEC_SEM_CLI = 0, FNEC_SEM_WORKER = 1, FNEC_SEM_PY = 2 };

typedef struct fnec_deck fnec_deck;

/* Parse deck text. NULL on a fatal parse error (message in err). */
fnec_deck *fnec_deck_parse(const char *text, char *err, size_t errlen);
void fnec_deck_free(fnec_deck *d);

/* Frequencies in Hz: the first FR card (CLI, worker) or all FR cards (Python
 * binding). Returns the count; writes at most max values. */
size_t fnec_deck_frequencies(const fnec_deck *d, int semantics, double *out, size_t max);

/* Segment count of the built geometry, or 0 with the error in
 * fnec_hip_last_error() when the geometry cannot be built. */
size_t fnec_deck_segment_count(fnec_deck *d);

/* Hallén inputs for one frequency, with the caller semantics of
 * FNEC_SEM_CLI (merged collinear wires, cross-component junctions, or the
 * conductor-path route for bent / reversed degree-2 chains) or
 * FNEC_SEM_WORKER / FNEC_SEM_PY (raw wire ranges, all junctions). All
 * pointers stay valid until the next call on the same deck. */
typedef struct {
    int route;                 /* 0: solve_hallen, 1: solve_hallen_paths */
    fnec_request req;          /* segments, rhs, cos, wires, junctions, ground,
                                  first feed (index + voltage), RP grid */
    const size_t *path_of;     /* route 1 */
    const size_t *free_ends;
    size_t n_free;
    size_t n_feeds;            /* feedpoint rows (CLI: every non plane-wave EX) */
    const size_t *feed_idx;
    const fnec_c64 *feed_v;    /* V/dl * dl (CLI, Python) or the EX voltage (worker) */
    const uint32_t *feed_tag;
    const uint32_t *feed_seg;
    const uint32_t *seg_tag;   /* per segment tag and 1-based index */
    const uint32_t *seg_index;
    /* Plane-wave and current-source decks (CLI semantics only; the first
     * plane-wave EX card wins over current sources, solve_session.rs:
     * 1327-1337): variant 1 plane wave, 2 current source, 0 delta gap. Route
     * 0 then uses the raw wire ranges (wire_endpoints_from_segs) with no KCL
     * rows, route 1 the conductor paths. */
    int variant;
    /* variant 1: req.rhs / req.cos_vec / sin_vec are build_planewave_hallen(_paths)'s
     * at the card's incidence; rhs_all holds n_rhs forcing vectors of n, the
     * first that one, then one per incidence of the receive sweep
     * (solve_session.rs:313-371) at (recv_theta, recv_phi), theta outer. */
    const double *sin_vec;
    size_t n_rhs;
    const fnec_c64 *rhs_all;
    size_t n_recv;
    const double *recv_theta;
    const double *recv_phi;
    /* variant 2: req.rhs / req.cos_vec are build_current_source_shape(_paths)'s,
     * src the source segment and i0 the forced current of the first EX 4
     * card; feed_port[i] marks the feed rows (every EX 4 card) whose voltage
     * is the solved port voltage (solve_session.rs:893-897). */
    size_t src;
    fnec_c64 i0;
    const unsigned char *feed_port;
} fnec_deck_problem;

/* A large receive sweep (4M or more elementary operations in its forcing)
 * is formed on up to 16 short-lived host threads, joined before the call
 * returns; the values are those of the one-thread loop. */
int fnec_deck_build(fnec_deck *d, int semantics, double freq_hz, fnec_deck_problem *out);

/* Solve one built problem on the GPU: route 0 or 1 as fnec_hip_solve_hallen or
 * fnec_hip_solve_hallen_paths, with the request's LD/TL/NT stamps added to Z
 * after the fill (solve_session.rs:1285-1309). x_out holds n currents followed
 * by the route's homogeneous constants: route 0 max(n_wires, 1), route 1 one
 * per conductor path (max path_of + 1). Variant 1 solves as
 * fnec_hip_solve_hallen_planewave (x_out: n_rhs blocks of n + 2 G values, G
 * ranges or paths), variant 2 as fnec_hip_solve_hallen_current_source (n + G +
 * 1 values, the port voltage last); both need FNEC_PREC_FP64. */
int fnec_deck_solve(const fnec_deck_problem *pr, int precision, fnec_c64 *x_out,
                    fnec_solve_stats *stats);

/* CLI pipeline for one frequency of the deck on the GPU: solve, feedpoint
 * rows, currents, RP with normalisation, finite-ground efficiency, and the
 * report contract v1 text (FREQ_MHZ ... RADIATION_PATTERN ...). Writes up to
 * cap bytes (NUL-terminated) and the full length to *needed. */
int fnec_deck_report(fnec_deck *d, size_t freq_index, int precision, char *out, size_t cap,
                     size_t *needed);

#ifdef __cplusplus
}
#endif
#endif
