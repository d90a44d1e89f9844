/*
 * fnec_hip — MI355X (gfx950) backend for the Hallén method-of-moments path of
 * fnec-rust: Z-matrix fill, the lambda-regularised least-squares solve
 * (Gram matrix + blocked complex LU with partial pivoting + triangular solve),
 * feedpoint impedance / currents, and the RP far-field pattern.
 *
 * Plain C ABI: caller-owned buffers, no pointer retained after a call returns,
 * no exceptions across the boundary. Thread-safe: calls are serialised by a
 * process-wide lock onto one device stream; the device context and its
 * memory arenas are created once per process and reused.
 *
 * Status codes map onto the reference's error enums:
 *   FNEC_OK                  Ok(..)
 *   FNEC_SINGULAR            SolveError::Singular            (linear.rs:23-34)
 *   FNEC_DIM_MISMATCH        SolveError::*DimensionMismatch  (linear.rs:23-34)
 *   FNEC_UNAVAILABLE         Option::None -> caller falls back to the CPU
 *                            (nec_accel wgpu_device.rs:832,1294,461 return None)
 *   FNEC_RESOURCE_EXHAUSTED  ErrorCode::ResourceExhausted   (nec_worker protocol.rs:167-174)
 *   FNEC_INTERNAL            ErrorCode::Internal
 *   FNEC_INVALID_ARG         out-of-range index / null pointer
 */
#ifndef FNEC_HIP_H
#define FNEC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FNEC_OK = 0,
    FNEC_SINGULAR = 1,
    FNEC_DIM_MISMATCH = 2,
    FNEC_UNAVAILABLE = 3,
    FNEC_RESOURCE_EXHAUSTED = 4,
    FNEC_INTERNAL = 5,
    FNEC_INVALID_ARG = 6,
};

/* FNEC_PREC_FP64       plain: the reference's algorithm and operation order in
 *                      FP64, bit for bit (see fnec_hip_solve_hallen).
 * FNEC_PREC_MIXED      FP32 LU of Z on the matrix cores preconditioning FP64
 *                      conjugate gradients on the same normal equations.
 * FNEC_PREC_FP64_MFMA  the same FP64 normal equations formed and factored on
 *                      the FP64 matrix cores (blocked, FMA accumulation): faster
 *                      than FNEC_PREC_FP64, but its summation order differs
 *                      from the reference's, so on ill-conditioned normal
 *                      equations its results move by about cond(G) * 1e-16. */
enum { FNEC_PREC_FP64 = 0, FNEC_PREC_MIXED = 1, FNEC_PREC_FP64_MFMA = 2 };

enum { FNEC_GROUND_FREE = 0, FNEC_GROUND_PEC = 1, FNEC_GROUND_FINITE = 2 };

/* Complex double, layout-compatible with num_complex::Complex64. */
typedef struct {
    double re, im;
} fnec_c64;

/* Segment input: nec_accel::ZSegmentInput (wgpu_device.rs:811-816). */
typedef struct {
    double mid[3];
    double dir[3];
    double length;
    double radius;
} fnec_seg;

/* Ground model (nec_solver geometry.rs:24-54). FREE also covers Deferred.
 * FINITE is the scalar-Gamma image model of GN 0/2 (matrix.rs:285-290 for the
 * fill, farfield.rs:333-388 for the pattern). */
typedef struct {
    int kind;
    double eps_r;
    double sigma;
} fnec_ground;

/* One RP row: FarFieldResult gains and axial ratio (farfield.rs:42-57). */
typedef struct {
    double gain_total_dbi;
    double gain_theta_dbi;
    double gain_phi_dbi;
    double axial_ratio;
} fnec_rp_row;

/* Per-solve diagnostics: the `diag:` values (solve_session.rs:118-207) plus
 * timings in microseconds measured with HIP events on the solve's stream. */
typedef struct {
    double abs_res;        /* ||Z I - C cos - rhs||_2          */
    double rel_res;        /* abs_res / ||rhs||_2              */
    double diag_spread;    /* max|Z_ii| / min|Z_ii|            */
    double backward_error; /* mixed: final normwise backward error of the LS normal equations */
    int refine_iters;      /* mixed: refinement iterations used */
    int n_swaps;           /* row interchanges taken by partial pivoting */
    int exact_panels;      /* LU panels that needed the interchange path */
    int rows;              /* R = N + constraint rows          */
    int unknowns;          /* S = N + homogeneous columns      */
    double t_fill_us, t_gram_us, t_lu_us, t_solve_us, t_total_us;
} fnec_solve_stats;

/* ---- device probes ----
 * Replaces enumerate_compute_adapters / run_noop_compute_pipeline
 * (wgpu_device.rs:24,67): number of usable gfx950 devices (0 => callers use
 * the CPU path), and an explicit one-time context initialisation. */
int fnec_hip_device_count(void);
int fnec_hip_init(void);
const char *fnec_hip_last_error(void);

/* ---- Z-matrix fill ----
 * Replaces fill_zmatrix_wgpu (wgpu_device.rs:832), in FP64 and with the PEC /
 * scalar-Gamma image term (matrix.rs:186-218). z_out: n*n row-major,
 * z_out[i*n+j] = A[i][j] (observation i, source j). */
int fnec_hip_fill_z(const fnec_seg *segs, size_t n, double freq_hz, const fnec_ground *ground,
                    fnec_c64 *z_out);

/* ---- fused fill -> Gram -> LU -> triangular solve ----
 * Replaces solve_hallen_gpu_resident (wgpu_device.rs:1294) and reproduces
 * nec_solver::solve_hallen (linear.rs:672-797): same augmented system (one
 * homogeneous column per wire range, I=0 rows at free wire ends, KCL rows
 * I[a] + sign I[b] = 0), same normal equations with lambda = 1e-8, solved by
 * LU with partial pivoting. Z never leaves the device.
 * n_wires == 0 means one range (0, n-1), as in the reference.
 * x_out: n + max(n_wires,1) values (currents, then one constant per wire).
 * precision: FNEC_PREC_FP64 (plain: the fill, the normal equations and their
 * Gaussian elimination in the reference's operation order, with the host
 * libm's sin / cos / hypot reproduced on the device, so Z, the normal matrix
 * and its factors equal the reference CPU path's bit for bit; the back
 * substitution runs blocked, about 1e-14 relative from the reference's
 * ascending-j sums, or in their order with the tunable "exact_backsub"),
 * FNEC_PREC_FP64_MFMA (the same system on the FP64 matrix cores) or
 * FNEC_PREC_MIXED (an FP32 LU of Z on the
 * matrix cores, its trailing updates on f16 hi + lo operand pairs by default,
 * used to precondition FP64 conjugate gradients on the same regularised
 * normal equations, converged to a relative residual of 1e-14; stats report
 * the iterations and the normwise backward error). stats may be NULL. */
int fnec_hip_solve_hallen(const fnec_seg *segs, size_t n, const fnec_c64 *rhs, const double *cos_vec,
                          const size_t *wire_first, const size_t *wire_last, size_t n_wires,
                          const size_t *j_a, const size_t *j_b, const double *j_sign, size_t n_junc,
                          double freq_hz, const fnec_ground *ground, int precision, fnec_c64 *x_out,
                          fnec_solve_stats *stats);

/* ---- conductor-path variant ----
 * Reproduces nec_solver::solve_hallen_paths (linear.rs:823-894), the CLI's
 * route for bent / reversed degree-2 junctions (solve_session.rs:1338-1378):
 * one homogeneous column per path (path_of[m] in 0..P-1), I=0 rows at the
 * listed free ends, no KCL rows. x_out: n + P values. */
int fnec_hip_solve_hallen_paths(const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                                const double *cos_vec, const size_t *path_of,
                                const size_t *free_ends, size_t n_free, double freq_hz,
                                const fnec_ground *ground, int precision, fnec_c64 *x_out,
                                fnec_solve_stats *stats);

/* ---- RP far field ----
 * Replaces run_rp_farfield_batch_wgpu (wgpu_device.rs:461) together with the
 * normalisation integral the reference computes on the CPU
 * (integrate_power_for_ground, farfield.rs:422-464, called from
 * compute_radiation_pattern, farfield.rs:563-647): same unit vectors, image
 * conventions, trapezoid weights, -999.99 sentinels and Stokes axial ratio.
 * total_out (may be NULL) receives the normalisation integral. */
int fnec_hip_rp(const fnec_seg *segs, const fnec_c64 *currents, size_t n, double freq_hz,
                const fnec_ground *ground, const double *theta_deg, const double *phi_deg,
                size_t npts, fnec_rp_row *out, double *total_out);

/* ---- RP far field without device allocation ----
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the same
 * computation and results as fnec_hip_rp, with every device buffer of the
 * job (inputs, fields, partial sums of the normalisation integral, totals,
 * rows, work list) in a fixed 32 MiB array each device holds from library
 * initialisation until the process ends. It calls no hipMalloc, hipFree or
 * pinned-memory allocation; inputs and outputs are copied through pageable
 * host memory. A job that needs more than 32 MiB (80 B per segment and 80 B
 * per pattern point, plus about 4 KB) returns FNEC_RESOURCE_EXHAUSTED with
 * the sizes in fnec_hip_last_error(). Calls share the array: the
 * process-wide lock of this ABI serialises them. */
int fnec_hip_rp_noalloc(const fnec_seg *segs, const fnec_c64 *currents, size_t n, double freq_hz,
                        const fnec_ground *ground, const double *theta_deg, const double *phi_deg,
                        size_t npts, fnec_rp_row *out, double *total_out);
/* One Z-matrix stamp, Z[row][col] += delta (ZMatrix::add_to_entry /
 * add_to_diagonal, matrix.rs:131-147; row = observation segment, col =
 * source segment): the LD loads (loads.rs:103-206, one diagonal stamp per
 * loaded segment holding the sum of its loads), the TL two-port stamps
 * (tl.rs:47-184) and the NT network stamps (network.rs:55-140). */
typedef struct {
    size_t row, col;
    fnec_c64 delta;
} fnec_stamp;

/* ---- plane-wave (receive) variant ----
 * Reproduces nec_solver::solve_hallen_planewave (linear.rs:910-985; wire
 * ranges as in fnec_hip_solve_hallen, n_wires == 0 meaning one range) or, when
 * path_of is not NULL, solve_hallen_planewave_paths (linear.rs:1017-1091; the
 * CLI's conductor paths, solve_session.rs:262-301): two homogeneous columns
 * (-cos, -sin) per range or path, I = 0 rows at both ends of every range or at
 * the listed free ends, no KCL rows; inputs as build_planewave_hallen(_paths)
 * returns them (planewave.rs:153-347). n_rhs incidences share the matrix: rhs
 * holds n_rhs forcing vectors of n (incidence k at rhs + k n) and one
 * elimination of the normal matrix serves them all, as the receive sweep
 * (solve_session.rs:313-371) solves every incidence on the same Z; each
 * incidence's result is bitwise what a solve of it alone gives. x_out: n_rhs
 * blocks of n + 2 G values (currents, then the cos and sin constants of each
 * range or path in turn). stamps (may be NULL) are added to Z after the fill as
 * fnec_request.stamps are. precision: FNEC_PREC_FP64 only (the plain path;
 * FNEC_UNAVAILABLE otherwise). */
int fnec_hip_solve_hallen_planewave(const fnec_seg *segs, size_t n, const fnec_c64 *rhs, size_t n_rhs,
                                    const double *cos_vec, const double *sin_vec,
                                    const size_t *wire_first, const size_t *wire_last, size_t n_wires,
                                    const size_t *path_of, const size_t *free_ends, size_t n_free,
                                    const fnec_stamp *stamps, size_t n_stamps, double freq_hz,
                                    const fnec_ground *ground, int precision, fnec_c64 *x_out,
                                    fnec_solve_stats *stats);

/* ---- current-source variant ----
 * Reproduces nec_solver::solve_hallen_current_source (linear.rs:1124-1210) or,
 * when path_of is not NULL, solve_hallen_current_source_paths (linear.rs:
 * 1241-1319): one homogeneous column (-cos) per range or path, the
 * port-voltage column -shape, and rows weighted by 1e6: I = 0 at both ends of
 * every range (or at the listed free ends), then I[src] = i0. shape and cos_vec
 * are build_current_source_shape(_paths)'s (excitation.rs:350-447). x_out:
 * n + G + 1 values: currents, the G constants, and the port voltage V
 * (Z_in = V / i0, solve_session.rs:893-897). precision: FNEC_PREC_FP64 only. */
int fnec_hip_solve_hallen_current_source(const fnec_seg *segs, size_t n, const fnec_c64 *shape,
                                         const double *cos_vec, size_t src, fnec_c64 i0,
                                         const size_t *wire_first, const size_t *wire_last,
                                         size_t n_wires, const size_t *path_of, const size_t *free_ends,
                                         size_t n_free, const fnec_stamp *stamps, size_t n_stamps,
                                         double freq_hz, const fnec_ground *ground, int precision,
                                         fnec_c64 *x_out, fnec_solve_stats *stats);

/* ---- batched solves (sweeps and many geometries in one submission) ----
 * One request = one Hallén solve (+ optional RP grid). Arrays are borrowed for
 * the duration of the call. */
typedef struct {
    const fnec_seg *segs;
    size_t n;
    const fnec_c64 *rhs;
    const double *cos_vec;
    const size_t *wire_first;
    const size_t *wire_last;
    size_t n_wires;
    const size_t *j_a;
    const size_t *j_b;
    const double *j_sign;
    size_t n_junc;
    double freq_hz;
    fnec_ground ground;
    size_t feed_index;       /* segment whose current gives Z_in = v_feed / I */
    fnec_c64 v_feed;         /* feed voltage (worker semantics: the EX voltage) */
    const double *rp_theta;  /* optional RP grid (NULL: none) */
    const double *rp_phi;
    size_t rp_npts;
    uint64_t tag;            /* caller's request id, echoed in the result */
    /* Optional output buffers (NULL: none): the currents and constants
     * (n + max(n_wires, 1)) and the RP rows (rp_npts). A session writes them
     * before the request's result is polled, so they must stay valid until
     * then; a batch call uses them where the result's x / rp are NULL. */
    fnec_c64 *x_out;
    fnec_rp_row *rp_out;
    /* Optional Z stamps (NULL / 0: none), added in this order to the filled Z
     * before anything reads it, as the reference adds the LD, TL and NT
     * stamps after assembling Z (solve_session.rs:1285-1309, nec_worker
     * solve.rs:179-185): the solve, the residual and the diagonal spread all
     * see the stamped matrix. row and col must be below n. */
    const fnec_stamp *stamps;
    size_t n_stamps;
} fnec_request;

typedef struct {
    uint64_t tag;
    int status;
    fnec_c64 z_in;
    fnec_c64 feed_current;
    fnec_c64 *x;             /* caller buffer of n + max(n_wires,1), or NULL */
    fnec_rp_row *rp;         /* caller buffer of rp_npts rows, or NULL */
    double rp_total;
    /* session clock (steady, microseconds): submit, start of the request's
     * batch (taken off the queue), and completion of that batch */
    double t_submit_us, t_start_us, t_done_us;
    uint64_t batch;          /* session batch the request ran in (0, 1, ...) */
    fnec_solve_stats stats;
} fnec_result;

/* Solve every request in one batched submission: each kernel launch walks a
 * work list of (problem, tile) items over the whole batch, on one engine lane
 * (its solve stream plus the LU's look-ahead stream). results[i] corresponds
 * to reqs[i]; a result's x / rp (or, where NULL, the request's x_out / rp_out)
 * receive the currents and the pattern rows. When the call returns non-zero,
 * every request it admitted carries that status (its outputs are not valid);
 * requests refused before admission carry their own. */
int fnec_hip_solve_batch(const fnec_request *reqs, size_t count, int precision, fnec_result *results);

/* ---- serving session (continuous batching, nec_worker serving mode) ----
 * The session's worker threads (one per engine lane, tunable "lanes", default
 * 2) take queued requests as batches and solve them; each batch runs on a lane
 * of its own (streams, device arena), so a batch formed from requests that
 * arrived while another is on the device runs beside it and their kernels
 * overlap. Requests are admitted in arrival order (within a size class, see
 * the tunable "batch_split_n") while the device footprint of all batches in
 * flight stays within budget_bytes; a batch holds Z, the LU
 * factors and the currents on the device until it completes. A batch closes
 * when no request has been submitted for 500 us, when it holds max_batch
 * requests, or when its oldest request has waited 20 ms. submit copies the
 * request's inputs and returns at once; poll returns finished results (up to
 * max), waiting up to wait_us for the first one while requests are queued or
 * running; pending counts queued, running and unpolled requests. A result's
 * status is that of its request; after a failed one, fnec_hip_last_error()
 * on the polling thread gives the batch's message. submit, poll and pending
 * may be called from several threads at once; destroy waits for the batches
 * on the device, drops queued requests, and must not overlap any other call
 * on the same session. */
typedef struct fnec_session fnec_session;
fnec_session *fnec_hip_session_create(size_t budget_bytes, int precision, int max_batch);
int fnec_hip_session_submit(fnec_session *s, const fnec_request *req);
int fnec_hip_session_poll(fnec_session *s, fnec_result *out, int max, int wait_us);
int fnec_hip_session_pending(fnec_session *s);
void fnec_hip_session_destroy(fnec_session *s);
/* Device bytes a request needs while resident (the admission unit). */
size_t fnec_hip_request_footprint(size_t n, size_t n_wires, size_t n_junc, size_t rp_npts, int precision);

/* ---- diagnostics (no counterpart in the reference) ----
 * Kernel-level parity hooks for the test suite: the device-assembled
 * augmented normal matrix [G | M^H y] of a solve_hallen problem (S x (S+1)
 * row-major), and a probe of the FP64 MFMA lane layout (A 16x4 and B 4x16
 * row-major in, 64 lanes x 4 accumulator values out). */
int fnec_hip_debug_normal_eqs(const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                              const double *cos_vec, const size_t *wire_first,
                              const size_t *wire_last, size_t n_wires, const size_t *j_a,
                              const size_t *j_b, const double *j_sign, size_t n_junc,
                              double freq_hz, const fnec_ground *ground, fnec_c64 *aug_out);
/* The mixed path's FP32 LU of Z for one solve_hallen problem, with the
 * current tunables: the packed factors L\U (unit L below the diagonal; n x n
 * row-major, real and imaginary parts interleaved) and ipiv (row i was
 * interchanged with row ipiv[i], in order of i), so that P Z32 = L U for the
 * FP32 copy Z32 of the fill. */
int fnec_hip_debug_lu_f32(const fnec_seg *segs, size_t n, const fnec_c64 *rhs, const double *cos_vec,
                          const size_t *wire_first, const size_t *wire_last, size_t n_wires,
                          const size_t *j_a, const size_t *j_b, const double *j_sign, size_t n_junc,
                          double freq_hz, const fnec_ground *ground, float *lu_out, int32_t *ipiv_out);
int fnec_hip_debug_mfma_probe(const double *a16x4, const double *b4x16, double *lanes_out);
/* The FP64 Z of the mixed and FP64-MFMA paths' fill (k_fill_z), on
 * fnec_hip_fill_z's inputs and layout; fnec_hip_fill_z itself returns the
 * plain path's fill, the reference's Z bit for bit. */
int fnec_hip_debug_fill_fast(const fnec_seg *segs, size_t n, double freq_hz, const fnec_ground *ground,
                             fnec_c64 *z_out);
/* The plain path (FNEC_PREC_FP64) of fnec_hip_solve_hallen on a caller-supplied
 * Z (n x n row-major, z[i*n+j] = A[i][j]) instead of the device fill: normal
 * equations, elimination, back substitution and residual. Parity hook: on the
 * oracle's Z it isolates the solve stages from the fill. */
int fnec_hip_debug_solve_z(const fnec_c64 *z, size_t n, const fnec_c64 *rhs, const double *cos_vec,
                           const size_t *wire_first, const size_t *wire_last, size_t n_wires,
                           const size_t *j_a, const size_t *j_b, const double *j_sign, size_t n_junc,
                           fnec_c64 *x_out, fnec_solve_stats *stats);
/* The plain path's elimination and back substitution on a caller-supplied
 * square system a x = b (a: n x n row-major), as solve_square_in_place does
 * it (linear.rs:1321-1383): ipiv_out[c] (may be NULL) is the row interchanged
 * with row c at elimination step c; FNEC_SINGULAR when a pivot's modulus is
 * below 1e-30. stats (may be NULL) gets n_swaps and exact_panels. */
int fnec_hip_debug_solve_square(const fnec_c64 *a, const fnec_c64 *b, size_t n, fnec_c64 *x_out,
                                int32_t *ipiv_out, fnec_solve_stats *stats);
/* The plain path's device elementary functions over an array: fn 0 sin(x),
 * 1 cos(x), 2 hypot(x, y), 3 sqrt(x), 4 x / y, 5 and 6 the sine and cosine
 * of sincos(x) (y may be NULL except for 2 and 4). */
int fnec_hip_debug_refmath(int fn, const double *x, const double *y, size_t n, double *out);
/* The same probe for the FP32 MFMA (v_mfma_f32_16x16x4f32) used by the mixed path. */
int fnec_hip_debug_mfma_probe_f32(const float *a16x4, const float *b4x16, float *lanes_out);

/* ---- tunables (no counterpart in the reference) ----
 * Knobs of the batched engine, set by the autotuner (autotune() in the
 * fnec-rust_amd package: timed batches of a representative workload) or by
 * hand. A knob is read when a batch is planned. They change speed, and the rounding of a mixed-path solve
 * within its accuracy contract (the summation order of the CG solves follows
 * the gang size); a result never depends on what else is in the batch.
 *   "solve1_max_n"    CG solves of problems with n <= value run in k_zsolve
 *                     (one workgroup or one gang per problem); larger n use
 *                     the flag-chained one-block-per-workgroup solve. 0 .. 8192.
 *   "gang_n", "gang_max"  gang size of a CG solve:
 *                     G = clamp(floor(n^2 / gang_n^2), 1, min(gang_max, ceil(n / 64)));
 *                     gang_n = 0 keeps one workgroup per problem.
 *                     gang_n 0 .. 65536 (default 512), gang_max 1 .. 64
 *                     (default 4).
 *   "mix_first_chunk" CG iterations launched before the first convergence
 *                     read-back. 1 .. 12.
 *   "update_tile_m"   row tile of the mixed path's FP32 trailing-matrix
 *                     update (k_lu_update_f32 work items): 64 (default; three
 *                     real MFMA products per complex multiply-add) or 128
 *                     (four). The factors differ at FP32 rounding level.
 *   "update_k128"     1 (default): with update_tile_m 64, every other 64-column
 *                     step leaves its trailing update to the next, which
 *                     applies both with one K = 128 pass on the f16 matrix
 *                     cores (each FP32 operand split into two f16 parts;
 *                     product error about 2^-21 relative), reading and writing
 *                     the trailing matrix once per 128 columns. 0: one K = 64
 *                     FP32 pass per step.
 *   "update_ring"     K = 128 passes: 1 streams the split operands through
 *                     an LDS ring by DMA on a persistent grid (one workgroup
 *                     per CU, all 160 KiB of its LDS; fastest alone); 0
 *                     (default) loads them into registers, one workgroup per
 *                     tile, which leaves room for kernels of other streams
 *                     (the next step's panels) to run beside it. Same bits.
 *   "mix_tol_exp"     mixed path: the CG stops when the normal-equation
 *                     residual is at most 10^-value of its right-hand side.
 *                     8 .. 15 (default 14). Not tuned automatically: it trades
 *                     accuracy, not speed alone.
 *   "batch_split_n"   serving sessions: while the queue holds less than
 *                     max_batch requests, a batch takes only requests of its
 *                     oldest request's size class (N <= value or N > value),
 *                     so that small solves do not wait for a large one's
 *                     factorisation; with a full batch queued, batches are
 *                     formed across classes. 0 .. 65536 (default 1024; 0: one
 *                     class).
 *   "fill_centred"    mixed and FP64-MFMA paths' fill: 1 (default) takes each
 *                     Gauss point's phase relative to the source centre's, a
 *                     short polynomial without range reduction, and the full
 *                     phase's sine and cosine once per element (segments with
 *                     k h above 0.51 keep the per-point form); 0 reduces every
 *                     point's phase.
 *   "exact_backsub"   plain path: 1 runs the back substitution in the
 *                     reference's order (ascending j in every row sum; a
 *                     serial chain of S^2/2 subtractions), so x equals the
 *                     reference's bit for bit; 0 (default) runs it blocked.
 *   "lanes"           batches that may run on the device at once (each on a
 *                     lane: two streams, a device arena, pinned staging);
 *                     also the worker count of a session created afterwards.
 *                     1 .. 4, default 2.
 * set: FNEC_OK or FNEC_INVALID_ARG; get: the value, -1 for an unknown name. */
int fnec_hip_set_tunable(const char *name, long value);
long fnec_hip_get_tunable(const char *name);

/* Per-kernel timing with HIP events on the solve stream (off by default).
 * Kernel classes: "fill", "gram", "lu_panel", "lu_trsm", "lu_update",
 * "backsub", "rp", and on the mixed path "zsolve" (Z / Z^H solves with the
 * FP32 factors), "matvec" (M p and M^H t in FP64), "pcg_ctl", "fsetup"
 * (Z^H solves of the constraint rows), "lu_update_narrow" (the K = 32 update
 * of the second 32-column panel inside each 64-column step of the FP32 LU;
 * "lu_update" is then the K = 64 trailing update), and the host-side phases
 * "host_prep" (problem and plan building), "host_stage" (staging copies),
 * "host_rp" (far-field set-up), "host_post" (result processing),
 * "host_wait" (time blocked on the stream) and "host_enqueue" (queueing a
 * solve's kernels up to its first wait), timed with a steady clock (flops
 * and bytes 0). out4 = {launches, total microseconds, algorithmic flops,
 * algorithmic bytes} accumulated since the last reset. */
void fnec_hip_profile_enable(int on);
/* Restrict event recording to one kernel class (NULL or "": every class), so
 * that a timed run carries events only around the launches it reports.
 * FNEC_INVALID_ARG for an unknown class name. */
int fnec_hip_profile_only(const char *kernel);
void fnec_hip_profile_reset(void);
int fnec_hip_profile_read(const char *kernel, double *out4);

#ifdef __cplusplus
}
#endif
#endif
