// Internal interface between the C ABI (abi.hip) and the engine (engine.hip).
#pragma once

#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fnec_hip.h"

namespace fnec {

// Host-side description of one Hallén problem (augmented system layout).
struct HostProblem {
    int n = 0, s = 0, r = 0, w = 0, ncons = 0;
    int ground = 0;
    double k = 0.0, freq = 0.0, gam_re = 0.0, gam_im = 0.0;
    std::vector<double> seg;   // n * 8
    std::vector<double> selfln; // n: 2 ln((h + R_end) / a), the self term's analytic part
    std::vector<double> cosv;  // n
    std::vector<double> rhs;   // 2n
    std::vector<int> runs;     // 3 per run: row begin, row end, column
    std::vector<int> rowcol;   // n: homogeneous column of each row
    std::vector<int> cons;     // 2 per constraint row
    std::vector<double> cons_sign;
    std::vector<double> stamps; // 4 per Z stamp, applied in order: row, col, re, im
    double phase_max = 0.0;     // bound on every fill phase k R (direct and image)
    // plane-wave (1) and current-source (2) systems, plain path only: rhs
    // holds nrhs forcing vectors of n; M's -sin / -shape entries as xruns
    // (row begin, row end, column, kind 1 or 2); rowent the extra columns of
    // each row (3 per row: -cos, -sin, -shape, -1 for none); the constraint
    // rows' value at a (default 1) and their y (default 0)
    int variant = 0;
    int nrhs = 1;
    std::vector<double> sinv;     // n
    std::vector<double> shape;    // 2n
    std::vector<int> xruns;
    std::vector<int> rowent;      // 3n
    std::vector<double> cons_aval; // ncons
    std::vector<double> cons_y;    // 2 ncons
};

// Host-side description of one far-field job. The grid is borrowed from
// the caller for the duration of the call that runs the job.
struct RpHost {
    int n = 0, ground = 0, ndir = 0;
    double k = 0.0, eps_r = 0.0, sigma = 0.0;
    std::vector<double> seg, cur;
    const double *theta = nullptr, *phi = nullptr;
    size_t npts = 0;
    const double *cur_dev = nullptr; // currents already on the device
};

// A batch of far-field jobs whose inputs are on their way to the device
// (rp_stage), waiting for its kernels and read-back (rp_finish).
struct RpStaged {
    size_t nj = 0;
    char *base = nullptr;
    bool fixed = false;                // base is the device scratch array
    std::vector<char> hstage, hback;   // pageable staging when fixed
    std::vector<size_t> o_tot, o_out, npts;
    size_t out_begin = 0, rows_begin = 0, out_end = 0;
    size_t pts_off = 0, pts_cnt = 0, norm_off = 0, norm_cnt = 0, all_off = 0;
    size_t o_probs = 0, o_ints = 0;
    double rp_terms = 0.0;
};

// Host-phase timing for the per-class profile (a no-op while profiling is
// off): problem and plan building, staging copies, far-field preparation,
// result processing, time blocked on the stream, and kernel enqueueing.
enum { kHostPrep = 14, kHostStage, kHostRp, kHostPost, kHostWait, kHostEnq };
class HostTimer {
  public:
    explicit HostTimer(int cls);
    ~HostTimer();
    HostTimer(const HostTimer &) = delete;
    HostTimer &operator=(const HostTimer &) = delete;

  private:
    int cls_;
    double t0_;
};

// Binds one of the engine's lanes (streams, events, device arenas, pinned
// staging) to this thread for the scope's lifetime. The default form takes a
// free lane, waiting while every lane runs a batch, and is a no-op when the
// thread already holds one; the other binds the lane of the thread that
// started this one (a helper thread working on that thread's batch). The
// device context must be initialised (ensure_ready) first.
class LaneScope {
  public:
    LaneScope();
    explicit LaneScope(void *lane);
    ~LaneScope();
    LaneScope(const LaneScope &) = delete;
    LaneScope &operator=(const LaneScope &) = delete;
    int status() const { return st_; }
    void *lane() const;

  private:
    int st_ = 0;
    void *owned_ = nullptr, *prev_ = nullptr;
    bool bound_ = false;
};

void set_error(const std::string &e);
const char *last_error();
int ensure_ready();
std::mutex &global_mutex();
size_t problem_bytes(int n, int s, int ncons, int nruns, int precision);

int build_problem(HostProblem &hp, const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                  const double *cos_vec, const std::vector<int> &row_col, int ncols,
                  const std::vector<int> &cons, const std::vector<double> &cons_sign,
                  double freq_hz, const fnec_ground *ground);
int problem_from_hallen(HostProblem &hp, const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                        const double *cos_vec, const size_t *wf, const size_t *wl, size_t nw,
                        const size_t *ja, const size_t *jb, const double *js, size_t nj,
                        double freq_hz, const fnec_ground *ground);
int problem_from_paths(HostProblem &hp, const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                       const double *cos_vec, const size_t *path_of, const size_t *free_ends,
                       size_t nfree, double freq_hz, const fnec_ground *ground);
// The plane-wave systems of solve_hallen_planewave (linear.rs:910-985; per
// wire range, I = 0 at both ends of each) and solve_hallen_planewave_paths
// (linear.rs:1017-1091; per conductor path when path_of is not null, I = 0
// at the free ends): two homogeneous columns (-cos, -sin) per wire or path;
// rhs holds nrhs forcing vectors of n (several incidences share the matrix).
int problem_from_planewave(HostProblem &hp, const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                           size_t nrhs, const double *cos_vec, const double *sin_vec,
                           const size_t *wf, const size_t *wl, size_t nw, const size_t *path_of,
                           const size_t *free_ends, size_t nfree, double freq_hz,
                           const fnec_ground *ground);
// The current-source systems of solve_hallen_current_source (linear.rs:
// 1124-1210) and solve_hallen_current_source_paths (linear.rs:1241-1319): one
// homogeneous column per wire or path, the port-voltage column -shape, the
// I = 0 rows and the source row I[src] = i0, all weighted by 1e6.
int problem_from_current_source(HostProblem &hp, const fnec_seg *segs, size_t n,
                                const fnec_c64 *shape, const double *cos_vec, size_t src,
                                fnec_c64 i0, const size_t *wf, const size_t *wl, size_t nw,
                                const size_t *path_of, const size_t *free_ends, size_t nfree,
                                double freq_hz, const fnec_ground *ground);
// Impedance stamps (row, column, delta) added to Z after the fill, before the
// solve; rows and columns are checked against the n segments.
int problem_add_stamps(HostProblem &hp, const fnec_stamp *stamps, size_t n_stamps, size_t n);

// overlap (optional) is called once, before any of the solve's kernels is
// queued, with the device addresses of the solution vectors; it may start host
// work that runs while the solve is queued and computed (the caller joins it).
int solve_problems(const std::vector<const HostProblem *> &hps, std::vector<fnec_c64 *> &x_out,
                   std::vector<fnec_solve_stats> &stats, std::vector<int> &status,
                   std::vector<double *> *device_x, int precision,
                   const std::function<void(const std::vector<double *> &)> &overlap = {});
// Z alone: the plain path's fill (exact) or the fast paths' (k_fill_z).
int fill_only(const HostProblem &hp, fnec_c64 *z_out, bool exact);
int normal_eqs_only(const HostProblem &hp, fnec_c64 *aug_out);
int solve_given_z(const HostProblem &hp, const fnec_c64 *z, fnec_c64 *x_out, fnec_solve_stats *stats);
int solve_square_exact(const fnec_c64 *a, const fnec_c64 *b, int n, fnec_c64 *x_out, int *ipiv_out,
                       fnec_solve_stats *stats);
int refmath_probe(int fn, const double *x, const double *y, double *out, size_t n);
// Fill and the mixed path's FP32 LU of Z only: the packed factors (n x n
// row-major, re and im interleaved) and the interchange of each row.
int lu_f32_only(const HostProblem &hp, float *lu_out, int *ipiv_out);
int rp_host_from(RpHost &h, const fnec_seg *segs, const fnec_c64 *cur, size_t n, double freq_hz,
                 const fnec_ground *ground, const double *theta, const double *phi, size_t npts);
// Far-field jobs in one batch: rp_stage uploads the inputs on the copy
// stream; rp_finish runs the kernels after the solve stream's work and reads
// back the totals, and the rows of the jobs whose outs[j] is not null.
// With fixed_base, the job lives at that device address (fixed_cap bytes;
// FNEC_RESOURCE_EXHAUSTED beyond) instead of the lane's arena, and stages
// through pageable host memory.
int rp_stage(const std::vector<RpHost> &jobs, RpStaged &st, char *fixed_base = nullptr,
             size_t fixed_cap = 0);
// Device scratch of fnec_hip_rp_noalloc.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
int rp_run_noalloc(const RpHost &job, fnec_rp_row *out, double &total);
int rp_finish(RpStaged &st, const std::vector<fnec_rp_row *> &outs, std::vector<double> &totals);
int rp_run(const std::vector<RpHost> &jobs, std::vector<fnec_rp_row *> &outs,
           std::vector<double> &totals);
int mfma_probe(const double *a, const double *b, double *out);
int mfma_probe32(const float *a, const float *b, float *out);
int set_tunable(const char *name, long value);
long get_tunable(const char *name);
void profile_enable(int on);
int profile_only(const char *name);
void profile_reset();
int profile_read(const char *name, double *out4);

} // namespace fnec
