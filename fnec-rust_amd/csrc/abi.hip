// C ABI of include/fnec_hip.h on top of the engine.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/fnec_hip.h"
#include "../../include/fnec_host.h"
#include "engine.h"

using namespace fnec;

namespace {

int check_precision(int precision)
{
    if (precision != FNEC_PREC_FP64 && precision != FNEC_PREC_MIXED && precision != FNEC_PREC_FP64_MFMA) {
        set_error("unknown precision");
        return FNEC_INVALID_ARG;
    }
    return FNEC_OK;
}

int solve_one(const HostProblem &hp, int precision, fnec_c64 *x_out, fnec_solve_stats *stats)
{
    std::vector<const HostProblem *> v{&hp};
    std::vector<fnec_c64 *> xo{x_out};
    std::vector<fnec_solve_stats> st;
    std::vector<int> status;
    int rc = solve_problems(v, xo, st, status, nullptr, precision);
    if (rc)
        return rc;
    if (stats)
        *stats = st[0];
    return status[0];
}

} // namespace

extern "C" {

int fnec_hip_device_count(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess)
        return 0;
    int ok = 0;
    for (int d = 0; d < count; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess &&
            std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
            ++ok;
    }
    return ok;
}

int fnec_hip_init(void)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    return ensure_ready();
}

const char *fnec_hip_last_error(void) { return last_error(); }

int fnec_hip_fill_z(const fnec_seg *segs, size_t n, double freq_hz, const fnec_ground *ground,
                    fnec_c64 *z_out)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if (!z_out) {
        set_error("null output");
        return FNEC_INVALID_ARG;
    }
    std::vector<fnec_c64> rhs(n, fnec_c64{0.0, 0.0});
    std::vector<double> cosv(n, 0.0);
    HostProblem hp;
    std::vector<int> row_col(n, 0);
    st = build_problem(hp, segs, n, rhs.data(), cosv.data(), row_col, 1, {}, {}, freq_hz, ground);
    if (st)
        return st;
    return fill_only(hp, z_out, true);
}

int fnec_hip_debug_fill_fast(const fnec_seg *segs, size_t n, double freq_hz, const fnec_ground *ground,
                             fnec_c64 *z_out)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if (!z_out) {
        set_error("null output");
        return FNEC_INVALID_ARG;
    }
    std::vector<fnec_c64> rhs(n, fnec_c64{0.0, 0.0});
    std::vector<double> cosv(n, 0.0);
    HostProblem hp;
    std::vector<int> row_col(n, 0);
    st = build_problem(hp, segs, n, rhs.data(), cosv.data(), row_col, 1, {}, {}, freq_hz, ground);
    if (st)
        return st;
    return fill_only(hp, z_out, false);
}

int fnec_hip_solve_hallen(const fnec_seg *segs, size_t n, const fnec_c64 *rhs, const double *cos_vec,
                          const size_t *wire_first, const size_t *wire_last, size_t n_wires,
                          const size_t *j_a, const size_t *j_b, const double *j_sign, size_t n_junc,
                          double freq_hz, const fnec_ground *ground, int precision, fnec_c64 *x_out,
                          fnec_solve_stats *stats)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if ((st = check_precision(precision)))
        return st;
    if (!x_out) {
        set_error("null output");
        return FNEC_INVALID_ARG;
    }
    HostProblem hp;
    st = problem_from_hallen(hp, segs, n, rhs, cos_vec, wire_first, wire_last, n_wires, j_a, j_b,
                             j_sign, n_junc, freq_hz, ground);
    if (st)
        return st;
    return solve_one(hp, precision, x_out, stats);
}

int fnec_hip_solve_hallen_paths(const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                                const double *cos_vec, const size_t *path_of,
                                const size_t *free_ends, size_t n_free, double freq_hz,
                                const fnec_ground *ground, int precision, fnec_c64 *x_out,
                                fnec_solve_stats *stats)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if ((st = check_precision(precision)))
        return st;
    if (!x_out) {
        set_error("null output");
        return FNEC_INVALID_ARG;
    }
    HostProblem hp;
    st = problem_from_paths(hp, segs, n, rhs, cos_vec, path_of, free_ends, n_free, freq_hz, ground);
    if (st)
        return st;
    return solve_one(hp, precision, x_out, stats);
}

int fnec_hip_solve_hallen_planewave(const fnec_seg *segs, size_t n, const fnec_c64 *rhs, size_t n_rhs,
                                    const double *cos_vec, const double *sin_vec,
                                    const size_t *wire_first, const size_t *wire_last, size_t n_wires,
                                    const size_t *path_of, const size_t *free_ends, size_t n_free,
                                    const fnec_stamp *stamps, size_t n_stamps, double freq_hz,
                                    const fnec_ground *ground, int precision, fnec_c64 *x_out,
                                    fnec_solve_stats *stats)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if ((st = check_precision(precision)))
        return st;
    if (!x_out) {
        set_error("null output");
        return FNEC_INVALID_ARG;
    }
    HostProblem hp;
    st = problem_from_planewave(hp, segs, n, rhs, n_rhs, cos_vec, sin_vec, wire_first, wire_last, n_wires,
                                path_of, free_ends, n_free, freq_hz, ground);
    if (!st)
        st = problem_add_stamps(hp, stamps, n_stamps, n);
    if (st)
        return st;
    return solve_one(hp, precision, x_out, stats);
}

int fnec_hip_solve_hallen_current_source(const fnec_seg *segs, size_t n, const fnec_c64 *shape,
                                         const double *cos_vec, size_t src, fnec_c64 i0,
                                         const size_t *wire_first, const size_t *wire_last,
                                         size_t n_wires, const size_t *path_of, const size_t *free_ends,
                                         size_t n_free, const fnec_stamp *stamps, size_t n_stamps,
                                         double freq_hz, const fnec_ground *ground, int precision,
                                         fnec_c64 *x_out, fnec_solve_stats *stats)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if ((st = check_precision(precision)))
        return st;
    if (!x_out) {
        set_error("null output");
        return FNEC_INVALID_ARG;
    }
    HostProblem hp;
    st = problem_from_current_source(hp, segs, n, shape, cos_vec, src, i0, wire_first, wire_last, n_wires,
                                     path_of, free_ends, n_free, freq_hz, ground);
    if (!st)
        st = problem_add_stamps(hp, stamps, n_stamps, n);
    if (st)
        return st;
    return solve_one(hp, precision, x_out, stats);
}

int fnec_deck_solve(const fnec_deck_problem *pr, int precision, fnec_c64 *x_out,
                    fnec_solve_stats *stats)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if ((st = check_precision(precision)))
        return st;
    if (!pr || !x_out) {
        set_error("null argument");
        return FNEC_INVALID_ARG;
    }
    const fnec_request &r = pr->req;
    HostProblem hp;
    const size_t *pof = pr->route == 1 ? pr->path_of : nullptr;
    if (pr->variant == 1)
        st = problem_from_planewave(hp, r.segs, r.n, pr->rhs_all, pr->n_rhs, r.cos_vec, pr->sin_vec,
                                    r.wire_first, r.wire_last, r.n_wires, pof, pr->free_ends, pr->n_free,
                                    r.freq_hz, &r.ground);
    else if (pr->variant == 2)
        st = problem_from_current_source(hp, r.segs, r.n, r.rhs, r.cos_vec, pr->src, pr->i0, r.wire_first,
                                         r.wire_last, r.n_wires, pof, pr->free_ends, pr->n_free, r.freq_hz,
                                         &r.ground);
    else if (pr->route == 1)
        st = problem_from_paths(hp, r.segs, r.n, r.rhs, r.cos_vec, pr->path_of, pr->free_ends,
                                pr->n_free, r.freq_hz, &r.ground);
    else
        st = problem_from_hallen(hp, r.segs, r.n, r.rhs, r.cos_vec, r.wire_first, r.wire_last,
                                 r.n_wires, r.j_a, r.j_b, r.j_sign, r.n_junc, r.freq_hz, &r.ground);
    if (!st)
        st = problem_add_stamps(hp, r.stamps, r.n_stamps, r.n);
    if (st)
        return st;
    return solve_one(hp, precision, x_out, stats);
}

int fnec_hip_rp(const fnec_seg *segs, const fnec_c64 *currents, size_t n, double freq_hz,
                const fnec_ground *ground, const double *theta_deg, const double *phi_deg,
                size_t npts, fnec_rp_row *out, double *total_out)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if (npts && !out) {
        set_error("null output");
        return FNEC_INVALID_ARG;
    }
    std::vector<RpHost> jobs(1);
    st = rp_host_from(jobs[0], segs, currents, n, freq_hz, ground, theta_deg, phi_deg, npts);
    if (st)
        return st;
    std::vector<fnec_rp_row *> outs{out};
    std::vector<double> totals;
    st = rp_run(jobs, outs, totals);
    if (st)
        return st;
    if (total_out)
        *total_out = totals[0];
    return FNEC_OK;
}

int fnec_hip_rp_noalloc(const fnec_seg *segs, const fnec_c64 *currents, size_t n, double freq_hz,
                        const fnec_ground *ground, const double *theta_deg, const double *phi_deg,
                        size_t npts, fnec_rp_row *out, double *total_out)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if (npts && !out) {
        set_error("null output");
        return FNEC_INVALID_ARG;
    }
    RpHost job;
    st = rp_host_from(job, segs, currents, n, freq_hz, ground, theta_deg, phi_deg, npts);
    if (st)
        return st;
    double total = 0.0;
    st = rp_run_noalloc(job, out, total);
    if (st)
        return st;
    if (total_out)
        *total_out = total;
    return FNEC_OK;
}

int fnec_hip_debug_normal_eqs(const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                              const double *cos_vec, const size_t *wire_first,
                              const size_t *wire_last, size_t n_wires, const size_t *j_a,
                              const size_t *j_b, const double *j_sign, size_t n_junc,
                              double freq_hz, const fnec_ground *ground, fnec_c64 *aug_out)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    HostProblem hp;
    st = problem_from_hallen(hp, segs, n, rhs, cos_vec, wire_first, wire_last, n_wires, j_a, j_b,
                             j_sign, n_junc, freq_hz, ground);
    if (st)
        return st;
    return normal_eqs_only(hp, aug_out);
}

int fnec_hip_debug_solve_z(const fnec_c64 *z, size_t n, const fnec_c64 *rhs, const double *cos_vec,
                           const size_t *wire_first, const size_t *wire_last, size_t n_wires,
                           const size_t *j_a, const size_t *j_b, const double *j_sign, size_t n_junc,
                           fnec_c64 *x_out, fnec_solve_stats *stats)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if (!z || !x_out || n == 0) {
        set_error("null or empty input");
        return FNEC_INVALID_ARG;
    }
    // the segments only size the problem: the fill does not run
    std::vector<fnec_seg> segs(n, fnec_seg{{0.0, 0.0, 0.0}, {0.0, 0.0, 1.0}, 1.0, 1e-3});
    HostProblem hp;
    st = problem_from_hallen(hp, segs.data(), n, rhs, cos_vec, wire_first, wire_last, n_wires, j_a, j_b,
                             j_sign, n_junc, 1.0, nullptr);
    if (st)
        return st;
    return solve_given_z(hp, z, x_out, stats);
}

int fnec_hip_debug_solve_square(const fnec_c64 *a, const fnec_c64 *b, size_t n, fnec_c64 *x_out,
                                int32_t *ipiv_out, fnec_solve_stats *stats)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    if (!a || !b || !x_out || n == 0 || n > (1u << 16)) {
        set_error("null input or order out of range (1..65536)");
        return FNEC_INVALID_ARG;
    }
    return solve_square_exact(a, b, (int)n, x_out, ipiv_out, stats);
}

int fnec_hip_debug_refmath(int fn, const double *x, const double *y, size_t n, double *out)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    return refmath_probe(fn, x, y, out, n);
}

int fnec_hip_debug_lu_f32(const fnec_seg *segs, size_t n, const fnec_c64 *rhs, const double *cos_vec,
                          const size_t *wire_first, const size_t *wire_last, size_t n_wires, const size_t *j_a,
                          const size_t *j_b, const double *j_sign, size_t n_junc, double freq_hz,
                          const fnec_ground *ground, float *lu_out, int32_t *ipiv_out)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    HostProblem hp;
    st = problem_from_hallen(hp, segs, n, rhs, cos_vec, wire_first, wire_last, n_wires, j_a, j_b,
                             j_sign, n_junc, freq_hz, ground);
    if (st)
        return st;
    return lu_f32_only(hp, lu_out, ipiv_out);
}

int fnec_hip_debug_mfma_probe(const double *a16x4, const double *b4x16, double *lanes_out)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    return mfma_probe(a16x4, b4x16, lanes_out);
}

int fnec_hip_debug_mfma_probe_f32(const float *a16x4, const float *b4x16, float *lanes_out)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    int st = ensure_ready();
    if (st)
        return st;
    return mfma_probe32(a16x4, b4x16, lanes_out);
}

int fnec_hip_set_tunable(const char *name, long value)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    return set_tunable(name, value);
}

long fnec_hip_get_tunable(const char *name)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    return get_tunable(name);
}

void fnec_hip_profile_enable(int on)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    profile_enable(on);
}

int fnec_hip_profile_only(const char *kernel)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    const int st = profile_only(kernel);
    if (st)
        set_error("unknown kernel class");
    return st;
}

void fnec_hip_profile_reset(void)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    profile_reset();
}

int fnec_hip_profile_read(const char *kernel, double *out4)
{
    std::lock_guard<std::mutex> lk(global_mutex());
    if (!kernel || !out4) {
        set_error("null argument");
        return FNEC_INVALID_ARG;
    }
    return profile_read(kernel, out4);
}

} // extern "C"
