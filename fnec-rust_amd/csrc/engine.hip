// Host engine of fnec_hip: device context, batch planning (problem tables and
// work lists), the launch sequence fill -> Gram -> LU -> back substitution ->
// residual, the far-field launches, and the C ABI of include/fnec_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fnec_hip.h"
#include "common.h"
#include "engine.h"
#include "lu.h"

namespace fnec {

// kernels.hip
__global__ void k_fill_z(const Prob *, const int4 *);
__global__ void k_fill_z_large(const Prob *, const int4 *);
__global__ void k_stamp_z(const Prob *, const int *);
__global__ void k_gram(const Prob *, const int4 *);
__global__ void k_gram_cols(const Prob *, const int2 *);
__global__ void k_gram_fix(const Prob *, const int *);
__global__ void k_backsub(const Prob *, const int2 *);
__global__ void k_residual_rows(const Prob *, const int2 *);
__global__ void k_residual_finish(const Prob *, const int *);
__global__ void k_rp_points(const RpProb *, const int2 *);
__global__ void k_rp_norm(const RpProb *, const int2 *);
__global__ void k_rp_finish(const RpProb *, const int *);
__global__ void k_zero(double *, size_t);
// mixed.hip
__global__ void k_trsv0(const Prob *, const int4 *, int, unsigned long long *);
__global__ void k_trsv1(const Prob *, const int4 *, int, unsigned long long *);
__global__ void k_perm(const Prob *, const int *);
__global__ void k_zsolve(const Prob *, const int4 *, int, int, unsigned long long *);
__global__ void k_gen_items(const int4 *, int4 *);
__global__ void k_fsetup(const Prob *, const int4 *, int, unsigned long long *);
__global__ void k_diag_inv(const Prob *, const int2 *);
__global__ void k_pre_setup(const Prob *, const int *);
__global__ void k_prec_mid(const Prob *, const int *);
__global__ void k_mv_z(const Prob *, const int4 *);
__global__ void k_mv_fin(const Prob *, const int2 *);
__global__ void k_mv_zh(const Prob *, const int2 *, int);
__global__ void k_mix_init(const Prob *, const int *);
__global__ void k_mix_start(const Prob *, const int *);
__global__ void k_mix_dir(const Prob *, const int *);
__global__ void k_mix_step(const Prob *, const int *);
__global__ void k_mix_flags(const Prob *, unsigned, int *);
__global__ void k_mfma_probe32(const float *, const float *, float *);
__global__ void k_mfma_probe(const double *, const double *, double *);
// exact.hip (plain path in the reference's operation order)
__global__ void k_fill_exact(const Prob *, const int4 *);
__global__ void k_gx_tile(const Prob *, const int4 *);
__global__ void k_gx_cols(const Prob *, const int2 *);
__global__ void k_gx_rhs(const Prob *, const int4 *);
__global__ void k_gx_extra(const Prob *, const int2 *);
__global__ void k_gx_fix(const Prob *, const int *);
__global__ void k_ge_diag(const Prob *, const int *, int);
__global__ void k_ge_rows(const Prob *, const int2 *, int);
__global__ void k_ge_exact(const Prob *, const int *, int);
__global__ void k_ge_u12(const Prob *, const int2 *, int);
__global__ void k_ge_update(const Prob *, const int4 *, int);
__global__ void k_backsub_exact(const Prob *, const int2 *);
__global__ void k_refmath_probe(int, const double *, const double *, double *, int);

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

static thread_local std::string g_err;
static std::mutex g_mu;

void set_error(const std::string &e) { g_err = e; }
const char *last_error() { return g_err.c_str(); }

#define HIPCHK(x)                                                                               \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) {                                                                 \
            set_error(std::string(#x) + ": " + hipGetErrorString(e_));                          \
            return FNEC_INTERNAL;                                                               \
        }                                                                                       \
    } while (0)

struct Context {
    bool ready = false;
    int status = FNEC_UNAVAILABLE;
    int device = -1;
    // device address of `scratch` on each device, from hipGetSymbolAddress
    // with that device current (the library drives g_ctx.device only)
    std::vector<char *> scratch_dev;
};

// Working memory of fnec_hip_rp_noalloc: its far-field job (inputs, fields,
// partial sums, totals, rows, work list) lives here instead of an arena the
// library allocates, so that entry point calls no hipMalloc. Each device
// holds its own copy, at its own address (Context::scratch_dev), until the
// process ends; every such call on a device uses the same bytes, and the
// process-wide lock of the C ABI keeps calls from overlapping.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static Context g_ctx;

struct ProfPending {
    int cls;
    hipEvent_t a, b;
    double flops, bytes;
};

// One lane runs one batch at a time: its own streams, events, device arenas,
// pinned staging and host caches. Batches on different lanes overlap on the
// device (the serving session keeps two in flight). A thread binds a lane with
// a LaneScope for the duration of an engine call.
struct Lane {
    bool init = false, busy = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8];
    // cached device arena, grown on demand
    char *arena = nullptr;
    size_t arena_bytes = 0;
    // far-field jobs read the currents in place from the solve arena, so they
    // get a device arena of their own
    char *rp_arena = nullptr;
    size_t rp_arena_bytes = 0;
    // pinned host staging for the batch upload and the result read-back, and
    // for far-field inputs staged while a solve runs
    char *h_up = nullptr, *h_back = nullptr, *h_rp = nullptr;
    size_t h_up_bytes = 0, h_back_bytes = 0, h_rp_bytes = 0;
    // the LU's panel chain runs one step ahead on this stream (default
    // priority), ordered against the solve stream with events from the pool;
    // far-field inputs are uploaded on it while the solve runs (ev_rp)
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_rp = nullptr;
    std::vector<hipEvent_t> sync_events;
    // per-kernel timing records of this lane's launches
    std::vector<ProfPending> prof_pending;
    // work lists keep their capacity across batches
    std::vector<int> plan_ints, rp_ints;
};

constexpr int kMaxLanes = 4;
static Lane g_lanes[kMaxLanes];
static int g_lanes_on = 2; // lanes in use (tunable "lanes")
static std::mutex g_lane_mu;
static std::condition_variable g_lane_cv;
static thread_local Lane *t_lane = nullptr;

// The lane bound to this thread (every engine entry point binds one).
static inline Lane &LN() { return *t_lane; }

static int lane_init(Lane &L)
{
    if (L.init)
        return FNEC_OK;
    if (hipStreamCreate(&L.stream) != hipSuccess || hipStreamCreate(&L.side_stream) != hipSuccess ||
        hipEventCreate(&L.ev_rp) != hipSuccess) {
        set_error("stream creation failed");
        return FNEC_INTERNAL;
    }
    for (auto &e : L.ev)
        if (hipEventCreate(&e) != hipSuccess) {
            set_error("event creation failed");
            return FNEC_INTERNAL;
        }
    L.init = true;
    return FNEC_OK;
}

LaneScope::LaneScope()
{
    if (t_lane)
        return; // this thread already holds a lane (nested engine call)
    if (!g_ctx.ready || g_ctx.status != FNEC_OK) {
        set_error("device not initialised");
        st_ = FNEC_UNAVAILABLE;
        return;
    }
    Lane *got = nullptr;
    {
        std::unique_lock<std::mutex> lk(g_lane_mu);
        g_lane_cv.wait(lk, [&] {
            for (int i = 0; i < g_lanes_on; ++i)
                if (!g_lanes[i].busy) {
                    got = &g_lanes[i];
                    return true;
                }
            return false;
        });
        got->busy = true;
    }
    (void)hipSetDevice(g_ctx.device);
    st_ = lane_init(*got);
    if (st_) {
        std::lock_guard<std::mutex> lk(g_lane_mu);
        got->busy = false;
        g_lane_cv.notify_all();
        return;
    }
    t_lane = got;
    owned_ = got;
}

LaneScope::LaneScope(void *lane) : prev_(t_lane)
{
    (void)hipSetDevice(g_ctx.device);
    t_lane = (Lane *)lane;
    bound_ = true;
}

LaneScope::~LaneScope()
{
    if (bound_)
        t_lane = (Lane *)prev_;
    if (!owned_)
        return;
    t_lane = nullptr;
    std::lock_guard<std::mutex> lk(g_lane_mu);
    ((Lane *)owned_)->busy = false;
    g_lane_cv.notify_all();
}

void *LaneScope::lane() const { return t_lane; }

// Wait for a stream by polling it: a blocking synchronisation can sleep the
// host thread and wake it a few milliseconds after the stream drained, which
// leaves the device idle between a solve's dependent phases.
static hipError_t stream_wait(hipStream_t st)
{
    for (;;) {
        const hipError_t e = hipStreamQuery(st);
        if (e != hipErrorNotReady)
            return e;
    }
}

// Optional per-kernel timing: HIP events around the launches of each tracked
// kernel class on the solve stream, with the algorithmic work of each launch.
struct ProfClass {
    double launches = 0, total_us = 0, flops = 0, bytes = 0;
};
// Classes from PC_HOST_PREP on are host-side phases timed with a steady clock
// (launches = occurrences, no flops or bytes).
enum { PC_FILL, PC_GRAM, PC_LU_PANEL, PC_LU_TRSM, PC_LU_UPDATE, PC_BACKSUB, PC_RP, PC_ZSOLVE,
       PC_MV, PC_PCG, PC_FSETUP, PC_LU_UPDN, PC_LU_UPD128, PC_LU_SPLIT, PC_HOST_PREP, PC_HOST_STAGE, PC_HOST_RP,
       PC_HOST_POST, PC_HOST_WAIT, PC_HOST_ENQ, PC_N };
static const char *kProfNames[PC_N] = {"fill", "gram", "lu_panel", "lu_trsm", "lu_update",
                                       "backsub", "rp", "zsolve", "matvec", "pcg_ctl", "fsetup",
                                       "lu_update_narrow", "lu_update_k128", "lu_split", "host_prep", "host_stage",
                                       "host_rp", "host_post", "host_wait", "host_enqueue"};
static bool g_prof = false;
static ProfClass g_prof_acc[PC_N];
static std::vector<hipEvent_t> g_prof_pool;
static std::mutex g_prof_mu; // accumulators and the event pool (lanes record concurrently)

static hipEvent_t prof_event()
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_pool.empty()) {
        hipEvent_t e = g_prof_pool.back();
        g_prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

static int g_prof_only = -1; // record only this class (-1: every class)

static_assert(PC_HOST_PREP == kHostPrep && PC_HOST_WAIT == kHostWait && PC_HOST_ENQ == kHostEnq, "host classes");

static double host_now_us()
{
    using namespace std::chrono;
    return (double)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() * 1e-3;
}

HostTimer::HostTimer(int cls) : cls_(cls), t0_(g_prof ? host_now_us() : 0.0) {}
HostTimer::~HostTimer()
{
    if (!g_prof || (g_prof_only >= 0 && cls_ != g_prof_only))
        return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    ProfClass &c = g_prof_acc[cls_];
    c.launches += 1;
    c.total_us += host_now_us() - t0_;
}

static hipEvent_t prof_begin(int cls, hipStream_t st = nullptr)
{
    if (!g_prof || (g_prof_only >= 0 && cls != g_prof_only))
        return nullptr;
    hipEvent_t e = prof_event();
    (void)hipEventRecord(e, st ? st : LN().stream);
    return e;
}

// Returns the index of the pending record (-1 when profiling is off) so that
// work known only after the run can be attributed to it.
static long prof_end(int cls, hipEvent_t a, double flops, double bytes, hipStream_t st = nullptr)
{
    if (!g_prof || !a)
        return -1;
    hipEvent_t b = prof_event();
    (void)hipEventRecord(b, st ? st : LN().stream);
    LN().prof_pending.push_back({cls, a, b, flops, bytes});
    return (long)LN().prof_pending.size() - 1;
}

static void prof_set_work(long idx, double flops, double bytes)
{
    std::vector<ProfPending> &pend = LN().prof_pending;
    if (idx >= 0 && idx < (long)pend.size()) {
        pend[idx].flops = flops;
        pend[idx].bytes = bytes;
    }
}

// Called after the lane's stream synchronisation: fold its pending event pairs.
static void prof_collect()
{
    std::vector<ProfPending> &pend = LN().prof_pending;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto &p : pend) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ProfClass &c = g_prof_acc[p.cls];
            c.launches += 1;
            c.total_us += ms * 1000.0;
            c.flops += p.flops;
            c.bytes += p.bytes;
        }
        g_prof_pool.push_back(p.a);
        g_prof_pool.push_back(p.b);
    }
    pend.clear();
}

void profile_enable(int on) { g_prof = on != 0; }
int profile_only(const char *name)
{
    if (!name || !*name) {
        g_prof_only = -1;
        return FNEC_OK;
    }
    for (int i = 0; i < PC_N; ++i)
        if (!std::strcmp(name, kProfNames[i])) {
            g_prof_only = i;
            return FNEC_OK;
        }
    return FNEC_INVALID_ARG;
}
void profile_reset()
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto &c : g_prof_acc)
        c = ProfClass{};
}
int profile_read(const char *name, double *out4)
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (int i = 0; i < PC_N; ++i)
        if (!std::strcmp(name, kProfNames[i])) {
            out4[0] = g_prof_acc[i].launches;
            out4[1] = g_prof_acc[i].total_us;
            out4[2] = g_prof_acc[i].flops;
            out4[3] = g_prof_acc[i].bytes;
            return FNEC_OK;
        }
    set_error("unknown kernel class");
    return FNEC_INVALID_ARG;
}

static int ctx_init()
{
    if (g_ctx.ready)
        return g_ctx.status;
    g_ctx.ready = true;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        set_error("no HIP device");
        g_ctx.status = FNEC_UNAVAILABLE;
        return g_ctx.status;
    }
    for (int d = 0; d < count; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) != hipSuccess)
            continue;
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) {
            g_ctx.device = d;
            break;
        }
    }
    if (g_ctx.device < 0) {
        set_error("no gfx950 device");
        g_ctx.status = FNEC_UNAVAILABLE;
        return g_ctx.status;
    }
    if (hipSetDevice(g_ctx.device) != hipSuccess) {
        set_error("device initialisation failed");
        g_ctx.status = FNEC_INTERNAL;
        return g_ctx.status;
    }
    g_ctx.scratch_dev.assign(count, nullptr);
    void *sp = nullptr;
    const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (se != hipSuccess || !sp) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se));
        g_ctx.status = FNEC_INTERNAL;
        return g_ctx.status;
    }
    g_ctx.scratch_dev[g_ctx.device] = (char *)sp;
    // the lanes in use create their streams and events here, so that later
    // calls (fnec_hip_rp_noalloc among them) create nothing; a lane enabled
    // later by the tunable creates its own when first bound
    for (int i = 0; i < g_lanes_on; ++i)
        if (lane_init(g_lanes[i]) != FNEC_OK) {
            g_ctx.status = FNEC_INTERNAL;
            return g_ctx.status;
        }
    g_ctx.status = FNEC_OK;
    return g_ctx.status;
}

int ensure_ready()
{
    int st = ctx_init();
    if (st == FNEC_OK)
        (void)hipSetDevice(g_ctx.device);
    return st;
}

// Pinned host buffer of at least `bytes`, grown on demand (the stream is
// idle whenever a batch is planned, so the old buffer is free to release).
static char *pinned(char *&buf, size_t &have, size_t bytes)
{
    if (bytes <= have)
        return buf;
    if (buf) {
        (void)hipStreamSynchronize(LN().stream);
        (void)hipStreamSynchronize(LN().side_stream);
        (void)hipHostFree(buf);
        buf = nullptr;
        have = 0;
    }
    const size_t want = bytes + bytes / 4 + (1u << 20);
    if (hipHostMalloc((void **)&buf, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        buf = nullptr;
        return nullptr;
    }
    have = want;
    return buf;
}

static int grow_arena(char *&arena, size_t &have, size_t bytes, char **out)
{
    if (bytes > have) {
        if (arena) {
            HIPCHK(hipStreamSynchronize(LN().stream));
            HIPCHK(hipStreamSynchronize(LN().side_stream));
            HIPCHK(hipFree(arena));
            arena = nullptr;
            have = 0;
        }
        size_t want = bytes + bytes / 4 + (64u << 20);
        hipError_t e = hipMalloc(&arena, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            e = hipMalloc(&arena, bytes);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                set_error("device memory exhausted");
                return FNEC_RESOURCE_EXHAUSTED;
            }
            want = bytes;
        }
        have = want;
    }
    *out = arena;
    return FNEC_OK;
}

static int arena_reserve(size_t bytes, char **out)
{
    return grow_arena(LN().arena, LN().arena_bytes, bytes, out);
}

// ---------------------------------------------------------------------------
// host-side problem construction
// ---------------------------------------------------------------------------

static inline int roundup(int v, int m) { return (v + m - 1) / m * m; }

// num-complex 0.4 sqrt (polar form) and the scalar reflection coefficient of
// matrix.rs:285-290.
static void csqrt_nc(double re, double im, double &sr, double &si)
{
#pragma clang fp contract(off)
    if (im == 0.0) {
        if (!std::signbit(re)) {
            sr = std::sqrt(re);
            si = im;
        } else {
            sr = 0.0;
            si = std::signbit(im) ? -std::sqrt(-re) : std::sqrt(-re);
        }
        return;
    }
    if (re == 0.0) {
        double x = std::sqrt(std::fabs(im) / 2.0);
        sr = x;
        si = std::signbit(im) ? -x : x;
        return;
    }
    double r = std::hypot(re, im), th = std::atan2(im, re);
    double rs = std::sqrt(r);
    sr = rs * std::cos(th / 2.0);
    si = rs * std::sin(th / 2.0);
}

static void fresnel_scalar(double freq, double eps_r, double sigma, double &gr, double &gi)
{
#pragma clang fp contract(off)
    double omega = 2.0 * kPi * freq;
    double er = std::isnan(eps_r) ? 1e-6 : std::max(eps_r, 1e-6);
    double sg = std::isnan(sigma) ? 0.0 : std::max(sigma, 0.0);
    double sr, si;
    csqrt_nc(er, -sg / (omega * kEps0Matrix), sr, si);
    double ar = sr - 1.0, ai = si, br = sr + 1.0, bi = si;
    double ns = br * br + bi * bi;
    gr = (ar * br + ai * bi) / ns;
    gi = (ai * br - ar * bi) / ns;
}

// The self term's analytic part 2 ln((h + R_end) / a), R_end = sqrt(h^2 + a^2),
// h = length / 2 (matrix.rs:302-316), one per segment with the host libm, in
// the reference's operation order.
static void self_log_terms(HostProblem &hp)
{
#pragma clang fp contract(off)
    const size_t n = hp.seg.size() / kSegStride;
    hp.selfln.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const double half = hp.seg[i * kSegStride + 6] * 0.5, a = hp.seg[i * kSegStride + 7];
        const double r_end = std::sqrt(half * half + a * a);
        hp.selfln[i] = 2.0 * std::log((half + r_end) / a);
    }
}

int build_problem(HostProblem &hp, const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                  const double *cos_vec, const std::vector<int> &row_col, int ncols,
                  const std::vector<int> &cons, const std::vector<double> &cons_sign,
                  double freq_hz, const fnec_ground *ground)
{
    if (n == 0 || !segs || !rhs || !cos_vec) {
        set_error("empty or null input");
        return FNEC_INVALID_ARG;
    }
    if (n > (size_t)1 << 20) {
        set_error("too many segments");
        return FNEC_INVALID_ARG;
    }
    if (!(freq_hz > 0.0) || !std::isfinite(freq_hz)) {
        set_error("frequency must be positive");
        return FNEC_INVALID_ARG;
    }
    hp.stamps.clear();
    hp.variant = 0;
    hp.nrhs = 1;
    hp.sinv.clear();
    hp.shape.clear();
    hp.xruns.clear();
    hp.rowent.clear();
    hp.cons_aval.clear();
    hp.cons_y.clear();
    hp.n = (int)n;
    hp.w = ncols;
    hp.s = (int)n + ncols;
    hp.ncons = (int)cons_sign.size();
    hp.r = (int)n + hp.ncons;
    hp.k = 2.0 * kPi * freq_hz / kC0;
    hp.freq = freq_hz;
    hp.ground = 0;
    hp.gam_re = hp.gam_im = 0.0;
    if (ground) {
        if (ground->kind == FNEC_GROUND_PEC)
            hp.ground = 1;
        else if (ground->kind == FNEC_GROUND_FINITE) {
            hp.ground = 2;
            fresnel_scalar(freq_hz, ground->eps_r, ground->sigma, hp.gam_re, hp.gam_im);
        } else if (ground->kind != FNEC_GROUND_FREE) {
            set_error("unknown ground kind");
            return FNEC_INVALID_ARG;
        }
    }
    hp.seg.resize(n * kSegStride);
    for (size_t i = 0; i < n; ++i) {
        double *o = &hp.seg[i * kSegStride];
        for (int c = 0; c < 3; ++c) {
            o[c] = segs[i].mid[c];
            o[3 + c] = segs[i].dir[c];
        }
        o[6] = segs[i].length;
        o[7] = segs[i].radius;
    }
    // Any two quadrature points p, q (or p and q's image) are at most
    // |p| + |q| apart, and R adds at most the radius: k (2 max|p| + max a)
    // bounds every phase of the fill (NaN inputs give NaN: the large form).
    {
        double pmax = 0.0, amax = 0.0;
        for (size_t i = 0; i < n; ++i) {
            const double *o = &hp.seg[i * kSegStride];
            const double m = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
            const double d = std::sqrt(o[3] * o[3] + o[4] * o[4] + o[5] * o[5]);
            pmax = std::max(pmax, m + 0.5 * std::fabs(o[6]) * d);
            amax = std::max(amax, std::fabs(o[7]));
            if (!std::isfinite(m + d + o[6] + o[7]))
                pmax = NAN;
        }
        hp.phase_max = hp.k * (2.0 * pmax + amax);
    }
    self_log_terms(hp);
    hp.cosv.assign(cos_vec, cos_vec + n);
    hp.rhs.resize(2 * n);
    for (size_t i = 0; i < n; ++i) {
        hp.rhs[2 * i] = rhs[i].re;
        hp.rhs[2 * i + 1] = rhs[i].im;
    }
    hp.runs.clear();
    for (size_t r = 0; r < n;) {
        size_t e = r + 1;
        while (e < n && row_col[e] == row_col[r])
            ++e;
        hp.runs.push_back((int)r);
        hp.runs.push_back((int)e);
        hp.runs.push_back(row_col[r]);
        r = e;
    }
    hp.rowcol.assign(row_col.begin(), row_col.begin() + n);
    hp.cons = cons;
    hp.cons_sign = cons_sign;
    return FNEC_OK;
}

// Augmented-system layout of solve_hallen (linear.rs:688-762).
int problem_from_hallen(HostProblem &hp, const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                        const double *cos_vec, const size_t *wf, const size_t *wl, size_t nw,
                        const size_t *ja, const size_t *jb, const double *js, size_t nj,
                        double freq_hz, const fnec_ground *ground)
{
    std::vector<size_t> ef, el;
    if (nw == 0 || n == 0) {
        ef.push_back(0);
        el.push_back(n ? n - 1 : 0);
    } else {
        if (!wf || !wl) {
            set_error("null wire range");
            return FNEC_INVALID_ARG;
        }
        ef.assign(wf, wf + nw);
        el.assign(wl, wl + nw);
    }
    if (nj && (!ja || !jb || !js)) {
        set_error("null junction arrays");
        return FNEC_INVALID_ARG;
    }
    for (size_t i = 0; i < ef.size(); ++i)
        if (ef[i] >= n || el[i] >= n) {
            set_error("wire range outside the segment list");
            return FNEC_INVALID_ARG;
        }
    std::vector<char> injunc(n, 0);
    for (size_t i = 0; i < nj; ++i) {
        if (ja[i] >= n || jb[i] >= n) {
            set_error("junction index outside the segment list");
            return FNEC_INVALID_ARG;
        }
        injunc[ja[i]] = injunc[jb[i]] = 1;
    }
    std::vector<int> row_col(n, 0);
    for (size_t w = 0; w < ef.size(); ++w)
        for (size_t r = ef[w]; r <= el[w] && r < n; ++r)
            row_col[r] = (int)w;
    std::vector<int> cons;
    std::vector<double> sign;
    for (size_t w = 0; w < ef.size(); ++w) {
        if (!injunc[ef[w]]) {
            cons.push_back((int)ef[w]);
            cons.push_back(-1);
            sign.push_back(1.0);
        }
        if (!injunc[el[w]]) {
            cons.push_back((int)el[w]);
            cons.push_back(-1);
            sign.push_back(1.0);
        }
    }
    for (size_t i = 0; i < nj; ++i) {
        cons.push_back((int)ja[i]);
        cons.push_back((int)jb[i]);
        sign.push_back(js[i]);
    }
    return build_problem(hp, segs, n, rhs, cos_vec, row_col, (int)ef.size(), cons, sign, freq_hz,
                         ground);
}

// Conductor-path layout of solve_hallen_paths (linear.rs:839-860).
int problem_from_paths(HostProblem &hp, const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                       const double *cos_vec, const size_t *path_of, const size_t *free_ends,
                       size_t nfree, double freq_hz, const fnec_ground *ground)
{
    if (!path_of || (nfree && !free_ends)) {
        set_error("null path arrays");
        return FNEC_INVALID_ARG;
    }
    size_t np = 0;
    std::vector<int> row_col(n);
    for (size_t i = 0; i < n; ++i) {
        if (path_of[i] >= n) {
            set_error("path index out of range");
            return FNEC_INVALID_ARG;
        }
        row_col[i] = (int)path_of[i];
        np = std::max(np, path_of[i] + 1);
    }
    std::vector<int> cons;
    std::vector<double> sign;
    for (size_t i = 0; i < nfree; ++i) {
        if (free_ends[i] >= n) {
            set_error("free end outside the segment list");
            return FNEC_INVALID_ARG;
        }
        cons.push_back((int)free_ends[i]);
        cons.push_back(-1);
        sign.push_back(1.0);
    }
    return build_problem(hp, segs, n, rhs, cos_vec, row_col, (int)np, cons, sign, freq_hz, ground);
}

// Shared by the plane-wave and current-source layouts: the homogeneous column
// (-cos) of each row, from the wire ranges (later ranges win, uncovered rows
// take range 0, as the reference's row_wire does) or from path_of; fills the
// per-range or per-path count and the I = 0 rows (both ends of every range,
// or the listed free ends).
static int variant_rows(size_t n, const size_t *wf, const size_t *wl, size_t nw, const size_t *path_of,
                        const size_t *free_ends, size_t nfree, std::vector<int> &row_col, int &groups,
                        std::vector<int> &ends)
{
    row_col.assign(n, 0);
    ends.clear();
    if (path_of) {
        if (nfree && !free_ends) {
            set_error("null free-end list");
            return FNEC_INVALID_ARG;
        }
        size_t np = 0;
        for (size_t i = 0; i < n; ++i) {
            if (path_of[i] >= n) {
                set_error("path index out of range");
                return FNEC_INVALID_ARG;
            }
            row_col[i] = (int)path_of[i];
            np = std::max(np, path_of[i] + 1);
        }
        for (size_t i = 0; i < nfree; ++i) {
            if (free_ends[i] >= n) {
                set_error("free end outside the segment list");
                return FNEC_INVALID_ARG;
            }
            ends.push_back((int)free_ends[i]);
        }
        groups = (int)np;
        return FNEC_OK;
    }
    std::vector<size_t> ef, el;
    if (nw == 0 || n == 0) {
        ef.push_back(0);
        el.push_back(n ? n - 1 : 0);
    } else {
        if (!wf || !wl) {
            set_error("null wire range");
            return FNEC_INVALID_ARG;
        }
        ef.assign(wf, wf + nw);
        el.assign(wl, wl + nw);
    }
    for (size_t w = 0; w < ef.size(); ++w) {
        if (ef[w] >= n || el[w] >= n) {
            set_error("wire range outside the segment list");
            return FNEC_INVALID_ARG;
        }
        for (size_t r = ef[w]; r <= el[w]; ++r)
            row_col[r] = (int)w;
        ends.push_back((int)ef[w]);
        ends.push_back((int)el[w]);
    }
    groups = (int)ef.size();
    return FNEC_OK;
}

// Runs (row begin, row end, column, kind) of equal column along the rows.
static void xruns_from(const std::vector<int> &col, int kind, std::vector<int> &out)
{
    const size_t n = col.size();
    for (size_t r = 0; r < n;) {
        size_t e = r + 1;
        while (e < n && col[e] == col[r])
            ++e;
        out.push_back((int)r);
        out.push_back((int)e);
        out.push_back(col[r]);
        out.push_back(kind);
        r = e;
    }
}

int problem_from_planewave(HostProblem &hp, const fnec_seg *segs, size_t n, const fnec_c64 *rhs,
                           size_t nrhs, const double *cos_vec, const double *sin_vec,
                           const size_t *wf, const size_t *wl, size_t nw, const size_t *path_of,
                           const size_t *free_ends, size_t nfree, double freq_hz,
                           const fnec_ground *ground)
{
    if (!sin_vec || nrhs == 0 || nrhs > ((size_t)1 << 20)) {
        set_error("plane wave: null sin vector or no right-hand side");
        return FNEC_INVALID_ARG;
    }
    std::vector<int> grp, ends;
    int groups = 0;
    int st = variant_rows(n, wf, wl, nw, path_of, free_ends, nfree, grp, groups, ends);
    if (st)
        return st;
    // columns n + 2 g (-cos) and n + 2 g + 1 (-sin) of group g
    std::vector<int> ccol(n), scol(n);
    for (size_t r = 0; r < n; ++r) {
        ccol[r] = 2 * grp[r];
        scol[r] = 2 * grp[r] + 1;
    }
    // I = 0 rows: per range, first then last end (linear.rs:946-949); per
    // path, the free ends in order (linear.rs:1069-1071)
    std::vector<int> cons;
    std::vector<double> sign;
    for (int e : ends) {
        cons.push_back(e);
        cons.push_back(-1);
        sign.push_back(1.0);
    }
    st = build_problem(hp, segs, n, rhs, cos_vec, ccol, 2 * groups, cons, sign, freq_hz, ground);
    if (st)
        return st;
    hp.variant = 1;
    hp.nrhs = (int)nrhs;
    hp.sinv.assign(sin_vec, sin_vec + n);
    hp.rhs.resize(2 * n * nrhs);
    for (size_t i = 0; i < n * nrhs; ++i) {
        hp.rhs[2 * i] = rhs[i].re;
        hp.rhs[2 * i + 1] = rhs[i].im;
    }
    xruns_from(scol, 1, hp.xruns);
    hp.rowent.resize(3 * n);
    for (size_t r = 0; r < n; ++r) {
        hp.rowent[3 * r] = ccol[r];
        hp.rowent[3 * r + 1] = scol[r];
        hp.rowent[3 * r + 2] = -1;
    }
    return FNEC_OK;
}

int problem_from_current_source(HostProblem &hp, const fnec_seg *segs, size_t n,
                                const fnec_c64 *shape, const double *cos_vec, size_t src,
                                fnec_c64 i0, const size_t *wf, const size_t *wl, size_t nw,
                                const size_t *path_of, const size_t *free_ends, size_t nfree,
                                double freq_hz, const fnec_ground *ground)
{
#pragma clang fp contract(off)
    if (!shape || src >= n) {
        set_error("current source: null shape or source segment outside the segment list");
        return FNEC_INVALID_ARG;
    }
    std::vector<int> grp, ends;
    int groups = 0;
    int st = variant_rows(n, wf, wl, nw, path_of, free_ends, nfree, grp, groups, ends);
    if (st)
        return st;
    // constraint rows, each weighted by 1e6: the I = 0 ends, then the source
    // row (linear.rs:1163-1172, 1296-1302)
    const double cw = 1.0e6;
    std::vector<int> cons;
    std::vector<double> sign;
    for (int e : ends) {
        cons.push_back(e);
        cons.push_back(-1);
        sign.push_back(1.0);
    }
    cons.push_back((int)src);
    cons.push_back(-1);
    sign.push_back(1.0);
    // the forcing of the segment rows is zero; -shape is the port-voltage
    // column n + groups
    std::vector<fnec_c64> zero(n, fnec_c64{0.0, 0.0});
    st = build_problem(hp, segs, n, zero.data(), cos_vec, grp, groups + 1, cons, sign, freq_hz, ground);
    if (st)
        return st;
    hp.variant = 2;
    hp.shape.resize(2 * n);
    for (size_t i = 0; i < n; ++i) {
        hp.shape[2 * i] = shape[i].re;
        hp.shape[2 * i + 1] = shape[i].im;
    }
    hp.xruns = {0, (int)n, groups, 2};
    hp.rowent.resize(3 * n);
    for (size_t r = 0; r < n; ++r) {
        hp.rowent[3 * r] = grp[r];
        hp.rowent[3 * r + 1] = -1;
        hp.rowent[3 * r + 2] = groups;
    }
    hp.cons_aval.assign(hp.ncons, cw);
    hp.cons_y.assign(2 * (size_t)hp.ncons, 0.0);
    // y = (1e6, 0) * i0 in num-complex's product form
    hp.cons_y[2 * (hp.ncons - 1)] = cw * i0.re - 0.0 * i0.im;
    hp.cons_y[2 * (hp.ncons - 1) + 1] = cw * i0.im + 0.0 * i0.re;
    return FNEC_OK;
}

int problem_add_stamps(HostProblem &hp, const fnec_stamp *stamps, size_t n_stamps, size_t n)
{
    hp.stamps.clear();
    if (!n_stamps)
        return FNEC_OK;
    if (!stamps) {
        set_error("null stamp list");
        return FNEC_INVALID_ARG;
    }
    // grouped by entry, in call order within an entry (k_stamp_z adds each
    // group in order; distinct entries are independent)
    std::vector<size_t> order(n_stamps);
    for (size_t q = 0; q < n_stamps; ++q) {
        if (stamps[q].row >= n || stamps[q].col >= n) {
            set_error("stamp outside the segment list");
            return FNEC_INVALID_ARG;
        }
        order[q] = q;
    }
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return stamps[a].col != stamps[b].col ? stamps[a].col < stamps[b].col : stamps[a].row < stamps[b].row;
    });
    hp.stamps.reserve(4 * n_stamps);
    for (size_t q : order) {
        const fnec_stamp &t = stamps[q];
        hp.stamps.push_back((double)t.row);
        hp.stamps.push_back((double)t.col);
        hp.stamps.push_back(t.delta.re);
        hp.stamps.push_back(t.delta.im);
    }
    return FNEC_OK;
}

// ---------------------------------------------------------------------------
// batch planning
// ---------------------------------------------------------------------------

struct Plan {
    std::vector<Prob> probs;
    int upd_tm = kTile32; // row tile of the FP32 trailing update (64 or 128)
    // work lists, concatenated into one int buffer
    std::vector<int> ints;
    size_t fill_off = 0, fill_cnt = 0; // fill items in the generated region (item units)
    size_t stamp_off = 0, stamp_cnt = 0; // problems with Z stamps (k_stamp_z)
    // Regular work lists (problem, y, z) generated on the device by
    // k_gen_items from one segment (first item, problem, ny, nz) per problem
    // and launch: the host builds O(problems) records instead of O(tiles) items.
    std::vector<int4> gensegs;
    size_t gen_items = 0, off_gen = 0, gensegs_off = 0;
    size_t gram_off = 0, gram_cnt = 0;
    size_t cols_off = 0, cols_cnt = 0;
    // plain path in the reference's operation order (FNEC_PREC_FP64); the
    // FP64 MFMA path (FNEC_PREC_FP64_MFMA) leaves it false
    bool exact = false;
    bool fill_large = false; // some fill phase may reach 9e4: k_fill_z_large
    bool given_z = false;  // Z uploaded by the caller (debug entry): no fill launch
    bool given_a = false;  // [A | b] uploaded by the caller: elimination only
    size_t xcols_off = 0, xcols_cnt = 0;
    size_t xrhs_off = 0, xrhs_cnt = 0;   // k_gx_rhs tiles (problems with several right-hand sides)
    size_t xext_off = 0, xext_cnt = 0;   // k_gx_extra task blocks
    size_t bs_off = 0, bs_cnt = 0;       // back substitution items (problem, right-hand side)
    size_t all_off = 0, all_cnt = 0;
    size_t res_off = 0, res_cnt = 0;
    std::vector<LuStep> steps;
    // algorithmic work per launch (profiling)
    double fill_flops = 0, gram_flops = 0, backsub_flops = 0;
    std::vector<double> step_upd_flops, step_panel_flops, step_trsm_flops;
    // arena layout
    size_t bytes = 0;
    size_t a_begin = 0, a_end = 0; // byte range of the A planes (zeroed per run)
    std::vector<size_t> off_seg, off_z, off_a, off_cos, off_rhs, off_runs, off_cons, off_sign,
        off_ipiv, off_info, off_x, off_stats, off_scr, off_panel, off_pmeta, off_selfln, off_stamps;
    // plane-wave / current-source inputs: sinv, shape, xruns, rowent, cons_aval, cons_y
    struct VarOff {
        size_t sinv, shape, xruns, rowent, caval, cy;
    };
    std::vector<VarOff> var;
    size_t off_probs = 0, off_ints = 0, off_lud = 0;
    std::vector<LuDesc<double>> lud;
    // mixed-precision path
    bool mixed = false;
    bool lu_only = false; // mixed path: stop after the FP32 factorisation
    struct MixOff {
        size_t fr, fpr, rowcol, fh, lk, ym, ls, vr, vz, vp, vq, vt, vv, part, colsq, pcg, pstate;
        size_t perm, tws, tflags, dinv, gx, linv, h16;
        int nchunk;
    };
    std::vector<MixOff> mix;
    struct Outer {                 // one 64-column step of the two-level FP32 LU
        int P;
        LuStep a, b;               // inner panels at P and P + 32
        bool has_b;
        size_t ua_off, ua_cnt;     // narrow update of columns [P + 32, P + 64)
        size_t li_off, li_cnt;     // problems whose L11^{-1} this step needs (order > P + 64)
        size_t ot_off, ot_cnt;     // interchanges left of P, and of + L11 solve for columns [P + 64, n)
        size_t on_off, on_cnt;     // outer update (K = 64) of the next step's columns [P + 64, P + 128)
        size_t of_off, of_cnt;     // outer update of the columns [P + 128, n)
        // K = 128 passes: a deferring step leaves its update of [P + 128, n)
        // to the next step; that combining step updates its own U12 rows
        // [P, P + 64) with the deferred factors (thin list, K = 64) between
        // its interchanges and its L11^{-1} product, then runs on and of with
        // K = 128 from P - 64
        bool defer, combine;
        size_t th_off, th_cnt;
        size_t sp_off, sp_cnt;     // combining step: (problem, 64-block, chunk * 2 + kind) of k_split_f16x2
        double f_panel, f_trsm, f_ua, b_ua, f_on, b_on, f_of, b_of, f_th, b_th;
    };
    std::vector<Outer> outer32;
    std::vector<LuDesc<float>> lud32;
    size_t off_lud32 = 0;
    size_t fsol_off = 0, fsol_cnt = 0, vsol_off = 0, vsol_cnt = 0, mvz_off = 0, mvz_cnt = 0;
    size_t flags_begin = 0, flags_end = 0, off_tickets = 0; // zeroed at the start of a batch
    size_t dinv_off = 0, dinv_cnt = 0;
    size_t zs_off = 0, zs_cnt = 0; // k_zsolve items (problem, member, G) for n <= g_solve1_max_n
    size_t zs_lds = 0;             // dynamic LDS of a k_zsolve launch
    size_t in_begin = 0, io_begin = 0, io_end = 0; // staged upload [in_begin, io_end), read-back [io_begin, io_end)
    size_t mvfin_off = 0, mvfin_cnt = 0, mvzh_off = 0, mvzh_cnt = 0;
    double zsolve_flops = 0, mv_flops = 0, zsolve_bytes = 0, mv_bytes = 0;
    std::vector<long> it_zs1, it_zs0, it_mv; // profiling records per CG iteration
    int iters_launched = 0;
    double fsol_flops = 0, fsol_bytes = 0;
    // k_fsetup launches: one per group of 16 constraint rows (group g covers
    // every problem with more than 16 g rows); items (problem, member, G, g)
    std::vector<size_t> fs_off, fs_cnt;
    size_t fs_lds = 0;
    size_t off_flags = 0;
};

constexpr int kMixMaxIt = 12;        // CG iterations at most
constexpr int kMixFirstChunk = 4;    // iterations launched before the first convergence check
constexpr int kMixNextChunk = 2;
constexpr int kFsRhsHost = 16; // constraint rows per k_fsetup gang
// launches with a ticket and an epoch: the set-up groups, then two kernels x
// two solves per CG iteration
constexpr int kMaxTrsvLaunches = kMixMaxCons / kFsRhsHost + 4 * kMixMaxIt + 1;
constexpr int kSolve1MaxNHost = 8192;               // largest n the one-workgroup solve holds in LDS

// Tunables (fnec_hip_set_tunable): the autotuner's knobs.
static int g_solve1_max_n = kSolve1MaxNHost; // CG solves of n <= this run one workgroup per problem
static int g_mix_first_chunk = kMixFirstChunk;
// Gang size of a CG solve: G = clamp(floor(n^2 / gang_n^2), 1, min(gang_max, blocks)),
// so a problem's solve time shrinks towards the chain latency of its blocks
// instead of one workgroup streaming the whole factor.
static int g_update_tile_m = 64; // row tile of the FP32 trailing update (64 or 128)
static int g_update_k128 = 1;    // K = 128 trailing-update passes (with 64-row tiles)
// K = 128 passes through the LDS ring (k_lu_update_f16x2s, fastest alone) or
// with register-direct operands (k_lu_update_f16x2d, the default: the ring's
// persistent grid holds every CU's LDS, which delays the other stream's
// panel kernels that overlap the update in a serving window)
static int g_update_ring = 0;
static int g_fill_centred = 1;
static int g_mix_tol_exp = 14; // CG stops at |h - (M^H M + lambda) x| <= 10^-exp |h|
static int g_batch_split_n = 1024; // serving batches: N classes split at this size (0: one class)
static int g_gang_n = 512;
static int g_exact_backsub = 0; // plain path: back substitution in the reference's order
static int g_gang_max = 4;
static int gang_size(int n)
{
    if (g_gang_n <= 0)
        return 1;
    const long q = (long)n * n / ((long)g_gang_n * g_gang_n);
    const int nb = (n + 63) / 64;
    return (int)std::max<long>(1, std::min<long>(q, std::min(g_gang_max, nb)));
}

static size_t take(size_t &cur, size_t bytes)
{
    size_t o = (cur + 255) & ~(size_t)255;
    cur = o + bytes;
    return o;
}

size_t problem_bytes(int n, int s, int ncons, int nruns, int precision)
{
    const size_t ldz = (size_t)roundup(n, 64), lda = (size_t)roundup(s, 64);
    const size_t nn = n, ss = s, nc = std::max(ncons, 1), w = s - n;
    size_t b = 0;
    b += 2 * 8 * ldz * nn + 256;                                     // Z (FP64)
    b += nn * 8 * kSegStride + 8 * nn + 16 * nn + 12 * nruns + 16 * nc + 4 * ss + 2048;
    b += 16 * ss + 64 + 32 + 32 * ((nn + 255) / 256 + 1);            // x, info, stats, residual
    b += 16 * kPanel + 16 + 256;                                     // panel metadata
    if (precision == FNEC_PREC_MIXED) {
        const size_t nchunk = (nn + kMvChunk - 1) / kMvChunk;
        b += 2 * 4 * ldz * nn + 2 * 4 * ldz * kPanel;                // FP32 factors + panel
        b += 4 * nn + 16 * nn * nc + 16 * nc * nc + 16 * nc * w + 16 * w * w;
        b += 4 * 16 * ss + 16 * (nn + nc) + 16 * nn + 16 * nchunk * nn + 8 * nn + 80;
        b += 4 * nn + 16 * nn * (nc + 1) + 8 * (nc + 1) * ((nn + 63) / 64);  // perm, solves
        b += ((nn + 63) / 64) * 2 * 64 * 64 * 8 + (2 * 64 * 64 + 2 * 32 * 32) * 4; // block inverses
        b += 8 * 2 * 256 * 16 * ((nn + 63) / 64);                    // gang-published blocks
        b += 2048 * ldz;                                             // split K = 128 operands
        b += 24 * 256;                                               // alignment of 24 buffers
    } else {
        b += 2 * 8 * lda * (ss + 1) + 256;                           // [G | M^H y]
        b += 2 * 8 * lda * kPanel + 256;                             // panel workspace
    }
    return b;
}

// Work lists of the batched LU of matrices with orders ms[b] carrying ncs[b]
// columns: one LuStep per panel position, plus algorithmic flops per launch.
static void lu_plan(std::vector<int> &ints, const std::vector<int> &ms, const std::vector<int> &ncs,
                    std::vector<LuStep> &steps, std::vector<double> &upd_flops,
                    std::vector<double> &panel_flops, std::vector<double> &trsm_flops,
                    bool swap_left = false, int utile = kTile, std::vector<double> *upd_bytes = nullptr,
                    double elem_bytes = 16.0, int utile_n = 0)
{
    if (utile_n <= 0)
        utile_n = utile;
    auto align4 = [&]() {
        while (ints.size() % 4)
            ints.push_back(0);
    };
    auto push2 = [&](int a, int b) {
        ints.push_back(a);
        ints.push_back(b);
    };
    const size_t np = ms.size();
    int smax = 0;
    for (int m : ms)
        smax = std::max(smax, m);
    steps.clear();
    upd_flops.clear();
    panel_flops.clear();
    trsm_flops.clear();
    for (int p = 0; p < smax; p += kPanel) {
        LuStep st;
        st.p = p;
        align4();
        st.panel_off = ints.size();
        for (size_t b = 0; b < np; ++b)
            if (ms[b] > p)
                ints.push_back((int)b);
        st.panel_cnt = ints.size() - st.panel_off;
        align4();
        st.trsm_off = ints.size();
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (s <= p)
                continue;
            const int jb = std::min(kPanel, s - p);
            const int ncols = ncs[b] - p - jb;
            for (int y = 0; y < (ncols + 63) / 64; ++y)
                push2((int)b, y);
            if (swap_left)
                for (int y = 0; y < (p + 63) / 64; ++y)
                    push2((int)b, -1 - y);
        }
        st.trsm_cnt = (ints.size() - st.trsm_off) / 2;
        align4();
        st.upd_off = ints.size();
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (s <= p)
                continue;
            const int jb = std::min(kPanel, s - p);
            const int nrows = s - p - jb, ncols = ncs[b] - p - jb;
            if (nrows <= 0 || ncols <= 0)
                continue;
            for (int y = 0; y < (nrows + utile - 1) / utile; ++y)
                for (int z = 0; z < (ncols + utile_n - 1) / utile_n; ++z) {
                    ints.push_back((int)b);
                    ints.push_back(y);
                    ints.push_back(z);
                    ints.push_back(0);
                }
        }
        st.upd_cnt = (ints.size() - st.upd_off) / 4;
        align4();
        st.rows_off = ints.size();
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (s <= p)
                continue;
            const int jb = std::min(kPanel, s - p);
            for (int y = 0; y < (s - p - jb + 255) / 256; ++y)
                push2((int)b, y);
        }
        st.rows_cnt = (ints.size() - st.rows_off) / 2;
        align4();
        st.commit_off = ints.size();
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (s <= p)
                continue;
            for (int y = 0; y < (s - p + 255) / 256; ++y)
                push2((int)b, y);
        }
        st.commit_cnt = (ints.size() - st.commit_off) / 2;
        double fu = 0, fp = 0, ft = 0, bu = 0;
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (s <= p)
                continue;
            const int jb = std::min(kPanel, s - p);
            const double nrows = s - p - jb, ncols = ncs[b] - p - jb, m = s - p;
            fu += 8.0 * nrows * ncols * jb;
            fp += 8.0 * m * jb * jb / 2.0;
            ft += 8.0 * ncols * jb * jb / 2.0;
            // trailing block read and written once, panel and U12 read once
            if (nrows > 0 && ncols > 0)
                bu += elem_bytes * (2.0 * nrows * ncols + nrows * jb + jb * ncols);
        }
        if (upd_bytes)
            upd_bytes->push_back(bu);
        upd_flops.push_back(fu);
        panel_flops.push_back(fp);
        trsm_flops.push_back(ft);
        steps.push_back(st);
    }
}

// Two-level plan of the FP32 LU (mixed path): 64-column outer steps made of
// two 32-column panels, so that the trailing update runs with K = 64 and reads
// and writes the trailing matrix once per 64 columns.
// Trailing-update items run in groups of this many row tiles, column tiles
// outer within a group: the group's L21 rows and each U12 column tile are
// read from HBM about once per group instead of once per row tile.
constexpr int kUpdGroupRows = 8;

static void lu_plan_outer(std::vector<int> &ints, const std::vector<int> &ms, std::vector<Plan::Outer> &out,
                          std::vector<int4> &gensegs, size_t &gen_items, int tm, bool k128)
{
    auto align4 = [&]() {
        while (ints.size() % 4)
            ints.push_back(0);
    };
    auto push2 = [&](int a, int b) {
        ints.push_back(a);
        ints.push_back(b);
    };
    auto push4 = [&](int a, int b, int c, int d) {
        ints.push_back(a);
        ints.push_back(b);
        ints.push_back(c);
        ints.push_back(d);
    };
    const size_t np = ms.size();
    int smax = 0;
    for (int m : ms)
        smax = std::max(smax, m);
    out.clear();
    auto panel_lists = [&](LuStep &st, int p) {
        align4();
        st.panel_off = ints.size();
        for (size_t b = 0; b < np; ++b)
            if (ms[b] > p)
                ints.push_back((int)b);
        st.panel_cnt = ints.size() - st.panel_off;
        align4();
        st.rows_off = ints.size();
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (s <= p)
                continue;
            const int jb = std::min(kPanel, s - p);
            for (int y = 0; y < (s - p - jb + 255) / 256; ++y)
                push2((int)b, y);
        }
        st.rows_cnt = (ints.size() - st.rows_off) / 2;
        align4();
        st.commit_off = ints.size();
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (s <= p)
                continue;
            for (int y = 0; y < (s - p + 255) / 256; ++y)
                push2((int)b, y);
        }
        st.commit_cnt = (ints.size() - st.commit_off) / 2;
    };
    const int nsteps = (smax + 2 * kPanel - 1) / (2 * kPanel);
    for (int P = 0; P < smax; P += 2 * kPanel) {
        Plan::Outer o{};
        o.P = P;
        const int k = P / (2 * kPanel);
        // K = 128 pairs: even steps defer (when a next step exists), odd steps combine
        o.combine = k128 && (k & 1);
        o.defer = k128 && !(k & 1) && k + 1 < nsteps;
        o.a.p = P;
        o.b.p = P + kPanel;
        o.has_b = false;
        // inner panel a; its interchanges reach the columns left of P in the
        // outer step's launch (ot), after the previous trailing update
        panel_lists(o.a, P);
        align4();
        o.a.trsm_off = ints.size();
        for (size_t b = 0; b < np; ++b)
            if (ms[b] > P + kPanel)
                push2((int)b, 0); // columns [P + 32, min(P + 64, s))
        o.a.trsm_cnt = (ints.size() - o.a.trsm_off) / 2;
        align4();
        o.ua_off = ints.size();
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (s <= P + kPanel)
                continue;
            for (int y = 0; y < (s - P - kPanel + kTile32 - 1) / kTile32; ++y)
                push4((int)b, y, 0, 0);
        }
        o.ua_cnt = (ints.size() - o.ua_off) / 4;
        // inner panel b; its interchanges on panel a's columns [P, P + 32)
        o.has_b = smax > P + kPanel;
        if (o.has_b) {
            panel_lists(o.b, P + kPanel);
            align4();
            o.b.trsm_off = ints.size();
            for (size_t b = 0; b < np; ++b)
                if (ms[b] > P + kPanel)
                    push2((int)b, -1 - P / 64);
            o.b.trsm_cnt = (ints.size() - o.b.trsm_off) / 2;
        }
        align4();
        o.li_off = ints.size();
        for (size_t b = 0; b < np; ++b)
            if (ms[b] > P + 2 * kPanel)
                ints.push_back((int)b);
        o.li_cnt = ints.size() - o.li_off;
        // outer step: interchanges left of P, interchanges + L11 solve right of P + 64
        align4();
        o.ot_off = ints.size();
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (s <= P)
                continue;
            for (int y = 0; y < P / 64; ++y)
                push2((int)b, -1 - y);
            for (int y = 0; y < (s - P - 2 * kPanel + kTrsm64Cols - 1) / kTrsm64Cols && s > P + 2 * kPanel; ++y)
                push2((int)b, y);
        }
        o.ot_cnt = (ints.size() - o.ot_off) / 2;
        // combining step: L21 rows and U12 columns [P + 64, n) in blocks of
        // 64, four 32-deep chunks, two kinds
        align4();
        o.sp_off = ints.size();
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (!o.combine || s <= P + 2 * kPanel)
                continue;
            for (int y = 0; y < (s - P - 2 * kPanel + 63) / 64; ++y)
                for (int z = 0; z < 8; ++z)
                    push4((int)b, y, z, 0);
        }
        o.sp_cnt = (ints.size() - o.sp_off) / 4;
        // outer update items (problem, row tile, column tile), generated on the
        // device: the next step's 64 columns, then the rest
        o.on_off = gen_items;
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (s <= P + 2 * kPanel)
                continue;
            const int ny = (s - P - 2 * kPanel + tm - 1) / tm;
            gensegs.push_back(make_int4((int)gen_items, (int)b, ny, 1));
            gen_items += (size_t)ny;
        }
        o.on_cnt = gen_items - o.on_off;
        o.of_off = gen_items;
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (s <= P + 4 * kPanel || o.defer)
                continue;
            const int ny = (s - P - 2 * kPanel + tm - 1) / tm;
            const int nz = (s - P - 4 * kPanel + kTile32N - 1) / kTile32N;
            gensegs.push_back(make_int4((int)gen_items, (int)b, ny, nz | kUpdGroupRows << 24));
            gen_items += (size_t)ny * nz;
        }
        o.of_cnt = gen_items - o.of_off;
        // combining step: rows [P, P + 64) of the columns [P + 64, n), in
        // 64-row tiles whatever tm
        o.th_off = gen_items;
        for (size_t b = 0; b < np; ++b) {
            const int s = ms[b];
            if (!o.combine || s <= P + 2 * kPanel)
                continue;
            const int nz = (s - P - 2 * kPanel + kTile32N - 1) / kTile32N;
            gensegs.push_back(make_int4((int)gen_items, (int)b, 1, nz));
            gen_items += (size_t)nz;
        }
        o.th_cnt = gen_items - o.th_off;
        // algorithmic work of the launches (8 flops per complex multiply-add;
        // 8 bytes per FP32 complex element)
        const double kd = o.combine ? 128.0 : 64.0; // depth of the on / of passes
        for (size_t b = 0; b < np; ++b) {
            const double s = ms[b];
            if (s <= P)
                continue;
            const double ja = std::min<double>(kPanel, s - P), jbb = std::max(0.0, std::min<double>(kPanel, s - P - kPanel));
            o.f_panel += 8.0 * (s - P) * ja * ja / 2.0 + 8.0 * (s - P - ja) * jbb * jbb / 2.0;
            const double rows_a = s - P - ja, cols_a = jbb;
            o.f_ua += 8.0 * rows_a * cols_a * ja;
            o.b_ua += 8.0 * (2.0 * rows_a * cols_a + rows_a * ja + ja * cols_a);
            const double rem = std::max(0.0, s - P - 2.0 * kPanel);
            const double cn = std::min(64.0, rem), cf = std::max(0.0, rem - 64.0);
            o.f_trsm += 8.0 * cols_a * ja * ja / 2.0 + 8.0 * rem * 64.0 * 64.0 / 2.0;
            o.f_on += 8.0 * rem * cn * kd;
            o.b_on += 8.0 * (2.0 * rem * cn + rem * kd + kd * cn);
            if (!o.defer) {
                o.f_of += 8.0 * rem * cf * kd;
                o.b_of += 8.0 * (2.0 * rem * cf + rem * kd + kd * cf);
            }
            if (o.combine && s > P + 2.0 * kPanel) {
                const double ct = s - P - 2.0 * kPanel; // rows [P, P + 64) x these columns, K = 64
                o.f_th += 8.0 * 64.0 * ct * 64.0;
                o.b_th += 8.0 * (2.0 * 64.0 * ct + 64.0 * 64.0 + 64.0 * ct);
            }
        }
        out.push_back(o);
    }
}

static void plan_build(Plan &pl, const std::vector<const HostProblem *> &hps, bool with_solve)
{
    const size_t np = hps.size();
    pl.probs.resize(np);
    size_t cur = 0;
    pl.off_seg.resize(np);
    pl.off_z.resize(np);
    pl.off_a.resize(np);
    pl.off_cos.resize(np);
    pl.off_rhs.resize(np);
    pl.off_runs.resize(np);
    pl.off_cons.resize(np);
    pl.off_sign.resize(np);
    pl.off_ipiv.resize(np);
    pl.off_info.resize(np);
    pl.off_x.resize(np);
    pl.off_stats.resize(np);
    pl.off_scr.resize(np);
    pl.off_panel.resize(np);
    pl.off_pmeta.resize(np);
    pl.off_selfln.resize(np);
    pl.off_stamps.resize(np);
    pl.var.assign(np, Plan::VarOff{0, 0, 0, 0, 0, 0});
    pl.fill_large = false;
    for (const HostProblem *h : hps)
        if (!(h->phase_max < 9e4))
            pl.fill_large = true;
    pl.mix.clear();
    // A planes first, contiguous, so one zero-fill covers them
    pl.a_begin = 0;
    for (size_t b = 0; b < np; ++b) {
        const HostProblem &h = *hps[b];
        size_t lda = (size_t)roundup(h.s, 64);
        pl.off_a[b] = with_solve && !pl.mixed ? take(cur, 2 * 8 * lda * (h.s + h.nrhs)) : 0;
    }
    pl.a_end = cur;
    for (size_t b = 0; b < np; ++b) {
        const HostProblem &h = *hps[b];
        size_t ldz = (size_t)roundup(h.n, 64);
        pl.off_z[b] = take(cur, 2 * 8 * ldz * h.n);
        pl.off_ipiv[b] = take(cur, 4 * (size_t)h.s);
        pl.off_scr[b] = take(cur, 32 * ((size_t)(h.n + 255) / 256 + 1));
        if (with_solve) {
            if (!pl.mixed)
                pl.off_panel[b] = take(cur, 2 * 8 * (size_t)roundup(h.s, 64) * kPanel);
            pl.off_pmeta[b] = take(cur, 8 * kPanel + 8 * kPanel + 16);
        }
        if (with_solve && pl.mixed) {
            Plan::MixOff m;
            const size_t n = h.n, nc = std::max(h.ncons, 1), w = h.w, sz = h.s;
            m.nchunk = (h.n + kMvChunk - 1) / kMvChunk;
            m.fr = take(cur, 2 * 4 * ldz * n);
            m.fpr = take(cur, 2 * 4 * ldz * kPanel);
            m.fh = take(cur, 16 * n * nc);
            m.lk = take(cur, 16 * nc * nc);
            m.ym = take(cur, 16 * nc * w);
            m.ls = take(cur, 16 * w * w);
            m.vr = take(cur, 16 * sz);
            m.vz = take(cur, 16 * sz);
            m.vp = take(cur, 16 * sz);
            m.vq = take(cur, 16 * sz);
            m.vt = take(cur, 16 * (n + nc));
            m.vv = take(cur, 16 * n);
            m.part = take(cur, 16 * (size_t)m.nchunk * n);
            m.colsq = take(cur, 8 * n);
            m.perm = take(cur, 4 * n);
            m.dinv = take(cur, ((n + 63) / 64) * 2 * 64 * 64 * 8);
            m.linv = take(cur, (2 * 64 * 64 + 2 * 32 * 32) * 4); // L11^{-1}, then the panel's U_D^{-1}
            m.h16 = take(cur, 2048 * ldz);                          // split L21 / U12 of a combining step
            m.tws = take(cur, 16 * n * (nc + 1));
            pl.mix.push_back(m);
        }
    }
    if (with_solve && pl.mixed) {
        // block-completion flags of the triangular solves and the launch
        // tickets, contiguous so that one fill clears them per batch
        pl.flags_begin = cur = (cur + 255) & ~(size_t)255;
        for (size_t b = 0; b < np; ++b) {
            const size_t nb = (hps[b]->n + 63) / 64, nc = std::max(hps[b]->ncons, 0);
            pl.mix[b].tflags = take(cur, 4 * (nc + 1) * 2 * nb);
            pl.mix[b].gx = take(cur, 8 * 2 * nb * 64 * kFsRhsHost * 4); // 16 vectors (set-up), 1 (CG)
        }
        pl.off_tickets = take(cur, 8 * kMaxTrsvLaunches);
        pl.flags_end = cur = (cur + 7) & ~(size_t)7;
    }
    // host inputs, contiguous: one staged copy uploads them (with the
    // problem table, the work lists and the zeroed result words below)
    pl.in_begin = cur = (cur + 255) & ~(size_t)255;
    for (size_t b = 0; b < np; ++b) {
        const HostProblem &h = *hps[b];
        pl.off_seg[b] = take(cur, (size_t)h.n * 8 * kSegStride);
        pl.off_selfln[b] = take(cur, 8 * (size_t)h.n);
        pl.off_stamps[b] = take(cur, 8 * h.stamps.size() + 8);
        pl.off_cos[b] = take(cur, 8 * (size_t)h.n);
        pl.off_rhs[b] = take(cur, 16 * (size_t)h.n * h.nrhs);
        pl.off_runs[b] = take(cur, 4 * h.runs.size() + 4);
        pl.off_cons[b] = take(cur, 4 * h.cons.size() + 4);
        pl.off_sign[b] = take(cur, 8 * h.cons_sign.size() + 8);
        if (h.variant) {
            Plan::VarOff &v = pl.var[b];
            v.sinv = take(cur, 8 * h.sinv.size() + 8);
            v.shape = take(cur, 8 * h.shape.size() + 8);
            v.xruns = take(cur, 4 * h.xruns.size() + 4);
            v.rowent = take(cur, 4 * h.rowent.size() + 4);
            v.caval = take(cur, 8 * h.cons_aval.size() + 8);
            v.cy = take(cur, 8 * h.cons_y.size() + 8);
        }
        if (with_solve && pl.mixed)
            pl.mix[b].rowcol = take(cur, 4 * (size_t)h.n);
    }
    // work lists
    auto push4 = [&](int a, int b, int c, int d) {
        pl.ints.push_back(a);
        pl.ints.push_back(b);
        pl.ints.push_back(c);
        pl.ints.push_back(d);
    };
    auto push2 = [&](int a, int b) {
        pl.ints.push_back(a);
        pl.ints.push_back(b);
    };
    // every list starts on a 16-byte boundary so int2 / int4 loads are aligned
    auto align4 = [&]() {
        while (pl.ints.size() % 4)
            pl.ints.push_back(0);
    };
    pl.ints.clear();
    pl.gensegs.clear();
    pl.gen_items = 0;
    align4();
    pl.fill_off = pl.gen_items;
    for (size_t b = 0; b < np; ++b) {
        const int rt = (hps[b]->n + 63) / 64, ct = (hps[b]->n + 15) / 16;
        pl.gensegs.push_back(make_int4((int)pl.gen_items, (int)b, rt, ct));
        pl.gen_items += (size_t)rt * ct;
    }
    pl.fill_cnt = pl.gen_items - pl.fill_off;
    align4();
    pl.stamp_off = pl.ints.size();
    for (size_t b = 0; b < np; ++b)
        if (!hps[b]->stamps.empty())
            pl.ints.push_back((int)b);
    pl.stamp_cnt = pl.ints.size() - pl.stamp_off;
    pl.fill_flops = pl.gram_flops = pl.backsub_flops = 0;
    for (size_t b = 0; b < np; ++b) {
        const double n = hps[b]->n, s = hps[b]->s;
        // fill unit: one Gauss-Legendre point = 22 flops + 5 special ops,
        // 8 points per element and image (counted as 27 operations)
        pl.fill_flops += n * n * 8.0 * (hps[b]->ground ? 2.0 : 1.0) * 27.0;
        pl.gram_flops += 4.0 * n * n * n;      // Hermitian-minimal Z^H Z
        pl.backsub_flops += 4.0 * s * s;       // 8 S^2 / 2 for the upper triangle
    }
    if (with_solve && pl.mixed) {
        align4();
        pl.all_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            pl.ints.push_back((int)b);
        pl.all_cnt = np;
        std::vector<int> ms(np);
        for (size_t b = 0; b < np; ++b)
            ms[b] = hps[b]->n;
        pl.upd_tm = g_update_tile_m;
        lu_plan_outer(pl.ints, ms, pl.outer32, pl.gensegs, pl.gen_items, pl.upd_tm,
                      g_update_k128 != 0 && pl.upd_tm == 64);
        // CG solves of problems above solve1_max_n (flag-chained k_trsv0/1):
        // items (problem, -1, phase, block), phase 0 blocks in forward order,
        // then phase 1 blocks in backward order, interleaved across problems
        // so that early tickets rarely wait
        auto trsv_list = [&](size_t &off, size_t &cnt) {
            align4();
            off = pl.ints.size();
            std::vector<int3> chains; // (problem, vector, blocks)
            int maxnb = 0;
            for (size_t b = 0; b < np; ++b) {
                const int nb = (hps[b]->n + 63) / 64;
                if (hps[b]->n <= g_solve1_max_n)
                    continue; // CG solves of this problem run in k_zsolve
                maxnb = std::max(maxnb, nb);
                chains.push_back(make_int3((int)b, -1, nb));
            }
            for (int ph = 0; ph < 2; ++ph)
                for (int st = 0; st < maxnb; ++st)
                    for (const int3 &c : chains)
                        if (st < c.z)
                            push4(c.x, c.y, ph, ph == 0 ? st : c.z - 1 - st);
            cnt = (pl.ints.size() - off) / 4;
        };
        align4();
        pl.dinv_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            for (int y = 0; y < (hps[b]->n + 63) / 64; ++y)
                push2((int)b, y);
        pl.dinv_cnt = (pl.ints.size() - pl.dinv_off) / 2;
        // set-up gangs: at most 8 owned blocks per member (16 vectors per row in LDS)
        pl.fs_off.clear();
        pl.fs_cnt.clear();
        pl.fs_lds = 0;
        pl.fsol_cnt = 0;
        {
            int maxg = 0;
            for (size_t b = 0; b < np; ++b)
                maxg = std::max(maxg, (hps[b]->ncons + kFsRhsHost - 1) / kFsRhsHost);
            std::vector<std::pair<int, int>> order; // (G, problem)
            for (size_t b = 0; b < np; ++b) {
                const int nb = (hps[b]->n + 63) / 64;
                order.push_back({std::max(gang_size(hps[b]->n), (nb + 7) / 8), (int)b});
            }
            std::stable_sort(order.begin(), order.end(), [&](const std::pair<int, int> &a, const std::pair<int, int> &c) {
                if (a.first != c.first)
                    return a.first > c.first;
                return hps[a.second]->n > hps[c.second]->n;
            });
            for (int g = 0; g < maxg; ++g) {
                align4();
                pl.fs_off.push_back(pl.ints.size());
                for (const auto &o : order) {
                    const HostProblem &h = *hps[o.second];
                    if (h.ncons <= g * kFsRhsHost)
                        continue;
                    const int G = o.first, nb = (h.n + 63) / 64;
                    for (int mbr = 0; mbr < G; ++mbr)
                        push4(o.second, mbr, G, g);
                    pl.fs_lds = std::max(pl.fs_lds, (size_t)16 * 64 * kFsRhsHost * ((nb + G - 1) / G));
                }
                pl.fs_cnt.push_back((pl.ints.size() - pl.fs_off.back()) / 4);
                pl.fsol_cnt += pl.fs_cnt.back();
            }
        }
        trsv_list(pl.vsol_off, pl.vsol_cnt);
        align4();
        // k_zsolve items, gangs first and larger problems first: a gang's
        // members take consecutive tickets and the longest solves start first
        pl.zs_off = pl.ints.size();
        pl.zs_lds = 0;
        {
            std::vector<std::pair<int, int>> order; // (G, problem)
            for (size_t b = 0; b < np; ++b)
                if (hps[b]->n <= g_solve1_max_n)
                    order.push_back({gang_size(hps[b]->n), (int)b});
            std::stable_sort(order.begin(), order.end(), [&](const std::pair<int, int> &a, const std::pair<int, int> &c) {
                if (a.first != c.first)
                    return a.first > c.first;
                return hps[a.second]->n > hps[c.second]->n;
            });
            for (const auto &o : order) {
                const int n = hps[o.second]->n, G = o.first;
                for (int mbr = 0; mbr < G; ++mbr)
                    push4(o.second, mbr, G, 0);
                const size_t lds = G == 1 ? 16 * (size_t)n : 16 * 64 * (size_t)(((n + 63) / 64 + G - 1) / G);
                pl.zs_lds = std::max(pl.zs_lds, lds);
            }
        }
        pl.zs_cnt = (pl.ints.size() - pl.zs_off) / 4;
        align4();
        pl.mvz_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            for (int y = 0; y < (hps[b]->n + kMvRows - 1) / kMvRows; ++y)
                for (int z = 0; z < pl.mix[b].nchunk; ++z)
                    push4((int)b, y, z, 0);
        pl.mvz_cnt = (pl.ints.size() - pl.mvz_off) / 4;
        align4();
        pl.mvfin_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            for (int y = 0; y < (hps[b]->n + kMvRows - 1) / kMvRows; ++y)
                push2((int)b, y);
        pl.mvfin_cnt = (pl.ints.size() - pl.mvfin_off) / 2;
        align4();
        pl.mvzh_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            for (int y = 0; y < (hps[b]->n + kMvCols - 1) / kMvCols; ++y)
                push2((int)b, y);
        pl.mvzh_cnt = (pl.ints.size() - pl.mvzh_off) / 2;
        align4();
        pl.res_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            for (int y = 0; y < (hps[b]->n + 255) / 256; ++y)
                push2((int)b, y);
        pl.res_cnt = (pl.ints.size() - pl.res_off) / 2;
        pl.off_flags = take(cur, 4 * np + 16);
        pl.zsolve_flops = pl.mv_flops = pl.zsolve_bytes = pl.mv_bytes = pl.fsol_flops = pl.fsol_bytes = 0;
        for (size_t b = 0; b < np; ++b) {
            const double n = hps[b]->n, r = hps[b]->r;
            pl.zsolve_flops += 8.0 * n * n;    // L and U (or U^H and L^H): n^2 CMAC
            pl.zsolve_bytes += 8.0 * n * n;    // both FP32 complex triangles read once
            pl.mv_flops += 2.0 * 8.0 * n * r;  // M p and M^H t
            pl.mv_bytes += 2.0 * 16.0 * n * n; // FP64 Z read twice
            pl.fsol_flops += 8.0 * n * n * hps[b]->ncons;
            pl.fsol_bytes += 8.0 * n * n * ((hps[b]->ncons + kFsRhsHost - 1) / kFsRhsHost);
        }
    } else if (with_solve) {
        align4();
        pl.gram_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b) {
            int nt = (hps[b]->n + kTile - 1) / kTile;
            for (int y = 0; y < nt; ++y)
                for (int z = 0; z <= y; ++z)
                    push4((int)b, y, z, 0);
        }
        pl.gram_cnt = (pl.ints.size() - pl.gram_off) / 4;
        align4();
        pl.cols_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            for (int y = 0; y < (hps[b]->n + 3) / 4; ++y)
                push2((int)b, y);
        pl.cols_cnt = (pl.ints.size() - pl.cols_off) / 2;
        align4();
        pl.xcols_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            for (int y = 0; y < (hps[b]->n + 255) / 256; ++y)
                push2((int)b, y);
        pl.xcols_cnt = (pl.ints.size() - pl.xcols_off) / 2;
        align4();
        pl.xrhs_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            if (hps[b]->nrhs > 1)
                for (int y = 0; y < (hps[b]->n + 63) / 64; ++y)
                    for (int z = 0; z < (hps[b]->nrhs + 63) / 64; ++z)
                        push4((int)b, y, z, 0);
        pl.xrhs_cnt = (pl.ints.size() - pl.xrhs_off) / 4;
        align4();
        pl.xext_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b) {
            const long h = hps[b]->s - hps[b]->n;
            const long tasks = h * (h + 1) / 2 + h * hps[b]->nrhs;
            for (long y = 0; y < (tasks + 255) / 256; ++y)
                push2((int)b, (int)y);
        }
        pl.xext_cnt = (pl.ints.size() - pl.xext_off) / 2;
        align4();
        pl.all_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            pl.ints.push_back((int)b);
        pl.all_cnt = np;
        std::vector<int> ms(np), ncs(np);
        for (size_t b = 0; b < np; ++b) {
            ms[b] = hps[b]->s;
            ncs[b] = hps[b]->s + hps[b]->nrhs;
        }
        lu_plan(pl.ints, ms, ncs, pl.steps, pl.step_upd_flops, pl.step_panel_flops,
                pl.step_trsm_flops);
        align4();
        pl.bs_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            for (int k = 0; k < hps[b]->nrhs; ++k)
                push2((int)b, k);
        pl.bs_cnt = (pl.ints.size() - pl.bs_off) / 2;
        align4();
        pl.res_off = pl.ints.size();
        for (size_t b = 0; b < np; ++b)
            for (int y = 0; y < (hps[b]->n + 255) / 256; ++y)
                push2((int)b, y);
        pl.res_cnt = (pl.ints.size() - pl.res_off) / 2;
    }
    pl.off_probs = take(cur, sizeof(Prob) * np);
    if (with_solve && !pl.mixed)
        pl.off_lud = take(cur, sizeof(LuDesc<double>) * np);
    if (with_solve && pl.mixed)
        pl.off_lud32 = take(cur, sizeof(LuDesc<float>) * np);
    align4();
    pl.gensegs_off = pl.ints.size();
    for (const int4 &g : pl.gensegs)
        push4(g.x, g.y, g.z, g.w);
    pl.off_ints = take(cur, 4 * pl.ints.size() + 16);
    // result words, packed (16-byte alignment): uploaded as zeros, read back
    // with one copy after the run
    pl.io_begin = cur = (cur + 255) & ~(size_t)255;
    auto take16 = [&](size_t bytes) {
        const size_t o = (cur + 15) & ~(size_t)15;
        cur = o + bytes;
        return o;
    };
    for (size_t b = 0; b < np; ++b) {
        pl.off_info[b] = take16(16);
        pl.off_stats[b] = take16(32);
        if (with_solve && pl.mixed) {
            pl.mix[b].pcg = take16(64);
            pl.mix[b].pstate = take16(16);
        }
        pl.off_x[b] = take16(16 * (size_t)hps[b]->s * hps[b]->nrhs);
    }
    pl.io_end = cur = (cur + 255) & ~(size_t)255;
    // generated work items (device only, written by k_gen_items)
    pl.off_gen = take(cur, 16 * pl.gen_items + 16);
    pl.bytes = cur + 256;
}

static int plan_upload(Plan &pl, const std::vector<const HostProblem *> &hps, char *base,
                       bool with_solve)
{
    hipStream_t st = LN().stream;
    const size_t np = hps.size();
    pl.lud.clear();
    pl.lud32.clear();
    for (size_t b = 0; b < np; ++b) {
        const HostProblem &h = *hps[b];
        Prob &P = pl.probs[b];
        std::memset(&P, 0, sizeof(P));
        P.n = h.n;
        P.s = h.s;
        P.r = h.r;
        P.w = h.w;
        P.ncons = h.ncons;
        P.nruns = (int)h.runs.size() / 3;
        P.ldz = roundup(h.n, 64);
        P.lda = roundup(h.s, 64);
        P.ground = h.ground;
        P.fill_centred = g_fill_centred;
        P.k = h.k;
        P.gam_re = h.gam_re;
        P.gam_im = h.gam_im;
        P.seg = (const double *)(base + pl.off_seg[b]);
        P.selfln = (const double *)(base + pl.off_selfln[b]);
        P.stamps = (const double *)(base + pl.off_stamps[b]);
        P.nstamps = (int)(h.stamps.size() / 4);
        P.zr = (double *)(base + pl.off_z[b]);
        P.zi = P.zr + (size_t)P.ldz * h.n;
        if (with_solve && !pl.mixed) {
            P.ar = (double *)(base + pl.off_a[b]);
            P.ai = P.ar + (size_t)P.lda * (h.s + h.nrhs);
        }
        P.nrhs = h.nrhs;
        P.variant = h.variant;
        if (h.variant) {
            const Plan::VarOff &v = pl.var[b];
            P.sinv = (const double *)(base + v.sinv);
            P.shape = (const double *)(base + v.shape);
            P.xruns = (const int *)(base + v.xruns);
            P.nxruns = (int)h.xruns.size() / 4;
            P.rowent = (const int *)(base + v.rowent);
            P.cons_aval = h.cons_aval.empty() ? nullptr : (const double *)(base + v.caval);
            P.cons_y = h.cons_y.empty() ? nullptr : (const double *)(base + v.cy);
        }
        P.cosv = (const double *)(base + pl.off_cos[b]);
        P.rhs = (const double *)(base + pl.off_rhs[b]);
        P.runs = (const int *)(base + pl.off_runs[b]);
        P.cons = (const int *)(base + pl.off_cons[b]);
        P.cons_sign = (const double *)(base + pl.off_sign[b]);
        P.ipiv = (int *)(base + pl.off_ipiv[b]);
        P.info = (int *)(base + pl.off_info[b]);
        P.x = (double *)(base + pl.off_x[b]);
        P.stats = (double *)(base + pl.off_stats[b]);
        P.scratch = (double *)(base + pl.off_scr[b]);
        if (with_solve) {
            P.colmax = (unsigned long long *)(base + pl.off_pmeta[b]);
            P.diagmag = (double *)(base + pl.off_pmeta[b] + 8 * kPanel);
            P.fast_ok = (int *)(base + pl.off_pmeta[b] + 16 * kPanel);
        }
        if (with_solve && pl.mixed) {
            const Plan::MixOff &m = pl.mix[b];
            const size_t plane = (size_t)P.ldz * h.n;
            P.fr = (float *)(base + m.fr);
            P.fi = P.fr + plane;
            P.fpr = (float *)(base + m.fpr);
            P.fpi = P.fpr + (size_t)P.ldz * kPanel;
            P.rowcol = (const int *)(base + m.rowcol);
            P.fh = (double *)(base + m.fh);
            P.lk = (double *)(base + m.lk);
            P.ym = (double *)(base + m.ym);
            P.ls = (double *)(base + m.ls);
            P.vr = (double *)(base + m.vr);
            P.vz = (double *)(base + m.vz);
            P.vp = (double *)(base + m.vp);
            P.vq = (double *)(base + m.vq);
            P.vt = (double *)(base + m.vt);
            P.vv = (double *)(base + m.vv);
            P.part = (double *)(base + m.part);
            P.colsq = (double *)(base + m.colsq);
            P.pcg = (double *)(base + m.pcg);
            P.pstate = (int *)(base + m.pstate);
            P.perm = (int *)(base + m.perm);
            P.tws = (double *)(base + m.tws);
            P.tflags = (int *)(base + m.tflags);
            P.gx = (unsigned long long *)(base + m.gx);
            P.dinv = (float *)(base + m.dinv);
            P.nchunk = m.nchunk;
            P.maxit = kMixMaxIt;
            P.tol = std::pow(10.0, -(double)g_mix_tol_exp);
            LuDesc<float> D;
            std::memset(&D, 0, sizeof(D));
            D.ar = P.fr;
            D.ai = P.fi;
            D.pr = P.fpr;
            D.pi = P.fpi;
            D.lda = P.ldz;
            D.m = h.n;
            D.ncols = h.n;
            D.ipiv = P.ipiv;
            D.info = P.info;
            D.colmax = P.colmax;
            D.diagmag = P.diagmag;
            D.fast_ok = P.fast_ok;
            D.linv = (float *)(base + m.linv);
            D.h16 = base + m.h16;
            pl.lud32.push_back(D);

        } else if (with_solve) {
            P.pr = (double *)(base + pl.off_panel[b]);
            P.pi = P.pr + (size_t)P.lda * kPanel;
            LuDesc<double> D;
            std::memset(&D, 0, sizeof(D));
            D.ar = P.ar;
            D.ai = P.ai;
            D.pr = P.pr;
            D.pi = P.pi;
            D.lda = P.lda;
            D.m = P.s;
            D.ncols = P.s + 1;
            D.ipiv = P.ipiv;
            D.info = P.info;
            D.colmax = P.colmax;
            D.diagmag = P.diagmag;
            D.fast_ok = P.fast_ok;
            pl.lud.push_back(D);
        }
    }
    // one staged copy: host inputs, problem table, work lists, zeroed results
    const size_t nbytes = pl.io_end - pl.in_begin;
    char *hs = pinned(LN().h_up, LN().h_up_bytes, nbytes);
    if (!hs) {
        set_error("pinned host memory exhausted");
        return FNEC_RESOURCE_EXHAUSTED;
    }
    auto put = [&](size_t off, const void *src, size_t len) {
        if (len)
            std::memcpy(hs + (off - pl.in_begin), src, len);
    };
    for (size_t b = 0; b < np; ++b) {
        const HostProblem &h = *hps[b];
        put(pl.off_seg[b], h.seg.data(), 8 * h.seg.size());
        put(pl.off_selfln[b], h.selfln.data(), 8 * h.selfln.size());
        put(pl.off_stamps[b], h.stamps.data(), 8 * h.stamps.size());
        put(pl.off_cos[b], h.cosv.data(), 8 * h.cosv.size());
        put(pl.off_rhs[b], h.rhs.data(), 8 * h.rhs.size());
        put(pl.off_runs[b], h.runs.data(), 4 * h.runs.size());
        put(pl.off_cons[b], h.cons.data(), 4 * h.cons.size());
        put(pl.off_sign[b], h.cons_sign.data(), 8 * h.cons_sign.size());
        if (h.variant) {
            const Plan::VarOff &v = pl.var[b];
            put(v.sinv, h.sinv.data(), 8 * h.sinv.size());
            put(v.shape, h.shape.data(), 8 * h.shape.size());
            put(v.xruns, h.xruns.data(), 4 * h.xruns.size());
            put(v.rowent, h.rowent.data(), 4 * h.rowent.size());
            put(v.caval, h.cons_aval.data(), 8 * h.cons_aval.size());
            put(v.cy, h.cons_y.data(), 8 * h.cons_y.size());
        }
        if (with_solve && pl.mixed)
            put(pl.mix[b].rowcol, h.rowcol.data(), 4 * h.rowcol.size());
    }
    put(pl.off_probs, pl.probs.data(), sizeof(Prob) * np);
    if (with_solve && !pl.mixed)
        put(pl.off_lud, pl.lud.data(), sizeof(LuDesc<double>) * np);
    if (with_solve && pl.mixed)
        put(pl.off_lud32, pl.lud32.data(), sizeof(LuDesc<float>) * np);
    put(pl.off_ints, pl.ints.data(), 4 * pl.ints.size());
    // result words start at zero; only device code writes them after this
    std::memset(hs + (pl.io_begin - pl.in_begin), 0, pl.io_end - pl.io_begin);
    HIPCHK(hipMemcpyAsync(base + pl.in_begin, hs, nbytes, hipMemcpyHostToDevice, st));
    return FNEC_OK;
}

// Generated work items as ints (item i at 4 i).
static inline const int *gen_ptr(char *base, const Plan &pl) { return (const int *)(base + pl.off_gen); }

// Expand the segment records into the generated work lists (before any
// kernel that reads them, on the same stream).
static void gen_launch(const Plan &pl, char *base, hipStream_t st)
{
    if (pl.gensegs.empty())
        return;
    hipLaunchKernelGGL(k_gen_items, dim3((unsigned)pl.gensegs.size()), dim3(256), 0, st,
                       (const int4 *)((const int *)(base + pl.off_ints) + pl.gensegs_off),
                       (int4 *)(base + pl.off_gen));
}

static inline const int *wl_ptr(char *base, const Plan &pl, size_t off)
{
    return (const int *)(base + pl.off_ints) + off;
}

// Mixed path: FP32 LU of Z, preconditioner set-up, CG iterations, residual.
static int mixed_launch(Plan &pl, char *base)
{
    HostTimer *enq = new HostTimer(kHostEnq); // until the first wait on the stream
    std::unique_ptr<HostTimer> enq_own(enq);
    hipStream_t st = LN().stream;
    const Prob *dp = (const Prob *)(base + pl.off_probs);
    const int *ints = wl_ptr(base, pl, 0);
    const int *all = ints + pl.all_off;
    const unsigned np = (unsigned)pl.all_cnt;
    HIPCHK(hipEventRecord(LN().ev[2], st));
    const LuDesc<float> *lud = (const LuDesc<float> *)(base + pl.off_lud32);
    // Look-ahead: the panel chain of step k + 1 (panels, their column solves
    // and the narrow update, all inside the columns [P + 64, P + 128)) runs on
    // the side stream once the main stream has updated those columns, while
    // the main stream updates the columns right of them. The interchanges of a
    // step reach the columns left of it in the main stream's outer launch,
    // after the trailing update that reads those columns.
    const size_t nsteps = pl.outer32.size();
    while (LN().sync_events.size() < 2 * nsteps + 1) {
        hipEvent_t e = nullptr;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        LN().sync_events.push_back(e);
    }
    hipEvent_t *ev_near = LN().sync_events.data();          // [k]: main updated step k + 1's columns
    hipEvent_t *ev_chain = LN().sync_events.data() + nsteps; // [k]: step k's panel chain done
    hipStream_t side = LN().side_stream;
    auto chain = [&](const Plan::Outer &o) {
        hipEvent_t e = prof_begin(PC_LU_PANEL, side);
        lu_launch_panel_f32(lud, o.a, ints, side);
        prof_end(PC_LU_PANEL, e, o.has_b ? 0.0 : o.f_panel, 0, side);
        if (o.a.trsm_cnt) {
            e = prof_begin(PC_LU_TRSM, side);
            lu_launch_trsm_f32(lud, o.a.trsm_off, o.a.trsm_cnt, ints, o.P, o.P + 2 * kPanel, side);
            prof_end(PC_LU_TRSM, e, 0, 0, side);
        }
        if (o.ua_cnt) {
            e = prof_begin(PC_LU_UPDN, side);
            lu_launch_update_narrow_f32(lud, o.ua_off, o.ua_cnt, ints, o.P, side);
            prof_end(PC_LU_UPDN, e, o.f_ua, o.b_ua, side);
        }
        if (o.has_b) {
            e = prof_begin(PC_LU_PANEL, side);
            lu_launch_panel_f32(lud, o.b, ints, side);
            prof_end(PC_LU_PANEL, e, o.f_panel, 0, side);
            if (o.b.trsm_cnt) {
                e = prof_begin(PC_LU_TRSM, side);
                lu_launch_trsm_f32(lud, o.b.trsm_off, o.b.trsm_cnt, ints, o.P + kPanel, 0, side);
                prof_end(PC_LU_TRSM, e, 0, 0, side);
            }
        }
        if (o.li_cnt) {
            e = prof_begin(PC_LU_TRSM, side);
            lu_launch_l11_inv(lud, o.li_off, o.li_cnt, ints, o.P, side);
            prof_end(PC_LU_TRSM, e, 0, 0, side);
        }
    };
    HIPCHK(hipEventRecord(ev_chain[nsteps], st)); // the fill is done
    HIPCHK(hipStreamWaitEvent(side, ev_chain[nsteps], 0));
    for (size_t k = 0; k < nsteps; ++k) {
        const Plan::Outer &o = pl.outer32[k];
        if (k == 0) {
            chain(o);
            HIPCHK(hipEventRecord(ev_chain[0], side));
        }
        HIPCHK(hipStreamWaitEvent(st, ev_chain[k], 0));
        if (o.ot_cnt && o.combine) {
            // interchanges, then the deferred step's factors on this step's
            // U12 rows, then the L11^{-1} product
            hipEvent_t e = prof_begin(PC_LU_TRSM);
            lu_launch_trsm64_f32(lud, o.ot_off, o.ot_cnt, ints, o.P, kTrsmSwap, st);
            prof_end(PC_LU_TRSM, e, 0, 0);
            if (o.th_cnt) {
                e = prof_begin(PC_LU_UPDN);
                lu_launch_update_f32(lud, 4 * o.th_off, o.th_cnt, gen_ptr(base, pl), o.P - 2 * kPanel, 2, o.P,
                                     o.P + 2 * kPanel, 0x7fffffff, 64, st);
                prof_end(PC_LU_UPDN, e, o.f_th, o.b_th);
            }
            e = prof_begin(PC_LU_TRSM);
            lu_launch_trsm64_f32(lud, o.ot_off, o.ot_cnt, ints, o.P, kTrsmSolve, st);
            prof_end(PC_LU_TRSM, e, o.f_trsm, 0);
            if (o.sp_cnt) {
                e = prof_begin(PC_LU_SPLIT);
                lu_launch_split_f16x2(lud, o.sp_off, o.sp_cnt, ints, o.P - 2 * kPanel, o.P + 2 * kPanel, st);
                prof_end(PC_LU_SPLIT, e, 0, 0);
            }
        } else if (o.ot_cnt) {
            hipEvent_t e = prof_begin(PC_LU_TRSM);
            lu_launch_trsm64_f32(lud, o.ot_off, o.ot_cnt, ints, o.P, kTrsmSwap | kTrsmSolve, st);
            prof_end(PC_LU_TRSM, e, o.f_trsm, 0);
        }
        if (o.on_cnt) {
            const int pc = o.combine ? PC_LU_UPD128 : PC_LU_UPDATE;
            hipEvent_t e = prof_begin(pc);
            if (o.combine)
                lu_launch_update_f16x2(lud, 4 * o.on_off, o.on_cnt, gen_ptr(base, pl), o.P - 2 * kPanel,
                                       o.P + 2 * kPanel, o.P + 2 * kPanel, o.P + 4 * kPanel, g_update_ring != 0, st);
            else
                lu_launch_update_f32(lud, 4 * o.on_off, o.on_cnt, gen_ptr(base, pl), o.P, 2, o.P + 2 * kPanel,
                                     o.P + 2 * kPanel, o.P + 4 * kPanel, pl.upd_tm, st);
            prof_end(pc, e, o.f_on, o.b_on);
        }
        if (k + 1 < nsteps) {
            HIPCHK(hipEventRecord(ev_near[k], st));
            HIPCHK(hipStreamWaitEvent(side, ev_near[k], 0));
            chain(pl.outer32[k + 1]);
            HIPCHK(hipEventRecord(ev_chain[k + 1], side));
        }
        if (o.of_cnt) {
            const int pc = o.combine ? PC_LU_UPD128 : PC_LU_UPDATE;
            hipEvent_t e = prof_begin(pc);
            if (o.combine)
                lu_launch_update_f16x2(lud, 4 * o.of_off, o.of_cnt, gen_ptr(base, pl), o.P - 2 * kPanel,
                                       o.P + 2 * kPanel, o.P + 4 * kPanel, 0x7fffffff, g_update_ring != 0, st);
            else
                lu_launch_update_f32(lud, 4 * o.of_off, o.of_cnt, gen_ptr(base, pl), o.P, 2, o.P + 2 * kPanel,
                                     o.P + 4 * kPanel, 0x7fffffff, pl.upd_tm, st);
            prof_end(pc, e, o.f_of, o.b_of);
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(LN().ev[3], st));
    if (pl.lu_only)
        return FNEC_OK;
    const int4 *vsol = (const int4 *)(ints + pl.vsol_off);
    unsigned long long *tickets = (unsigned long long *)(base + pl.off_tickets);
    int ntrsv = 0; // launches of k_trsv so far: epoch ntrsv + 1, ticket counter ntrsv
    // CG solves: one workgroup per problem with its vector in LDS for
    // n <= kSolve1MaxNHost; the flag-chained multi-workgroup solve above that.
    // The choice depends on the problem only, so batching never changes a result.
    if (pl.zs_cnt && pl.zs_lds > 64 * 1024)
        HIPCHK(hipFuncSetAttribute((const void *)k_zsolve, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)pl.zs_lds));
    auto zsolve = [&](int mode) {
        if (pl.zs_cnt) {
            hipLaunchKernelGGL(k_zsolve, dim3((unsigned)pl.zs_cnt), dim3(1024), pl.zs_lds, st, dp,
                               (const int4 *)(ints + pl.zs_off), mode, ntrsv + 1, tickets + ntrsv);
            ++ntrsv;
        }
        if (pl.vsol_cnt) {
            if (mode == 0)
                hipLaunchKernelGGL(k_trsv0, dim3((unsigned)pl.vsol_cnt), dim3(256), 0, st, dp, vsol,
                                   ntrsv + 1, tickets + ntrsv);
            else
                hipLaunchKernelGGL(k_trsv1, dim3((unsigned)pl.vsol_cnt), dim3(256), 0, st, dp, vsol,
                                   ntrsv + 1, tickets + ntrsv);
            ++ntrsv;
        }
    };
    {
        const size_t count = (pl.flags_end - pl.flags_begin) / 8;
        hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>(1024, (count + 255) / 256)), dim3(256), 0,
                           st, (double *)(base + pl.flags_begin), count);
    }
    hipLaunchKernelGGL(k_perm, dim3(np), dim3(256), 0, st, dp, all);
    hipLaunchKernelGGL(k_diag_inv, dim3((unsigned)pl.dinv_cnt), dim3(64), 0, st, dp,
                       (const int2 *)(ints + pl.dinv_off));
    hipLaunchKernelGGL(k_mix_init, dim3(np), dim3(1024), 0, st, dp, all);
    if (pl.fsol_cnt) {
        hipEvent_t e = prof_begin(PC_FSETUP);
        if (pl.fs_lds > 64 * 1024)
            HIPCHK(hipFuncSetAttribute((const void *)k_fsetup, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)pl.fs_lds));
        for (size_t g = 0; g < pl.fs_off.size(); ++g)
            if (pl.fs_cnt[g]) {
                hipLaunchKernelGGL(k_fsetup, dim3((unsigned)pl.fs_cnt[g]), dim3(1024), pl.fs_lds, st, dp,
                                   (const int4 *)(ints + pl.fs_off[g]), ntrsv + 1, tickets + ntrsv);
                ++ntrsv;
            }
        prof_end(PC_FSETUP, e, pl.fsol_flops, pl.fsol_bytes);
    }
    hipLaunchKernelGGL(k_pre_setup, dim3(np), dim3(1024), 0, st, dp, all);
    hipLaunchKernelGGL(k_mv_zh, dim3((unsigned)pl.mvzh_cnt), dim3(256), 0, st, dp,
                       (const int2 *)(ints + pl.mvzh_off), 1);
    hipLaunchKernelGGL(k_mix_start, dim3(np), dim3(1024), 0, st, dp, all);
    // CG iterations in chunks; between chunks one 4-byte-per-problem read-back
    // tells whether any problem still iterates.
    int *flags_dev = (int *)(base + pl.off_flags);
    std::vector<int> flags(np);
    pl.it_zs1.clear();
    pl.it_zs0.clear();
    pl.it_mv.clear();
    int it = 0;
    while (it < kMixMaxIt) {
        const int chunk = it == 0 ? g_mix_first_chunk : kMixNextChunk;
        for (int c = 0; c < chunk && it < kMixMaxIt; ++c, ++it) {
            hipEvent_t e = prof_begin(PC_ZSOLVE);
            zsolve(1);
            pl.it_zs1.push_back(prof_end(PC_ZSOLVE, e, pl.zsolve_flops, pl.zsolve_bytes));
            e = prof_begin(PC_PCG);
            hipLaunchKernelGGL(k_prec_mid, dim3(np), dim3(1024), 0, st, dp, all);
            prof_end(PC_PCG, e, 0, 0);
            e = prof_begin(PC_ZSOLVE);
            zsolve(0);
            pl.it_zs0.push_back(prof_end(PC_ZSOLVE, e, pl.zsolve_flops, pl.zsolve_bytes));
            e = prof_begin(PC_PCG);
            hipLaunchKernelGGL(k_mix_dir, dim3(np), dim3(1024), 0, st, dp, all);
            prof_end(PC_PCG, e, 0, 0);
            e = prof_begin(PC_MV);
            hipLaunchKernelGGL(k_mv_z, dim3((unsigned)pl.mvz_cnt), dim3(kMvRows), 0, st, dp,
                               (const int4 *)(ints + pl.mvz_off));
            hipLaunchKernelGGL(k_mv_fin, dim3((unsigned)pl.mvfin_cnt), dim3(kMvRows), 0, st, dp,
                               (const int2 *)(ints + pl.mvfin_off));
            hipLaunchKernelGGL(k_mv_zh, dim3((unsigned)pl.mvzh_cnt), dim3(256), 0, st, dp,
                               (const int2 *)(ints + pl.mvzh_off), 0);
            pl.it_mv.push_back(prof_end(PC_MV, e, pl.mv_flops, pl.mv_bytes));
            e = prof_begin(PC_PCG);
            hipLaunchKernelGGL(k_mix_step, dim3(np), dim3(1024), 0, st, dp, all);
            prof_end(PC_PCG, e, 0, 0);
        }
        if (it >= kMixMaxIt)
            break;
        hipLaunchKernelGGL(k_mix_flags, dim3((np + 255) / 256), dim3(256), 0, st, dp, np, flags_dev);
        HIPCHK(hipMemcpyAsync(flags.data(), flags_dev, 4 * np, hipMemcpyDeviceToHost, st));
        enq_own.reset();
        {
            HostTimer ht(kHostWait);
            HIPCHK(stream_wait(st));
        }
        bool all_done = true;
        for (unsigned b = 0; b < np; ++b)
            all_done = all_done && flags[b] != 0;
        if (all_done)
            break;
    }
    pl.iters_launched = it;
    hipLaunchKernelGGL(k_residual_rows, dim3((unsigned)pl.res_cnt), dim3(256), 0, st, dp,
                       (const int2 *)(ints + pl.res_off));
    hipLaunchKernelGGL(k_residual_finish, dim3(np), dim3(64), 0, st, dp, all);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(LN().ev[4], st));
    return FNEC_OK;
}

// Launch the whole solve on the context stream; records phase events.
// Plain path in the reference's operation order: normal equations (unless
// the caller supplied [A | b]), elimination in 32-column panels, back
// substitution. Events 2 and 3 bracket the elimination.
// [G | M^H Y] of the plain path in the reference's summation order: Z^H Z
// tiles, the extra columns against Z and the single right-hand side, the
// other right-hand sides, the extra-column block, then the constraint rows and
// lambda (which k_gx_fix adds last).
static void gx_launch(const Plan &pl, char *base, hipStream_t st)
{
    const Prob *dp = (const Prob *)(base + pl.off_probs);
    if (pl.gram_cnt)
        hipLaunchKernelGGL(k_gx_tile, dim3((unsigned)pl.gram_cnt), dim3(256), 0, st, dp,
                           (const int4 *)wl_ptr(base, pl, pl.gram_off));
    if (pl.xcols_cnt)
        hipLaunchKernelGGL(k_gx_cols, dim3((unsigned)pl.xcols_cnt), dim3(256), 0, st, dp,
                           (const int2 *)wl_ptr(base, pl, pl.xcols_off));
    if (pl.xrhs_cnt)
        hipLaunchKernelGGL(k_gx_rhs, dim3((unsigned)pl.xrhs_cnt), dim3(256), 0, st, dp,
                           (const int4 *)wl_ptr(base, pl, pl.xrhs_off));
    if (pl.xext_cnt)
        hipLaunchKernelGGL(k_gx_extra, dim3((unsigned)pl.xext_cnt), dim3(256), 0, st, dp,
                           (const int2 *)wl_ptr(base, pl, pl.xext_off));
    hipLaunchKernelGGL(k_gx_fix, dim3((unsigned)pl.all_cnt), dim3(64), 0, st, dp,
                       wl_ptr(base, pl, pl.all_off));
}

static int exact_launch(Plan &pl, char *base)
{
    hipStream_t st = LN().stream;
    const Prob *dp = (const Prob *)(base + pl.off_probs);
    const int *all = wl_ptr(base, pl, pl.all_off);
    if (!pl.given_a) {
        hipEvent_t e = prof_begin(PC_GRAM);
        gx_launch(pl, base, st);
        prof_end(PC_GRAM, e, pl.gram_flops, 0);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(LN().ev[2], st));
    const int *ints = wl_ptr(base, pl, 0);
    for (size_t si = 0; si < pl.steps.size(); ++si) {
        const LuStep &s = pl.steps[si];
        if (s.panel_cnt) {
            hipEvent_t e = prof_begin(PC_LU_PANEL);
            hipLaunchKernelGGL(k_ge_diag, dim3((unsigned)s.panel_cnt), dim3(64), 0, st, dp,
                               ints + s.panel_off, s.p);
            if (s.rows_cnt)
                hipLaunchKernelGGL(k_ge_rows, dim3((unsigned)s.rows_cnt), dim3(256), 0, st, dp,
                                   (const int2 *)(ints + s.rows_off), s.p);
            hipLaunchKernelGGL(k_ge_exact, dim3((unsigned)s.panel_cnt), dim3(1024), 0, st, dp,
                               ints + s.panel_off, s.p);
            prof_end(PC_LU_PANEL, e, pl.step_panel_flops[si], 0);
        }
        if (s.trsm_cnt) {
            hipEvent_t e = prof_begin(PC_LU_TRSM);
            hipLaunchKernelGGL(k_ge_u12, dim3((unsigned)s.trsm_cnt), dim3(64), 0, st, dp,
                               (const int2 *)(ints + s.trsm_off), s.p);
            prof_end(PC_LU_TRSM, e, pl.step_trsm_flops[si], 0);
        }
        if (s.upd_cnt) {
            hipEvent_t e = prof_begin(PC_LU_UPDATE);
            hipLaunchKernelGGL(k_ge_update, dim3((unsigned)s.upd_cnt), dim3(256), 0, st, dp,
                               (const int4 *)(ints + s.upd_off), s.p);
            prof_end(PC_LU_UPDATE, e, pl.step_upd_flops[si], 0);
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(LN().ev[3], st));
    {
        hipEvent_t e = prof_begin(PC_BACKSUB);
        const int2 *bs = (const int2 *)wl_ptr(base, pl, pl.bs_off);
        if (g_exact_backsub)
            hipLaunchKernelGGL(k_backsub_exact, dim3((unsigned)pl.bs_cnt), dim3(256), 0, st, dp, bs);
        else
            hipLaunchKernelGGL(k_backsub, dim3((unsigned)pl.bs_cnt), dim3(kBsThreads), 0, st, dp, bs);
        prof_end(PC_BACKSUB, e, pl.backsub_flops, 0);
    }
    if (!pl.given_a) {
        hipLaunchKernelGGL(k_residual_rows, dim3((unsigned)pl.res_cnt), dim3(256), 0, st, dp,
                           (const int2 *)wl_ptr(base, pl, pl.res_off));
        hipLaunchKernelGGL(k_residual_finish, dim3((unsigned)pl.all_cnt), dim3(64), 0, st, dp, all);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(LN().ev[4], st));
    return FNEC_OK;
}

static int plan_launch(Plan &pl, char *base, bool with_solve)
{
    hipStream_t st = LN().stream;
    const Prob *dp = (const Prob *)(base + pl.off_probs);
    HIPCHK(hipEventRecord(LN().ev[0], st));
    gen_launch(pl, base, st);
    if (pl.fill_cnt && !pl.given_z && !pl.given_a) {
        hipEvent_t e = prof_begin(PC_FILL);
        hipLaunchKernelGGL(pl.exact ? k_fill_exact : pl.fill_large ? k_fill_z_large : k_fill_z,
                           dim3((unsigned)pl.fill_cnt), dim3(256), 0,
                           st, dp, (const int4 *)gen_ptr(base, pl) + pl.fill_off);
        prof_end(PC_FILL, e, pl.fill_flops, 0);
        if (pl.stamp_cnt)
            hipLaunchKernelGGL(k_stamp_z, dim3((unsigned)pl.stamp_cnt), dim3(256), 0, st, dp,
                               wl_ptr(base, pl, pl.stamp_off));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(LN().ev[1], st));
    if (!with_solve)
        return FNEC_OK;
    if (pl.mixed)
        return mixed_launch(pl, base);
    if (pl.exact)
        return exact_launch(pl, base);
    {
        size_t count = (pl.a_end - pl.a_begin) / 8;
        if (count)
            hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, st, (double *)(base + pl.a_begin), count);
    }
    {
        hipEvent_t e = prof_begin(PC_GRAM);
        hipLaunchKernelGGL(k_gram, dim3((unsigned)pl.gram_cnt), dim3(256), 0, st, dp,
                           (const int4 *)wl_ptr(base, pl, pl.gram_off));
        prof_end(PC_GRAM, e, pl.gram_flops, 0);
    }
    hipLaunchKernelGGL(k_gram_cols, dim3((unsigned)pl.cols_cnt), dim3(256), 0, st, dp,
                       (const int2 *)wl_ptr(base, pl, pl.cols_off));
    hipLaunchKernelGGL(k_gram_fix, dim3((unsigned)pl.all_cnt), dim3(256), 0, st, dp,
                       wl_ptr(base, pl, pl.all_off));
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(LN().ev[2], st));
    const LuDesc<double> *lud = (const LuDesc<double> *)(base + pl.off_lud);
    const int *ints = wl_ptr(base, pl, 0);
    for (size_t si = 0; si < pl.steps.size(); ++si) {
        const auto &s = pl.steps[si];
        if (s.panel_cnt) {
            hipEvent_t e = prof_begin(PC_LU_PANEL);
            lu_launch_panel(lud, s, ints, st);
            prof_end(PC_LU_PANEL, e, pl.step_panel_flops[si], 0);
        }
        if (s.trsm_cnt) {
            hipEvent_t e = prof_begin(PC_LU_TRSM);
            lu_launch_trsm(lud, s, ints, st);
            prof_end(PC_LU_TRSM, e, pl.step_trsm_flops[si], 0);
        }
        if (s.upd_cnt) {
            hipEvent_t e = prof_begin(PC_LU_UPDATE);
            lu_launch_update(lud, s, ints, st);
            prof_end(PC_LU_UPDATE, e, pl.step_upd_flops[si], 0);
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(LN().ev[3], st));
    {
        hipEvent_t e = prof_begin(PC_BACKSUB);
        hipLaunchKernelGGL(k_backsub, dim3((unsigned)pl.bs_cnt), dim3(kBsThreads), 0, st, dp,
                           (const int2 *)wl_ptr(base, pl, pl.bs_off));
        prof_end(PC_BACKSUB, e, pl.backsub_flops, 0);
    }
    hipLaunchKernelGGL(k_residual_rows, dim3((unsigned)pl.res_cnt), dim3(256), 0, st, dp,
                       (const int2 *)wl_ptr(base, pl, pl.res_off));
    hipLaunchKernelGGL(k_residual_finish, dim3((unsigned)pl.all_cnt), dim3(64), 0, st, dp,
                       wl_ptr(base, pl, pl.all_off));
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(LN().ev[4], st));
    return FNEC_OK;
}

static float ev_ms(int a, int b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, LN().ev[a], LN().ev[b]) != hipSuccess)
        return 0.f;
    return ms;
}

// Solve a batch of problems; x_out[b] receives s complex values (may be null).
int solve_problems(const std::vector<const HostProblem *> &hps, std::vector<fnec_c64 *> &x_out,
                   std::vector<fnec_solve_stats> &stats, std::vector<int> &status,
                   std::vector<double *> *device_x, int precision,
                   const std::function<void(const std::vector<double *> &)> &overlap)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    const bool mixed = precision == FNEC_PREC_MIXED;
    for (const HostProblem *h : hps)
        if ((h->variant || h->nrhs != 1) && precision != FNEC_PREC_FP64) {
            set_error("plane-wave and current-source systems are solved on the plain FP64 path");
            return FNEC_UNAVAILABLE;
        }
    if (mixed)
        for (const HostProblem *h : hps)
            if (h->w < 1 || h->w > kMixMaxW || h->ncons > kMixMaxCons) {
                set_error("mixed precision accepts 1.." + std::to_string(kMixMaxW) +
                          " homogeneous columns and at most " + std::to_string(kMixMaxCons) +
                          " constraint rows");
                return FNEC_INVALID_ARG;
            }
    Plan pl;
    pl.mixed = mixed;
    pl.exact = precision == FNEC_PREC_FP64;
    // the work lists keep their capacity across batches
    std::vector<int> &ints_cache = LN().plan_ints;
    pl.ints.swap(ints_cache);
    struct IntsBack {
        Plan &p;
        std::vector<int> &c;
        ~IntsBack() { p.ints.swap(c); }
    } ints_back{pl, ints_cache};
    {
        HostTimer ht(kHostPrep);
        plan_build(pl, hps, true);
    }
    char *base = nullptr;
    int st = arena_reserve(pl.bytes, &base);
    if (st)
        return st;
    {
        HostTimer ht(kHostStage);
        st = plan_upload(pl, hps, base, true);
    }
    if (st)
        return st;
    std::vector<double *> dev_x(hps.size());
    for (size_t b = 0; b < hps.size(); ++b)
        dev_x[b] = (double *)(base + pl.off_x[b]);
    // the caller's concurrent work starts before the launches: queueing them
    // blocks on the device queue, so the host has no lead to spend later
    if (overlap)
        overlap(dev_x);
    st = plan_launch(pl, base, true);
    if (st)
        return st;
    // all result words in one copy
    const size_t nback = pl.io_end - pl.io_begin;
    char *hb = pinned(LN().h_back, LN().h_back_bytes, nback);
    if (!hb) {
        set_error("pinned host memory exhausted");
        return FNEC_RESOURCE_EXHAUSTED;
    }
    HIPCHK(hipMemcpyAsync(hb, base + pl.io_begin, nback, hipMemcpyDeviceToHost, LN().stream));
    {
        HostTimer ht(kHostWait);
        HIPCHK(stream_wait(LN().stream));
    }
    HostTimer ht_post(kHostPost);
    auto back = [&](size_t off) { return hb + (off - pl.io_begin); };
    if (mixed && g_prof) {
        // the work of CG iteration i is done only by problems that ran it
        std::vector<int> iters(hps.size(), 0);
        for (size_t b = 0; b < hps.size(); ++b)
            iters[b] = ((const int *)back(pl.mix[b].pstate))[1];
        for (int i = 0; i < pl.iters_launched; ++i) {
            double zf = 0, zb = 0, mf = 0, mb = 0;
            for (size_t b = 0; b < hps.size(); ++b)
                if (iters[b] > i) {
                    const double n = hps[b]->n, r = hps[b]->r;
                    zf += 8.0 * n * n;
                    zb += 8.0 * n * n;
                    mf += 16.0 * n * r;
                    mb += 32.0 * n * n;
                }
            prof_set_work(pl.it_zs1[i], zf, zb);
            prof_set_work(pl.it_zs0[i], zf, zb);
            prof_set_work(pl.it_mv[i], mf, mb);
        }
    }
    prof_collect();
    const float t_fill = ev_ms(0, 1) * 1000.f, t_gram = ev_ms(1, 2) * 1000.f,
                t_lu = ev_ms(2, 3) * 1000.f, t_sol = ev_ms(3, 4) * 1000.f;
    const size_t np = hps.size();
    status.assign(np, FNEC_OK);
    stats.assign(np, fnec_solve_stats{});
    for (size_t b = 0; b < np; ++b) {
        const int *info = (const int *)back(pl.off_info[b]);
        const double *sv = (const double *)back(pl.off_stats[b]);
        status[b] = info[0] ? FNEC_SINGULAR : FNEC_OK;
        fnec_solve_stats &s = stats[b];
        if (mixed && !info[0]) {
            const double *pcg = (const double *)back(pl.mix[b].pcg);
            const int *ps = (const int *)back(pl.mix[b].pstate);
            s.refine_iters = ps[1];
            const double den = pcg[3] * pcg[4] + pcg[1];
            s.backward_error = den > 0.0 ? pcg[2] / den : 0.0;
            if (ps[2] != 0 || !(pcg[2] <= 1e-8 * pcg[1])) {
                set_error(ps[2] ? "mixed-precision iteration broke down"
                                : "mixed-precision iteration did not converge");
                status[b] = FNEC_INTERNAL;
            }
        }
        s.abs_res = sv[0];
        s.rel_res = sv[1];
        s.diag_spread = sv[2];
        s.n_swaps = info[1];
        s.exact_panels = info[2];
        s.rows = hps[b]->r;
        s.unknowns = hps[b]->s;
        s.t_fill_us = t_fill;
        s.t_gram_us = t_gram;
        s.t_lu_us = t_lu;
        s.t_solve_us = t_sol;
        s.t_total_us = t_fill + t_gram + t_lu + t_sol;
        if (x_out.size() > b && x_out[b] && status[b] == FNEC_OK)
            std::memcpy(x_out[b], back(pl.off_x[b]), 16 * (size_t)hps[b]->s * hps[b]->nrhs);
        if (device_x)
            device_x->push_back(dev_x[b]);
    }
    return FNEC_OK;
}

int fill_only(const HostProblem &hp, fnec_c64 *z_out, bool exact)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    Plan pl;
    pl.exact = exact;
    std::vector<const HostProblem *> v{&hp};
    plan_build(pl, v, false);
    char *base = nullptr;
    int st = arena_reserve(pl.bytes, &base);
    if (st)
        return st;
    if ((st = plan_upload(pl, v, base, false)))
        return st;
    if ((st = plan_launch(pl, base, false)))
        return st;
    const int n = hp.n, ldz = roundup(n, 64);
    std::vector<double> zr((size_t)ldz * n), zi((size_t)ldz * n);
    HIPCHK(hipMemcpyAsync(zr.data(), base + pl.off_z[0], 8 * zr.size(), hipMemcpyDeviceToHost, LN().stream));
    HIPCHK(hipMemcpyAsync(zi.data(), base + pl.off_z[0] + 8 * zr.size(), 8 * zi.size(), hipMemcpyDeviceToHost, LN().stream));
    HIPCHK(hipStreamSynchronize(LN().stream));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            z_out[(size_t)i * n + j].re = zr[(size_t)j * ldz + i];
            z_out[(size_t)i * n + j].im = zi[(size_t)j * ldz + i];
        }
    return FNEC_OK;
}

int normal_eqs_only(const HostProblem &hp, fnec_c64 *aug_out)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    Plan pl;
    pl.exact = true;
    std::vector<const HostProblem *> v{&hp};
    plan_build(pl, v, true);
    char *base = nullptr;
    int st = arena_reserve(pl.bytes, &base);
    if (st)
        return st;
    if ((st = plan_upload(pl, v, base, true)))
        return st;
    hipStream_t s = LN().stream;
    const Prob *dp = (const Prob *)(base + pl.off_probs);
    gen_launch(pl, base, s);
    hipLaunchKernelGGL(k_fill_exact, dim3((unsigned)pl.fill_cnt), dim3(256), 0, s, dp,
                       (const int4 *)gen_ptr(base, pl) + pl.fill_off);
    gx_launch(pl, base, s);
    HIPCHK(hipGetLastError());
    const int S = hp.s, lda = roundup(S, 64);
    std::vector<double> ar((size_t)lda * (S + 1)), ai((size_t)lda * (S + 1));
    HIPCHK(hipMemcpyAsync(ar.data(), base + pl.off_a[0], 8 * ar.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(ai.data(), base + pl.off_a[0] + 8 * ar.size(), 8 * ai.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i < S; ++i)
        for (int j = 0; j <= S; ++j) {
            aug_out[(size_t)i * (S + 1) + j].re = ar[(size_t)j * lda + i];
            aug_out[(size_t)i * (S + 1) + j].im = ai[(size_t)j * lda + i];
        }
    return FNEC_OK;
}

// Plain path on a caller-supplied Z (n x n row-major): normal equations,
// elimination and back substitution of the problem hp, without the fill.
int solve_given_z(const HostProblem &hp, const fnec_c64 *z, fnec_c64 *x_out, fnec_solve_stats *stats)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    Plan pl;
    pl.exact = true;
    pl.given_z = true;
    std::vector<const HostProblem *> v{&hp};
    plan_build(pl, v, true);
    char *base = nullptr;
    int st = arena_reserve(pl.bytes, &base);
    if (st)
        return st;
    if ((st = plan_upload(pl, v, base, true)))
        return st;
    const int n = hp.n, ldz = roundup(n, 64);
    std::vector<double> zp(2 * (size_t)ldz * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            zp[(size_t)j * ldz + i] = z[(size_t)i * n + j].re;
            zp[(size_t)ldz * n + (size_t)j * ldz + i] = z[(size_t)i * n + j].im;
        }
    hipStream_t s = LN().stream;
    HIPCHK(hipMemcpyAsync(base + pl.off_z[0], zp.data(), 8 * zp.size(), hipMemcpyHostToDevice, s));
    if ((st = plan_launch(pl, base, true)))
        return st;
    std::vector<char> back(pl.io_end - pl.io_begin);
    HIPCHK(hipMemcpyAsync(back.data(), base + pl.io_begin, back.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const int *info = (const int *)(back.data() + (pl.off_info[0] - pl.io_begin));
    const double *sv = (const double *)(back.data() + (pl.off_stats[0] - pl.io_begin));
    if (stats) {
        *stats = fnec_solve_stats{};
        stats->abs_res = sv[0];
        stats->rel_res = sv[1];
        stats->diag_spread = sv[2];
        stats->n_swaps = info[1];
        stats->exact_panels = info[2];
        stats->rows = hp.r;
        stats->unknowns = hp.s;
    }
    if (info[0])
        return FNEC_SINGULAR;
    std::memcpy(x_out, back.data() + (pl.off_x[0] - pl.io_begin), 16 * (size_t)hp.s);
    return FNEC_OK;
}

// Plain elimination (partial pivoting in the reference's order) and back
// substitution of a caller-supplied n x n system a x = b (a row-major).
// ipiv_out[c] is the row interchanged with row c at elimination step c.
int solve_square_exact(const fnec_c64 *a, const fnec_c64 *b, int n, fnec_c64 *x_out, int *ipiv_out,
                       fnec_solve_stats *stats)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    HostProblem hp;
    hp.n = n;
    hp.s = n;
    hp.r = n;
    hp.w = 0;
    hp.seg.assign((size_t)n * kSegStride, 1.0);
    hp.selfln.assign(n, 0.0);
    hp.cosv.assign(n, 0.0);
    hp.rhs.assign(2 * (size_t)n, 0.0);
    hp.runs = {0, n, 0};
    hp.k = 1.0;
    Plan pl;
    pl.exact = true;
    pl.given_a = true;
    std::vector<const HostProblem *> v{&hp};
    plan_build(pl, v, true);
    char *base = nullptr;
    int st = arena_reserve(pl.bytes, &base);
    if (st)
        return st;
    if ((st = plan_upload(pl, v, base, true)))
        return st;
    const int lda = roundup(n, 64);
    const size_t plane = (size_t)lda * (n + 1);
    std::vector<double> ap(2 * plane, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            ap[(size_t)j * lda + i] = a[(size_t)i * n + j].re;
            ap[plane + (size_t)j * lda + i] = a[(size_t)i * n + j].im;
        }
        ap[(size_t)n * lda + i] = b[i].re;
        ap[plane + (size_t)n * lda + i] = b[i].im;
    }
    hipStream_t s = LN().stream;
    HIPCHK(hipMemcpyAsync(base + pl.off_a[0], ap.data(), 8 * ap.size(), hipMemcpyHostToDevice, s));
    if ((st = plan_launch(pl, base, true)))
        return st;
    std::vector<char> back(pl.io_end - pl.io_begin);
    HIPCHK(hipMemcpyAsync(back.data(), base + pl.io_begin, back.size(), hipMemcpyDeviceToHost, s));
    std::vector<int> piv(n);
    HIPCHK(hipMemcpyAsync(piv.data(), base + pl.off_ipiv[0], 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const int *info = (const int *)(back.data() + (pl.off_info[0] - pl.io_begin));
    if (stats) {
        *stats = fnec_solve_stats{};
        stats->n_swaps = info[1];
        stats->exact_panels = info[2];
        stats->rows = n;
        stats->unknowns = n;
    }
    if (ipiv_out)
        std::memcpy(ipiv_out, piv.data(), 4 * (size_t)n);
    if (info[0])
        return FNEC_SINGULAR;
    std::memcpy(x_out, back.data() + (pl.off_x[0] - pl.io_begin), 16 * (size_t)n);
    return FNEC_OK;
}

int lu_f32_only(const HostProblem &hp, float *lu_out, int *ipiv_out)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    if (hp.w < 1 || hp.w > kMixMaxW || hp.ncons > kMixMaxCons) {
        set_error("mixed precision accepts 1.." + std::to_string(kMixMaxW) + " homogeneous columns and at most " +
                  std::to_string(kMixMaxCons) + " constraint rows");
        return FNEC_INVALID_ARG;
    }
    Plan pl;
    pl.mixed = true;
    pl.lu_only = true;
    std::vector<const HostProblem *> v{&hp};
    plan_build(pl, v, true);
    char *base = nullptr;
    int st = arena_reserve(pl.bytes, &base);
    if (st)
        return st;
    if ((st = plan_upload(pl, v, base, true)))
        return st;
    if ((st = plan_launch(pl, base, true)))
        return st;
    hipStream_t s = LN().stream;
    const int n = hp.n, ldz = roundup(n, 64);
    const size_t plane = (size_t)ldz * n;
    std::vector<float> f(2 * plane);
    std::vector<int> info(4), ipiv(n);
    HIPCHK(hipMemcpyAsync(f.data(), base + pl.mix[0].fr, 4 * f.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(ipiv.data(), base + pl.off_ipiv[0], 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(info.data(), base + pl.off_info[0], 4 * info.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (info[0]) {
        set_error("FP32 factorisation met a singular pivot");
        return FNEC_SINGULAR;
    }
    for (int i = 0; i < n; ++i) {
        ipiv_out[i] = ipiv[i];
        for (int j = 0; j < n; ++j) {
            lu_out[2 * ((size_t)i * n + j)] = f[(size_t)j * ldz + i];
            lu_out[2 * ((size_t)i * n + j) + 1] = f[plane + (size_t)j * ldz + i];
        }
    }
    return FNEC_OK;
}

// ---------------------------------------------------------------------------
// far field
// ---------------------------------------------------------------------------

int rp_stage(const std::vector<RpHost> &jobs, RpStaged &S, char *fixed_base, size_t fixed_cap)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    HostTimer ht(kHostRp);
    const size_t nj = jobs.size();
    S = RpStaged{};
    S.nj = nj;
    S.fixed = fixed_base != nullptr;
    size_t cur = 0;
    std::vector<size_t> o_seg(nj), o_cur(nj), o_th(nj), o_ph(nj), o_field(nj), o_part(nj);
    S.o_tot.resize(nj);
    S.o_out.resize(nj);
    S.npts.resize(nj);
    std::vector<int> &ints = LN().rp_ints; // capacity kept across batches
    ints.clear();
    // device-only work buffers
    for (size_t j = 0; j < nj; ++j) {
        const RpHost &h = jobs[j];
        S.npts[j] = h.npts;
        o_field[j] = take(cur, 32 * h.npts + 32);
        o_part[j] = take(cur, 8 * (size_t)((h.ndir + 255) / 256) + 8);
    }
    S.pts_off = ints.size();
    for (size_t j = 0; j < nj; ++j)
        for (int y = 0; y < ((int)jobs[j].npts + 255) / 256; ++y) {
            ints.push_back((int)j);
            ints.push_back(y);
        }
    S.pts_cnt = (ints.size() - S.pts_off) / 2;
    S.norm_off = ints.size();
    for (size_t j = 0; j < nj; ++j)
        for (int y = 0; y < (jobs[j].ndir + 255) / 256; ++y) {
            ints.push_back((int)j);
            ints.push_back(y);
        }
    S.norm_cnt = (ints.size() - S.norm_off) / 2;
    S.all_off = ints.size();
    for (size_t j = 0; j < nj; ++j)
        ints.push_back((int)j);
    // inputs (one staged copy), then outputs: the totals, then the rows
    const size_t in_begin = cur = (cur + 255) & ~(size_t)255;
    for (size_t j = 0; j < nj; ++j) {
        const RpHost &h = jobs[j];
        o_seg[j] = take(cur, 8 * h.seg.size());
        o_cur[j] = h.cur_dev ? 0 : take(cur, 8 * h.cur.size());
        o_th[j] = take(cur, 8 * h.npts + 8);
        o_ph[j] = take(cur, 8 * h.npts + 8);
    }
    S.o_probs = take(cur, sizeof(RpProb) * nj);
    S.o_ints = take(cur, 4 * ints.size() + 16);
    S.out_begin = cur = (cur + 255) & ~(size_t)255;
    for (size_t j = 0; j < nj; ++j)
        S.o_tot[j] = S.out_begin + 16 * j;
    S.rows_begin = cur = (S.out_begin + 16 * nj + 255) & ~(size_t)255;
    for (size_t j = 0; j < nj; ++j) {
        S.o_out[j] = cur;
        cur += 32 * jobs[j].npts;
    }
    S.out_end = cur = (cur + 255) & ~(size_t)255;
    if (S.fixed) {
        if (cur + 256 > fixed_cap) {
            set_error("far-field job needs " + std::to_string(cur + 256) + " bytes of device scratch; " +
                      std::to_string(fixed_cap) + " are available");
            return FNEC_RESOURCE_EXHAUSTED;
        }
        S.base = fixed_base;
    } else {
        int st = grow_arena(LN().rp_arena, LN().rp_arena_bytes, cur + 256, &S.base);
        if (st)
            return st;
    }
    char *base = S.base;
    std::vector<RpProb> probs(nj);
    // staging in pinned memory the lane grows on demand; with a fixed base,
    // in pageable host memory (no allocation by the HIP runtime)
    if (S.fixed)
        S.hstage.resize(S.out_begin - in_begin);
    char *hs = S.fixed ? S.hstage.data() : pinned(LN().h_rp, LN().h_rp_bytes, S.out_begin - in_begin);
    if (!hs) {
        set_error("pinned host memory exhausted");
        return FNEC_RESOURCE_EXHAUSTED;
    }
    auto put = [&](size_t off, const void *src, size_t len) {
        if (len)
            std::memcpy(hs + (off - in_begin), src, len);
    };
    for (size_t j = 0; j < nj; ++j) {
        const RpHost &h = jobs[j];
        RpProb &R = probs[j];
        std::memset(&R, 0, sizeof(R));
        R.n = h.n;
        R.npts = (int)h.npts;
        R.ground = h.ground;
        R.ndir = h.ndir;
        R.k = h.k;
        R.eps_r = h.eps_r;
        R.sigma = h.sigma;
        R.seg = (const double *)(base + o_seg[j]);
        R.cur = h.cur_dev ? h.cur_dev : (const double *)(base + o_cur[j]);
        R.theta = (const double *)(base + o_th[j]);
        R.phi = (const double *)(base + o_ph[j]);
        R.field = (double *)(base + o_field[j]);
        R.partial = (double *)(base + o_part[j]);
        R.npartial = (h.ndir + 255) / 256;
        R.total = (double *)(base + S.o_tot[j]);
        R.out = (double *)(base + S.o_out[j]);
        put(o_seg[j], h.seg.data(), 8 * h.seg.size());
        if (!h.cur_dev)
            put(o_cur[j], h.cur.data(), 8 * h.cur.size());
        put(o_th[j], h.theta, 8 * h.npts);
        put(o_ph[j], h.phi, 8 * h.npts);
        S.rp_terms += (double)h.n * (h.ground ? 2.0 : 1.0) * ((double)h.npts + h.ndir);
    }
    put(S.o_probs, probs.data(), sizeof(RpProb) * nj);
    put(S.o_ints, ints.data(), 4 * ints.size());
    // the copy stream carries the upload so that it overlaps the solve's kernels
    HIPCHK(hipMemcpyAsync(base + in_begin, hs, S.out_begin - in_begin, hipMemcpyHostToDevice, LN().side_stream));
    HIPCHK(hipEventRecord(LN().ev_rp, LN().side_stream));
    return FNEC_OK;
}

int rp_finish(RpStaged &S, const std::vector<fnec_rp_row *> &outs, std::vector<double> &totals)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    const size_t nj = S.nj;
    totals.assign(nj, 0.0);
    if (!nj)
        return FNEC_OK;
    hipStream_t s = LN().stream;
    char *base = S.base;
    HIPCHK(hipStreamWaitEvent(s, LN().ev_rp, 0));
    const RpProb *dp = (const RpProb *)(base + S.o_probs);
    const int *wi = (const int *)(base + S.o_ints);
    hipEvent_t e_rp = prof_begin(PC_RP);
    if (S.pts_cnt)
        hipLaunchKernelGGL(k_rp_points, dim3((unsigned)S.pts_cnt), dim3(256), 0, s, dp,
                           (const int2 *)(wi + S.pts_off));
    if (S.norm_cnt)
        hipLaunchKernelGGL(k_rp_norm, dim3((unsigned)S.norm_cnt), dim3(256), 0, s, dp,
                           (const int2 *)(wi + S.norm_off));
    // RP unit: one (direction, segment[, image]) term = 32 flops + sin + cos
    prof_end(PC_RP, e_rp, S.rp_terms * 34.0, 0);
    hipLaunchKernelGGL(k_rp_finish, dim3((unsigned)nj), dim3(256), 0, s, dp, wi + S.all_off);
    HIPCHK(hipGetLastError());
    bool rows = false;
    for (size_t j = 0; j < nj; ++j)
        rows = rows || (outs.size() > j && outs[j] && S.npts[j]);
    const size_t end = rows ? S.out_end : S.rows_begin;
    if (S.fixed)
        S.hback.resize(end - S.out_begin);
    char *hb = S.fixed ? S.hback.data() : pinned(LN().h_back, LN().h_back_bytes, end - S.out_begin);
    if (!hb) {
        set_error("pinned host memory exhausted");
        return FNEC_RESOURCE_EXHAUSTED;
    }
    HIPCHK(hipMemcpyAsync(hb, base + S.out_begin, end - S.out_begin, hipMemcpyDeviceToHost, s));
    {
        HostTimer ht(kHostWait);
        HIPCHK(stream_wait(s));
    }
    HostTimer ht_post(kHostPost);
    prof_collect();
    for (size_t j = 0; j < nj; ++j) {
        totals[j] = *(const double *)(hb + (S.o_tot[j] - S.out_begin));
        if (outs.size() > j && outs[j] && S.npts[j])
            std::memcpy(outs[j], hb + (S.o_out[j] - S.out_begin), 32 * S.npts[j]);
    }
    return FNEC_OK;
}

int rp_run(const std::vector<RpHost> &jobs, std::vector<fnec_rp_row *> &outs, std::vector<double> &totals)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    RpStaged S;
    int st = rp_stage(jobs, S);
    if (st)
        return st;
    return rp_finish(S, outs, totals);
}

// One far-field job in the device scratch array: no device or pinned memory
// is allocated.
int rp_run_noalloc(const RpHost &job, fnec_rp_row *out, double &total)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    char *sp = g_ctx.device >= 0 && (size_t)g_ctx.device < g_ctx.scratch_dev.size()
                   ? g_ctx.scratch_dev[g_ctx.device] : nullptr;
    if (!sp) {
        set_error("device scratch not set up");
        return FNEC_INTERNAL;
    }
    RpStaged S;
    std::vector<RpHost> jobs{job};
    int st = rp_stage(jobs, S, sp, kNoAllocScratchBytes);
    if (st)
        return st;
    std::vector<fnec_rp_row *> outs{out};
    std::vector<double> totals;
    st = rp_finish(S, outs, totals);
    if (st)
        return st;
    total = totals[0];
    return FNEC_OK;
}

int rp_host_from(RpHost &h, const fnec_seg *segs, const fnec_c64 *cur, size_t n, double freq_hz,
                 const fnec_ground *ground, const double *theta, const double *phi, size_t npts)
{
    if (n == 0 || !segs || (!cur && !h.cur_dev) || (npts && (!theta || !phi))) {
        set_error("empty or null input");
        return FNEC_INVALID_ARG;
    }
    if (!(freq_hz > 0.0)) {
        set_error("frequency must be positive");
        return FNEC_INVALID_ARG;
    }
    h.n = (int)n;
    h.seg.resize(n * kSegStride);
    for (size_t i = 0; i < n; ++i) {
        double *o = &h.seg[i * kSegStride];
        for (int c = 0; c < 3; ++c) {
            o[c] = segs[i].mid[c];
            o[3 + c] = segs[i].dir[c];
        }
        o[6] = segs[i].length;
        // slot 7 (the far field has no use for the radius): 1 when the
        // segment continues a uniform run, i.e. has the previous segment's
        // length and direction and its midpoint is the previous midpoint plus
        // length * dir to within rounding (a straight wire's equal segments);
        // the far-field sum then advances the phase factor by a rotation
        double cont = 0.0;
        if (i > 0) {
            const fnec_seg &a = segs[i - 1], &b = segs[i];
            bool same = a.length == b.length;
            double scale = b.length, dev = 0.0;
            for (int c = 0; c < 3 && same; ++c) {
                same = a.dir[c] == b.dir[c];
                scale += std::fabs(a.mid[c]) + std::fabs(b.mid[c]);
                dev = std::max(dev, std::fabs(b.mid[c] - a.mid[c] - b.length * b.dir[c]));
            }
            if (same && dev <= 64.0 * 2.220446049250313e-16 * scale)
                cont = 1.0;
        }
        o[7] = cont;
    }
    if (cur) {
        h.cur.resize(2 * n);
        for (size_t i = 0; i < n; ++i) {
            h.cur[2 * i] = cur[i].re;
            h.cur[2 * i + 1] = cur[i].im;
        }
    }
    h.theta = theta;
    h.phi = phi;
    h.npts = npts;
    h.ground = 0;
    h.eps_r = h.sigma = 0.0;
    if (ground && ground->kind == FNEC_GROUND_PEC)
        h.ground = 1;
    else if (ground && ground->kind == FNEC_GROUND_FINITE) {
        h.ground = 2;
        h.eps_r = ground->eps_r;
        h.sigma = ground->sigma;
    }
    h.ndir = (h.ground ? 91 : 181) * 360;
    const double lambda = kC0 / freq_hz;
    h.k = 2.0 * kPi / lambda;
    return FNEC_OK;
}

int refmath_probe(int fn, const double *x, const double *y, double *out, size_t n)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    if (fn < 0 || fn > 6 || !x || !out || ((fn == 2 || fn == 4) && !y) || n > ((size_t)1 << 26)) {
        set_error("refmath probe: bad arguments");
        return FNEC_INVALID_ARG;
    }
    char *base = nullptr;
    const size_t bytes = 8 * n;
    int st = arena_reserve(3 * bytes + 1024, &base);
    if (st || n == 0)
        return st;
    hipStream_t s = LN().stream;
    double *dx = (double *)base, *dy = (double *)(base + ((bytes + 255) & ~(size_t)255)),
           *dout = (double *)(base + 2 * ((bytes + 255) & ~(size_t)255));
    HIPCHK(hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, s));
    if (y)
        HIPCHK(hipMemcpyAsync(dy, y, bytes, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_refmath_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fn, dx,
                       y ? dy : dx, dout, (int)n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return FNEC_OK;
}

int mfma_probe(const double *a, const double *b, double *out)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    char *base = nullptr;
    int st = arena_reserve(4096 * 4, &base);
    if (st)
        return st;
    hipStream_t s = LN().stream;
    HIPCHK(hipMemcpyAsync(base, a, 64 * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(base + 1024, b, 64 * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, (const double *)base,
                       (const double *)(base + 1024), (double *)(base + 2048));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, base + 2048, 256 * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return FNEC_OK;
}

int mfma_probe32(const float *a, const float *b, float *out)
{
    LaneScope lane_scope;
    if (lane_scope.status())
        return lane_scope.status();
    char *base = nullptr;
    int st = arena_reserve(4096 * 4, &base);
    if (st)
        return st;
    hipStream_t s = LN().stream;
    HIPCHK(hipMemcpyAsync(base, a, 64 * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(base + 1024, b, 64 * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_mfma_probe32, dim3(1), dim3(64), 0, s, (const float *)base,
                       (const float *)(base + 1024), (float *)(base + 2048));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, base + 2048, 256 * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return FNEC_OK;
}

int set_tunable(const char *name, long value)
{
    const std::string k = name ? name : "";
    if (k == "solve1_max_n" && value >= 0 && value <= kSolve1MaxNHost) {
        g_solve1_max_n = (int)value;
        return FNEC_OK;
    }
    if (k == "mix_first_chunk" && value >= 1 && value <= kMixMaxIt) {
        g_mix_first_chunk = (int)value;
        return FNEC_OK;
    }
    if (k == "gang_n" && value >= 0 && value <= 1 << 16) {
        g_gang_n = (int)value;
        return FNEC_OK;
    }
    if (k == "gang_max" && value >= 1 && value <= 64) {
        g_gang_max = (int)value;
        return FNEC_OK;
    }
    if (k == "mix_tol_exp" && value >= 8 && value <= 15) {
        g_mix_tol_exp = (int)value;
        return FNEC_OK;
    }
    if (k == "batch_split_n" && value >= 0 && value <= 1 << 16) {
        g_batch_split_n = (int)value;
        return FNEC_OK;
    }
    if (k == "update_tile_m" && (value == 64 || value == 128)) {
        g_update_tile_m = (int)value;
        return FNEC_OK;
    }
    if (k == "update_ring" && (value == 0 || value == 1)) {
        g_update_ring = (int)value;
        return FNEC_OK;
    }
    if (k == "update_k128" && (value == 0 || value == 1)) {
        g_update_k128 = (int)value;
        return FNEC_OK;
    }
    if (k == "exact_backsub" && (value == 0 || value == 1)) {
        g_exact_backsub = (int)value;
        return FNEC_OK;
    }
    if (k == "fill_centred" && (value == 0 || value == 1)) {
        g_fill_centred = (int)value;
        return FNEC_OK;
    }
    if (k == "lanes" && value >= 1 && value <= kMaxLanes) {
        std::lock_guard<std::mutex> lk(g_lane_mu);
        g_lanes_on = (int)value;
        g_lane_cv.notify_all();
        return FNEC_OK;
    }
    set_error("unknown tunable or value out of range");
    return FNEC_INVALID_ARG;
}

long get_tunable(const char *name)
{
    const std::string k = name ? name : "";
    if (k == "solve1_max_n")
        return g_solve1_max_n;
    if (k == "mix_first_chunk")
        return g_mix_first_chunk;
    if (k == "gang_n")
        return g_gang_n;
    if (k == "gang_max")
        return g_gang_max;
    if (k == "lanes")
        return g_lanes_on;
    if (k == "update_tile_m")
        return g_update_tile_m;
    if (k == "update_ring")
        return g_update_ring;
    if (k == "update_k128")
        return g_update_k128;
    if (k == "batch_split_n")
        return g_batch_split_n;
    if (k == "mix_tol_exp")
        return g_mix_tol_exp;
    if (k == "exact_backsub")
        return g_exact_backsub;
    if (k == "fill_centred")
        return g_fill_centred;
    return -1;
}

std::mutex &global_mutex() { return g_mu; }

} // namespace fnec
