// C ABI: contexts, evolution, observables, checkpointing, FFT sub-API.
//
// Host orchestration of the reference's driver functions
// (programs/bec_wigner_3d.cuda/wigner.cu): initEvolution, propagate,
// calculateEvolution and the observables, on the two-pass kernel scheme of
// passes.hpp.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "multi.hpp"

struct bec_ctx {
    int prec = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<bec::Group> groups;
    double t = 0.0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double *d_samples = nullptr;
    size_t sample_bytes = 0;
    // hipGraph of STEP_BLOCK full steps [K N] over all groups; kernel
    // arguments are device pointers only, so it stays valid when dt changes
    hipGraph_t body_graph = nullptr;
    hipGraphExec_t body_exec = nullptr;
    // whole segments K/2 N (K N)^(L-1) K/2 of up to SEGMENT_GRAPH_MAX steps,
    // one graph per segment length L, captured on first use or by
    // bec_propagate_prepare (a few lengths at most: the sampling period and
    // the remainder of a call)
    struct SegGraph {
        int steps;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<SegGraph> seg;
    bool use_graphs = true;
    // one stream per group: independent groups run as parallel graph branches
    std::vector<hipStream_t> aux;
    std::vector<hipEvent_t> aux_ev;
    hipEvent_t fork_ev = nullptr;
    // bec_set_multi_launch: every group in one launch per pass (multi.hip);
    // the plan holds the groups' device pointers and is rebuilt after a change
    bool multi = false;
    bool multi_stale = true;
    bec::MultiPlan multi_plan{};
};

static const int STEP_BLOCK = 32;
static const int SEGMENT_GRAPH_MAX = 64;   // longest segment captured whole
static const size_t SEGMENT_GRAPHS_KEPT = 8;

// Drops every captured graph: they hold the kernels' launch configuration
// and arguments, which schedules and the loss-noise switch change.
static void drop_graphs(bec_ctx *ctx)
{
    if (ctx->body_exec) (void)hipGraphExecDestroy(ctx->body_exec);
    if (ctx->body_graph) (void)hipGraphDestroy(ctx->body_graph);
    ctx->body_exec = nullptr;
    ctx->body_graph = nullptr;
    for (auto &s : ctx->seg) {
        (void)hipGraphExecDestroy(s.exec);
        (void)hipGraphDestroy(s.graph);
    }
    ctx->seg.clear();
    ctx->multi_stale = true;
}

namespace bec {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

int hip_status(hipError_t e)
{
    if (e == hipSuccess) return BEC_OK;
    if (e == hipErrorOutOfMemory) return BEC_ERR_ALLOC;
    if (e == hipErrorInvalidValue || e == hipErrorInvalidConfiguration) return BEC_ERR_SHAPE;
    return BEC_ERR_HIP;
}

hipError_t dev_alloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    return hipMalloc(p, bytes);
}

void dev_free(void *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

double k_of(int i, double dk, int n) { return (2 * i > n) ? dk * (i - n) : dk * i; }  // getK, wigner_kernel.cu:15-21

TrajParams make_traj_params(const bec_calc_params &c, double dt)
{
    TrajParams p{};
    p.xmax = c.xmax;
    p.ymax = c.ymax;
    p.zmax = c.zmax;
    p.dx = c.dx;
    p.dy = c.dy;
    p.dz = c.dz;
    p.inv_lambda2 = 1.0 / (c.lambda * c.lambda);
    p.g11 = c.g11;
    p.g12 = c.g12;
    p.g22 = c.g22;
    p.l111 = c.l111;
    p.l12 = c.l12;
    p.l22 = c.l22;
    p.detuning = c.detuning;
    p.dt = dt;
    p.dt_steady = c.dt_steady;
    p.N0 = (double)c.N;
    p.vac = c.V / (2.0 * c.dx * c.dy * c.dz);
    p.itmax = c.itmax;
    return p;
}

// (cos a, sin a) in FP32 as the float pair within two ulps of each whose
// modulus error c^2 + s^2 - 1 is smallest (ties: the smaller angle error), the
// rule of the in-register twiddles (fft_reg.hpp).  Rounded separately to the
// nearest floats, a table entry keeps a modulus error of up to 6e-8 that every
// step applies again to the same mode, a systematic drift of the FP32 norm.
static void unit_f32(double a, float *cr, float *sr)
{
    const double c0 = cos(a), s0 = sin(a);
    float cs[5], ss[5];
    cs[2] = (float)c0;
    ss[2] = (float)s0;
    for (int i = 1; i <= 2; ++i) {
        cs[2 + i] = nextafterf(cs[1 + i], INFINITY);
        cs[2 - i] = nextafterf(cs[3 - i], -INFINITY);
        ss[2 + i] = nextafterf(ss[1 + i], INFINITY);
        ss[2 - i] = nextafterf(ss[3 - i], -INFINITY);
    }
    double be = 1e300, bp = 1e300;
    for (float c : cs)
        for (float s : ss) {
            const double e = fabs((double)c * c + (double)s * s - 1.0);
            const double pe = fabs(c0 * (double)s - s0 * (double)c);
            if (e < be || (e == be && pe < bp)) {
                be = e;
                bp = pe;
                *cr = c;
                *sr = s;
            }
        }
}

// exp(i a) * scale into entry idx: FP32 from unit_f32 (scale a power of two,
// so the scaling is exact), FP64 from cos and sin directly
static void put_unit(std::vector<char> &buf, long idx, int prec, double a, double scale)
{
    if (prec == 0) {
        float c, s;
        unit_f32(a, &c, &s);
        float *f = (float *)buf.data();
        f[2 * idx] = (float)(c * scale);
        f[2 * idx + 1] = (float)(s * scale);
    } else {
        double *d = (double *)buf.data();
        d[2 * idx] = cos(a) * scale;
        d[2 * idx + 1] = sin(a) * scale;
    }
}

static void put_complex(std::vector<char> &buf, long idx, int prec, double re, double im)
{
    if (prec == 0) {
        float *f = (float *)buf.data();
        f[2 * idx] = (float)re;
        f[2 * idx + 1] = (float)im;
    } else {
        double *d = (double *)buf.data();
        d[2 * idx] = re;
        d[2 * idx + 1] = im;
    }
}

// Kinetic tables of variant v for record 0: separable factors, computed in
// double and rounded once.  KT_RT_*: exp(+i c dt k^2/2) with c = 1/2, 1
// (wigner_kernel.cu:137-151, sign quirk Q1 kept); KT_IM_HALF:
// exp(-dt_steady/2 * k^2/2) (wigner_kernel.cu:128-134).  1/nz is folded into
// the z factor.
int group_build_tables(Group &g, double dt, hipStream_t s)
{
    const bec_calc_params &c = g.calc;
    const int nx = c.nvx, ny = c.nvy, nz = c.nvz, nxy = nx * ny, prec = g.dev.prec;
    const size_t es = elem_size(prec);
    const int R = g.dev.n_records;
    std::vector<char> hxy((size_t)KT_COUNT * R * nxy * es), hz((size_t)KT_COUNT * R * nz * es);
    for (int v = 0; v < KT_COUNT; ++v)
        for (int r = 0; r < R; ++r) {
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    double kx = k_of(x, c.dkx, nx), ky = k_of(y, c.dky, ny);
                    double e = (kx * kx + ky * ky) / 2.0;
                    long idx = ((long)v * R + r) * nxy + y * nx + x;
                    if (v == KT_IM_HALF) {
                        put_complex(hxy, idx, prec, exp(-c.dt_steady / 2.0 * e), 0.0);
                    } else {
                        double ph = e * dt * (v == KT_RT_HALF ? 0.5 : 1.0);
                        put_unit(hxy, idx, prec, ph, 1.0);
                    }
                }
            for (int z = 0; z < nz; ++z) {
                double kz = k_of(z, c.dkz, nz);
                double e = kz * kz / 2.0;
                long idx = ((long)v * R + r) * nz + z;
                if (v == KT_IM_HALF) {
                    put_complex(hz, idx, prec, exp(-c.dt_steady / 2.0 * e) / nz, 0.0);
                } else {
                    double ph = e * dt * (v == KT_RT_HALF ? 0.5 : 1.0);
                    put_unit(hz, idx, prec, ph, 1.0 / nz);   // nz is a power of two
                }
            }
        }
    BEC_HIP(hipMemcpyAsync(g.dev.kxy, hxy.data(), hxy.size(), hipMemcpyHostToDevice, s));
    BEC_HIP(hipMemcpyAsync(g.dev.kz, hz.data(), hz.size(), hipMemcpyHostToDevice, s));
    g.hp = make_traj_params(c, dt);
    BEC_HIP(hipMemcpyAsync(g.dev.params, &g.hp, sizeof(TrajParams), hipMemcpyHostToDevice, s));
    BEC_HIP(hipStreamSynchronize(s));
    g.tables_dt = dt;
    g.tables_dts = c.dt_steady;
    return BEC_OK;
}

int group_setup(Group &g, const bec_calc_params &c, int prec, int ntraj, bool alloc_state)
{
    if (!shape_supported(c.nvx, c.nvy, c.nvz)) {
        set_last_error("unsupported lattice: need nvx == nvy in {16, 32}, nvz in {64, 128, 256}");
        return BEC_ERR_SHAPE;
    }
    if (ntraj <= 0) return BEC_ERR_INVALID;
    g.calc = c;
    g.cells = (long)c.nvx * c.nvy * c.nvz;
    DevGroup &d = g.dev;
    d.prec = prec;
    d.nx = c.nvx;
    d.ny = c.nvy;
    d.nz = c.nvz;
    d.ntraj = ntraj;
    d.rec_shift = 30;
    d.n_records = 1;
    // default schedule: the autotuned choice on MI355X for 128x16x16 batches
    // (one component per z workgroup, streaming accesses, XCD-aware placement)
    d.pw = 0;
    d.z_r1 = 0;
    d.z_split = 1;
    d.nt = prec == 0 ? 1 : 0;
    d.xcd = ntraj % 8 == 0 ? 1 : 0;
    d.xy_minw = 1;
    d.z_half = 0;
    d.xy_split = 0;
    d.itmax1 = c.itmax == 1 ? 1 : 0;
    const size_t es = elem_size(prec);
    if (alloc_state) {
        BEC_HIP(dev_alloc(&d.a, es * g.cells * ntraj));
        BEC_HIP(dev_alloc(&d.b, es * g.cells * ntraj));
    }
    BEC_HIP(dev_alloc((void **)&d.params, sizeof(TrajParams)));
    BEC_HIP(dev_alloc(&d.kxy, es * KT_COUNT * c.nvx * c.nvy));
    BEC_HIP(dev_alloc(&d.kz, es * KT_COUNT * c.nvz));
    BEC_HIP(dev_alloc(&d.tw, es * c.nvz));
    std::vector<char> tw(es * c.nvz);
    for (int j = 0; j < c.nvz; ++j) {
        double ang = 2.0 * M_PI * j / c.nvz;
        put_unit(tw, j, prec, ang, 1.0);
    }
    BEC_HIP(hipMemcpy(d.tw, tw.data(), tw.size(), hipMemcpyHostToDevice));
    g.partial_slots = xy_workgroups(d, ntraj);
    BEC_HIP(dev_alloc((void **)&g.d_partial, sizeof(double) * 8 * g.partial_slots));
    BEC_HIP(dev_alloc((void **)&g.d_sums, sizeof(double) * 4 * ntraj));
    return BEC_OK;
}

void group_release(Group &g)
{
    DevGroup &d = g.dev;
    dev_free(d.a);
    dev_free(d.b);
    void *p = d.params;
    dev_free(p);
    d.params = nullptr;
    dev_free(d.kxy);
    dev_free(d.kz);
    dev_free(d.tw);
    void *q = g.d_partial;
    dev_free(q);
    g.d_partial = nullptr;
    void *r = g.d_sums;
    dev_free(r);
    g.d_sums = nullptr;
    void *n = d.noise_step;
    dev_free(n);
    d.noise_step = nullptr;
    void *pj = g.d_proj;
    dev_free(pj);
    g.d_proj = nullptr;
    dev_free(g.d_avg);
}

int upload_complex(const double *h, void *d, long n, int prec, hipStream_t s)
{
    if (prec == 1) {
        BEC_HIP(hipMemcpyAsync(d, h, sizeof(double) * 2 * n, hipMemcpyHostToDevice, s));
        BEC_HIP(hipStreamSynchronize(s));
        return BEC_OK;
    }
    std::vector<float> f(2 * n);
    for (long i = 0; i < 2 * n; ++i) f[i] = (float)h[i];
    BEC_HIP(hipMemcpyAsync(d, f.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, s));
    BEC_HIP(hipStreamSynchronize(s));
    return BEC_OK;
}

int download_complex(const void *d, double *h, long n, int prec, hipStream_t s)
{
    if (prec == 1) {
        BEC_HIP(hipMemcpyAsync(h, d, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s));
        BEC_HIP(hipStreamSynchronize(s));
        return BEC_OK;
    }
    std::vector<float> f(2 * n);
    BEC_HIP(hipMemcpyAsync(f.data(), d, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, s));
    BEC_HIP(hipStreamSynchronize(s));
    for (long i = 0; i < 2 * n; ++i) h[i] = f[i];
    return BEC_OK;
}

} // namespace bec

using namespace bec;

#define CHECK_CTX(ctx)                                                                                 \
    do {                                                                                               \
        if (!(ctx)) return BEC_ERR_INVALID;                                                            \
    } while (0)
#define CHECK_GROUP(ctx, g)                                                                            \
    do {                                                                                               \
        CHECK_CTX(ctx);                                                                                \
        if ((g) < 0 || (g) >= (int)(ctx)->groups.size()) return BEC_ERR_INVALID;                      \
    } while (0)

extern "C" const char *bec_strerror(int status)
{
    switch (status) {
    case BEC_OK: return "ok";
    case BEC_ERR_INVALID: return "invalid argument";
    case BEC_ERR_SHAPE: return "unsupported grid shape or batch";
    case BEC_ERR_ALLOC: return "allocation failed";
    case BEC_ERR_HIP: return g_last_error.empty() ? "HIP runtime error" : g_last_error.c_str();
    case BEC_ERR_NOT_CONVERGED: return "ground state did not converge within max_iter";
    case BEC_ERR_NONFINITE: return "non-finite value detected in the state";
    case BEC_ERR_STATE: return "call out of order";
    case BEC_ERR_UNSUPPORTED: return "unsupported option";
    default: return "unknown status";
    }
}

extern "C" int bec_device_count(int *count)
{
    if (!count) return BEC_ERR_INVALID;
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return BEC_ERR_HIP;
    }
    return BEC_OK;
}

extern "C" int bec_create_groups(const bec_calc_params *groups, int n_groups, int precision, void *stream,
                                 bec_ctx **out)
{
    if (!groups || n_groups <= 0 || !out) return BEC_ERR_INVALID;
    if (precision != BEC_FP32 && precision != BEC_FP64) return BEC_ERR_INVALID;
    *out = nullptr;
    // validate every group before touching the device
    for (int i = 0; i < n_groups; ++i) {
        const bec_calc_params &c = groups[i];
        if (!shape_supported(c.nvx, c.nvy, c.nvz)) {
            set_last_error("unsupported lattice: need nvx == nvy in {16, 32}, nvz in {64, 128, 256}");
            return BEC_ERR_SHAPE;
        }
        if (c.ne <= 0 || c.cells != c.nvx * c.nvy * c.nvz || !(c.dx > 0) || !(c.dy > 0) || !(c.dz > 0))
            return BEC_ERR_INVALID;
    }
    bec_ctx *ctx = new bec_ctx();
    ctx->prec = precision;
    if (stream) {
        ctx->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete ctx;
            return hip_status(e);
        }
        ctx->own_stream = true;
    }
    ctx->groups.resize(n_groups);
    for (int i = 0; i < n_groups; ++i) {
        int st = group_setup(ctx->groups[i], groups[i], precision, groups[i].ne, true);
        if (st == BEC_OK) st = group_build_tables(ctx->groups[i], groups[i].dt_evo, ctx->stream);
        if (st != BEC_OK) {
            bec_destroy(ctx);
            return st;
        }
    }
    if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) != hipSuccess) {
        bec_destroy(ctx);
        return BEC_ERR_HIP;
    }
    if (n_groups > 1) {
        ctx->aux.assign(n_groups, nullptr);
        ctx->aux_ev.assign(n_groups, nullptr);
        for (int i = 0; i < n_groups; ++i)
            if (hipStreamCreateWithFlags(&ctx->aux[i], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&ctx->aux_ev[i], hipEventDisableTiming) != hipSuccess) {
                bec_destroy(ctx);
                return BEC_ERR_HIP;
            }
    }
    *out = ctx;
    return BEC_OK;
}

extern "C" int bec_create(const bec_calc_params *p, int precision, void *stream, bec_ctx **out)
{
    return bec_create_groups(p, 1, precision, stream, out);
}

extern "C" int bec_destroy(bec_ctx *ctx)
{
    CHECK_CTX(ctx);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    drop_graphs(ctx);
    for (auto &g : ctx->groups) group_release(g);
    void *p = ctx->d_samples;
    dev_free(p);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    for (auto e : ctx->aux_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto s : ctx->aux)
        if (s) (void)hipStreamDestroy(s);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BEC_OK;
}

extern "C" int bec_synchronize(bec_ctx *ctx)
{
    CHECK_CTX(ctx);
    BEC_HIP(hipStreamSynchronize(ctx->stream));
    return BEC_OK;
}

extern "C" int bec_group_count(const bec_ctx *ctx) { return ctx ? (int)ctx->groups.size() : 0; }
extern "C" double bec_get_time(const bec_ctx *ctx) { return ctx ? ctx->t : 0.0; }
extern "C" int bec_set_time(bec_ctx *ctx, double t)
{
    CHECK_CTX(ctx);
    ctx->t = t;
    return BEC_OK;
}

// ------------------------------------------------------------ evolution

static int ensure_tables(bec_ctx *ctx, double dt)
{
    for (auto &g : ctx->groups)
        if (g.tables_dt != dt) {
            int st = group_build_tables(g, dt, ctx->stream);
            if (st != BEC_OK) return st;
        }
    return BEC_OK;
}

static int sample_group(bec_ctx *ctx, Group &g, double *dst)
{
    BEC_HIP(run_xy_sample(g.dev, g.d_partial, ctx->stream));
    BEC_HIP(run_reduce_partials(g.d_partial, xy_wg_per_traj(g.dev), g.dev.ntraj, 4, dst, ctx->stream));
    return BEC_OK;
}

static int total_traj(const bec_ctx *ctx)
{
    int n = 0;
    for (auto &g : ctx->groups) n += g.dev.ntraj;
    return n;
}

static int ensure_multi_plan(bec_ctx *ctx)
{
    if (!ctx->multi_stale) return BEC_OK;
    std::vector<const DevGroup *> devs;
    for (auto &g : ctx->groups) devs.push_back(&g.dev);
    if (!multi_plan(devs.data(), (int)devs.size(), &ctx->multi_plan)) {
        set_last_error("these groups cannot share one launch (FP32 states, at most 8 groups, the same itmax == 1 "
                       "and loss-noise setting)");
        return BEC_ERR_UNSUPPORTED;
    }
    ctx->multi_stale = false;
    return BEC_OK;
}

// The kinetic z pass (table variant) of every group on stream s: one launch
// over all of them when the context steps its groups together, else one each.
static int z_pass_all(bec_ctx *ctx, int variant, hipStream_t s)
{
    if (ctx->multi) {
        int st = ensure_multi_plan(ctx);
        if (st != BEC_OK) return st;
        BEC_HIP(run_z_multi(ctx->multi_plan, variant, s));
        return BEC_OK;
    }
    for (auto &g : ctx->groups) BEC_HIP(run_z_kinetic(g.dev, variant, s));
    return BEC_OK;
}

// the nonlinear xy pass of every group, as z_pass_all
static int xy_pass_all(bec_ctx *ctx, hipStream_t s)
{
    if (ctx->multi) {
        int st = ensure_multi_plan(ctx);
        if (st != BEC_OK) return st;
        BEC_HIP(run_xy_multi(ctx->multi_plan, s));
        return BEC_OK;
    }
    for (auto &g : ctx->groups) BEC_HIP(run_xy_nonlinear(g.dev, s));
    return BEC_OK;
}

#define BEC_TRY(x)                                                                                                     \
    do {                                                                                                               \
        const int st_ = (x);                                                                                           \
        if (st_ != BEC_OK) return st_;                                                                                 \
    } while (0)

// count full steps [K N] (merged kinetic step + nonlinear step) on every group
static int launch_full_steps(bec_ctx *ctx, int count)
{
    for (int i = 0; i < count; ++i) {
        BEC_TRY(z_pass_all(ctx, KT_RT_FULL, ctx->stream));
        BEC_TRY(xy_pass_all(ctx, ctx->stream));
    }
    return BEC_OK;
}

// Capture STEP_BLOCK full steps.  With several groups each group's kernels go
// on its own stream, forked from and joined back into the capture stream, so
// the graph holds one independent branch per group (uneven grids overlap).
static int capture_branches(bec_ctx *ctx)
{
    if (ctx->groups.size() == 1 || ctx->multi) return launch_full_steps(ctx, STEP_BLOCK);
    BEC_HIP(hipEventRecord(ctx->fork_ev, ctx->stream));
    for (size_t i = 0; i < ctx->groups.size(); ++i) {
        hipStream_t s = ctx->aux[i];
        BEC_HIP(hipStreamWaitEvent(s, ctx->fork_ev, 0));
        Group &g = ctx->groups[i];
        for (int k = 0; k < STEP_BLOCK; ++k) {
            BEC_HIP(run_z_kinetic(g.dev, KT_RT_FULL, s));
            BEC_HIP(run_xy_nonlinear(g.dev, s));
        }
        BEC_HIP(hipEventRecord(ctx->aux_ev[i], s));
    }
    for (size_t i = 0; i < ctx->groups.size(); ++i) BEC_HIP(hipStreamWaitEvent(ctx->stream, ctx->aux_ev[i], 0));
    return BEC_OK;
}

static int ensure_body_graph(bec_ctx *ctx)
{
    if (ctx->body_exec) return BEC_OK;
    if (ctx->multi) BEC_TRY(ensure_multi_plan(ctx));   // outside the capture
    hipGraph_t graph = nullptr;
    BEC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int st = capture_branches(ctx);
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (st != BEC_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    BEC_HIP(e);
    e = hipGraphInstantiate(&ctx->body_exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        ctx->body_exec = nullptr;
        BEC_HIP(e);
    }
    ctx->body_graph = graph;
    return BEC_OK;
}

static int run_full_steps(bec_ctx *ctx, int count)
{
    if (ctx->use_graphs && count >= STEP_BLOCK) {
        int st = ensure_body_graph(ctx);
        if (st != BEC_OK) return st;
        for (int i = 0; i < count / STEP_BLOCK; ++i) BEC_HIP(hipGraphLaunch(ctx->body_exec, ctx->stream));
        count %= STEP_BLOCK;
    }
    return launch_full_steps(ctx, count);
}

// The kernels of one segment of L steps of one group: K/2 N (K N)^(L-1) K/2.
static int launch_segment(Group &g, int L, hipStream_t s)
{
    BEC_HIP(run_z_kinetic(g.dev, KT_RT_HALF, s));
    BEC_HIP(run_xy_nonlinear(g.dev, s));
    for (int k = 1; k < L; ++k) {
        BEC_HIP(run_z_kinetic(g.dev, KT_RT_FULL, s));
        BEC_HIP(run_xy_nonlinear(g.dev, s));
    }
    BEC_HIP(run_z_kinetic(g.dev, KT_RT_HALF, s));
    return BEC_OK;
}

// the same segment on every group, one launch per pass over all of them
static int launch_segment_multi(bec_ctx *ctx, int L, hipStream_t s)
{
    BEC_TRY(z_pass_all(ctx, KT_RT_HALF, s));
    BEC_TRY(xy_pass_all(ctx, s));
    for (int k = 1; k < L; ++k) {
        BEC_TRY(z_pass_all(ctx, KT_RT_FULL, s));
        BEC_TRY(xy_pass_all(ctx, s));
    }
    return z_pass_all(ctx, KT_RT_HALF, s);
}

// The graph of a whole segment of L steps (L <= SEGMENT_GRAPH_MAX), captured
// on first use: one launch per call of up to that many steps.  Several groups
// run as parallel branches, as in the step-block graph, or, when the context
// steps them together, as one launch per pass.
static int segment_graph(bec_ctx *ctx, int L, hipGraphExec_t *out)
{
    for (auto &s : ctx->seg)
        if (s.steps == L) {
            *out = s.exec;
            return BEC_OK;
        }
    if (ctx->multi) BEC_TRY(ensure_multi_plan(ctx));   // outside the capture
    hipGraph_t graph = nullptr;
    BEC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int st = BEC_OK;
    if (ctx->multi) {
        st = launch_segment_multi(ctx, L, ctx->stream);
    } else if (ctx->groups.size() == 1) {
        st = launch_segment(ctx->groups[0], L, ctx->stream);
    } else {
        hipError_t e = hipEventRecord(ctx->fork_ev, ctx->stream);
        for (size_t i = 0; i < ctx->groups.size() && e == hipSuccess && st == BEC_OK; ++i) {
            e = hipStreamWaitEvent(ctx->aux[i], ctx->fork_ev, 0);
            if (e == hipSuccess) st = launch_segment(ctx->groups[i], L, ctx->aux[i]);
            if (e == hipSuccess && st == BEC_OK) e = hipEventRecord(ctx->aux_ev[i], ctx->aux[i]);
        }
        for (size_t i = 0; i < ctx->groups.size() && e == hipSuccess && st == BEC_OK; ++i)
            e = hipStreamWaitEvent(ctx->stream, ctx->aux_ev[i], 0);
        if (st == BEC_OK && e != hipSuccess) {
            set_last_error(hipGetErrorString(e));
            st = hip_status(e);
        }
    }
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (st != BEC_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    BEC_HIP(e);
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        BEC_HIP(e);
    }
    if (ctx->seg.size() >= SEGMENT_GRAPHS_KEPT) {   // oldest out
        (void)hipGraphExecDestroy(ctx->seg.front().exec);
        (void)hipGraphDestroy(ctx->seg.front().graph);
        ctx->seg.erase(ctx->seg.begin());
    }
    ctx->seg.push_back({L, graph, exec});
    *out = exec;
    return BEC_OK;
}

// One call: segments of sample_every steps, each K/2 N (K N)^(L-1) K/2 and a
// read-only sample at its end; the trailing half step of one step and the
// leading half step of the next are merged into one K inside a segment.
static int propagate_impl(bec_ctx *ctx, int n_steps, double dt, int sample_every, double *d_samples)
{
    (void)dt;
    const int ntot = total_traj(ctx);
    int pos = 0, sample_idx = 0;
    while (pos < n_steps) {
        const int L = sample_every > 0 ? std::min(sample_every, n_steps - pos) : n_steps - pos;
        int st = BEC_OK;
        if (ctx->use_graphs && L <= SEGMENT_GRAPH_MAX) {
            hipGraphExec_t exec = nullptr;
            st = segment_graph(ctx, L, &exec);
            if (st != BEC_OK) return st;
            BEC_HIP(hipGraphLaunch(exec, ctx->stream));
        } else {
            BEC_TRY(z_pass_all(ctx, KT_RT_HALF, ctx->stream));
            BEC_TRY(xy_pass_all(ctx, ctx->stream));
            BEC_TRY(run_full_steps(ctx, L - 1));
            BEC_TRY(z_pass_all(ctx, KT_RT_HALF, ctx->stream));
        }
        if (sample_every > 0 && L == sample_every && d_samples) {
            long off = (long)sample_idx * ntot * 4;
            for (auto &g : ctx->groups) {
                st = sample_group(ctx, g, d_samples + off);
                if (st != BEC_OK) return st;
                off += (long)g.dev.ntraj * 4;
            }
            ++sample_idx;
        }
        pos += L;
    }
    return BEC_OK;
}

extern "C" int bec_propagate(bec_ctx *ctx, int n_steps, double dt, int sample_every, double *sums_out)
{
    CHECK_CTX(ctx);
    if (n_steps < 0) return BEC_ERR_INVALID;
    for (auto &g : ctx->groups)
        if (!g.initialized) return BEC_ERR_STATE;
    if (n_steps == 0) return BEC_OK;
    if (dt <= 0) dt = ctx->groups[0].calc.dt_evo;
    int st = ensure_tables(ctx, dt);
    if (st != BEC_OK) return st;
    const int n_samples = sample_every > 0 ? n_steps / sample_every : 0;
    const size_t bytes = sizeof(double) * 4 * (size_t)total_traj(ctx) * n_samples;
    if (n_samples > 0 && sums_out && bytes > ctx->sample_bytes) {
        void *p = ctx->d_samples;
        dev_free(p);
        ctx->d_samples = nullptr;
        BEC_HIP(dev_alloc((void **)&ctx->d_samples, bytes));
        ctx->sample_bytes = bytes;
    }
    st = propagate_impl(ctx, n_steps, dt, sums_out ? sample_every : 0, sums_out ? ctx->d_samples : nullptr);
    if (st != BEC_OK) return st;
    ctx->t += n_steps * dt;
    if (n_samples > 0 && sums_out) {
        BEC_HIP(hipMemcpyAsync(sums_out, ctx->d_samples, bytes, hipMemcpyDeviceToHost, ctx->stream));
    }
    BEC_HIP(hipStreamSynchronize(ctx->stream));
    return BEC_OK;
}

extern "C" int bec_propagate_prepare(bec_ctx *ctx, int n_steps, double dt, int sample_every)
{
    CHECK_CTX(ctx);
    if (n_steps < 0) return BEC_ERR_INVALID;
    for (auto &g : ctx->groups)
        if (!g.initialized) return BEC_ERR_STATE;
    if (n_steps == 0 || !ctx->use_graphs) return BEC_OK;
    if (dt <= 0) dt = ctx->groups[0].calc.dt_evo;
    int st = ensure_tables(ctx, dt);
    if (st != BEC_OK) return st;
    // the segment lengths the call runs: the period, and the remainder
    const int first = sample_every > 0 ? std::min(sample_every, n_steps) : n_steps;
    const int last = sample_every > 0 && n_steps % sample_every ? n_steps % sample_every : first;
    for (int L : {first, last}) {
        if (L <= SEGMENT_GRAPH_MAX) {
            hipGraphExec_t exec = nullptr;
            st = segment_graph(ctx, L, &exec);
        } else if (L - 1 >= STEP_BLOCK) {
            st = ensure_body_graph(ctx);
        }
        if (st != BEC_OK) return st;
    }
    return BEC_OK;
}

extern "C" int bec_propagate_graph(bec_ctx *ctx, int n_steps, double dt, int steps_per_graph)
{
    CHECK_CTX(ctx);
    if (n_steps < 0 || steps_per_graph <= 0 || n_steps % steps_per_graph) return BEC_ERR_INVALID;
    for (auto &g : ctx->groups)
        if (!g.initialized) return BEC_ERR_STATE;
    if (n_steps == 0) return BEC_OK;
    if (dt <= 0) dt = ctx->groups[0].calc.dt_evo;
    int st = ensure_tables(ctx, dt);
    if (st != BEC_OK) return st;
    if (ctx->multi) BEC_TRY(ensure_multi_plan(ctx));   // outside the capture
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    BEC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const bool ug = ctx->use_graphs;
    ctx->use_graphs = false;
    st = propagate_impl(ctx, steps_per_graph, dt, 0, nullptr);
    ctx->use_graphs = ug;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (st != BEC_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    BEC_HIP(e);
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        BEC_HIP(e);
    }
    for (int r = 0; r < n_steps / steps_per_graph; ++r) {
        e = hipGraphLaunch(exec, ctx->stream);
        if (e != hipSuccess) break;
    }
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    BEC_HIP(e);
    BEC_HIP(e2);
    ctx->t += n_steps * dt;
    return BEC_OK;
}

extern "C" int bec_pulse(bec_ctx *ctx, int kind)
{
    CHECK_CTX(ctx);
    if (kind != BEC_PULSE_HALF_PI && kind != BEC_PULSE_PI) return BEC_ERR_INVALID;
    for (auto &g : ctx->groups) BEC_HIP(run_pulse(g.dev, kind, ctx->stream));
    BEC_HIP(hipStreamSynchronize(ctx->stream));
    return BEC_OK;
}

extern "C" int bec_pulse_rotation(bec_ctx *ctx, double theta, double phi)
{
    CHECK_CTX(ctx);
    if (!std::isfinite(theta) || !std::isfinite(phi)) return BEC_ERR_INVALID;
    for (auto &g : ctx->groups)
        if (!g.initialized) return BEC_ERR_STATE;
    const double C = cos(0.5 * theta), S = sin(0.5 * theta), cp = cos(phi), sp = sin(phi);
    // U = [[C, -i e^{-i phi} S], [-i e^{i phi} S, C]]
    const double u[8] = {C, 0.0, -sp * S, -cp * S, sp * S, -cp * S, C, 0.0};
    for (auto &g : ctx->groups) BEC_HIP(run_rotation(g.dev, u, ctx->stream));
    BEC_HIP(hipStreamSynchronize(ctx->stream));
    return BEC_OK;
}

// The equilibration loop of initEvolution (wigner.cu:222-223) is written in
// float: for(t = 0; t <= t_equilib; t += dt_evo).  It runs at least once.
static int equilibration_steps(const bec_calc_params &c)
{
    float te = (float)c.t_equilib, dt = (float)c.dt_evo;
    if (!(dt > 0.0f)) return 1;
    int n = 0;
    for (float t = 0.0f; t <= te; t += dt) ++n;
    return n;
}

extern "C" int bec_init_evolution(bec_ctx *ctx, int group, const double *gs_x, uint64_t seed, int equilibrate)
{
    CHECK_GROUP(ctx, group);
    if (!gs_x) return BEC_ERR_INVALID;
    Group &g = ctx->groups[group];
    const size_t es = elem_size(g.dev.prec);
    void *d_gs = nullptr;
    BEC_HIP(dev_alloc(&d_gs, es * g.cells));
    int st = upload_complex(gs_x, d_gs, g.cells, g.dev.prec, ctx->stream);
    if (st == BEC_OK) {
        const bec_calc_params &c = g.calc;
        hipError_t e = run_noise_init(g.dev, d_gs, seed, c.Va, c.Vb, 1.0 / sqrt(c.dx * c.dy * c.dz), ctx->stream);
        if (e == hipSuccess) {
            // x-space -> hybrid: transform along x, y with 1/(nx ny)
            const double sc = 1.0 / (c.nvx * c.nvy);
            e = run_xy_fft(g.dev, g.dev.a, g.dev.a, g.dev.ntraj, +1, sc, ctx->stream);
            if (e == hipSuccess) e = run_xy_fft(g.dev, g.dev.b, g.dev.b, g.dev.ntraj, +1, sc, ctx->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            set_last_error(hipGetErrorString(e));
            st = hip_status(e);
        }
    }
    dev_free(d_gs);
    if (st != BEC_OK) return st;
    g.initialized = true;
    if (equilibrate) {
        // equilibration runs on this group only
        const int n = equilibration_steps(g.calc);
        st = ensure_tables(ctx, g.calc.dt_evo);
        if (st != BEC_OK) return st;
        for (int s = 0; s < n; ++s) {
            BEC_HIP(run_z_kinetic(g.dev, s == 0 ? KT_RT_HALF : KT_RT_FULL, ctx->stream));
            BEC_HIP(run_xy_nonlinear(g.dev, ctx->stream));
        }
        BEC_HIP(run_z_kinetic(g.dev, KT_RT_HALF, ctx->stream));
    }
    BEC_HIP(run_pulse(g.dev, BEC_PULSE_HALF_PI, ctx->stream));
    BEC_HIP(hipStreamSynchronize(ctx->stream));
    return BEC_OK;
}

// ---------------------------------------------------------- observables

extern "C" int bec_trajectory_sums(bec_ctx *ctx, int group, double *sums)
{
    CHECK_GROUP(ctx, group);
    if (!sums) return BEC_ERR_INVALID;
    Group &g = ctx->groups[group];
    if (!g.initialized) return BEC_ERR_STATE;
    int st = sample_group(ctx, g, g.d_sums);
    if (st != BEC_OK) return st;
    BEC_HIP(hipMemcpyAsync(sums, g.d_sums, sizeof(double) * 4 * g.dev.ntraj, hipMemcpyDeviceToHost, ctx->stream));
    BEC_HIP(hipStreamSynchronize(ctx->stream));
    return BEC_OK;
}

// Observables from the per-trajectory sums S_aa, S_bb, S_ab = sum conj(a) b.
// halfPiRotate(alpha) (wigner_kernel.cu:344-360) gives per cell
//   |a'|^2 = (|a|^2 + |b|^2)/2 + Re(conj(a) b (sin - i cos)) ,
//   |b'|^2 = (|a|^2 + |b|^2)/2 - Re(conj(a) b (sin - i cos)) ... with
// Re(conj(a) b (s - i c)) = s Re(conj(a) b) + c Im(conj(a) b), so the sums
// after any virtual pulse follow from the four sums exactly.
static void obs_from_sums(const bec_calc_params &c, const double *sums, int ne, double t, bec_obs *o)
{
    double saa = 0, sbb = 0, sre = 0, sim = 0;
    for (int e = 0; e < ne; ++e) {
        saa += sums[4 * e];
        sbb += sums[4 * e + 1];
        sre += sums[4 * e + 2];
        sim += sums[4 * e + 3];
    }
    const double dV = c.dx * c.dy * c.dz;
    const double vac_total = (double)ne * c.cells * c.V / (2.0 * dV);
    auto rotated = [&](double alpha, double *n1, double *n2) {
        const double cr = sin(alpha) * sre + cos(alpha) * sim;
        *n1 = 0.5 * (saa + sbb) + cr - vac_total;
        *n2 = 0.5 * (saa + sbb) - cr - vac_total;
    };
    double n1, n2;
    rotated(0.0, &n1, &n2);
    o->t = t;
    o->Na = n1 * dV;
    o->Nb = n2 * dV;
    o->N_total = (n1 + n2) * dV / ne;
    double vmax = 0;
    for (float alpha = 0.0f; alpha < 2.0f * (float)M_PI; alpha += 0.5f) {   // getVisibility, wigner.cu:267-277
        double a1, a2;
        rotated((double)alpha, &a1, &a2);
        a1 = fabs(a1);
        a2 = fabs(a2);
        double r = (a1 - a2) / (a1 + a2);
        if (fabs(r) > vmax) vmax = fabs(r);
    }
    o->visibility_grid = vmax;
    const double N1 = saa - vac_total, N2 = sbb - vac_total;
    o->N1 = N1 * dV;
    o->N2 = N2 * dV;
    o->visibility = 2.0 * sqrt(sre * sre + sim * sim) / (N1 + N2);
    o->phase = atan2(sim, sre);
    // spread of the per-trajectory populations after the virtual pulse (alpha = 0)
    // (two passes over the trajectories: mean, then squared deviations)
    const double vac1 = (double)c.cells * c.V / (2.0 * dV);
    auto pop = [&](int e, int which) {
        const double *s4 = sums + 4 * e;
        return (0.5 * (s4[0] + s4[1]) + (which ? -s4[3] : s4[3]) - vac1) * dV;
    };
    double m1 = 0, m2 = 0, q1 = 0, q2 = 0;
    for (int e = 0; e < ne; ++e) {
        m1 += pop(e, 0);
        m2 += pop(e, 1);
    }
    m1 /= ne;
    m2 /= ne;
    for (int e = 0; e < ne; ++e) {
        q1 += (pop(e, 0) - m1) * (pop(e, 0) - m1);
        q2 += (pop(e, 1) - m2) * (pop(e, 1) - m2);
    }
    o->Na_stderr = ne > 1 ? sqrt(q1 / ne / (ne - 1)) : 0.0;
    o->Nb_stderr = ne > 1 ? sqrt(q2 / ne / (ne - 1)) : 0.0;
}

extern "C" int bec_observables(bec_ctx *ctx, int group, bec_obs *out)
{
    CHECK_GROUP(ctx, group);
    if (!out) return BEC_ERR_INVALID;
    Group &g = ctx->groups[group];
    std::vector<double> sums(4 * (size_t)g.dev.ntraj);
    int st = bec_trajectory_sums(ctx, group, sums.data());
    if (st != BEC_OK) return st;
    for (double v : sums)
        if (!std::isfinite(v)) return BEC_ERR_NONFINITE;
    obs_from_sums(g.calc, sums.data(), g.dev.ntraj, ctx->t, out);
    return BEC_OK;
}

extern "C" int bec_calculate_evolution(bec_ctx *ctx, double dt, bec_obs *out)
{
    CHECK_CTX(ctx);
    int st = bec_propagate(ctx, 1, dt, 0, nullptr);
    if (st != BEC_OK) return st;
    if (out)
        for (int i = 0; i < (int)ctx->groups.size(); ++i) {
            st = bec_observables(ctx, i, &out[i]);
            if (st != BEC_OK) return st;
        }
    return BEC_OK;
}

namespace bec {
// The projections of wigner.cu:279-303,388-410 of group `group`, computed on
// the device into the workspace `buf` of `cap` doubles (BEC_ERR_SHAPE when
// projection_doubles() exceeds it), then copied into the caller's arrays (each
// may be NULL): xy [nvy][nvx] (sum over z), zy [nvy][nvz] (sum over x),
// axial [nvz].  buf NULL selects the group's own workspace, allocated on first use.
int ctx_projections(bec_ctx *ctx, int group, double alpha, double *buf, size_t cap, double *xy_a, double *xy_b,
                    double *zy_a, double *zy_b, double *axial_a, double *axial_b)
{
    CHECK_GROUP(ctx, group);
    if (!std::isfinite(alpha)) return BEC_ERR_INVALID;
    Group &g = ctx->groups[group];
    if (!g.initialized) return BEC_ERR_STATE;
    const long need = projection_doubles(g.dev);
    if (!buf) {
        if (!g.d_proj) BEC_HIP(dev_alloc((void **)&g.d_proj, sizeof(double) * need));
        buf = g.d_proj;
        cap = (size_t)need;
    }
    if (need > (long)cap) {
        set_last_error("projections: the lattice needs " + std::to_string(need) + " doubles of workspace, " +
                       std::to_string(cap) + " are available");
        return BEC_ERR_SHAPE;
    }
    const long nx = g.calc.nvx, ny = g.calc.nvy, nz = g.calc.nvz;
    const double vac = g.calc.V / (2.0 * g.calc.dx * g.calc.dy * g.calc.dz);
    BEC_HIP(run_projections(g.dev, alpha, vac, buf, ctx->stream));
    const double *xy = buf + 2 * g.cells, *zy = xy + 2 * nx * ny, *ax = zy + 2 * ny * nz;
    double *dst[6] = {xy_a, xy_b, zy_a, zy_b, axial_a, axial_b};
    const double *src[6] = {xy, xy + nx * ny, zy, zy + ny * nz, ax, ax + nz};
    const long len[6] = {nx * ny, nx * ny, ny * nz, ny * nz, nz, nz};
    for (int i = 0; i < 6; ++i)
        if (dst[i])
            BEC_HIP(hipMemcpyAsync(dst[i], src[i], sizeof(double) * len[i], hipMemcpyDeviceToHost, ctx->stream));
    BEC_HIP(hipStreamSynchronize(ctx->stream));
    return BEC_OK;
}
} // namespace bec

extern "C" int bec_projections(bec_ctx *ctx, int group, double alpha, double *xy_a, double *xy_b, double *zy_a,
                               double *zy_b, double *axial_a, double *axial_b)
{
    return ctx_projections(ctx, group, alpha, nullptr, 0, xy_a, xy_b, zy_a, zy_b, axial_a, axial_b);
}

// ---------------------------------------------------------- checkpoint

extern "C" int bec_get_state(bec_ctx *ctx, int group, int space, double *a, double *b)
{
    CHECK_GROUP(ctx, group);
    if (space != BEC_SPACE_X && space != BEC_SPACE_K && space != BEC_SPACE_HYBRID) return BEC_ERR_INVALID;
    Group &g = ctx->groups[group];
    if (!g.initialized) return BEC_ERR_STATE;
    const long n = g.cells * g.dev.ntraj;
    if (space == BEC_SPACE_HYBRID) {   // the stored representation, exact (FP32 values widen exactly)
        int st = a ? download_complex(g.dev.a, a, n, g.dev.prec, ctx->stream) : BEC_OK;
        if (st == BEC_OK && b) st = download_complex(g.dev.b, b, n, g.dev.prec, ctx->stream);
        return st;
    }
    void *tmp = nullptr;
    BEC_HIP(dev_alloc(&tmp, elem_size(g.dev.prec) * n));
    int st = BEC_OK;
    for (int c = 0; c < 2 && st == BEC_OK; ++c) {
        double *h = c ? b : a;
        if (!h) continue;
        void *src = c ? g.dev.b : g.dev.a;
        hipError_t e = space == BEC_SPACE_X
                           ? run_xy_fft(g.dev, src, tmp, g.dev.ntraj, -1, 1.0, ctx->stream)
                           : run_z_fft(g.dev, src, tmp, g.dev.ntraj, +1, 1.0 / g.calc.nvz, ctx->stream);
        if (e != hipSuccess) {
            set_last_error(hipGetErrorString(e));
            st = hip_status(e);
            break;
        }
        st = download_complex(tmp, h, n, g.dev.prec, ctx->stream);
    }
    dev_free(tmp);
    return st;
}

extern "C" int bec_set_state(bec_ctx *ctx, int group, int space, const double *a, const double *b)
{
    CHECK_GROUP(ctx, group);
    if (space != BEC_SPACE_X && space != BEC_SPACE_K && space != BEC_SPACE_HYBRID) return BEC_ERR_INVALID;
    if (!a || !b) return BEC_ERR_INVALID;
    Group &g = ctx->groups[group];
    const long n = g.cells * g.dev.ntraj;
    for (int c = 0; c < 2; ++c) {
        void *dst = c ? g.dev.b : g.dev.a;
        int st = upload_complex(c ? b : a, dst, n, g.dev.prec, ctx->stream);
        if (st != BEC_OK) return st;
        if (space == BEC_SPACE_HYBRID) continue;
        hipError_t e = space == BEC_SPACE_X
                           ? run_xy_fft(g.dev, dst, dst, g.dev.ntraj, +1, 1.0 / (g.calc.nvx * g.calc.nvy), ctx->stream)
                           : run_z_fft(g.dev, dst, dst, g.dev.ntraj, -1, 1.0, ctx->stream);
        if (e != hipSuccess) {
            set_last_error(hipGetErrorString(e));
            return hip_status(e);
        }
    }
    BEC_HIP(hipStreamSynchronize(ctx->stream));
    g.initialized = true;
    return BEC_OK;
}

extern "C" int bec_calculate_averages(bec_ctx *ctx, int group, bec_avg_report *out)
{
    CHECK_GROUP(ctx, group);
    if (!out) return BEC_ERR_INVALID;
    Group &g = ctx->groups[group];
    if (!g.initialized) return BEC_ERR_STATE;
    const long n = g.cells * g.dev.ntraj;
    const size_t sbytes = elem_size(g.dev.prec) * n;
    if (!g.d_avg) BEC_HIP(dev_alloc(&g.d_avg, 2 * sbytes + sizeof(double) * 6 * g.cells));
    void *ax = g.d_avg, *bx = (char *)g.d_avg + sbytes;
    double *d_out = (double *)((char *)g.d_avg + 2 * sbytes);
    BEC_HIP(run_xy_fft(g.dev, g.dev.a, ax, g.dev.ntraj, -1, 1.0, ctx->stream));
    BEC_HIP(run_xy_fft(g.dev, g.dev.b, bx, g.dev.ntraj, -1, 1.0, ctx->stream));
    BEC_HIP(run_averages(g.dev, ax, bx, d_out, ctx->stream));
    std::vector<double> h(6 * (size_t)g.cells);
    BEC_HIP(hipMemcpyAsync(h.data(), d_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    BEC_HIP(hipStreamSynchronize(ctx->stream));
    double t[6] = {0, 0, 0, 0, 0, 0};
    for (long i = 0; i < g.cells; ++i)
        for (int k = 0; k < 6; ++k) t[k] += h[6 * i + k];
    const bec_calc_params &c = g.calc;
    const int ne = g.dev.ntraj;
    const double dV = c.dx * c.dy * c.dz;
    const double norm = (double)c.N / ((t[0] + t[1]) / ne);
    out->N = (t[0] + t[1]) * dV / ne;
    out->norm = norm;
    out->var_re_a = t[2] / ((double)g.cells * ne) * norm;
    out->var_im_a = t[3] / ((double)g.cells * ne) * norm;
    out->var_re_b = t[4] / ((double)g.cells * ne) * norm;
    out->var_im_b = t[5] / ((double)g.cells * ne) * norm;
    return BEC_OK;
}

// -------------------------------------------------------------- timing

extern "C" int bec_time_passes(bec_ctx *ctx, int group, int iters, double *ms)
{
    CHECK_GROUP(ctx, group);
    if (iters <= 0 || !ms) return BEC_ERR_INVALID;
    Group &g = ctx->groups[group];
    if (!g.initialized) return BEC_ERR_STATE;
    int st = ensure_tables(ctx, g.tables_dt > 0 ? g.tables_dt : g.calc.dt_evo);
    if (st != BEC_OK) return st;
    float t = 0;
    // z-pass (full kinetic step)
    BEC_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for (int i = 0; i < iters; ++i) BEC_HIP(run_z_kinetic(g.dev, KT_RT_FULL, ctx->stream));
    BEC_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    BEC_HIP(hipEventSynchronize(ctx->ev1));
    BEC_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    ms[0] = t / iters;
    BEC_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for (int i = 0; i < iters; ++i) BEC_HIP(run_xy_nonlinear(g.dev, ctx->stream));
    BEC_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    BEC_HIP(hipEventSynchronize(ctx->ev1));
    BEC_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    ms[1] = t / iters;
    BEC_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    for (int i = 0; i < iters; ++i) BEC_HIP(run_xy_sample(g.dev, g.d_partial, ctx->stream));
    BEC_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    BEC_HIP(hipEventSynchronize(ctx->ev1));
    BEC_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    ms[2] = t / iters;
    return BEC_OK;
}

// ------------------------------------------------------------- FFT API

struct bec_fft_plan {
    bec::DevGroup dev{};
    void *tw = nullptr;
};

extern "C" int bec_fft_plan3d(bec_fft_plan **plan, int nz, int ny, int nx, int precision, int batch)
{
    if (!plan || batch <= 0) return BEC_ERR_INVALID;
    if (precision != BEC_FP32 && precision != BEC_FP64) return BEC_ERR_INVALID;
    *plan = nullptr;
    if (!shape_supported(nx, ny, nz)) return BEC_ERR_SHAPE;
    bec_fft_plan *p = new bec_fft_plan();
    p->dev.prec = precision;
    p->dev.nx = nx;
    p->dev.ny = ny;
    p->dev.nz = nz;
    p->dev.ntraj = batch;
    const size_t es = elem_size(precision);
    std::vector<char> tw(es * nz);
    for (int j = 0; j < nz; ++j) put_unit(tw, j, precision, 2.0 * M_PI * j / nz, 1.0);
    if (dev_alloc(&p->dev.tw, tw.size()) != hipSuccess ||
        hipMemcpy(p->dev.tw, tw.data(), tw.size(), hipMemcpyHostToDevice) != hipSuccess) {
        dev_free(p->dev.tw);
        delete p;
        return BEC_ERR_ALLOC;
    }
    *plan = p;
    return BEC_OK;
}

extern "C" int bec_fft_execute(bec_fft_plan *plan, const void *d_in, void *d_out, int sign, void *stream)
{
    if (!plan || !d_in || !d_out) return BEC_ERR_INVALID;
    if (sign != BEC_FFT_FORWARD && sign != BEC_FFT_INVERSE) return BEC_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    BEC_HIP(run_xy_fft(plan->dev, d_in, d_out, plan->dev.ntraj, sign, 1.0, s));
    BEC_HIP(run_z_fft(plan->dev, d_out, d_out, plan->dev.ntraj, sign, 1.0, s));
    BEC_HIP(hipStreamSynchronize(s));
    return BEC_OK;
}

extern "C" int bec_fft_destroy(bec_fft_plan *plan)
{
    if (!plan) return BEC_ERR_INVALID;
    dev_free(plan->dev.tw);
    delete plan;
    return BEC_OK;
}

// --------------------------------------------------------- device memory

extern "C" int bec_device_alloc(void **d_ptr, size_t bytes)
{
    if (!d_ptr) return BEC_ERR_INVALID;
    BEC_HIP(hipMalloc(d_ptr, bytes));
    return BEC_OK;
}
extern "C" int bec_device_free(void *d_ptr)
{
    BEC_HIP(hipFree(d_ptr));
    return BEC_OK;
}
extern "C" int bec_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes)
{
    BEC_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
    return BEC_OK;
}
extern "C" int bec_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes)
{
    BEC_HIP(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
    return BEC_OK;
}

namespace bec {
const bec_calc_params *ctx_group_params(bec_ctx *ctx, int group, hipStream_t *stream)
{
    if (!ctx || group < 0 || group >= (int)ctx->groups.size()) return nullptr;
    if (stream) *stream = ctx->stream;
    return &ctx->groups[group].calc;
}
} // namespace bec

extern "C" int bec_set_loss_noise(bec_ctx *ctx, int enable, uint64_t seed, unsigned first_step)
{
    CHECK_CTX(ctx);
    BEC_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < ctx->groups.size(); ++i) {
        DevGroup &d = ctx->groups[i].dev;
        if (!enable) {
            void *q = d.noise_step;
            dev_free(q);
            d.noise_step = nullptr;
            continue;
        }
        if (!d.noise_step) BEC_HIP(dev_alloc((void **)&d.noise_step, sizeof(unsigned)));
        BEC_HIP(hipMemcpy(d.noise_step, &first_step, sizeof(unsigned), hipMemcpyHostToDevice));
        d.noise_k0 = (uint32_t)seed;
        d.noise_k1 = (uint32_t)(seed >> 32);
        d.group_index = (int)i;
    }
    // a captured step graph holds the launches without (or with) the noise
    drop_graphs(ctx);
    return BEC_OK;
}

extern "C" int bec_get_loss_noise_step(bec_ctx *ctx, int group, unsigned *step)
{
    CHECK_GROUP(ctx, group);
    if (!step) return BEC_ERR_INVALID;
    const DevGroup &d = ctx->groups[group].dev;
    if (!d.noise_step) return BEC_ERR_STATE;
    BEC_HIP(hipStreamSynchronize(ctx->stream));
    BEC_HIP(hipMemcpy(step, d.noise_step, sizeof(unsigned), hipMemcpyDeviceToHost));
    return BEC_OK;
}

extern "C" int bec_set_graphs(bec_ctx *ctx, int enable)
{
    CHECK_CTX(ctx);
    ctx->use_graphs = enable != 0;
    return BEC_OK;
}

extern "C" int bec_set_multi_launch(bec_ctx *ctx, int enable)
{
    CHECK_CTX(ctx);
    drop_graphs(ctx);
    ctx->multi = false;
    if (!enable) return BEC_OK;
    int st = ensure_multi_plan(ctx);
    if (st != BEC_OK) return st;
    ctx->multi = true;
    return BEC_OK;
}

extern "C" int bec_device_copy(void *d_dst, const void *d_src, size_t bytes, void *stream)
{
    if (!d_dst || !d_src || bytes % 16) return BEC_ERR_INVALID;
    BEC_HIP(bec::run_copy(d_dst, d_src, bytes, (hipStream_t)stream));
    BEC_HIP(hipStreamSynchronize((hipStream_t)stream));
    return BEC_OK;
}

// ------------------------------------------------------------ autotuner

extern "C" int bec_set_schedule(bec_ctx *ctx, int group, const bec_schedule *sch)
{
    CHECK_GROUP(ctx, group);
    if (!sch) return BEC_ERR_INVALID;
    Group &g = ctx->groups[group];
    Schedule cand[64];
    const int n = schedule_candidates(g.dev.prec, g.dev.nx, g.dev.nz, g.dev.ntraj, cand, 64, true);
    bool ok = false;
    for (int i = 0; i < n && i < 64; ++i)
        if (cand[i].pw == sch->pw && cand[i].z_r1 == sch->z_r1 && cand[i].z_split == sch->z_split &&
            cand[i].nt == sch->nt && cand[i].xcd == sch->xcd && cand[i].xy_minw == sch->xy_minw &&
            cand[i].z_half == sch->z_half && cand[i].xy_split == sch->xy_split)
            ok = true;
    if (!ok) return BEC_ERR_UNSUPPORTED;
    g.dev.pw = sch->pw;
    g.dev.z_r1 = sch->z_r1;
    g.dev.z_split = sch->z_split;
    g.dev.nt = sch->nt;
    g.dev.xcd = sch->xcd;
    g.dev.xy_minw = sch->xy_minw;
    g.dev.z_half = sch->z_half;
    g.dev.xy_split = sch->xy_split;
    // a captured step graph holds the old launch configuration
    drop_graphs(ctx);
    return BEC_OK;
}

extern "C" int bec_get_schedule(bec_ctx *ctx, int group, bec_schedule *sch)
{
    CHECK_GROUP(ctx, group);
    if (!sch) return BEC_ERR_INVALID;
    const DevGroup &d = ctx->groups[group].dev;
    sch->pw = d.pw ? d.pw : default_pw(d.prec, d.nx, d.nz);
    sch->z_r1 = d.z_r1 ? d.z_r1 : default_r1(d.nz);
    sch->z_split = d.z_split;
    sch->nt = d.nt;
    sch->xcd = d.xcd;
    sch->xy_minw = d.xy_minw ? d.xy_minw : 1;
    sch->z_half = d.z_half;
    sch->xy_split = d.xy_split;
    sch->ms_per_step = 0;
    return BEC_OK;
}

extern "C" int bec_schedule_candidates(bec_ctx *ctx, int group, bec_schedule *out, int max_out, int *count)
{
    CHECK_GROUP(ctx, group);
    if (!count) return BEC_ERR_INVALID;
    const DevGroup &d = ctx->groups[group].dev;
    Schedule cand[64];
    const int n = schedule_candidates(d.prec, d.nx, d.nz, d.ntraj, cand, 64);
    for (int i = 0; i < n && i < max_out && i < 64; ++i) {
        out[i].pw = cand[i].pw;
        out[i].z_r1 = cand[i].z_r1;
        out[i].z_split = cand[i].z_split;
        out[i].nt = cand[i].nt;
        out[i].xcd = cand[i].xcd;
        out[i].xy_minw = cand[i].xy_minw;
        out[i].z_half = cand[i].z_half;
        out[i].xy_split = cand[i].xy_split;
        out[i].ms_per_step = 0;
    }
    *count = n;
    return BEC_OK;
}

// Time every candidate schedule on this context's own lattice and batch with
// alternating z/xy passes (the real access pattern), keep the fastest.  The
// state is saved before and restored after.
extern "C" int bec_autotune(bec_ctx *ctx, int group, int steps, bec_schedule *results, int max_results,
                            int *n_results, bec_schedule *best)
{
    CHECK_GROUP(ctx, group);
    if (steps <= 0) return BEC_ERR_INVALID;
    if (n_results) *n_results = 0;
    Group &g = ctx->groups[group];
    if (!g.initialized) return BEC_ERR_STATE;
    int st = ensure_tables(ctx, g.tables_dt > 0 ? g.tables_dt : g.calc.dt_evo);
    if (st != BEC_OK) return st;
    const size_t bytes = elem_size(g.dev.prec) * g.cells * g.dev.ntraj;
    void *sa = nullptr, *sb = nullptr;
    BEC_HIP(dev_alloc(&sa, bytes));
    hipError_t e = dev_alloc(&sb, bytes);
    if (e != hipSuccess) {
        dev_free(sa);
        BEC_HIP(e);
    }
    BEC_HIP(hipMemcpyAsync(sa, g.dev.a, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    BEC_HIP(hipMemcpyAsync(sb, g.dev.b, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    Schedule cand[64];
    const int n = std::min(64, schedule_candidates(g.dev.prec, g.dev.nx, g.dev.nz, g.dev.ntraj, cand, 64));
    const DevGroup saved = g.dev;
    double best_ms = 1e30;
    int best_i = -1;
    for (int i = 0; i < n && st == BEC_OK; ++i) {
        g.dev.pw = cand[i].pw;
        g.dev.z_r1 = cand[i].z_r1;
        g.dev.z_split = cand[i].z_split;
        g.dev.nt = cand[i].nt;
        g.dev.xcd = cand[i].xcd;
        g.dev.xy_minw = cand[i].xy_minw;
        g.dev.z_half = cand[i].z_half;
        g.dev.xy_split = cand[i].xy_split;
        // The steps run as propagate() will run them: replayed from a captured graph
        // when the context uses graphs (at small batches, kernels launched one by
        // one are bound by the launch rate and every candidate times alike).  The
        // first repetition warms up; the fastest of the other three is kept, so
        // that one slow repetition does not decide between candidates a percent apart.
        auto launch_steps = [&]() -> hipError_t {
            hipError_t r = hipSuccess;
            for (int k = 0; k < steps && r == hipSuccess; ++k) {
                r = run_z_kinetic(g.dev, KT_RT_FULL, ctx->stream);
                if (r == hipSuccess) r = run_xy_nonlinear(g.dev, ctx->stream);
            }
            return r;
        };
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (ctx->use_graphs) {
            e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                const hipError_t el = launch_steps();
                e = hipStreamEndCapture(ctx->stream, &graph);
                if (el != hipSuccess) e = el;
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        }
        float ms = 1e30f;
        for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {
            float r = 0;
            e = hipEventRecord(ctx->ev0, ctx->stream);
            if (e == hipSuccess) e = exec ? hipGraphLaunch(exec, ctx->stream) : launch_steps();
            if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
            if (e == hipSuccess) e = hipEventElapsedTime(&r, ctx->ev0, ctx->ev1);
            if (e == hipSuccess && rep > 0 && r < ms) ms = r;
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            set_last_error(hipGetErrorString(e));
            st = hip_status(e);
            break;
        }
        const double per = ms / steps;
        if (results && i < max_results) {
            results[i].pw = cand[i].pw;
            results[i].z_r1 = cand[i].z_r1;
            results[i].z_split = cand[i].z_split;
            results[i].nt = cand[i].nt;
            results[i].xcd = cand[i].xcd;
            results[i].xy_minw = cand[i].xy_minw;
            results[i].z_half = cand[i].z_half;
            results[i].xy_split = cand[i].xy_split;
            results[i].ms_per_step = per;
        }
        if (per < best_ms) {
            best_ms = per;
            best_i = i;
        }
    }
    g.dev = saved;
    if (st == BEC_OK && best_i >= 0) {
        bec_schedule b{cand[best_i].pw, cand[best_i].z_r1, cand[best_i].z_split, cand[best_i].nt,
                       cand[best_i].xcd, cand[best_i].xy_minw, cand[best_i].z_half, cand[best_i].xy_split,
                       best_ms};
        st = bec_set_schedule(ctx, group, &b);
        if (best) *best = b;
    }
    e = hipMemcpyAsync(g.dev.a, sa, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(g.dev.b, sb, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(sa);
    dev_free(sb);
    if (st != BEC_OK) return st;
    BEC_HIP(e);
    if (n_results) *n_results = n;
    return BEC_OK;
}

extern "C" int bec_time_copy(size_t bytes, int iters, double *ms_per_copy)
{
    if (bytes == 0 || bytes % 16 || iters <= 0 || !ms_per_copy) return BEC_ERR_INVALID;
    void *a = nullptr, *b = nullptr;
    BEC_HIP(dev_alloc(&a, bytes));
    hipError_t e = dev_alloc(&b, bytes);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = run_copy(b, a, bytes, nullptr);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (int i = 0; i < iters && e == hipSuccess; ++i) e = run_copy(b, a, bytes, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(a);
    dev_free(b);
    BEC_HIP(e);
    *ms_per_copy = ms / iters;
    return BEC_OK;
}
