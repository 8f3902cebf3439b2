// Ground state: calculateSteadyState (programs/bec_wigner_3d.cuda/wigner.cu:84-175)
// and its batched, residual-targeted extension.  All reductions, the
// convergence tests and the renormalisation run in FP64 on the GPU; the host
// only polls the per-trajectory done flags between blocks of iterations.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gs_backend.hpp"
#include "host_common.hpp"

namespace bec {

namespace {

struct GsWork {
    Group g;
    void *X = nullptr, *W = nullptr, *P = nullptr, *KC = nullptr;
    double *ph = nullptr, *pr = nullptr, *pm = nullptr, *pn = nullptr;
    double *sums = nullptr;        // [batch][8] reduced sums
    double *exy = nullptr, *ez = nullptr, *mu_tf = nullptr, *ones = nullptr;
    GsScalars sc{};
    int *zero_done = nullptr;
    double *target = nullptr;      // pair mode: per-trajectory atom numbers
    int pair = 0;                  // 1: trajectories 2j, 2j+1 are the components of set j
    hipStream_t stream = nullptr;  // the solve's stream: clears are ordered with its kernels
    std::vector<void *> owned;

    ~GsWork()
    {
        for (void *&p : owned) dev_free(p);
        group_release(g);
    }
    template <typename Ptr> hipError_t alloc(Ptr **p, size_t bytes)
    {
        void *q = nullptr;
        hipError_t e = dev_alloc(&q, bytes);
        if (e == hipSuccess) {
            owned.push_back(q);
            // on the solve's (non-blocking) stream: a null-stream memset is not
            // ordered with it and could land after the first uploads or kernels
            e = hipMemsetAsync(q, 0, bytes, stream);
        }
        *p = (Ptr *)q;
        return e;
    }
};

#define GS_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            set_last_error(std::string(#call) + ": " + hipGetErrorString(e_));                        \
            return hip_status(e_);                                                                     \
        }                                                                                              \
    } while (0)

void put_c(std::vector<char> &buf, long idx, int prec, double re, double im)
{
    if (prec == 0) {
        ((float *)buf.data())[2 * idx] = (float)re;
        ((float *)buf.data())[2 * idx + 1] = (float)im;
    } else {
        ((double *)buf.data())[2 * idx] = re;
        ((double *)buf.data())[2 * idx + 1] = im;
    }
}

// Per-record parameters and tables for a batch of parameter sets.  One record
// per set; with frac_a != nullptr (pair mode) each set has two trajectories,
// its components, with atom numbers frac_a N and (1 - frac_a) N.
int setup_work(GsWork &w, const bec_calc_params *params, int batch, int prec, hipStream_t s,
               const double *frac_a = nullptr)
{
    const bec_calc_params &c0 = params[0];
    for (int i = 1; i < batch; ++i)
        if (params[i].nvx != c0.nvx || params[i].nvy != c0.nvy || params[i].nvz != c0.nvz) return BEC_ERR_SHAPE;
    w.stream = s;
    w.pair = frac_a ? 1 : 0;
    const int ntraj = batch << w.pair;
    int st = group_setup(w.g, c0, prec, ntraj, false);
    if (st != BEC_OK) return st;
    DevGroup &d = w.g.dev;
    // small FP64 batches: 16 pencils per z workgroup (twice the workgroups of the
    // default 32, each with half the work); a solve of one set is latency-bound
    if (prec == BEC_FP64 && (long)ntraj * c0.nvx * c0.nvy / 32 < 256) d.pw = 16;
    // one parameter record per set
    void *p = d.params;
    dev_free(p);
    dev_free(d.kxy);
    dev_free(d.kz);
    d.params = nullptr;
    d.rec_shift = w.pair;
    d.n_records = batch;
    const int nx = c0.nvx, ny = c0.nvy, nz = c0.nvz, nxy = nx * ny;
    const long cells = w.g.cells;
    const size_t es = elem_size(prec);
    GS_HIP(dev_alloc((void **)&d.params, sizeof(TrajParams) * batch));
    GS_HIP(dev_alloc(&d.kxy, es * KT_COUNT * batch * nxy));
    GS_HIP(dev_alloc(&d.kz, es * KT_COUNT * batch * nz));
    std::vector<TrajParams> hp(batch);
    std::vector<char> hxy(es * KT_COUNT * batch * nxy, 0), hz(es * KT_COUNT * batch * nz, 0);
    std::vector<double> exy((size_t)batch * nxy), ez((size_t)batch * nz), mu(batch);
    for (int r = 0; r < batch; ++r) {
        const bec_calc_params &c = params[r];
        hp[r] = make_traj_params(c, c.dt_evo);
        mu[r] = c.mu;
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const double kx = k_of(x, c.dkx, nx), ky = k_of(y, c.dky, ny);
                const double e = (kx * kx + ky * ky) / 2.0;
                exy[(size_t)r * nxy + y * nx + x] = e;
                put_c(hxy, ((long)KT_IM_HALF * batch + r) * nxy + y * nx + x, prec, exp(-c.dt_steady / 2.0 * e), 0.0);
            }
        for (int z = 0; z < nz; ++z) {
            const double kz = k_of(z, c.dkz, nz);
            const double e = kz * kz / 2.0;
            ez[(size_t)r * nz + z] = e;
            put_c(hz, ((long)KT_IM_HALF * batch + r) * nz + z, prec, exp(-c.dt_steady / 2.0 * e) / nz, 0.0);
        }
    }
    GS_HIP(hipMemcpyAsync(d.params, hp.data(), sizeof(TrajParams) * batch, hipMemcpyHostToDevice, s));
    GS_HIP(hipMemcpyAsync(d.kxy, hxy.data(), hxy.size(), hipMemcpyHostToDevice, s));
    GS_HIP(hipMemcpyAsync(d.kz, hz.data(), hz.size(), hipMemcpyHostToDevice, s));
    GS_HIP(w.alloc(&w.exy, sizeof(double) * exy.size()));
    GS_HIP(w.alloc(&w.ez, sizeof(double) * ez.size()));
    GS_HIP(w.alloc(&w.mu_tf, sizeof(double) * batch));
    GS_HIP(hipMemcpyAsync(w.exy, exy.data(), sizeof(double) * exy.size(), hipMemcpyHostToDevice, s));
    GS_HIP(hipMemcpyAsync(w.ez, ez.data(), sizeof(double) * ez.size(), hipMemcpyHostToDevice, s));
    GS_HIP(hipMemcpyAsync(w.mu_tf, mu.data(), sizeof(double) * batch, hipMemcpyHostToDevice, s));
    const size_t sbytes = es * cells * ntraj;
    GS_HIP(w.alloc(&w.X, sbytes));
    GS_HIP(w.alloc(&w.W, sbytes));
    GS_HIP(w.alloc(&w.P, sbytes));
    GS_HIP(w.alloc(&w.KC, sbytes));
    const long slots = xy_workgroups(d, ntraj);
    GS_HIP(w.alloc(&w.ph, sizeof(double) * 5 * slots));
    GS_HIP(w.alloc(&w.pr, sizeof(double) * 1 * slots));
    GS_HIP(w.alloc(&w.pm, sizeof(double) * 4 * slots));
    GS_HIP(w.alloc(&w.pn, sizeof(double) * 1 * slots));
    GS_HIP(w.alloc(&w.sums, sizeof(double) * 16 * ntraj));
    GS_HIP(w.alloc(&w.ones, sizeof(double) * ntraj));
    GS_HIP(w.alloc(&w.sc.s, sizeof(double) * ntraj));
    GS_HIP(w.alloc(&w.sc.mu, sizeof(double) * ntraj));
    GS_HIP(w.alloc(&w.sc.e_prev, sizeof(double) * ntraj));
    GS_HIP(w.alloc(&w.sc.res, sizeof(double) * ntraj));
    GS_HIP(w.alloc(&w.sc.e_ref, sizeof(double) * ntraj));
    GS_HIP(w.alloc(&w.sc.iters, sizeof(int) * ntraj));
    GS_HIP(w.alloc(&w.sc.done, sizeof(int) * ntraj));
    GS_HIP(w.alloc(&w.zero_done, sizeof(int) * ntraj));
    w.sc.pair = w.pair;
    w.sc.target = nullptr;
    std::vector<double> ones(ntraj, 1.0), tgt(ntraj);
    if (w.pair) {
        for (int r = 0; r < batch; ++r) {
            tgt[2 * r] = frac_a[r] * params[r].N;
            tgt[2 * r + 1] = (1.0 - frac_a[r]) * params[r].N;
        }
        GS_HIP(w.alloc(&w.target, sizeof(double) * ntraj));
        GS_HIP(hipMemcpyAsync(w.target, tgt.data(), sizeof(double) * ntraj, hipMemcpyHostToDevice, s));
        w.sc.target = w.target;
    }
    GS_HIP(hipMemcpyAsync(w.ones, ones.data(), sizeof(double) * ntraj, hipMemcpyHostToDevice, s));
    GS_HIP(hipMemcpyAsync(w.sc.s, ones.data(), sizeof(double) * ntraj, hipMemcpyHostToDevice, s));
    GS_HIP(hipStreamSynchronize(s));
    return BEC_OK;
}

// Residual machinery on the x-space iterate X (scale sc.s pending).
hipError_t residual_pass(GsWork &w, const GsScalars &sc, double tol, int check, hipStream_t s)
{
    const DevGroup &d = w.g.dev;
    GsTables t{w.exy, w.ez};
    hipError_t e;
    if ((e = gs_xy_forward(d, w.X, w.W, sc, s)) != hipSuccess) return e;
    if ((e = gs_z_diag(d, w.W, sc, t, 0, 0.0, nullptr, s)) != hipSuccess) return e;
    if ((e = gs_xy_hamiltonian(d, w.X, w.W, sc, w.ph, s)) != hipSuccess) return e;
    if ((e = gs_scalar_mu(d, w.ph, sc, s)) != hipSuccess) return e;
    if ((e = gs_xy_residual(d, w.X, w.W, sc, w.pr, s)) != hipSuccess) return e;
    return gs_scalar_residual(d, w.ph, w.pr, sc, tol, check, s);
}

int split_step_iteration(GsWork &w, const bec_gs_options &o, int it, hipStream_t s)
{
    const DevGroup &d = w.g.dev;
    GsScalars once = w.sc;
    once.s = w.ones;
    GsTables t{w.exy, w.ez};
    GS_HIP(gs_z_imag(d, w.P, w.sc, nullptr, s));          // K/2 with the pending renormalisation
    GS_HIP(gs_xy_nonlinear_imag(d, w.P, w.sc, s));        // x-space step
    GS_HIP(gs_z_imag(d, w.P, once, w.KC, s));             // K/2, keep psi_k for the energy integral
    GS_HIP(gs_xy_measure_ref(d, w.P, w.KC, w.sc, t, w.pm, s));
    GS_HIP(gs_scalar_ref(d, w.pm, w.sc, o.tol, o.rule == BEC_GS_RULE_REFERENCE, s));
    if (o.rule == BEC_GS_RULE_RESIDUAL && (it + 1) % std::max(1, o.check_every) == 0) {
        GS_HIP(run_xy_fft(d, w.P, w.X, d.ntraj, -1, 1.0, s));
        GS_HIP(residual_pass(w, w.sc, o.tol, 1, s));
    }
    return BEC_OK;
}

// One gradient-flow iteration in six launches: the per-trajectory scalars
// (mu, the residual test, the renormalisation of the previous update) are
// evaluated inside the passes that consume them.  Iteration 0 starts from
// the normalised guess; later ones renormalise the previous update first.
int gradient_iteration(GsWork &w, const bec_gs_options &o, int it, double alpha, hipStream_t s)
{
    const DevGroup &d = w.g.dev;
    GsTables t{w.exy, w.ez};
    const int check = o.rule == BEC_GS_RULE_RESIDUAL && (it % std::max(1, o.check_every) == 0);
    GS_HIP(gs_xy_forward(d, w.X, w.W, w.sc, s, it > 0 ? w.pn : nullptr));
    GS_HIP(gs_z_diag(d, w.W, w.sc, t, 0, 0.0, nullptr, s));
    GS_HIP(gs_xy_hamiltonian(d, w.X, w.W, w.sc, w.ph, s));
    GS_HIP(gs_xy_residual(d, w.X, w.W, w.sc, w.pr, s, w.ph));
    // preconditioner shift: the option, else each record's Thomas-Fermi mu
    const GsResidualTest test{w.ph, w.pr, o.tol, check};
    GS_HIP(gs_z_diag(d, w.W, w.sc, t, 1, alpha, alpha > 0 ? nullptr : w.mu_tf, s, &test));
    GS_HIP(gs_xy_update(d, w.X, w.P, w.W, w.sc, o.tau, o.beta, w.pn, s));
    return BEC_OK;
}

// E, mu, N, residual and the reference integrals of the final x-space state X.
int final_report(GsWork &w, const bec_calc_params *params, int batch, hipStream_t s, bec_gs_report *reps)
{
    const DevGroup &d = w.g.dev;
    GsScalars sc = w.sc;
    sc.done = w.zero_done;
    sc.s = w.ones;
    GS_HIP(residual_pass(w, sc, 0.0, 0, s));
    const int per = xy_wg_per_traj(d);
    std::vector<double> h5(5 * (size_t)batch), r1(batch), m4(4 * (size_t)batch), mu(batch), res(batch);
    GS_HIP(run_reduce_partials(w.ph, per, batch, 5, w.sums, s));
    GS_HIP(hipMemcpyAsync(h5.data(), w.sums, sizeof(double) * 5 * batch, hipMemcpyDeviceToHost, s));
    GS_HIP(hipMemcpyAsync(res.data(), sc.res, sizeof(double) * batch, hipMemcpyDeviceToHost, s));
    GS_HIP(hipMemcpyAsync(mu.data(), sc.mu, sizeof(double) * batch, hipMemcpyDeviceToHost, s));
    GS_HIP(hipStreamSynchronize(s));
    // reference integrals: psi_k (normalised) and the hybrid of psi
    GsTables t{w.exy, w.ez};
    GS_HIP(gs_xy_forward(d, w.X, w.W, sc, s));
    GS_HIP(run_z_fft(d, w.W, w.KC, batch, +1, 1.0 / d.nz, s));
    GS_HIP(gs_xy_measure_ref(d, w.W, w.KC, sc, t, w.pm, s));
    GS_HIP(run_reduce_partials(w.pm, per, batch, 4, w.sums + 8 * batch, s));
    GS_HIP(hipMemcpyAsync(m4.data(), w.sums + 8 * batch, sizeof(double) * 4 * batch, hipMemcpyDeviceToHost, s));
    GS_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < batch; ++i) {
        const bec_calc_params &c = params[i];
        const double dV = c.dx * c.dy * c.dz;
        const double *h = &h5[5 * i];
        const double N = h[1] * dV;
        bec_gs_report &r = reps[i];
        r.N = N;
        r.mu = mu[i];
        r.E = (h[2] + h[3] + c.g11 * h[4] / 2.0) * dV / N;
        r.residual = res[i];
        const double *m = &m4[4 * i];
        r.E_ref = dV * (m[1] + m[3] + c.g11 * m[2] / 2.0) / c.N;
        r.mu_ref = dV * (m[1] + m[3] + c.g11 * m[2]) / c.N;
    }
    return BEC_OK;
}

// Pair mode: E, mu per component, N per component and the pair residual of
// the final x-space state X.  With K_c, P_c the kinetic and potential sums and
// N_c mu_c = K_c + P_c + (interaction of c), E = (K + P + sum_c N_c mu_c) / 2.
int final_report2(GsWork &w, const bec_calc_params *params, int batch, hipStream_t s, bec_gs2_report *reps)
{
    const DevGroup &d = w.g.dev;
    const int ntraj = d.ntraj;
    GsScalars sc = w.sc;
    sc.done = w.zero_done;
    sc.s = w.ones;
    GS_HIP(residual_pass(w, sc, 0.0, 0, s));
    const int per = xy_wg_per_traj(d);
    std::vector<double> h5(5 * (size_t)ntraj), mu(ntraj), res(ntraj);
    GS_HIP(run_reduce_partials(w.ph, per, ntraj, 5, w.sums, s));
    GS_HIP(hipMemcpyAsync(h5.data(), w.sums, sizeof(double) * 5 * ntraj, hipMemcpyDeviceToHost, s));
    GS_HIP(hipMemcpyAsync(res.data(), sc.res, sizeof(double) * ntraj, hipMemcpyDeviceToHost, s));
    GS_HIP(hipMemcpyAsync(mu.data(), sc.mu, sizeof(double) * ntraj, hipMemcpyDeviceToHost, s));
    GS_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < batch; ++i) {
        const bec_calc_params &c = params[i];
        const double dV = c.dx * c.dy * c.dz;
        const double *ha = &h5[5 * (2 * i)], *hb = &h5[5 * (2 * i + 1)];
        bec_gs2_report &r = reps[i];
        r.N_a = ha[1] * dV;
        r.N_b = hb[1] * dV;
        r.mu_a = mu[2 * i];
        r.mu_b = mu[2 * i + 1];
        r.E = (ha[2] + ha[3] + hb[2] + hb[3] + ha[0] + hb[0]) * dV / 2.0 / (r.N_a + r.N_b);
        r.residual = std::max(res[2 * i], res[2 * i + 1]);
    }
    return BEC_OK;
}

} // namespace

// Argument checks done before any device call.
int gs_validate(const bec_calc_params *params, int batch, const bec_gs_options *opt)
{
    if (!params || batch <= 0 || !opt) return BEC_ERR_INVALID;
    const bec_gs_options &o = *opt;
    if (o.method != BEC_GS_SPLIT_STEP && o.method != BEC_GS_GRADIENT) return BEC_ERR_INVALID;
    if (o.rule != BEC_GS_RULE_REFERENCE && o.rule != BEC_GS_RULE_RESIDUAL) return BEC_ERR_INVALID;
    if (o.method == BEC_GS_GRADIENT && o.rule != BEC_GS_RULE_RESIDUAL) return BEC_ERR_INVALID;
    if (o.max_iter <= 0 || !(o.tol > 0)) return BEC_ERR_INVALID;
    if (o.precision != BEC_FP32 && o.precision != BEC_FP64) return BEC_ERR_INVALID;
    if (o.method == BEC_GS_GRADIENT && !(o.tau > 0)) return BEC_ERR_INVALID;
    for (int i = 0; i < batch; ++i) {
        const bec_calc_params &c = params[i];
        if (!shape_supported(c.nvx, c.nvy, c.nvz)) return BEC_ERR_SHAPE;
        if (c.nvx != params[0].nvx || c.nvy != params[0].nvy || c.nvz != params[0].nvz) return BEC_ERR_SHAPE;
        if (c.N <= 0 || !(c.dt_steady > 0) || !(c.dx > 0) || !(c.dy > 0) || !(c.dz > 0)) return BEC_ERR_INVALID;
    }
    return BEC_OK;
}

// net != nullptr (two-component solves only): the start is the surrogate
// network's output on the device, net [batch][cells][2], with its spin weight
int gs_solve(const bec_calc_params *params, int batch, const double *guesses, const bec_gs_options *opt,
             hipStream_t s, double *out_x, bec_gs_report *reps, const double *frac_a, bec_gs2_report *reps2,
             const float *net = nullptr, double spin_weight = 1.0)
{
    int vst = gs_validate(params, batch, opt);
    if (vst != BEC_OK) return vst;
    const bec_gs_options o = *opt;
    GsWork w;
    int st = setup_work(w, params, batch, o.precision, s, frac_a);
    if (st != BEC_OK) return st;
    DevGroup &d = w.g.dev;
    const int ntraj = d.ntraj;
    const long n = w.g.cells * ntraj;

    // start: guesses, the network's output, or Thomas-Fermi (fillWithTFGroundState), normalised to N
    if (net) {
        // the amplitude unit sqrt(mu / g11) of each set, as the host conversion uses it
        std::vector<double> unit(batch);
        for (int j = 0; j < batch; ++j) unit[j] = std::sqrt(params[j].mu / params[j].g11);
        double *d_unit = nullptr;
        GS_HIP(dev_alloc((void **)&d_unit, sizeof(double) * batch));
        hipError_t e = hipMemcpyAsync(d_unit, unit.data(), sizeof(double) * batch, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = gs_net_start(d, w.X, net, d_unit, spin_weight, 1.0 / std::sqrt(2.0), s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        void *p = d_unit;
        dev_free(p);
        GS_HIP(e);
    } else if (guesses) {
        st = upload_complex(guesses, w.X, n, o.precision, s);
        if (st != BEC_OK) return st;
    } else {
        GS_HIP(gs_tf_start(d, w.X, w.mu_tf, s));
    }
    GS_HIP(gs_norm_only(d, w.X, w.sc, w.pn, s));
    const double alpha = o.alpha > 0 ? o.alpha : 0.0;   // 0: per-record mu
    if (o.method == BEC_GS_SPLIT_STEP) {
        GS_HIP(gs_xy_forward(d, w.X, w.P, w.sc, s));     // x -> hybrid, renormalised
        GS_HIP(gs_apply_scale(d, w.X, w.sc, s));          // keeps X consistent, s <- 1
    } else {
        GS_HIP(gs_apply_scale(d, w.X, w.sc, s));
        GS_HIP(hipMemcpyAsync(w.P, w.X, elem_size(o.precision) * n, hipMemcpyDeviceToDevice, s));
    }
    hipEvent_t e0, e1;
    GS_HIP(hipEventCreate(&e0));
    GS_HIP(hipEventCreate(&e1));
    GS_HIP(hipEventRecord(e0, s));
    std::vector<int> done(ntraj, 0);
    int it = 0;
    const int block = 16;
    while (it < o.max_iter) {
        const int nb = std::min(block, o.max_iter - it);
        for (int j = 0; j < nb; ++j, ++it) {
            st = o.method == BEC_GS_SPLIT_STEP ? split_step_iteration(w, o, it, s)
                                               : gradient_iteration(w, o, it, alpha, s);
            if (st != BEC_OK) break;
        }
        if (st != BEC_OK) break;
        GS_HIP(hipMemcpyAsync(done.data(), w.sc.done, sizeof(int) * ntraj, hipMemcpyDeviceToHost, s));
        GS_HIP(hipStreamSynchronize(s));
        if (std::all_of(done.begin(), done.end(), [](int v) { return v != 0; })) break;
    }
    GS_HIP(hipEventRecord(e1, s));
    GS_HIP(hipEventSynchronize(e1));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != BEC_OK) return st;
    if (o.method == BEC_GS_SPLIT_STEP) GS_HIP(run_xy_fft(d, w.P, w.X, batch, -1, 1.0, s));
    // the last update's renormalisation (trajectories still running; done ones skip)
    if (o.method == BEC_GS_GRADIENT && it > 0) GS_HIP(gs_scalar_norm(d, w.pn, w.sc, s));
    GS_HIP(gs_apply_scale(d, w.X, w.sc, s));
    std::vector<int> iters(ntraj);
    GS_HIP(hipMemcpyAsync(iters.data(), w.sc.iters, sizeof(int) * ntraj, hipMemcpyDeviceToHost, s));
    GS_HIP(hipMemcpyAsync(done.data(), w.sc.done, sizeof(int) * ntraj, hipMemcpyDeviceToHost, s));
    GS_HIP(hipStreamSynchronize(s));
    if (w.pair) {
        std::vector<bec_gs2_report> r2(batch);
        st = final_report2(w, params, batch, s, r2.data());
        if (st != BEC_OK) return st;
        if (out_x) {
            st = download_complex(w.X, out_x, n, o.precision, s);
            if (st != BEC_OK) return st;
        }
        bool all_conv = true, finite = true;
        for (int i = 0; i < batch; ++i) {
            r2[i].iterations = iters[2 * i];
            r2[i].converged = done[2 * i] == 1 && done[2 * i + 1] == 1;
            r2[i].elapsed_ms = ms;
            if (!r2[i].converged) all_conv = false;
            if (done[2 * i] == 2 || done[2 * i + 1] == 2 || !std::isfinite(r2[i].residual)) finite = false;
            if (reps2) reps2[i] = r2[i];
        }
        if (!finite) return BEC_ERR_NONFINITE;
        return all_conv ? BEC_OK : BEC_ERR_NOT_CONVERGED;
    }
    std::vector<bec_gs_report> rr(batch);
    st = final_report(w, params, batch, s, rr.data());
    if (st != BEC_OK) return st;
    if (out_x) {
        st = download_complex(w.X, out_x, n, o.precision, s);
        if (st != BEC_OK) return st;
    }
    bool all_conv = true, finite = true;
    for (int i = 0; i < batch; ++i) {
        rr[i].iterations = iters[i];
        rr[i].converged = done[i] == 1;
        rr[i].elapsed_ms = ms;
        if (done[i] != 1) all_conv = false;
        if (done[i] == 2 || !std::isfinite(rr[i].residual)) finite = false;
        if (reps) reps[i] = rr[i];
    }
    if (!finite) return BEC_ERR_NONFINITE;
    return all_conv ? BEC_OK : BEC_ERR_NOT_CONVERGED;
}

int gs_residual_of(const bec_calc_params &c, int prec, const double *psi_x, hipStream_t s, double *res, double *mu)
{
    GsWork w;
    int st = setup_work(w, &c, 1, prec, s);
    if (st != BEC_OK) return st;
    st = upload_complex(psi_x, w.X, w.g.cells, prec, s);
    if (st != BEC_OK) return st;
    GsScalars sc = w.sc;
    sc.done = w.zero_done;
    sc.s = w.ones;
    GS_HIP(residual_pass(w, sc, 0.0, 0, s));
    double r = 0, m = 0;
    GS_HIP(hipMemcpyAsync(&r, sc.res, sizeof(double), hipMemcpyDeviceToHost, s));
    GS_HIP(hipMemcpyAsync(&m, sc.mu, sizeof(double), hipMemcpyDeviceToHost, s));
    GS_HIP(hipStreamSynchronize(s));
    if (res) *res = r;
    if (mu) *mu = m;
    return BEC_OK;
}

} // namespace bec

struct bec_ctx;
extern "C" int bec_group_count(const bec_ctx *ctx);
namespace bec {
const bec_calc_params *ctx_group_params(bec_ctx *ctx, int group, hipStream_t *stream);
}

extern "C" int bec_default_gs_options(bec_gs_options *o)
{
    if (!o) return BEC_ERR_INVALID;
    memset(o, 0, sizeof(*o));
    o->method = BEC_GS_SPLIT_STEP;
    o->rule = BEC_GS_RULE_REFERENCE;
    o->tol = 1e-6;
    o->max_iter = 200000;
    o->check_every = 1;
    o->precision = BEC_FP64;
    o->alpha = 0.0;
    o->beta = 0.6;
    o->tau = 0.5;
    return BEC_OK;
}

extern "C" int bec_calculate_steady_state(bec_ctx *ctx, int group, const double *guess_x, const bec_gs_options *opt,
                                          double *out_x, bec_gs_report *rep)
{
    hipStream_t s = nullptr;
    const bec_calc_params *c = bec::ctx_group_params(ctx, group, &s);
    if (!c) return BEC_ERR_INVALID;
    bec_gs_options o;
    if (opt) o = *opt;
    else bec_default_gs_options(&o);
    bec_gs_report r{};
    int st = bec::gs_solve(c, 1, guess_x, &o, s, out_x, &r, nullptr, nullptr);
    if (rep) *rep = r;
    return st;
}

extern "C" int bec_ground_states_batched(const bec_calc_params *params, int batch, const double *guesses,
                                         const bec_gs_options *opt, void *stream, double *out_x, bec_gs_report *reps)
{
    bec_gs_options o;
    if (opt) o = *opt;
    else bec_default_gs_options(&o);
    int vst = bec::gs_validate(params, batch, &o);
    if (vst != BEC_OK) return vst;
    hipStream_t s = (hipStream_t)stream;
    bool own = false;
    if (!s) {
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return BEC_ERR_HIP;
        own = true;
    }
    int st = bec::gs_solve(params, batch, guesses, &o, s, out_x, reps, nullptr, nullptr);
    if (own) (void)hipStreamDestroy(s);
    return st;
}

static int two_component(const bec_calc_params *params, int batch, const double *frac_a, const double *guesses,
                         const float *net, double spin_weight, const bec_gs_options *opt, void *stream,
                         double *out_x, bec_gs2_report *reps);

extern "C" int bec_ground_states_two_component(const bec_calc_params *params, int batch, const double *frac_a,
                                               const double *guesses, const bec_gs_options *opt, void *stream,
                                               double *out_x, bec_gs2_report *reps)
{
    return two_component(params, batch, frac_a, guesses, nullptr, 1.0, opt, stream, out_x, reps);
}

extern "C" int bec_ground_states_two_component_net(const bec_calc_params *params, int batch, const double *frac_a,
                                                   const float *d_net, double spin_weight,
                                                   const bec_gs_options *opt, void *stream, double *out_x,
                                                   bec_gs2_report *reps)
{
    if (!d_net || !(spin_weight > 0.0)) return BEC_ERR_INVALID;
    return two_component(params, batch, frac_a, nullptr, d_net, spin_weight, opt, stream, out_x, reps);
}

static int two_component(const bec_calc_params *params, int batch, const double *frac_a, const double *guesses,
                         const float *net, double spin_weight, const bec_gs_options *opt, void *stream,
                         double *out_x, bec_gs2_report *reps)
{
    bec_gs_options o;
    if (opt) o = *opt;
    else {
        bec_default_gs_options(&o);
        o.method = BEC_GS_GRADIENT;
        o.rule = BEC_GS_RULE_RESIDUAL;
        // the coupled flow's step and momentum: over 40 parameter sets of the
        // learned start's ranges (N 5k-100k, trap 5-20 / 50-150 Hz, a12 95-100 a0)
        // 157 iterations on average to residual 1e-6, against 774 with the
        // one-component pair (0.5, 0.6)
        o.tau = BEC_GS2_DEFAULT_TAU;
        o.beta = BEC_GS2_DEFAULT_BETA;
    }
    if (o.method != BEC_GS_GRADIENT || o.rule != BEC_GS_RULE_RESIDUAL) return BEC_ERR_INVALID;
    int vst = bec::gs_validate(params, batch, &o);
    if (vst != BEC_OK) return vst;
    std::vector<double> fr(batch, 0.5);
    if (frac_a)
        for (int i = 0; i < batch; ++i) {
            if (!(frac_a[i] > 0.0 && frac_a[i] < 1.0)) return BEC_ERR_INVALID;
            fr[i] = frac_a[i];
        }
    hipStream_t s = (hipStream_t)stream;
    bool own = false;
    if (!s) {
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return BEC_ERR_HIP;
        own = true;
        if (net) {
            // the network's output was written on the null stream: order after it
            hipEvent_t ev = nullptr;
            hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventRecord(ev, nullptr);
            if (e == hipSuccess) e = hipStreamWaitEvent(s, ev, 0);
            if (ev) (void)hipEventDestroy(ev);
            if (e != hipSuccess) {
                (void)hipStreamDestroy(s);
                return BEC_ERR_HIP;
            }
        }
    }
    int st = bec::gs_solve(params, batch, guesses, &o, s, out_x, nullptr, fr.data(), reps, net, spin_weight);
    if (own) (void)hipStreamDestroy(s);
    return st;
}

extern "C" int bec_state_residual(bec_ctx *ctx, int group, const double *psi_x, double *residual, double *mu)
{
    hipStream_t s = nullptr;
    const bec_calc_params *c = bec::ctx_group_params(ctx, group, &s);
    if (!c || !psi_x) return BEC_ERR_INVALID;
    return bec::gs_residual_of(*c, BEC_FP64, psi_x, s, residual, mu);
}
