// C ABI of the 2D program on the GPU (bec_run_2d): validation, device buffers,
// the two launches (kernels2d.hip) and the ensemble statistics of
// programs/bec_wigner_2d.cpp/wigner.cpp:431-443 on the host.
#include <hip/hip_runtime.h>
#include <math.h>

#include <vector>

#include "backend2d.hpp"
#include "host_common.hpp"

namespace {

bool pow2_in(int v, int lo, int hi) { return v >= lo && v <= hi && (v & (v - 1)) == 0; }

int validate_2d(const bec_2d_params &p)
{
    if (!pow2_in(p.nvx, 4, bec::W2D_MAX_N) || !pow2_in(p.nvy, 4, bec::W2D_MAX_N)) return BEC_ERR_SHAPE;
    if (!(p.xmax > 0) || !(p.ymax > 0) || p.ne < 1 || p.n_meas < 1 || p.ntW < 0 || p.itmax < 0) return BEC_ERR_INVALID;
    if (!(p.dtG > 0) || !(p.dtW > 0)) return BEC_ERR_INVALID;
    if (p.gs_rule < 0 || p.gs_rule > 2) return BEC_ERR_INVALID;
    if (p.gs_rule == 0 && p.ntG < 1) return BEC_ERR_INVALID;
    if (p.gs_rule >= 1 && (p.gs_max_iter < 1 || !(p.gs_tol > 0))) return BEC_ERR_INVALID;
    if (p.gs_rule == 2 && (!(p.n_atoms > 0) || !(p.gs_tau > 0) || !(p.gs_beta >= 0) || !(p.g11 > 0)))
        return BEC_ERR_INVALID;
    if (p.ne > 65535) return BEC_ERR_SHAPE;
    return BEC_OK;
}

struct Buffers {
    std::vector<void *> p;
    ~Buffers()
    {
        for (void *q : p) bec::dev_free(q);
    }
    hipError_t alloc(void **q, size_t bytes)
    {
        hipError_t e = bec::dev_alloc(q, bytes);
        if (e == hipSuccess) p.push_back(*q);
        return e;
    }
};

} // namespace

extern "C" int bec_run_2d(const bec_2d_params *pp, double *m, double *var, double *vis, double *gs_x, double *avn,
                          bec_2d_report *rep)
{
    if (!pp || !m || !var) return BEC_ERR_INVALID;
    const bec_2d_params p = *pp;
    int st = validate_2d(p);
    if (st != BEC_OK) return st;
    const long cells = (long)p.nvx * p.nvy;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return BEC_ERR_HIP;
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    Buffers buf;
    double2 *d_gs = nullptr, *d_pv = nullptr;
    double *d_rep = nullptr, *d_meas = nullptr, *d_avn = nullptr;
    BEC_HIP(buf.alloc((void **)&d_gs, sizeof(double2) * cells));
    if (p.gs_rule == 2) BEC_HIP(buf.alloc((void **)&d_pv, sizeof(double2) * cells));
    BEC_HIP(buf.alloc((void **)&d_rep, sizeof(double) * 4));
    BEC_HIP(buf.alloc((void **)&d_meas, sizeof(double) * 3 * p.ne * p.n_meas));
    if (avn) BEC_HIP(buf.alloc((void **)&d_avn, sizeof(double) * cells * p.ne));
    hipEvent_t e0, e1;
    BEC_HIP(hipEventCreate(&e0));
    BEC_HIP(hipEventCreate(&e1));
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};
    BEC_HIP(hipEventRecord(e0, s));
    BEC_HIP(bec::run_w2d_ground_state(p, d_gs, d_pv, d_rep, s));
    BEC_HIP(bec::run_w2d_ensemble(p, d_gs, d_meas, d_avn, s));
    BEC_HIP(hipEventRecord(e1, s));
    std::vector<double> h_rep(4), h_meas(3 * (size_t)p.ne * p.n_meas), h_avn(avn ? cells * p.ne : 0);
    std::vector<double2> h_gs(cells);
    BEC_HIP(hipMemcpyAsync(h_rep.data(), d_rep, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    BEC_HIP(hipMemcpyAsync(h_meas.data(), d_meas, sizeof(double) * h_meas.size(), hipMemcpyDeviceToHost, s));
    BEC_HIP(hipMemcpyAsync(h_gs.data(), d_gs, sizeof(double2) * cells, hipMemcpyDeviceToHost, s));
    if (avn) BEC_HIP(hipMemcpyAsync(h_avn.data(), d_avn, sizeof(double) * h_avn.size(), hipMemcpyDeviceToHost, s));
    BEC_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);

    // ensemble statistics (wigner.cpp:431-443): means, then sqrt(|<N^2> - <N>^2| / (ne - 1))
    const int nm = p.n_meas, ne = p.ne;
    for (int i = 0; i < 2 * nm; ++i) m[i] = var[i] = 0.0;
    if (vis)
        for (int i = 0; i < nm; ++i) vis[i] = 0.0;
    for (int j = 0; j < ne; ++j)
        for (int k = 0; k < nm; ++k) {
            const double *q = &h_meas[((size_t)j * nm + k) * 3];
            m[k] += q[0];
            m[k + nm] += q[1];
            var[k] += q[0] * q[0];
            var[k + nm] += q[1] * q[1];
            if (vis) vis[k] += q[2];
        }
    for (int i = 0; i < 2 * nm; ++i) {
        m[i] /= ne;
        const double v = var[i] / ne - m[i] * m[i];
        var[i] = ne > 1 ? sqrt(fabs(v) / (ne - 1)) : 0.0;
    }
    if (vis)
        for (int i = 0; i < nm; ++i) vis[i] /= ne;
    if (avn)
        for (long c = 0; c < cells; ++c) {
            double t = 0.0;
            for (int j = 0; j < ne; ++j) t += h_avn[(size_t)j * cells + c];
            avn[c] = t / ne;
        }
    if (gs_x)
        for (long c = 0; c < cells; ++c) {
            gs_x[2 * c] = h_gs[c].x;
            gs_x[2 * c + 1] = h_gs[c].y;
        }
    if (rep) {
        rep->N0 = h_rep[0];
        rep->gs_iterations = (int)h_rep[1];
        rep->gs_residual = h_rep[2];
        rep->gs_mu = h_rep[3];
        rep->elapsed_ms = ms;
    }
    return BEC_OK;
}
