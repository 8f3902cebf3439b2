// C ABI of the 1D program on the GPU (bec_1d_cells, bec_run_1d): validation,
// device buffers, the three launches (kernels1d.hip) and the final statistics
// of programs/bec_wigner_2d.cpp/wigner_1d.cpp:405-413 on the host.
#include <hip/hip_runtime.h>
#include <math.h>

#include <cmath>

#include <vector>

#include "backend1d.hpp"
#include "host_common.hpp"

namespace {

// the largest samples x trajectories product whose per-trajectory samples
// (24 B each) are staged on the device
constexpr long long W1D_MAX_SAMPLES = 1LL << 26;

int cells_1d(const bec_1d_params &p)
{
    if (!(p.zmax > 0) || !(p.dz > 0) || !std::isfinite(p.zmax) || !(p.zmax / p.dz < 1e6)) return -BEC_ERR_INVALID;
    const int nv = 1 + 2 * (int)(p.zmax / p.dz);
    if (nv > bec::W1D_MAX_NV) return -BEC_ERR_SHAPE;
    return nv;
}

int validate_1d(const bec_1d_params &p, int *nv)
{
    const int n = cells_1d(p);
    if (n < 0) return -n;
    if (!(p.dt > 0) || p.ntG < 1 || p.itmax < 0 || p.ntW < 1 || p.ne < 1) return BEC_ERR_INVALID;
    if ((long long)p.ne * p.ntW > W1D_MAX_SAMPLES || p.ne > 0x7fffffff / 2) return BEC_ERR_SHAPE;
    *nv = n;
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

extern "C" int bec_1d_cells(const bec_1d_params *p)
{
    if (!p) return -BEC_ERR_INVALID;
    return cells_1d(*p);
}

extern "C" int bec_run_1d(const bec_1d_params *pp, double *m, double *var, double *vis, double *gs_x, double *avn,
                          double *ntot, bec_1d_report *rep)
{
    if (!pp || !m || !var) return BEC_ERR_INVALID;
    const bec_1d_params p = *pp;
    int n = 0;
    int st = validate_1d(p, &n);
    if (st != BEC_OK) return st;
    const int nt = p.ntW, ne = p.ne;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return BEC_ERR_HIP;
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    Buffers buf;
    double2 *d_gs = nullptr;
    double *d_rep = nullptr, *d_ntot = nullptr, *d_meas = nullptr, *d_avn = nullptr, *d_sums = nullptr,
           *d_avn_sum = nullptr;
    BEC_HIP(buf.alloc((void **)&d_gs, sizeof(double2) * n));
    BEC_HIP(buf.alloc((void **)&d_rep, sizeof(double) * 3));
    BEC_HIP(buf.alloc((void **)&d_ntot, sizeof(double) * p.ntG));
    BEC_HIP(buf.alloc((void **)&d_meas, sizeof(double) * 3 * (size_t)ne * nt));
    BEC_HIP(buf.alloc((void **)&d_avn, sizeof(double) * (size_t)ne * n));
    BEC_HIP(buf.alloc((void **)&d_sums, sizeof(double) * 5 * (size_t)nt));
    BEC_HIP(buf.alloc((void **)&d_avn_sum, sizeof(double) * n));
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
    BEC_HIP(bec::run_w1d_ground_state(p, n, d_gs, d_ntot, d_rep, s));
    BEC_HIP(bec::run_w1d_ensemble(p, n, d_gs, d_meas, d_avn, s));
    BEC_HIP(bec::run_w1d_reduce(p, n, d_meas, d_avn, d_sums, d_avn_sum, s));
    BEC_HIP(hipEventRecord(e1, s));
    std::vector<double> h_rep(3), h_sums(5 * (size_t)nt), h_avn(n), h_ntot(p.ntG);
    std::vector<double2> h_gs(n);
    BEC_HIP(hipMemcpyAsync(h_rep.data(), d_rep, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
    BEC_HIP(hipMemcpyAsync(h_sums.data(), d_sums, sizeof(double) * h_sums.size(), hipMemcpyDeviceToHost, s));
    BEC_HIP(hipMemcpyAsync(h_avn.data(), d_avn_sum, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    BEC_HIP(hipMemcpyAsync(h_ntot.data(), d_ntot, sizeof(double) * p.ntG, hipMemcpyDeviceToHost, s));
    BEC_HIP(hipMemcpyAsync(h_gs.data(), d_gs, sizeof(double2) * n, hipMemcpyDeviceToHost, s));
    BEC_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);

    // wigner_1d.cpp:405-413 as written: means, then sqrt((<N^2> - <N>^2) / (ne - 1))
    for (int i = 0; i < 2 * nt; ++i) {
        m[i] = h_sums[i] / ne;
        const double v = h_sums[2 * (size_t)nt + i] / ne - m[i] * m[i];
        var[i] = sqrt(v / (ne - 1));
    }
    if (vis)
        for (int i = 0; i < nt; ++i) vis[i] = h_sums[4 * (size_t)nt + i] / ne;
    if (avn)
        for (int c = 0; c < n; ++c) avn[c] = h_avn[c] / ne;
    if (ntot)
        for (int i = 0; i < p.ntG; ++i) ntot[i] = h_ntot[i];
    if (gs_x)
        for (int c = 0; c < n; ++c) {
            gs_x[2 * c] = h_gs[c].x;
            gs_x[2 * c + 1] = h_gs[c].y;
        }
    if (rep) {
        rep->N0 = h_rep[0];
        rep->nv = n;
        rep->gs_residual = h_rep[1];
        rep->gs_mu = h_rep[2];
        rep->elapsed_ms = ms;
    }
    return BEC_OK;
}
