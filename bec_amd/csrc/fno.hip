// Fourier neural operator for the learned ground-state start (include/bec_fno.h).
//
// Layout: activations [batch][z][y][x][channel] in FP32; cells S = nz ny nx,
// plane P = ny nx.  The truncated spectral path of layer l is five GEMMs
// (bec_gemm, BF16 operands, FP32 accumulation):
//   Zf = Fz h           [b][kz re/im][P W]       M = 2 mz, K = nz
//   Sp = Fxy Zf         [b][kz][re/im kxy][W]    M = 2 mxy, K = 2 P   (batch b kz)
//   T  = Sp E           per mode, M = batch, K = N = 2 W               (batch modes)
//   Zi = Fxy^T T        [b][kz][re/im y x][W]
//   u  = Fz^T Zi        [b][z][P W]
// with Fz (2 mz x nz) and Fxy (2 mxy x 2 P) the real forms of the truncated
// DFTs scaled by 1/sqrt(nz) and 1/sqrt(P) (so the inverse transforms are their
// transposes), and E the real 2W x 2W form of the complex per-mode matrix R,
// [[Re R, Im R], [-Im R, Re R]] in row-vector convention.  The per-cell linear
// map adds u += h W + b as a GEMM with beta = 1.  The backward pass applies
// the transposed maps in reverse order; weight gradients are GEMMs with split K.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/bec_fno.h"
#include "host_common.hpp"

namespace bec {

hipError_t run_gemm(const bec_gemm_desc &g, hipStream_t s);

namespace {

constexpr int KSPLIT = 256;

bec_map2 m1(long stride)
{
    bec_map2 m;
    m.hi = 0;
    m.lo = stride;
    m.lg = 40;
    m.pad_ = 0;
    return m;
}

bec_map2 m2(long lo_count, long lo_stride, long hi_stride)
{
    bec_map2 m;
    int lg = 0;
    while ((1L << lg) < lo_count) ++lg;
    m.hi = hi_stride;
    m.lo = lo_stride;
    m.lg = lg;
    m.pad_ = 0;
    return m;
}

bool pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }

// ---------------------------------------------------------------- kernels

__device__ __forceinline__ unsigned short to_bf16(float f)
{
    unsigned int u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__global__ void gelu_fwd(const float *u, float *y, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = u[i];
    y[i] = 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}

// d <- d * gelu'(u), gelu'(x) = Phi(x) + x phi(x)
__global__ void gelu_bwd(const float *u, float *d, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = u[i];
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    d[i] *= cdf + x * pdf;
}

// dy = 2 (y - t) / count; per-block partial sums of (y - t)^2 in FP64
__global__ void mse_grad(const float *y, const float *t, float *dy, long n, float scale, double *partial)
{
    __shared__ double sh[256];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        const float d = y[i] - t[i];
        dy[i] = scale * d;
        s = (double)d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__device__ __forceinline__ long map_off(const bec_map2 &m, long i)
{
    return (i >> m.lg) * m.hi + (i & ((1L << m.lg) - 1)) * m.lo;
}

// The per-mode channel mixing for a single sample (one row per mode) with
// 2W = 64 channels: C[t][n] = sum_k bf16(A[t][k]) E[t][k][n].  One wave per
// mode t; lane L owns the channels n = 8 (L & 7) + i (i < 8) and the rows
// k = (L >> 3) + 8 j, and reads E[t] in 16-byte pieces, so one load of the
// wave covers 8 whole rows.  The 8 partial sums of a channel are combined
// across lanes in a fixed butterfly order.  The operands are rounded as the
// GEMM rounds them (A to BF16 nearest even, E already BF16); the products are
// exact in FP32.
__global__ __launch_bounds__(256) void mode_row_mix64(const float *A, const unsigned short *E, float *C, bec_map2 ab,
                                                      bec_map2 ak, long ebatch, bec_map2 cb, bec_map2 cn, int K,
                                                      long modes)
{
    const long t = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= modes) return;   // t is the same for the whole wave
    const int lane = threadIdx.x & 63, nb = lane & 7, kp = lane >> 3;
    const float *a = A + map_off(ab, t);
    const unsigned short *e = E + t * ebatch + nb * 8;
    float acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = 0.f;
#pragma unroll 4
    for (int k = kp; k < K; k += 8) {
        const float x = a[map_off(ak, k)];
        const unsigned int u = __float_as_uint(x);
        const float xr = (u & 0x7f800000u) == 0x7f800000u ? x : __uint_as_float((unsigned int)to_bf16(x) << 16);
        const uint4 w = *(const uint4 *)(e + (long)k * 64);
        const unsigned int ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[2 * i] = fmaf(xr, __uint_as_float(ws[i] << 16), acc[2 * i]);
            acc[2 * i + 1] = fmaf(xr, __uint_as_float(ws[i] & 0xffff0000u), acc[2 * i + 1]);
        }
    }
#pragma unroll
    for (int off = 8; off < 64; off <<= 1)
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] += __shfl_xor(acc[i], off);
    if (kp == 0) {
        const long cbase = map_off(cb, t);
#pragma unroll
        for (int i = 0; i < 8; ++i) C[cbase + map_off(cn, nb * 8 + i)] = acc[i];
    }
}

struct Gs2Lattice {
    int nx, ny, nz;
    double xmax, ymax, zmax, dx, dy, dz, lambda, mu;
    float scaled[4];
};

// One cell of the learned start's input features per thread, in FP64 with the
// operations in the host construction's order (no contraction), then FP32.
__global__ void gs2_features(Gs2Lattice g, float *out, long cells)
{
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const int xi = (int)(i % g.nx);
    const long r = i / g.nx;
    const int yi = (int)(r % g.ny), zi = (int)(r / g.ny);
    const double X = -g.xmax + g.dx * xi, Y = -g.ymax + g.dy * yi, Z = -g.zmax + g.dz * zi;
    const double V = 0.5 * (X * X + Y * Y + Z * Z / (g.lambda * g.lambda));
    float4 *o = (float4 *)(out + i * 16);
    o[0] = make_float4((float)sqrt(fmax(g.mu - V, 0.0) / g.mu), (float)(V / g.mu), (float)(X / g.xmax),
                       (float)(Y / g.ymax));
    o[1] = make_float4((float)(Z / g.zmax), g.scaled[0], g.scaled[1], g.scaled[2]);
    o[2] = make_float4(g.scaled[3], 0.f, 0.f, 0.f);
    o[3] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// E[t] (2W x 2W, BF16) from R[t][i][o][re, im]
__global__ void expand_modes(const float *R, unsigned short *E, int W, long modes)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)W * W;
    if (i >= modes * per) return;
    const long t = i / per;
    const int io = (int)(i % per), a = io / W, o = io % W;
    const float re = R[2 * i], im = R[2 * i + 1];
    unsigned short *e = E + t * 4 * per;
    const int w2 = 2 * W;
    e[a * w2 + o] = to_bf16(re);
    e[a * w2 + W + o] = to_bf16(im);
    e[(W + a) * w2 + o] = to_bf16(-im);
    e[(W + a) * w2 + W + o] = to_bf16(re);
}

// dR[t][i][o] = (dE00 + dE11, dE01 - dE10)
__global__ void fold_modes(const float *dE, float *dR, int W, long modes)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)W * W;
    if (i >= modes * per) return;
    const long t = i / per;
    const int io = (int)(i % per), a = io / W, o = io % W;
    const float *e = dE + t * 4 * per;
    const int w2 = 2 * W;
    dR[2 * i] = e[a * w2 + o] + e[(W + a) * w2 + W + o];
    dR[2 * i + 1] = e[a * w2 + W + o] - e[(W + a) * w2 + o];
}

__global__ void adam_kernel(float *p, const float *g, float *m, float *v, long n, float lr, float b1, float b2,
                            float eps, float wd, float c1, float c2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    p[i] -= lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
}

dim3 grid1(long n) { return dim3((unsigned)((n + 255) / 256)); }

// splitmix64 -> uniform (-1, 1)
struct Rng {
    uint64_t s;
    float uni()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (float)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
    }
};

} // namespace
} // namespace bec

struct bec_fno {
    bec_fno_config c{};
    long S = 0, P = 0, mxy = 0, modes = 0;
    // parameter offsets (floats) into params / grads / m / v
    long oP = 0, op = 0, oQ1 = 0, oq1 = 0, oQ2 = 0, oq2 = 0, n_params = 0;
    std::vector<long> oW, ob, oR;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    unsigned short *E = nullptr;      // [layers][modes][2W][2W] BF16
    float *Fz = nullptr, *Fxy = nullptr, *one = nullptr;
    // activations
    std::vector<float *> h, u;        // h[0..L], u[0..L-1] (u[L-1] aliases h[L])
    float *zf = nullptr, *tm = nullptr, *zi = nullptr;   // spectral temporaries
    std::vector<float *> sp;          // per layer: DFT_xy DFT_z h (kept for dE)
    float *g1 = nullptr, *a1 = nullptr, *y = nullptr;
    float *dy = nullptr, *dA = nullptr, *dB = nullptr, *dH = nullptr, *dE = nullptr, *work = nullptr;
    double *loss_part = nullptr;
    std::vector<void *> owned;
    std::vector<std::pair<const char *, size_t>> extents;   // device buffers (bytes) for the host-side GEMM checks
    long adam_t = 0;
};

namespace bec {
namespace {

hipError_t falloc(bec_fno *f, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) {
        f->owned.push_back(*p);
        f->extents.push_back({(const char *)*p, bytes});
    }
    return e;
}

// Elements of size es from ptr to the end of the registered buffer holding it (0: none).
long room(const bec_fno *f, const void *ptr, size_t es)
{
    const char *q = (const char *)ptr;
    for (const auto &b : f->extents)
        if (q >= b.first && q < b.first + b.second) return (long)((b.first + b.second - q) / es);
    return 0;
}

// Registers caller buffers for the duration of one API call.
struct CallerBuffers {
    bec_fno *f;
    size_t n0;
    CallerBuffers(bec_fno *f_, std::initializer_list<std::pair<const void *, size_t>> bufs) : f(f_), n0(f_->extents.size())
    {
        for (const auto &b : bufs) f->extents.push_back({(const char *)b.first, b.second});
    }
    ~CallerBuffers() { f->extents.resize(n0); }
};

void layout(bec_fno *f)
{
    const bec_fno_config &c = f->c;
    const long W = c.width;
    long o = 0;
    f->oP = o; o += (long)c.c_in * W;
    f->op = o; o += W;
    f->oW.assign(c.layers, 0);
    f->ob.assign(c.layers, 0);
    f->oR.assign(c.layers, 0);
    for (int l = 0; l < c.layers; ++l) {
        f->oW[l] = o; o += W * W;
        f->ob[l] = o; o += W;
        f->oR[l] = o; o += f->modes * W * W * 2;
    }
    f->oQ1 = o; o += W * c.hidden;
    f->oq1 = o; o += c.hidden;
    f->oQ2 = o; o += (long)c.hidden * c.c_out;
    f->oq2 = o; o += c.c_out;
    f->n_params = o;
}

hipError_t expand_all(bec_fno *f, hipStream_t s)
{
    const long per = f->modes * f->c.width * f->c.width;
    for (int l = 0; l < f->c.layers; ++l) {
        hipLaunchKernelGGL(expand_modes, grid1(per), dim3(256), 0, s, f->params + f->oR[l],
                           f->E + (long)l * f->modes * 4 * f->c.width * f->c.width, f->c.width, f->modes);
    }
    return hipGetLastError();
}

// Largest offset a two-level map reaches over indices 0..n-1 (strides >= 0).
long max_off(const bec_map2 &m, long n)
{
    if (n <= 0) return 0;
    const long last = n - 1, mask = (1L << m.lg) - 1;
    const long h = last >> m.lg, l = last & mask;
    long best = h * (long)m.hi + l * (long)m.lo;
    if (h >= 1) best = std::max(best, (h - 1) * (long)m.hi + mask * (long)m.lo);
    return best;
}

// One GEMM with the network's conventions.  Before the launch, the largest
// offset each operand map reaches is checked on the host against the buffer
// that holds the operand (every device buffer the network touches is
// registered), so a layout error is reported instead of launched.
hipError_t G(const bec_fno *f, const void *A, int abf, const void *B, int bbf, float *C, int M, int N, int K,
             bec_map2 am, bec_map2 ak, bec_map2 bk, bec_map2 bn, bec_map2 cm, bec_map2 cn, int batch, bec_map2 ab,
             bec_map2 bb, bec_map2 cb, float beta, const float *bias, int ksplit, float *work, hipStream_t s)
{
    const long ea = max_off(ab, batch) + max_off(am, M) + max_off(ak, K);
    const long eb = max_off(bb, batch) + max_off(bk, K) + max_off(bn, N);
    const long ec = max_off(cb, batch) + max_off(cm, M) + max_off(cn, N);
    // A GEMM that fills few workgroups (small batches: the transforms of a
    // single-sample forward pass) walks its K tiles in one dependent chain
    // per workgroup; split K to about 1024 workgroups (4 per CU; 512 left the
    // single-sample pass at 0.57 ms against 0.49), reduced in a fixed order.
    const long wgs = ((long)(N + 63) / 64) * ((M + 63) / 64) * batch;
    if (ksplit == 1 && wgs < 256 && K >= 128) {
        ksplit = (int)std::min<long>(std::min<long>(8, (1024 + wgs - 1) / wgs), K / 64);
        work = f->work;
    }
    bool ok = ea < room(f, A, abf ? 2 : 4) && eb < room(f, B, bbf ? 2 : 4) && ec < room(f, C, 4);
    if (bias) ok = ok && N <= room(f, bias, 4);
    if (ksplit > 1) ok = ok && (long)ksplit * batch * M * N <= room(f, work, 4);
    if (!ok) {
        set_last_error("surrogate GEMM operand map exceeds its buffer");
        return hipErrorInvalidValue;
    }
    bec_gemm_desc d;
    memset(&d, 0, sizeof(d));
    d.A = A;
    d.B = B;
    d.C = C;
    d.bias = bias;
    d.work = work;
    d.a_bf16 = abf;
    d.b_bf16 = bbf;
    d.M = M;
    d.N = N;
    d.K = K;
    d.batch = batch;
    d.ksplit = ksplit;
    d.alpha = 1.0f;
    d.beta = beta;
    d.am = am; d.ak = ak; d.bk = bk; d.bn = bn; d.cm = cm; d.cn = cn;
    d.ab = ab; d.bb = bb; d.cb = cb;
    return run_gemm(d, s);
}

// C[rows][n] (+)= A[rows][k] B[k][n] for row-major per-cell activations
hipError_t rowgemm(const bec_fno *f, const float *A, const float *B, float *C, long rows, int K, int N, float beta,
                   const float *bias, hipStream_t s)
{
    return G(f, A, 0, B, 0, C, (int)rows, N, K, m1(K), m1(1), m1(N), m1(1), m1(N), m1(1), 1, m1(0), m1(0), m1(0), beta,
             bias, 1, nullptr, s);
}

// C[rows][n] = A[rows][k] B^T   (B stored [n][k])
hipError_t rowgemm_bt(const bec_fno *f, const float *A, const float *Bt, float *C, long rows, int K, int N,
                      float beta, hipStream_t s)
{
    return G(f, A, 0, Bt, 0, C, (int)rows, N, K, m1(K), m1(1), m1(1), m1(K), m1(N), m1(1), 1, m1(0), m1(0), m1(0), beta,
             nullptr, 1, nullptr, s);
}

// dWt[k][n] = sum_rows X[rows][k] D[rows][n]  (weight gradient, split K)
hipError_t wgrad(const bec_fno *f, const float *X, int K, const float *D, int N, long rows, float *out,
                 hipStream_t s)
{
    float *work = f->work;
    return G(f, X, 0, D, 0, out, K, N, (int)rows, m1(1), m1(K), m1(N), m1(1), m1(N), m1(1), 1, m1(0), m1(0), m1(0), 0.0f,
             nullptr, KSPLIT, work, s);
}

// db[n] = sum_rows D[rows][n]: a GEMM with a constant-one A operand (stride 0)
hipError_t bgrad(const bec_fno *f, const float *D, int N, long rows, float *out, hipStream_t s)
{
    const float *one = f->one;
    float *work = f->work;
    return G(f, one, 0, D, 0, out, 1, N, (int)rows, m1(0), m1(0), m1(N), m1(1), m1(N), m1(1), 1, m1(0), m1(0), m1(0),
             0.0f, nullptr, KSPLIT, work, s);
}

struct Spec {
    int B, nz, mz2, W;
    long P, PW, mxy;
};

// Zf = Fz h  (transpose = 0) or out = Fz^T zin (transpose = 1, beta)
hipError_t zdft(const bec_fno *f, const Spec &q, const float *in, float *out, int transpose, float beta, hipStream_t s)
{
    if (!transpose)
        return G(f, f->Fz, 0, in, 0, out, q.mz2, (int)q.PW, q.nz, m1(q.nz), m1(1), m1(q.PW), m1(1), m1(q.PW), m1(1), q.B,
                 m1(0), m1((long)q.nz * q.PW), m1((long)q.mz2 * q.PW), 0.0f, nullptr, 1, nullptr, s);
    return G(f, f->Fz, 0, in, 0, out, q.nz, (int)q.PW, q.mz2, m1(1), m1(q.nz), m1(q.PW), m1(1), m1(q.PW), m1(1), q.B,
             m1(0), m1((long)q.mz2 * q.PW), m1((long)q.nz * q.PW), beta, nullptr, 1, nullptr, s);
}

// Sp = Fxy Zf (transpose = 0, [b kz][2 mxy][W]) or Zi = Fxy^T T (transpose = 1, [b kz][2 P][W])
hipError_t xydft(const bec_fno *f, const Spec &q, const float *in, float *out, int transpose, hipStream_t s)
{
    const long rows_in = transpose ? 2 * q.mxy : 2 * q.P, rows_out = transpose ? 2 * q.P : 2 * q.mxy;
    const long lda = 2 * q.P;   // Fxy is [2 mxy][2 P]
    bec_map2 am = transpose ? m1(1) : m1(lda), ak = transpose ? m1(lda) : m1(1);
    return G(f, f->Fxy, 0, in, 0, out, (int)rows_out, q.W, (int)rows_in, am, ak, m1(q.W), m1(1), m1(q.W), m1(1),
             q.B * (q.mz2 / 2), m1(0), m1(rows_in * q.W), m1(rows_out * q.W), 0.0f, nullptr, 1, nullptr, s);
}

// per-mode maps of the [b][kz][re/im][kxy][W] spectral layout
struct ModeMaps {
    bec_map2 row, ch, batch;
};
ModeMaps mode_maps(const Spec &q)
{
    ModeMaps m;
    m.row = m1((long)(q.mz2 / 2) * 2 * q.mxy * q.W);       // sample
    m.ch = m2(q.W, 1, q.mxy * q.W);                          // (re/im, channel)
    m.batch = m2(q.mxy, q.W, 2 * q.mxy * q.W);               // (kz, kxy)
    return m;
}

// The channel mixing of a batch-1 forward pass with 2W = 64 on mode_row_mix64
// instead of 64 x 64 GEMM tiles that would hold one row each; the operand
// extents are checked as G checks them.
hipError_t mode_mix_row(const bec_fno *f, const float *A, const unsigned short *El, float *C, int W2,
                        const ModeMaps &mm, long modes, hipStream_t s)
{
    const long eb = (modes - 1) * (long)W2 * W2 + (long)W2 * W2 - 1;
    const long ea = max_off(mm.batch, modes) + max_off(mm.ch, W2);
    const long ec = max_off(mm.batch, modes) + max_off(mm.ch, W2);
    if (W2 != 64 || ((uintptr_t)El & 15) != 0 ||
        !(ea < room(f, A, 4) && eb < room(f, El, 2) && ec < room(f, C, 4))) {
        set_last_error("surrogate mode mixing exceeds its buffer");
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(mode_row_mix64, dim3((unsigned)((modes + 3) / 4)), dim3(256), 0, s, A, El, C, mm.batch, mm.ch,
                       (long)W2 * W2, mm.batch, mm.ch, W2, modes);
    return hipGetLastError();
}

hipError_t forward_impl(bec_fno *f, int B, const float *x, hipStream_t s)
{
    const bec_fno_config &c = f->c;
    const int W = c.width, L = c.layers;
    const long rows = (long)B * f->S;
    Spec q{B, c.nz, 2 * c.mz, W, f->P, f->P * W, f->mxy};
    const ModeMaps mm = mode_maps(q);
    const long modes = f->modes;
    hipError_t e;
    // lift
    if ((e = rowgemm(f, x, f->params + f->oP, f->h[0], rows, c.c_in, W, 0.0f, f->params + f->op, s)) != hipSuccess)
        return e;
    for (int l = 0; l < L; ++l) {
        const float *hin = f->h[l];
        float *ul = f->u[l];
        if ((e = zdft(f, q, hin, f->zf, 0, 0.0f, s)) != hipSuccess) return e;
        if ((e = xydft(f, q, f->zf, f->sp[l], 0, s)) != hipSuccess) return e;
        const unsigned short *El = f->E + (long)l * modes * 4 * W * W;
        if (B == 1 && 2 * W == 64) {
            if ((e = mode_mix_row(f, f->sp[l], El, f->tm, 2 * W, mm, modes, s)) != hipSuccess) return e;
        } else if ((e = G(f, f->sp[l], 0, El, 1, f->tm, B, 2 * W, 2 * W, mm.row, mm.ch, m1(2 * W), m1(1), mm.row,
                          mm.ch, (int)modes, mm.batch, m1(4L * W * W), mm.batch, 0.0f, nullptr, 1, nullptr, s)) !=
                   hipSuccess)
            return e;
        if ((e = xydft(f, q, f->tm, f->zi, 1, s)) != hipSuccess) return e;
        if ((e = zdft(f, q, f->zi, ul, 1, 0.0f, s)) != hipSuccess) return e;
        if ((e = rowgemm(f, hin, f->params + f->oW[l], ul, rows, W, W, 1.0f, f->params + f->ob[l], s)) != hipSuccess)
            return e;
        if (l < L - 1) {
            hipLaunchKernelGGL(gelu_fwd, grid1(rows * W), dim3(256), 0, s, ul, f->h[l + 1], rows * W);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    }
    // projection
    if ((e = rowgemm(f, f->h[L], f->params + f->oQ1, f->g1, rows, W, c.hidden, 0.0f, f->params + f->oq1, s)) !=
        hipSuccess)
        return e;
    hipLaunchKernelGGL(gelu_fwd, grid1(rows * c.hidden), dim3(256), 0, s, f->g1, f->a1, rows * c.hidden);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return rowgemm(f, f->a1, f->params + f->oQ2, f->y, rows, c.hidden, c.c_out, 0.0f, f->params + f->oq2, s);
}

hipError_t backward_impl(bec_fno *f, int B, const float *x, hipStream_t s)
{
    const bec_fno_config &c = f->c;
    const int W = c.width, L = c.layers, H = c.hidden;
    const long rows = (long)B * f->S;
    Spec q{B, c.nz, 2 * c.mz, W, f->P, f->P * W, f->mxy};
    const ModeMaps mm = mode_maps(q);
    const long modes = f->modes;
    float *g = f->grads;
    hipError_t e;
    // projection: y = a1 Q2 + q2, a1 = gelu(g1), g1 = hL Q1 + q1
    if ((e = wgrad(f, f->a1, H, f->dy, c.c_out, rows, g + f->oQ2, s)) != hipSuccess) return e;
    if ((e = bgrad(f, f->dy, c.c_out, rows, g + f->oq2, s)) != hipSuccess) return e;
    float *da1 = f->dB;   // [rows][H]
    if ((e = rowgemm_bt(f, f->dy, f->params + f->oQ2, da1, rows, c.c_out, H, 0.0f, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(gelu_bwd, grid1(rows * H), dim3(256), 0, s, f->g1, da1, rows * H);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = wgrad(f, f->h[L], W, da1, H, rows, g + f->oQ1, s)) != hipSuccess) return e;
    if ((e = bgrad(f, da1, H, rows, g + f->oq1, s)) != hipSuccess) return e;
    float *dh = f->dA;    // gradient w.r.t. the current layer output
    if ((e = rowgemm_bt(f, da1, f->params + f->oQ1, dh, rows, H, W, 0.0f, s)) != hipSuccess) return e;
    for (int l = L - 1; l >= 0; --l) {
        float *du = dh;   // in place: du = dh * gelu'(u_l)
        if (l < L - 1) {
            hipLaunchKernelGGL(gelu_bwd, grid1(rows * W), dim3(256), 0, s, f->u[l], du, rows * W);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        const float *hin = f->h[l];
        if ((e = wgrad(f, hin, W, du, W, rows, g + f->oW[l], s)) != hipSuccess) return e;
        if ((e = bgrad(f, du, W, rows, g + f->ob[l], s)) != hipSuccess) return e;
        float *dhin = (dh == f->dA) ? f->dH : f->dA;
        if ((e = rowgemm_bt(f, du, f->params + f->oW[l], dhin, rows, W, W, 0.0f, s)) != hipSuccess) return e;
        // spectral path, adjoint order
        if ((e = zdft(f, q, du, f->zi, 0, 0.0f, s)) != hipSuccess) return e;          // dZi = Fz du
        if ((e = xydft(f, q, f->zi, f->tm, 0, s)) != hipSuccess) return e;            // dT = Fxy dZi
        const unsigned short *El = f->E + (long)l * modes * 4 * W * W;
        // dE[t] = Sp[t]^T dT[t]  (K = batch)
        if ((e = G(f, f->sp[l], 0, f->tm, 0, f->dE, 2 * W, 2 * W, B, mm.ch, mm.row, mm.row, mm.ch, m1(2 * W), m1(1),
                   (int)modes, mm.batch, mm.batch, m1(4L * W * W), 0.0f, nullptr, 1, nullptr, s)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(fold_modes, grid1(modes * W * W), dim3(256), 0, s, f->dE, g + f->oR[l], W, modes);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // dSp = dT E^T
        if ((e = G(f, f->tm, 0, El, 1, f->zf, B, 2 * W, 2 * W, mm.row, mm.ch, m1(1), m1(2 * W), mm.row, mm.ch,
                   (int)modes, mm.batch, m1(4L * W * W), mm.batch, 0.0f, nullptr, 1, nullptr, s)) != hipSuccess)
            return e;
        if ((e = xydft(f, q, f->zf, f->zi, 1, s)) != hipSuccess) return e;            // dZf = Fxy^T dSp
        if ((e = zdft(f, q, f->zi, dhin, 1, 1.0f, s)) != hipSuccess) return e;        // dh += Fz^T dZf
        dh = dhin;
    }
    // lift
    if ((e = wgrad(f, x, c.c_in, dh, W, rows, g + f->oP, s)) != hipSuccess) return e;
    return bgrad(f, dh, W, rows, g + f->op, s);
}

int validate(const bec_fno_config &c)
{
    if (c.nz < 1 || c.ny < 1 || c.nx < 1 || c.width < 1 || c.layers < 1 || c.c_in < 1 || c.c_out < 1 ||
        c.hidden < 1 || c.max_batch < 1)
        return BEC_ERR_INVALID;
    if (c.mz < 2 || (c.mz & 1) || c.mz > c.nz || c.my < 1 || c.my > c.ny || c.mx < 1 || c.mx > c.nx)
        return BEC_ERR_INVALID;
    // two-level maps need power-of-two lo counts
    if (!pow2(c.width) || !pow2((long)c.my * c.mx)) return BEC_ERR_SHAPE;
    if ((long)c.max_batch * c.nz * c.ny * c.nx > (1L << 31) / 2) return BEC_ERR_SHAPE;
    if ((long)c.max_batch * (c.mz / 2) > 65535 || (long)c.mz / 2 * c.my * c.mx > 65535) return BEC_ERR_SHAPE;
    return BEC_OK;
}

// Fz: 2 mz x nz rows (re, im) of the retained frequencies 0..mz/2-1, -mz/2..-1;
// Fxy: 2 mxy x 2 P, rows (re/im, ky, kx), columns (re/im, y, x).
hipError_t build_tables(bec_fno *f)
{
    const bec_fno_config &c = f->c;
    const long P = f->P, mxy = f->mxy;
    std::vector<float> fz((size_t)2 * c.mz * c.nz), fxy((size_t)(2 * mxy) * (2 * P));
    const double sz = 1.0 / sqrt((double)c.nz), sp = 1.0 / sqrt((double)P);
    for (int ki = 0; ki < c.mz; ++ki) {
        const int k = ki < c.mz / 2 ? ki : ki - c.mz;
        for (int z = 0; z < c.nz; ++z) {
            const double th = 2.0 * M_PI * (double)k * z / c.nz;
            fz[(size_t)(2 * ki) * c.nz + z] = (float)(cos(th) * sz);
            fz[(size_t)(2 * ki + 1) * c.nz + z] = (float)(-sin(th) * sz);
        }
    }
    for (int kyi = 0; kyi < c.my; ++kyi)
        for (int kxi = 0; kxi < c.mx; ++kxi) {
            const int ky = kyi < (c.my + 1) / 2 ? kyi : kyi - c.my;
            const int kx = kxi < (c.mx + 1) / 2 ? kxi : kxi - c.mx;
            const long r = (long)kyi * c.mx + kxi;
            for (int y = 0; y < c.ny; ++y)
                for (int x = 0; x < c.nx; ++x) {
                    const double th = 2.0 * M_PI * ((double)ky * y / c.ny + (double)kx * x / c.nx);
                    const float cs = (float)(cos(th) * sp), sn = (float)(sin(th) * sp);
                    const long q = (long)y * c.nx + x;
                    fxy[(size_t)r * 2 * P + q] = cs;              // re <- re
                    fxy[(size_t)r * 2 * P + P + q] = sn;          // re <- im
                    fxy[(size_t)(mxy + r) * 2 * P + q] = -sn;     // im <- re
                    fxy[(size_t)(mxy + r) * 2 * P + P + q] = cs;  // im <- im
                }
        }
    hipError_t e = hipMemcpy(f->Fz, fz.data(), fz.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(f->Fxy, fxy.data(), fxy.size() * sizeof(float), hipMemcpyHostToDevice);
    return e;
}

void init_params(bec_fno *f, uint64_t seed, std::vector<float> &p)
{
    const bec_fno_config &c = f->c;
    const long W = c.width;
    p.assign(f->n_params, 0.0f);
    Rng r{seed};
    auto fill = [&](long off, long n, double scale) {
        for (long i = 0; i < n; ++i) p[off + i] = (float)(r.uni() * scale);
    };
    fill(f->oP, (long)c.c_in * W, 1.0 / sqrt((double)c.c_in));
    fill(f->op, W, 1.0 / sqrt((double)c.c_in));
    for (int l = 0; l < c.layers; ++l) {
        fill(f->oW[l], W * W, 1.0 / sqrt((double)W));
        fill(f->ob[l], W, 1.0 / sqrt((double)W));
        fill(f->oR[l], f->modes * W * W * 2, 1.0 / (double)W);
    }
    fill(f->oQ1, W * c.hidden, 1.0 / sqrt((double)W));
    fill(f->oq1, c.hidden, 1.0 / sqrt((double)W));
    fill(f->oQ2, (long)c.hidden * c.c_out, 1.0 / sqrt((double)c.hidden));
    fill(f->oq2, c.c_out, 1.0 / sqrt((double)c.hidden));
}

int create(const bec_fno_config &c, uint64_t seed, bec_fno *f)
{
    f->c = c;
    f->S = (long)c.nz * c.ny * c.nx;
    f->P = (long)c.ny * c.nx;
    f->mxy = (long)c.my * c.mx;
    f->modes = (long)c.mz * f->mxy;
    layout(f);
    const long W = c.width, L = c.layers, Bm = c.max_batch;
    const long rows = Bm * f->S;
    const long spec = Bm * 2L * c.mz * f->P * W;          // z-DFT output / xy-DFT input
    const long specm = Bm * 2L * c.mz * f->mxy * W;       // retained modes
    const size_t fb = sizeof(float);
#define FA(ptr, n) BEC_HIP(falloc(f, (void **)&(ptr), (size_t)(n)))
    FA(f->params, fb * f->n_params);
    FA(f->grads, fb * f->n_params);
    FA(f->adam_m, fb * f->n_params);
    FA(f->adam_v, fb * f->n_params);
    FA(f->E, sizeof(unsigned short) * L * f->modes * 4 * W * W);
    FA(f->Fz, fb * 2 * c.mz * c.nz);
    FA(f->Fxy, fb * 4 * f->mxy * f->P);
    FA(f->one, fb);
    f->h.assign(L + 1, nullptr);
    f->u.assign(L, nullptr);
    f->sp.assign(L, nullptr);
    for (int l = 0; l <= L; ++l) FA(f->h[l], fb * rows * W);
    for (int l = 0; l < L - 1; ++l) FA(f->u[l], fb * rows * W);
    f->u[L - 1] = f->h[L];
    for (int l = 0; l < L; ++l) FA(f->sp[l], fb * specm);
    FA(f->zf, fb * spec);
    FA(f->zi, fb * spec);
    FA(f->tm, fb * specm);
    FA(f->g1, fb * rows * c.hidden);
    FA(f->a1, fb * rows * c.hidden);
    FA(f->y, fb * rows * c.c_out);
    FA(f->dy, fb * rows * c.c_out);
    FA(f->dA, fb * rows * W);
    FA(f->dH, fb * rows * W);
    FA(f->dB, fb * rows * c.hidden);
    FA(f->dE, fb * f->modes * 4 * W * W);
    long mn = std::max(std::max((long)c.c_in * W, W * W), std::max(W * (long)c.hidden, (long)c.hidden * c.c_out));
    // split-K partials: the weight gradients (KSPLIT layers of at most mn) and the
    // small-batch splits of G (at most 8 layers of < 256 64 x 64 tiles)
    FA(f->work, fb * std::max((long)KSPLIT * mn, 8L * 256 * 64 * 64));
    FA(f->loss_part, sizeof(double) * ((rows * c.c_out + 255) / 256));
#undef FA
    std::vector<float> p;
    init_params(f, seed, p);
    const float one = 1.0f;
    BEC_HIP(hipMemcpy(f->params, p.data(), fb * f->n_params, hipMemcpyHostToDevice));
    BEC_HIP(hipMemcpy(f->one, &one, fb, hipMemcpyHostToDevice));
    BEC_HIP(hipMemset(f->adam_m, 0, fb * f->n_params));
    BEC_HIP(hipMemset(f->adam_v, 0, fb * f->n_params));
    BEC_HIP(hipMemset(f->grads, 0, fb * f->n_params));
    BEC_HIP(build_tables(f));
    BEC_HIP(expand_all(f, nullptr));
    BEC_HIP(hipDeviceSynchronize());
    return BEC_OK;
}

void release(bec_fno *f)
{
    for (void *p : f->owned) (void)hipFree(p);
    f->owned.clear();
}

} // namespace
} // namespace bec

extern "C" int bec_fno_create(const bec_fno_config *cfg, uint64_t seed, bec_fno **out)
{
    if (!cfg || !out) return BEC_ERR_INVALID;
    *out = nullptr;
    int st = bec::validate(*cfg);
    if (st != BEC_OK) return st;
    bec_fno *f = new bec_fno;
    st = bec::create(*cfg, seed, f);
    if (st != BEC_OK) {
        bec::release(f);
        delete f;
        return st;
    }
    *out = f;
    return BEC_OK;
}

extern "C" int bec_fno_destroy(bec_fno *f)
{
    if (!f) return BEC_ERR_INVALID;
    (void)hipDeviceSynchronize();
    bec::release(f);
    delete f;
    return BEC_OK;
}

extern "C" long long bec_fno_param_count(const bec_fno *f) { return f ? f->n_params : -1; }

extern "C" int bec_fno_get_params(const bec_fno *f, float *host_out)
{
    if (!f || !host_out) return BEC_ERR_INVALID;
    BEC_HIP(hipDeviceSynchronize());
    BEC_HIP(hipMemcpy(host_out, f->params, sizeof(float) * f->n_params, hipMemcpyDeviceToHost));
    return BEC_OK;
}

extern "C" int bec_fno_set_params(bec_fno *f, const float *host_in)
{
    if (!f || !host_in) return BEC_ERR_INVALID;
    BEC_HIP(hipDeviceSynchronize());
    BEC_HIP(hipMemcpy(f->params, host_in, sizeof(float) * f->n_params, hipMemcpyHostToDevice));
    BEC_HIP(bec::expand_all(f, nullptr));
    BEC_HIP(hipDeviceSynchronize());
    return BEC_OK;
}

extern "C" int bec_fno_get_grads(const bec_fno *f, float *host_out)
{
    if (!f || !host_out) return BEC_ERR_INVALID;
    BEC_HIP(hipDeviceSynchronize());
    BEC_HIP(hipMemcpy(host_out, f->grads, sizeof(float) * f->n_params, hipMemcpyDeviceToHost));
    return BEC_OK;
}

extern "C" int bec_fno_forward(bec_fno *f, int batch, const float *d_x, float *d_y, void *stream)
{
    if (!f || !d_x || !d_y || batch < 1 || batch > f->c.max_batch) return BEC_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    bec::CallerBuffers reg(f, {{d_x, sizeof(float) * batch * f->S * f->c.c_in}});
    BEC_HIP(bec::forward_impl(f, batch, d_x, s));
    BEC_HIP(hipMemcpyAsync(d_y, f->y, sizeof(float) * batch * f->S * f->c.c_out, hipMemcpyDeviceToDevice, s));
    return BEC_OK;
}

extern "C" int bec_fno_gs2_features(const bec_calc_params *p, const float scaled[4], float *d_x, void *stream)
{
    if (!p || !scaled || !d_x || ((uintptr_t)d_x & 15) != 0) return BEC_ERR_INVALID;
    if (p->nvx < 1 || p->nvy < 1 || p->nvz < 1 || (long)p->cells != (long)p->nvx * p->nvy * p->nvz) return BEC_ERR_SHAPE;
    if (!(p->mu > 0.0) || !(p->xmax > 0.0) || !(p->ymax > 0.0) || !(p->zmax > 0.0) || !(p->lambda > 0.0))
        return BEC_ERR_INVALID;
    bec::Gs2Lattice g{p->nvx, p->nvy, p->nvz, p->xmax, p->ymax, p->zmax, p->dx, p->dy, p->dz, p->lambda, p->mu,
                      {scaled[0], scaled[1], scaled[2], scaled[3]}};
    const long cells = p->cells;
    hipLaunchKernelGGL(bec::gs2_features, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, (hipStream_t)stream, g,
                       d_x, cells);
    BEC_HIP(hipGetLastError());
    return BEC_OK;
}

extern "C" int bec_fno_loss_backward(bec_fno *f, int batch, const float *d_x, const float *d_target, double *loss,
                                     void *stream)
{
    if (!f || !d_x || !d_target || batch < 1 || batch > f->c.max_batch) return BEC_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    bec::CallerBuffers reg(f, {{d_x, sizeof(float) * batch * f->S * f->c.c_in}});
    BEC_HIP(bec::forward_impl(f, batch, d_x, s));
    const long n = (long)batch * f->S * f->c.c_out;
    const long blocks = (n + 255) / 256;
    hipLaunchKernelGGL(bec::mse_grad, dim3((unsigned)blocks), dim3(256), 0, s, f->y, d_target, f->dy, n,
                       (float)(2.0 / n), f->loss_part);
    BEC_HIP(hipGetLastError());
    BEC_HIP(bec::backward_impl(f, batch, d_x, s));
    if (loss) {
        std::vector<double> part(blocks);
        BEC_HIP(hipMemcpyAsync(part.data(), f->loss_part, sizeof(double) * blocks, hipMemcpyDeviceToHost, s));
        BEC_HIP(hipStreamSynchronize(s));
        double t = 0.0;
        for (double v : part) t += v;
        *loss = t / (double)n;
    }
    return BEC_OK;
}

extern "C" int bec_fno_adam_step(bec_fno *f, float lr, float beta1, float beta2, float eps, float weight_decay,
                                 void *stream)
{
    if (!f || !(lr >= 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return BEC_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    f->adam_t += 1;
    const float c1 = (float)(1.0 - pow((double)beta1, (double)f->adam_t));
    const float c2 = (float)(1.0 - pow((double)beta2, (double)f->adam_t));
    hipLaunchKernelGGL(bec::adam_kernel, bec::grid1(f->n_params), dim3(256), 0, s, f->params, f->grads, f->adam_m,
                       f->adam_v, f->n_params, lr, beta1, beta2, eps, weight_decay, c1, c2);
    BEC_HIP(hipGetLastError());
    BEC_HIP(bec::expand_all(f, s));
    return BEC_OK;
}
