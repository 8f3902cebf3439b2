// Strided batched GEMM on BF16 operands with FP32 accumulation, on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).  The building block of the
// surrogate network (bec_fno.h): DFTs along the lattice axes, per-mode channel
// mixing and the per-cell linear layers are all instances of it.
//
// Workgroup tile 64 x 64 (4 waves, each 32 x 32 = 2 x 2 MFMA blocks), K tile
// 32: one MFMA per block and K tile.  Operands are read through two-level
// index maps (any strides), rounded to BF16 (nearest even) and staged in LDS
// as [row][k] for A and [col][k] for B, so that each lane's eight K values are
// one 16-byte LDS read.
//
// MFMA 16x16x32 operands (lane l, item i = 0..7): lane l supplies row
// m = l % 16 of A and column n = l % 16 of B at the eight K positions
// 8 (l / 16) + i.  A and B use the same K assignment, so the product is the
// full sum over the 32 K values whatever order the unit adds them in.
// Result: D[m = 4 (l / 16) + i][n = l % 16], i = 0..3.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bec_fno.h"
#include "host_common.hpp"

namespace bec {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;   // LDS row: 40 bf16 = 80 B

struct Map {
    long hi, lo;
    int lg;
};

struct Desc {
    const void *A, *B;
    float *C;
    const float *bias;
    float *work;
    int M, N, K, batch, ksplit;
    float alpha, beta;
    Map am, ak, bk, bn, cm, cn, ab, bb, cb;
    int a_m_fast, b_n_fast;   // thread order of the global loads (element path)
    // 8-wide loads: 0 element path, 1 eight consecutive K values per thread,
    // 2 eight consecutive rows of A (columns of B) per thread (host-checked:
    // unit stride, every offset a multiple of 8, size a multiple of 8)
    int a_vec, b_vec;
};

__device__ __forceinline__ long moff(const Map &m, long i)
{
    return (i >> m.lg) * m.hi + (i & ((1L << m.lg) - 1)) * m.lo;
}

__device__ __forceinline__ unsigned short to_bf16(float f)
{
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

template <bool BF> __device__ __forceinline__ unsigned short load_bf16(const void *base, long off)
{
    if constexpr (BF) return ((const unsigned short *)base)[off];
    else return to_bf16(((const float *)base)[off]);
}

typedef unsigned int uint4_t __attribute__((ext_vector_type(4)));

// eight consecutive values from element offset off (a multiple of 8), as BF16
template <bool BF> __device__ __forceinline__ uint4_t load8_bf16(const void *base, long off)
{
    if constexpr (BF) {
        return *(const uint4_t *)((const unsigned short *)base + off);
    } else {
        const float4_t lo = *(const float4_t *)((const float *)base + off);
        const float4_t hi = *(const float4_t *)((const float *)base + off + 4);
        uint4_t r;
        r.x = to_bf16(lo.x) | ((unsigned)to_bf16(lo.y) << 16);
        r.y = to_bf16(lo.z) | ((unsigned)to_bf16(lo.w) << 16);
        r.z = to_bf16(hi.x) | ((unsigned)to_bf16(hi.y) << 16);
        r.w = to_bf16(hi.z) | ((unsigned)to_bf16(hi.w) << 16);
        return r;
    }
}

// Stages the 64 x 32 tile of one operand into S[row][k]: rows r0.. of the
// operand (m of A, n of B), K values k0..; v selects the load form (Desc).
template <bool BF>
__device__ __forceinline__ void stage(unsigned short (*S)[LDK], int v, int fast, const void *X, long base,
                                      const Map &rm, const Map &km, int r0, int rows, int k0, int ke, int tid)
{
    if (v == 1) {   // eight consecutive K values: one 16-byte LDS store
        const int r = tid >> 2, kk = (tid & 3) * 8;
        uint4_t w = (uint4_t){0u, 0u, 0u, 0u};
        if (r0 + r < rows && k0 + kk < ke) w = load8_bf16<BF>(X, base + moff(rm, r0 + r) + moff(km, k0 + kk));
        *(uint4_t *)&S[r][kk] = w;
    } else if (v == 2) {   // eight consecutive rows at one K value
        const int r = (tid & 7) * 8, kk = tid >> 3;
        uint4_t w = (uint4_t){0u, 0u, 0u, 0u};
        if (r0 + r < rows && k0 + kk < ke) w = load8_bf16<BF>(X, base + moff(rm, r0 + r) + moff(km, k0 + kk));
        const unsigned e[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            S[r + 2 * i][kk] = (unsigned short)(e[i] & 0xffffu);
            S[r + 2 * i + 1][kk] = (unsigned short)(e[i] >> 16);
        }
    } else {
#pragma unroll
        for (int e = 0; e < (BM * BK) / 256; ++e) {
            const int idx = tid + e * 256;
            int r, kk;
            if (fast) { r = idx % BM; kk = idx / BM; }
            else { r = idx / BK; kk = idx % BK; }
            const int m = r0 + r, k = k0 + kk;
            unsigned short val = 0;
            if (m < rows && k < ke) val = load_bf16<BF>(X, base + moff(rm, m) + moff(km, k));
            S[r][kk] = val;
        }
    }
}

template <bool ABF, bool BBF>
__global__ __launch_bounds__(256) void gemm_kernel(Desc d)
{
    __shared__ __align__(16) unsigned short As[BM][LDK];
    __shared__ __align__(16) unsigned short Bs[BN][LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bz = blockIdx.z;
    const int bat = bz / d.ksplit, split = bz - bat * d.ksplit;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const long abase = moff(d.ab, bat), bbase = moff(d.bb, bat);

    // K range of this split (whole K tiles)
    const int ktiles = (d.K + BK - 1) / BK;
    const int per = (ktiles + d.ksplit - 1) / d.ksplit;
    const int kb = split * per * BK;
    const int ke = min(d.K, (split + 1) * per * BK);

    float4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};

    for (int k0 = kb; k0 < ke; k0 += BK) {
        stage<ABF>(As, d.a_vec, d.a_m_fast, d.A, abase, d.am, d.ak, m0, d.M, k0, ke, tid);
        stage<BBF>(Bs, d.b_vec, d.b_n_fast, d.B, bbase, d.bn, d.bk, n0, d.N, k0, ke, tid);
        __syncthreads();
        {
            const int kq = 8 * (lane >> 4);
            bf16x8_t af[2], bf[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) af[i] = *(const bf16x8_t *)&As[wm + i * 16 + (lane & 15)][kq];
#pragma unroll
            for (int j = 0; j < 2; ++j) bf[j] = *(const bf16x8_t *)&Bs[wn + j * 16 + (lane & 15)][kq];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }

    const long cbase = moff(d.cb, bat);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + q;
                const int n = n0 + wn + j * 16 + (lane & 15);
                if (m >= d.M || n >= d.N) continue;
                const float v = acc[i][j][q];
                if (d.ksplit > 1) {
                    d.work[(((long)split * d.batch + bat) * d.M + m) * d.N + n] = v;
                } else {
                    const long off = cbase + moff(d.cm, m) + moff(d.cn, n);
                    float r = d.alpha * v;
                    if (d.beta != 0.f) r += d.beta * d.C[off];
                    if (d.bias) r += d.bias[n];
                    d.C[off] = r;
                }
            }
}

// Sum of the split-K partials in split order, then the epilogue.
__global__ void gemm_split_reduce(Desc d)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)d.batch * d.M * d.N;
    if (i >= total) return;
    const int n = (int)(i % d.N);
    const int m = (int)((i / d.N) % d.M);
    const int bat = (int)(i / ((long)d.M * d.N));
    float s = 0.f;
    for (int k = 0; k < d.ksplit; ++k) s += d.work[(long)k * total + i];
    const long off = moff(d.cb, bat) + moff(d.cm, m) + moff(d.cn, n);
    float r = d.alpha * s;
    if (d.beta != 0.f) r += d.beta * d.C[off];
    if (d.bias) r += d.bias[n];
    d.C[off] = r;
}

Map to_map(const bec_map2 &m)
{
    Map r;
    r.hi = (long)m.hi;
    r.lo = (long)m.lo;
    r.lg = m.lg;
    return r;
}

bool map_ok(const bec_map2 &m) { return m.lg >= 0 && m.lg <= 40; }

// every offset of the map over indices [0, range) is a multiple of 8
bool mult8(const bec_map2 &m, long range)
{
    const bool lo_ok = m.lg == 0 || m.lo % 8 == 0;
    const bool hi_ok = m.hi % 8 == 0 || range <= (1L << m.lg);
    return lo_ok && hi_ok;
}

// runs of 8 consecutive indices (starting at multiples of 8) are 8 consecutive
// elements starting at a multiple of 8
bool unit8(const bec_map2 &m, long range)
{
    return m.lo == 1 && m.lg >= 3 && (m.hi % 8 == 0 || range <= (1L << m.lg));
}

// the 8-wide load form for one operand: rows rm over [0, rows), K map km,
// batch map bm
int vec_form(const bec_map2 &rm, long rows, const bec_map2 &km, long K, const bec_map2 &bm, long batch,
             const void *ptr)
{
    if ((uintptr_t)ptr % 16) return 0;
    if (!mult8(bm, batch)) return 0;
    if (K % 8 == 0 && unit8(km, K) && mult8(rm, rows)) return 1;
    if (rows % 8 == 0 && unit8(rm, rows) && mult8(km, K)) return 2;
    return 0;
}

} // namespace

hipError_t run_gemm(const bec_gemm_desc &g, hipStream_t s)
{
    Desc d;
    d.A = g.A;
    d.B = g.B;
    d.C = g.C;
    d.bias = g.bias;
    d.work = g.work;
    d.M = g.M;
    d.N = g.N;
    d.K = g.K;
    d.batch = g.batch;
    d.ksplit = g.ksplit < 1 ? 1 : g.ksplit;
    d.alpha = g.alpha;
    d.beta = g.beta;
    d.am = to_map(g.am);
    d.ak = to_map(g.ak);
    d.bk = to_map(g.bk);
    d.bn = to_map(g.bn);
    d.cm = to_map(g.cm);
    d.cn = to_map(g.cn);
    d.ab = to_map(g.ab);
    d.bb = to_map(g.bb);
    d.cb = to_map(g.cb);
    // coalesce the global loads along whichever operand index has unit stride
    d.a_m_fast = (g.am.lo == 1 && g.am.lg >= 6) ? 1 : 0;
    d.b_n_fast = (g.bn.lo == 1 && g.bn.lg >= 6) ? 1 : 0;
    d.a_vec = vec_form(g.am, g.M, g.ak, g.K, g.ab, g.batch, g.A);
    d.b_vec = vec_form(g.bn, g.N, g.bk, g.K, g.bb, g.batch, g.B);
    dim3 grid((unsigned)((g.N + BN - 1) / BN), (unsigned)((g.M + BM - 1) / BM), (unsigned)(g.batch * d.ksplit));
    if (g.a_bf16 && g.b_bf16) hipLaunchKernelGGL((gemm_kernel<true, true>), grid, dim3(256), 0, s, d);
    else if (g.a_bf16) hipLaunchKernelGGL((gemm_kernel<true, false>), grid, dim3(256), 0, s, d);
    else if (g.b_bf16) hipLaunchKernelGGL((gemm_kernel<false, true>), grid, dim3(256), 0, s, d);
    else hipLaunchKernelGGL((gemm_kernel<false, false>), grid, dim3(256), 0, s, d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || d.ksplit == 1) return e;
    const long total = (long)g.batch * g.M * g.N;
    hipLaunchKernelGGL(gemm_split_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, d);
    return hipGetLastError();
}

int gemm_validate(const bec_gemm_desc &g)
{
    if (!g.A || !g.B || !g.C) return BEC_ERR_INVALID;
    if (g.M < 1 || g.N < 1 || g.K < 1 || g.batch < 1 || g.ksplit < 0) return BEC_ERR_INVALID;
    if (g.batch > 65535 * 64 || (long)g.batch * (g.ksplit < 1 ? 1 : g.ksplit) > 65535) return BEC_ERR_INVALID;
    if ((long)(g.M + BM - 1) / BM > 65535) return BEC_ERR_INVALID;
    if (g.ksplit > 1 && !g.work) return BEC_ERR_INVALID;
    const bec_map2 *maps[] = {&g.am, &g.ak, &g.bk, &g.bn, &g.cm, &g.cn, &g.ab, &g.bb, &g.cb};
    for (const bec_map2 *m : maps)
        if (!map_ok(*m)) return BEC_ERR_INVALID;
    return BEC_OK;
}

} // namespace bec

extern "C" int bec_gemm(const bec_gemm_desc *d, void *stream)
{
    if (!d) return BEC_ERR_INVALID;
    int st = bec::gemm_validate(*d);
    if (st != BEC_OK) return st;
    hipError_t e = bec::run_gemm(*d, (hipStream_t)stream);
    return e == hipSuccess ? BEC_OK : BEC_ERR_HIP;
}

extern "C" int bec_gemm_time(const bec_gemm_desc *d, int iters, double *ms_per_gemm)
{
    if (!d || iters <= 0 || !ms_per_gemm) return BEC_ERR_INVALID;
    int st = bec::gemm_validate(*d);
    if (st != BEC_OK) return st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = bec::run_gemm(*d, nullptr);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (int i = 0; i < iters && e == hipSuccess; ++i) e = bec::run_gemm(*d, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return BEC_ERR_HIP;
    *ms_per_gemm = (double)ms / iters;
    return BEC_OK;
}

extern "C" int bec_device_synchronize(void)
{
    return hipDeviceSynchronize() == hipSuccess ? BEC_OK : BEC_ERR_HIP;
}
