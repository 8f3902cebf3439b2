// Strided batched GEMM on BF16 operands with FP32 accumulation, on the
// gfx950 matrix cores (v_mfma_f32_16x16x16_bf16).  The building block of the
// surrogate network (bec_fno.h): DFTs along the lattice axes, per-mode channel
// mixing and the per-cell linear layers are all instances of it.
//
// Workgroup tile 64 x 64 (4 waves, each 32 x 32 = 2 x 2 MFMA blocks), K tile
// 32.  Operands are read through two-level index maps (any strides), rounded
// to BF16 (nearest even) and staged in LDS as [row][k] for A and [col][k]
// for B, so that each lane's four K values are one 8-byte LDS read.
//
// MFMA 16x16x16 operand layout (lane l, item i = 0..3):
//   A[m = l % 16][k = 4 (l / 16) + i],  B[k = 4 (l / 16) + i][n = l % 16],
//   D[m = 4 (l / 16) + i][n = l % 16].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bec_fno.h"
#include "host_common.hpp"

namespace bec {

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
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
    int a_m_fast, b_n_fast;   // thread order of the global loads
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

template <bool ABF, bool BBF>
__global__ __launch_bounds__(256) void gemm_kernel(Desc d)
{
    __shared__ unsigned short As[BM][LDK];
    __shared__ unsigned short Bs[BN][LDK];
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
#pragma unroll
        for (int e = 0; e < (BM * BK) / 256; ++e) {
            const int idx = tid + e * 256;
            int r, kk;
            if (d.a_m_fast) { r = idx % BM; kk = idx / BM; }
            else { r = idx / BK; kk = idx % BK; }
            const int m = m0 + r, k = k0 + kk;
            unsigned short v = 0;
            if (m < d.M && k < ke) v = load_bf16<ABF>(d.A, abase + moff(d.am, m) + moff(d.ak, k));
            As[r][kk] = v;
        }
#pragma unroll
        for (int e = 0; e < (BN * BK) / 256; ++e) {
            const int idx = tid + e * 256;
            int c, kk;
            if (d.b_n_fast) { c = idx % BN; kk = idx / BN; }
            else { c = idx / BK; kk = idx % BK; }
            const int n = n0 + c, k = k0 + kk;
            unsigned short v = 0;
            if (n < d.N && k < ke) v = load_bf16<BBF>(d.B, bbase + moff(d.bk, k) + moff(d.bn, n));
            Bs[c][kk] = v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            const int kq = ks + 4 * (lane >> 4);
            short4_t af[2], bf[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) af[i] = *(const short4_t *)&As[wm + i * 16 + (lane & 15)][kq];
#pragma unroll
            for (int j = 0; j < 2; ++j) bf[j] = *(const short4_t *)&Bs[wn + j * 16 + (lane & 15)][kq];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bf[j], acc[i][j], 0, 0, 0);
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
