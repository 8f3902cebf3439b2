"""GPU: the BF16 MFMA GEMM (bec_gemm) against an FP64 numpy product of the
same BF16-rounded operands, for the layouts the surrogate network uses: row-
and column-major operands, ragged sizes, two-level strides, batches, split K
and the alpha / beta / bias epilogue."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _off(m, i):
    i = np.asarray(i, dtype=np.int64)
    return (i >> m.lg) * m.hi + (i & ((1 << m.lg) - 1)) * m.lo


def _reference(F, A, B, C0, M, N, K, am, ak, bk, bn, cm, cn, batch, ab, bb, cb, alpha, beta, bias):
    """C after the GEMM, computed in FP64 from the BF16-rounded operands via the index maps."""
    Ar = F.from_bf16_bits(F.to_bf16_bits(A)).astype(np.float64) if A.dtype != np.uint16 else \
        F.from_bf16_bits(A).astype(np.float64)
    Br = F.from_bf16_bits(F.to_bf16_bits(B)).astype(np.float64) if B.dtype != np.uint16 else \
        F.from_bf16_bits(B).astype(np.float64)
    C = C0.astype(np.float64).copy()
    mi, ki, ni = np.arange(M), np.arange(K), np.arange(N)
    for t in range(batch):
        Am = Ar.ravel()[_off(ab, t) + _off(am, mi)[:, None] + _off(ak, ki)[None, :]]
        Bm = Br.ravel()[_off(bb, t) + _off(bk, ki)[:, None] + _off(bn, ni)[None, :]]
        idx = _off(cb, t) + _off(cm, mi)[:, None] + _off(cn, ni)[None, :]
        r = alpha * (Am @ Bm) + beta * C.ravel()[idx]
        if bias is not None:
            r = r + bias[None, :]
        flat = C.ravel()
        flat[idx] = r
        C = flat.reshape(C.shape)
    return C


def _run(F, rng, A, B, Cshape, M, N, K, maps, batch=1, bmaps=None, alpha=1.0, beta=0.0, use_bias=False,
         ksplit=1, a_bf16=False, b_bf16=False):
    am, ak, bk, bn, cm, cn = maps
    ab, bb, cb = bmaps if bmaps else (F.m1(0), F.m1(0), F.m1(0))
    C0 = rng.normal(size=Cshape).astype(np.float32)
    bias = rng.normal(size=N).astype(np.float32) if use_bias else None
    Ah = F.to_bf16_bits(A) if a_bf16 else A.astype(np.float32)
    Bh = F.to_bf16_bits(B) if b_bf16 else B.astype(np.float32)
    dA = F.DeviceArray(Ah.shape, Ah.dtype, Ah)
    dB = F.DeviceArray(Bh.shape, Bh.dtype, Bh)
    dC = F.DeviceArray(C0.shape, np.float32, C0)
    dbias = F.DeviceArray(bias.shape, np.float32, bias) if use_bias else None
    work = F.DeviceArray((ksplit * batch * M * N,), np.float32) if ksplit > 1 else None
    F.gemm(dA, dB, dC, M, N, K, am, ak, bk, bn, cm, cn, batch=batch, ab=ab, bb=bb, cb=cb, alpha=alpha, beta=beta,
           bias=dbias, ksplit=ksplit, work=work)
    got = dC.download().astype(np.float64)
    ref = _reference(F, Ah, Bh, C0, M, N, K, am, ak, bk, bn, cm, cn, batch, ab, bb, cb, alpha, beta, bias)
    scale = np.abs(ref).max() + 1e-30
    err = np.abs(got - ref).max() / scale
    return err


def test_row_major_ragged(bec):
    from bec_amd import fno as F
    rng = np.random.default_rng(0)
    M, N, K = 100, 70, 45
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    err = _run(F, rng, A, B, (M, N), M, N, K, (F.m1(K), F.m1(1), F.m1(N), F.m1(1), F.m1(N), F.m1(1)))
    assert err < 1e-5, err


def test_column_major_bf16_operands_with_epilogue(bec):
    from bec_amd import fno as F
    rng = np.random.default_rng(1)
    M, N, K = 130, 96, 200
    A = rng.normal(size=(K, M)).astype(np.float32)      # A[m][k] at k * M + m
    B = rng.normal(size=(N, K)).astype(np.float32)      # B[k][n] at n * K + k
    err = _run(F, rng, A, B, (N, M), M, N, K, (F.m1(1), F.m1(M), F.m1(1), F.m1(K), F.m1(1), F.m1(M)),
               alpha=0.5, beta=-1.25, use_bias=True, a_bf16=True, b_bf16=True)
    assert err < 1e-5, err


def test_two_level_maps_and_batches(bec):
    """Per-mode channel mixing: rows = batch samples, K = (re/im, channel) with the
    re/im halves 8192 apart, one GEMM per mode (two-level batch map)."""
    from bec_amd import fno as F
    rng = np.random.default_rng(2)
    S, W = 4, 32                  # samples, channels
    modes_hi, modes_lo = 3, 16    # 48 modes as (hi, lo) with lo stride 32
    sample = 2 * 3 * 16 * 32 * 2  # generous sample stride
    X = rng.normal(size=(S * sample,)).astype(np.float32)
    E = rng.normal(size=(modes_hi * modes_lo, 2 * W, 2 * W)).astype(np.float32)
    Y = np.zeros(S * sample, dtype=np.float32)
    re_im = modes_hi * modes_lo * W   # distance between the re and im halves
    am = F.m1(sample)
    ak = F.m2(W, 1, re_im)
    bk, bn = F.m1(2 * W), F.m1(1)
    cm, cn = F.m1(sample), F.m2(W, 1, re_im)
    ab = F.m2(modes_lo, W, modes_lo * W)
    bb = F.m1(4 * W * W)
    cb = F.m2(modes_lo, W, modes_lo * W)
    err = _run(F, rng, X, E, Y.shape, S, 2 * W, 2 * W, (am, ak, bk, bn, cm, cn), batch=modes_hi * modes_lo,
               bmaps=(ab, bb, cb), b_bf16=True)
    assert err < 1e-5, err


def test_split_k_weight_gradient(bec):
    """dW = X^T dY over many cells: M = N = 32, K = 5000, split 8 ways."""
    from bec_amd import fno as F
    rng = np.random.default_rng(3)
    K, W = 5000, 32
    X = rng.normal(size=(K, W)).astype(np.float32)
    D = rng.normal(size=(K, W)).astype(np.float32)
    err = _run(F, rng, X, D, (W, W), W, W, K, (F.m1(1), F.m1(W), F.m1(W), F.m1(1), F.m1(W), F.m1(1)),
               beta=1.0, use_bias=True, ksplit=8)
    assert err < 2e-5, err


def test_split_k_is_deterministic(bec):
    from bec_amd import fno as F
    rng = np.random.default_rng(4)
    K, W = 20000, 32
    X = rng.normal(size=(K, W)).astype(np.float32)
    dX = F.DeviceArray(X.shape, np.float32, X)
    outs = []
    for _ in range(2):
        dC = F.DeviceArray((W, W), np.float32, np.zeros((W, W), np.float32))
        work = F.DeviceArray((16 * W * W,), np.float32)
        F.gemm(dX, dX, dC, W, W, K, F.m1(1), F.m1(W), F.m1(W), F.m1(1), F.m1(W), F.m1(1), ksplit=16, work=work)
        outs.append(dC.download())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("a_bf16,b_bf16", [(False, False), (True, True), (False, True)])
def test_eight_wide_loads_k_and_row_runs(bec, a_bf16, b_bf16):
    """Sizes and strides that select the 8-wide operand loads: A row-major with
    K a multiple of 8 (eight K values per load), B row-major with N a multiple of
    8 (eight columns per load); M ragged, split K over a partial last tile."""
    from bec_amd import fno as F
    rng = np.random.default_rng(5)
    M, N, K = 137, 72, 200
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    maps = (F.m1(K), F.m1(1), F.m1(N), F.m1(1), F.m1(N), F.m1(1))
    err = _run(F, rng, A, B, (M, N), M, N, K, maps, a_bf16=a_bf16, b_bf16=b_bf16)
    assert err < 1e-5, err
    err = _run(F, rng, A, B, (M, N), M, N, K, maps, ksplit=3, a_bf16=a_bf16, b_bf16=b_bf16, beta=1.0)
    assert err < 1e-5, err


def test_eight_wide_loads_transposed_and_batched(bec):
    """A column-major with M a multiple of 8 (eight rows per load), B column-major
    with K a multiple of 8 (eight K values per load), batches 8-aligned, a
    two-level K map on A whose runs are 8-aligned."""
    from bec_amd import fno as F
    rng = np.random.default_rng(6)
    M, N, K, batch = 96, 50, 64, 3
    A = rng.normal(size=(batch, K, M)).astype(np.float32)   # A[t][m][k] at t*K*M + k*M + m
    B = rng.normal(size=(batch, N, K)).astype(np.float32)   # B[t][k][n] at t*N*K + n*K + k
    maps = (F.m1(1), F.m1(M), F.m1(1), F.m1(K), F.m1(N), F.m1(1))
    err = _run(F, rng, A, B, (batch, M, N), M, N, K, maps, batch=batch,
               bmaps=(F.m1(K * M), F.m1(N * K), F.m1(M * N)), a_bf16=True)
    assert err < 1e-5, err
    # K as two runs of 32 (hi stride 40 * M keeps every run start 8-aligned)
    A2 = rng.normal(size=(80 * M,)).astype(np.float32)
    Bk = rng.normal(size=(K, N)).astype(np.float32)
    err = _run(F, rng, A2, Bk, (M, N), M, N, K, (F.m1(1), F.m2(32, M, 40 * M), F.m1(N), F.m1(1), F.m1(N), F.m1(1)))
    assert err < 1e-5, err
