"""Host side of the learned ground-state start (include/bec_fno.h): ctypes
bindings of the BF16 MFMA GEMM and the Fourier neural operator, the input
features, and the device buffers they work on."""
import ctypes as C

import numpy as np

from .api import OK, BecError, CalcParams, check, lib

_P = C.c_void_p

ONE_LEVEL = 40


class Map2(C.Structure):
    """Two-level index map: off(i) = (i >> lg) * hi + (i & (2^lg - 1)) * lo (elements)."""
    _fields_ = [("hi", C.c_longlong), ("lo", C.c_longlong), ("lg", C.c_int), ("pad_", C.c_int)]


def m1(stride):
    return Map2(0, stride, ONE_LEVEL, 0)


def m2(lo_count, lo_stride, hi_stride):
    lg = int(lo_count).bit_length() - 1
    if 1 << lg != lo_count:
        raise ValueError("lo_count must be a power of two")
    return Map2(hi_stride, lo_stride, lg, 0)


class GemmDesc(C.Structure):
    _fields_ = [("A", _P), ("B", _P), ("C", _P), ("bias", _P), ("work", _P),
                ("a_bf16", C.c_int), ("b_bf16", C.c_int),
                ("M", C.c_int), ("N", C.c_int), ("K", C.c_int), ("batch", C.c_int), ("ksplit", C.c_int),
                ("alpha", C.c_float), ("beta", C.c_float),
                ("am", Map2), ("ak", Map2), ("bk", Map2), ("bn", Map2), ("cm", Map2), ("cn", Map2),
                ("ab", Map2), ("bb", Map2), ("cb", Map2)]


class FnoConfig(C.Structure):
    _fields_ = [("nz", C.c_int), ("ny", C.c_int), ("nx", C.c_int), ("mz", C.c_int), ("my", C.c_int),
                ("mx", C.c_int), ("width", C.c_int), ("layers", C.c_int), ("c_in", C.c_int), ("c_out", C.c_int),
                ("hidden", C.c_int), ("max_batch", C.c_int)]


EXPORTS = {
    "bec_gemm": (C.c_int, [C.POINTER(GemmDesc), _P]),
    "bec_gemm_time": (C.c_int, [C.POINTER(GemmDesc), C.c_int, C.POINTER(C.c_double)]),
    "bec_device_synchronize": (C.c_int, []),
    "bec_fno_create": (C.c_int, [C.POINTER(FnoConfig), C.c_uint64, C.POINTER(_P)]),
    "bec_fno_destroy": (C.c_int, [_P]),
    "bec_fno_param_count": (C.c_longlong, [_P]),
    "bec_fno_get_params": (C.c_int, [_P, _P]),
    "bec_fno_set_params": (C.c_int, [_P, _P]),
    "bec_fno_forward": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "bec_fno_gs2_features": (C.c_int, [C.POINTER(CalcParams), C.POINTER(C.c_float), _P, _P]),
    "bec_fno_loss_backward": (C.c_int, [_P, C.c_int, _P, _P, C.POINTER(C.c_double), _P]),
    "bec_fno_get_grads": (C.c_int, [_P, _P]),
    "bec_fno_adam_step": (C.c_int, [_P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, _P]),
}

_BOUND = False


def flib():
    global _BOUND
    L = lib()
    if not _BOUND:
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _BOUND = True
    return L


class DeviceArray:
    """A device allocation with its host shape and dtype (float32 or uint16 for BF16)."""

    def __init__(self, shape, dtype=np.float32, host=None):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        check(lib().bec_device_alloc(C.byref(self.ptr), max(self.nbytes, 16)), "bec_device_alloc")
        if host is not None:
            self.upload(host)

    def upload(self, host):
        h = np.ascontiguousarray(host, dtype=self.dtype)
        if h.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(lib().bec_memcpy_h2d(self.ptr, h.ctypes.data_as(_P), h.nbytes), "bec_memcpy_h2d")

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().bec_memcpy_d2h(out.ctypes.data_as(_P), self.ptr, self.nbytes), "bec_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().bec_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_bf16_bits(x):
    """float32 -> BF16 bit patterns, round to nearest even (the kernels' rounding)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


def gemm(A, B, Cout, M, N, K, am, ak, bk, bn, cm, cn, batch=1, ab=None, bb=None, cb=None, alpha=1.0, beta=0.0,
         bias=None, ksplit=1, work=None, stream=None, time_iters=0):
    """C = alpha A B + beta C (+ bias) through bec_gemm; with time_iters > 0, bec_gemm_time
    instead, returning the mean milliseconds of time_iters launches."""
    d = GemmDesc()
    d.A, d.B, d.C = A.ptr, B.ptr, Cout.ptr
    d.bias = bias.ptr if bias is not None else None
    d.work = work.ptr if work is not None else None
    d.a_bf16 = 1 if A.dtype == np.uint16 else 0
    d.b_bf16 = 1 if B.dtype == np.uint16 else 0
    d.M, d.N, d.K, d.batch, d.ksplit = M, N, K, batch, ksplit
    d.alpha, d.beta = alpha, beta
    d.am, d.ak, d.bk, d.bn, d.cm, d.cn = am, ak, bk, bn, cm, cn
    d.ab = ab if ab is not None else m1(0)
    d.bb = bb if bb is not None else m1(0)
    d.cb = cb if cb is not None else m1(0)
    if time_iters > 0:
        ms = C.c_double()
        st = flib().bec_gemm_time(C.byref(d), int(time_iters), C.byref(ms))
        if st != OK:
            raise BecError(st, "bec_gemm_time")
        return ms.value
    st = flib().bec_gemm(C.byref(d), stream)
    if st != OK:
        raise BecError(st, "bec_gemm")


class FNO:
    """The surrogate network on the GPU (bec_fno_*).  cfg: dict of bec_fno_config fields."""

    def __init__(self, seed=0, **cfg):
        self.cfg = FnoConfig(**cfg)
        self._h = C.c_void_p()
        check(flib().bec_fno_create(C.byref(self.cfg), seed, C.byref(self._h)), "bec_fno_create")
        self.n_params = int(flib().bec_fno_param_count(self._h))
        self.cells = self.cfg.nz * self.cfg.ny * self.cfg.nx

    def close(self):
        if self._h:
            flib().bec_fno_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get_params(self):
        out = np.empty(self.n_params, dtype=np.float32)
        check(flib().bec_fno_get_params(self._h, out.ctypes.data_as(_P)), "bec_fno_get_params")
        return out

    def set_params(self, flat):
        a = np.ascontiguousarray(flat, dtype=np.float32)
        if a.size != self.n_params:
            raise ValueError("parameter count mismatch")
        check(flib().bec_fno_set_params(self._h, a.ctypes.data_as(_P)), "bec_fno_set_params")

    def get_grads(self):
        out = np.empty(self.n_params, dtype=np.float32)
        check(flib().bec_fno_get_grads(self._h, out.ctypes.data_as(_P)), "bec_fno_get_grads")
        return out

    def forward(self, d_x, d_y, batch, stream=None):
        check(flib().bec_fno_forward(self._h, batch, d_x.ptr, d_y.ptr, stream), "bec_fno_forward")

    def loss_backward(self, d_x, d_t, batch, stream=None):
        loss = C.c_double()
        check(flib().bec_fno_loss_backward(self._h, batch, d_x.ptr, d_t.ptr, C.byref(loss), stream),
              "bec_fno_loss_backward")
        return loss.value

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0, stream=None):
        check(flib().bec_fno_adam_step(self._h, lr, beta1, beta2, eps, weight_decay, stream), "bec_fno_adam_step")

    def predict(self, x):
        """Host in, host out: x [batch][cells][c_in] -> [batch][cells][c_out]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        b = x.shape[0]
        dx = DeviceArray(x.shape, np.float32, x)
        dy = DeviceArray((b, self.cells, self.cfg.c_out), np.float32)
        self.forward(dx, dy, b)
        return dy.download()


def device_synchronize():
    """Wait for the device (bec_device_synchronize)."""
    st = flib().bec_device_synchronize()
    if st != OK:
        raise BecError(st, "bec_device_synchronize")


def gs2_features(calc, scaled, d_x, stream=None):
    """The learned start's input features of one parameter set, written into the
    device buffer d_x ([cells][16] FP32) by bec_fno_gs2_features."""
    need = int(calc.cells) * 16 * 4
    if isinstance(d_x, DeviceView):
        nbytes = d_x.base.nbytes - (d_x.ptr.value - d_x.base.ptr.value)
    elif isinstance(d_x, DeviceArray):
        nbytes = d_x.nbytes
    else:
        raise ValueError("d_x must be a DeviceArray or DeviceView")
    if nbytes < need:
        raise ValueError(f"d_x holds {nbytes} bytes; the features of {calc.cells} cells need {need}")
    if len(scaled) != 4:
        raise ValueError("scaled must hold the four scaled parameters")
    s = (C.c_float * 4)(*[float(v) for v in scaled])
    check(flib().bec_fno_gs2_features(C.byref(calc), s, d_x.ptr, stream), "bec_fno_gs2_features")


class DeviceView:
    """A byte offset into a DeviceArray (no ownership), usable wherever .ptr is read."""

    def __init__(self, base, offset_bytes):
        if offset_bytes < 0 or offset_bytes >= base.nbytes:
            raise ValueError("view outside the allocation")
        self.base = base
        self.ptr = C.c_void_p(base.ptr.value + int(offset_bytes))


def main(argv=None):
    """Times the forward pass of the production network (python -m bec_amd.fno):
    host wall clock over --iters back-to-back passes on one set of device buffers."""
    import argparse
    import json
    import time

    from . import surrogate as S
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--batch", type=int, default=1)
    a = ap.parse_args(argv)
    net = FNO(seed=7, max_batch=max(16, a.batch), **S.NET)
    th = S.sample_params(a.batch, 3)
    x = np.stack([S.features(S.calc_for(t), t) for t in th])
    dx = DeviceArray(x.shape, np.float32, x)
    dy = DeviceArray((a.batch, net.cells, net.cfg.c_out), np.float32)
    for _ in range(5):
        net.forward(dx, dy, a.batch)
    dy.download()
    t0 = time.perf_counter()
    for _ in range(a.iters):
        net.forward(dx, dy, a.batch)
    dy.download()
    ms = (time.perf_counter() - t0) / a.iters * 1e3
    print(json.dumps({"batch": a.batch, "iters": a.iters, "ms_per_forward": ms}), flush=True)
    net.close()


if __name__ == "__main__":
    main()
