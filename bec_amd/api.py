"""ctypes binding of libbec_amd.so (the C ABI declared in include/bec_api.h).

The product path always runs the HIP kernels: if the shared library is
missing this module raises instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbec_amd.so")

OK = 0
ERR_INVALID, ERR_SHAPE, ERR_ALLOC, ERR_HIP, ERR_NOT_CONVERGED, ERR_NONFINITE, ERR_STATE, ERR_UNSUPPORTED = range(1, 9)
FP32, FP64 = 0, 1
SPACE_X, SPACE_K, SPACE_HYBRID = 0, 1, 2
PULSE_HALF_PI, PULSE_PI = 0, 1
FFT_FORWARD, FFT_INVERSE = -1, 1
GS_SPLIT_STEP, GS_GRADIENT = 0, 1
GS2_DEFAULT_TAU, GS2_DEFAULT_BETA = 1.0, 0.85   # the coupled flow's step and momentum (bec_api.h)
GS_RULE_REFERENCE, GS_RULE_RESIDUAL = 0, 1


class BecError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        msg = lib().bec_strerror(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{where}: {msg} (status {status})")


class ModelParams(C.Structure):
    """ModelParameters (programs/bec_wigner_3d.cuda/defines.h:25-54)."""
    _fields_ = [("a11", C.c_double), ("a12", C.c_double), ("a22", C.c_double), ("m", C.c_double),
                ("fx", C.c_double), ("fy", C.c_double), ("fz", C.c_double), ("N", C.c_int),
                ("detuning", C.c_double), ("gamma111", C.c_double), ("gamma12", C.c_double),
                ("gamma22", C.c_double), ("Va", C.c_double), ("Vb", C.c_double),
                ("nvx", C.c_int), ("nvy", C.c_int), ("nvz", C.c_int), ("itmax", C.c_int),
                ("dt_steady", C.c_double), ("t_equilib", C.c_double), ("dt_evo", C.c_double),
                ("ne", C.c_int)]


class CalcParams(C.Structure):
    """CalculationParameters (programs/bec_wigner_3d.cuda/defines.h:57-112)."""
    _fields_ = [("N", C.c_int), ("g11", C.c_double), ("g12", C.c_double), ("g22", C.c_double),
                ("mu", C.c_double), ("cells", C.c_int), ("Va", C.c_double), ("Vb", C.c_double),
                ("V", C.c_double), ("lambda_", C.c_double), ("l_rho", C.c_double), ("t_rho", C.c_double),
                ("detuning", C.c_double), ("l111", C.c_double), ("l12", C.c_double), ("l22", C.c_double),
                ("nvx", C.c_int), ("nvy", C.c_int), ("nvz", C.c_int),
                ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("dkx", C.c_double), ("dky", C.c_double), ("dkz", C.c_double),
                ("itmax", C.c_int), ("dt_steady", C.c_double), ("t_equilib", C.c_double),
                ("dt_evo", C.c_double), ("ne", C.c_int)]


class GsOptions(C.Structure):
    _fields_ = [("method", C.c_int), ("rule", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int),
                ("check_every", C.c_int), ("precision", C.c_int), ("alpha", C.c_double),
                ("beta", C.c_double), ("tau", C.c_double)]


class GsReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("E_ref", C.c_double),
                ("mu_ref", C.c_double), ("E", C.c_double), ("mu", C.c_double), ("N", C.c_double),
                ("residual", C.c_double), ("elapsed_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Gs2Report(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("E", C.c_double), ("mu_a", C.c_double),
                ("mu_b", C.c_double), ("N_a", C.c_double), ("N_b", C.c_double), ("residual", C.c_double),
                ("elapsed_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Params2D(C.Structure):
    """bec_2d_params: the file-scope constants of programs/bec_wigner_2d.cpp/wigner.cpp:106-136."""
    _fields_ = [("nvx", C.c_int), ("nvy", C.c_int), ("xmax", C.c_double), ("ymax", C.c_double),
                ("g11", C.c_double), ("g12", C.c_double), ("g22", C.c_double),
                ("E", C.c_double), ("mu", C.c_double), ("Va", C.c_double), ("Vb", C.c_double),
                ("dtG", C.c_double), ("ntG", C.c_int), ("itmax", C.c_int),
                ("dtW", C.c_double), ("ntW", C.c_int), ("n_meas", C.c_int), ("ne", C.c_int),
                ("n0", C.c_double), ("seed", C.c_uint64), ("noise_dev", C.c_double),
                ("gs_rule", C.c_int), ("gs_tol", C.c_double), ("gs_max_iter", C.c_int),
                ("n_atoms", C.c_double), ("gs_tau", C.c_double), ("gs_beta", C.c_double)]


class Report2D(C.Structure):
    _fields_ = [("N0", C.c_double), ("gs_iterations", C.c_int), ("gs_residual", C.c_double),
                ("gs_mu", C.c_double), ("elapsed_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Params1D(C.Structure):
    """bec_1d_params: the file-scope constants of programs/bec_wigner_2d.cpp/wigner_1d.cpp:93-130."""
    _fields_ = [("zmax", C.c_double), ("dz", C.c_double),
                ("g11", C.c_double), ("g12", C.c_double), ("g22", C.c_double),
                ("E", C.c_double), ("mu", C.c_double), ("Va", C.c_double), ("Vb", C.c_double),
                ("dt", C.c_double), ("ntG", C.c_int), ("itmax", C.c_int), ("ntW", C.c_int), ("ne", C.c_int),
                ("n0", C.c_double), ("seed", C.c_uint64), ("noise_dev", C.c_double)]


class Report1D(C.Structure):
    _fields_ = [("N0", C.c_double), ("nv", C.c_int), ("gs_residual", C.c_double), ("gs_mu", C.c_double),
                ("elapsed_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class AvgReport(C.Structure):
    _fields_ = [("N", C.c_double), ("norm", C.c_double), ("var_re_a", C.c_double), ("var_im_a", C.c_double),
                ("var_re_b", C.c_double), ("var_im_b", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Obs(C.Structure):
    _fields_ = [("t", C.c_double), ("Na", C.c_double), ("Nb", C.c_double), ("N_total", C.c_double),
                ("visibility_grid", C.c_double), ("visibility", C.c_double), ("phase", C.c_double),
                ("N1", C.c_double), ("N2", C.c_double), ("Na_stderr", C.c_double), ("Nb_stderr", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Schedule(C.Structure):
    _fields_ = [("pw", C.c_int), ("z_r1", C.c_int), ("z_split", C.c_int), ("nt", C.c_int), ("xcd", C.c_int),
                ("xy_minw", C.c_int), ("z_half", C.c_int), ("xy_split", C.c_int), ("ms_per_step", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_P = C.c_void_p
_D = C.POINTER(C.c_double)

EXPORTS = {
    "bec_default_model_params": (C.c_int, [C.POINTER(ModelParams)]),
    "bec_fill_derived_params": (C.c_int, [C.POINTER(ModelParams), C.POINTER(CalcParams)]),
    "bec_create": (C.c_int, [C.POINTER(CalcParams), C.c_int, _P, C.POINTER(_P)]),
    "bec_create_groups": (C.c_int, [C.POINTER(CalcParams), C.c_int, C.c_int, _P, C.POINTER(_P)]),
    "bec_destroy": (C.c_int, [_P]),
    "bec_strerror": (C.c_char_p, [C.c_int]),
    "bec_synchronize": (C.c_int, [_P]),
    "bec_group_count": (C.c_int, [_P]),
    "bec_default_gs_options": (C.c_int, [C.POINTER(GsOptions)]),
    "bec_calculate_steady_state": (C.c_int, [_P, C.c_int, _P, C.POINTER(GsOptions), _P, C.POINTER(GsReport)]),
    "bec_ground_states_batched": (C.c_int, [C.POINTER(CalcParams), C.c_int, _P, C.POINTER(GsOptions), _P, _P,
                                            C.POINTER(GsReport)]),
    "bec_ground_states_two_component": (C.c_int, [C.POINTER(CalcParams), C.c_int, _P, _P, C.POINTER(GsOptions),
                                                  _P, _P, C.POINTER(Gs2Report)]),
    "bec_ground_states_two_component_net": (C.c_int, [C.POINTER(CalcParams), C.c_int, _P, _P, C.c_double,
                                                      C.POINTER(GsOptions), _P, _P, C.POINTER(Gs2Report)]),
    "bec_state_residual": (C.c_int, [_P, C.c_int, _P, _D, _D]),
    "bec_run_2d": (C.c_int, [C.POINTER(Params2D), _P, _P, _P, _P, _P, C.POINTER(Report2D)]),
    "bec_fill_2d_params": (C.c_int, [C.POINTER(ModelParams), C.POINTER(Params2D)]),
    "bec_1d_cells": (C.c_int, [C.POINTER(Params1D)]),
    "bec_run_1d": (C.c_int, [C.POINTER(Params1D), _P, _P, _P, _P, _P, _P, C.POINTER(Report1D)]),
    "bec_init_evolution": (C.c_int, [_P, C.c_int, _P, C.c_uint64, C.c_int]),
    "bec_propagate": (C.c_int, [_P, C.c_int, C.c_double, C.c_int, _P]),
    "bec_propagate_graph": (C.c_int, [_P, C.c_int, C.c_double, C.c_int]),
    "bec_propagate_prepare": (C.c_int, [_P, C.c_int, C.c_double, C.c_int]),
    "bec_observables": (C.c_int, [_P, C.c_int, C.POINTER(Obs)]),
    "bec_calculate_evolution": (C.c_int, [_P, C.c_double, C.POINTER(Obs)]),
    "bec_trajectory_sums": (C.c_int, [_P, C.c_int, _P]),
    "bec_projections": (C.c_int, [_P, C.c_int, C.c_double, _P, _P, _P, _P, _P, _P]),
    "bec_projections_static": (C.c_int, [_P, C.c_int, C.c_double, _P, _P, _P, _P, _P, _P]),
    "bec_pulse": (C.c_int, [_P, C.c_int]),
    "bec_pulse_rotation": (C.c_int, [_P, C.c_double, C.c_double]),
    "bec_get_state": (C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    "bec_set_state": (C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    "bec_get_time": (C.c_double, [_P]),
    "bec_set_time": (C.c_int, [_P, C.c_double]),
    "bec_time_passes": (C.c_int, [_P, C.c_int, C.c_int, _P]),
    "bec_fft_plan3d": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bec_fft_execute": (C.c_int, [_P, _P, _P, C.c_int, _P]),
    "bec_fft_destroy": (C.c_int, [_P]),
    "bec_device_alloc": (C.c_int, [C.POINTER(_P), C.c_size_t]),
    "bec_device_free": (C.c_int, [_P]),
    "bec_memcpy_h2d": (C.c_int, [_P, _P, C.c_size_t]),
    "bec_memcpy_d2h": (C.c_int, [_P, _P, C.c_size_t]),
    "bec_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "bec_set_graphs": (C.c_int, [_P, C.c_int]),
    "bec_set_multi_launch": (C.c_int, [_P, C.c_int]),
    "bec_set_loss_noise": (C.c_int, [_P, C.c_int, C.c_uint64, C.c_uint]),
    "bec_calculate_averages": (C.c_int, [_P, C.c_int, C.POINTER(AvgReport)]),
    "bec_get_loss_noise_step": (C.c_int, [_P, C.c_int, C.POINTER(C.c_uint)]),
    "bec_set_schedule": (C.c_int, [_P, C.c_int, C.POINTER(Schedule)]),
    "bec_get_schedule": (C.c_int, [_P, C.c_int, C.POINTER(Schedule)]),
    "bec_schedule_candidates": (C.c_int, [_P, C.c_int, C.POINTER(Schedule), C.c_int, C.POINTER(C.c_int)]),
    "bec_autotune": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(Schedule), C.c_int, C.POINTER(C.c_int),
                               C.POINTER(Schedule)]),
    "bec_device_copy": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "bec_time_copy": (C.c_int, [C.c_size_t, C.c_int, _D]),
}

_LIB = None


def lib():
    """Load libbec_amd.so; raise if it is not built (no fallback path exists)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m bec_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(status, where):
    if status != OK:
        raise BecError(status, where)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def default_model(**kw):
    m = ModelParams()
    check(lib().bec_default_model_params(C.byref(m)), "bec_default_model_params")
    for k, v in kw.items():
        if not hasattr(m, k):
            raise KeyError(k)
        setattr(m, k, v)
    return m


def derive(model):
    c = CalcParams()
    check(lib().bec_fill_derived_params(C.byref(model), C.byref(c)), "bec_fill_derived_params")
    return c


def gs_options(**kw):
    o = GsOptions()
    check(lib().bec_default_gs_options(C.byref(o)), "bec_default_gs_options")
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def time_copy(nbytes, iters=10):
    """Event-timed device copy of nbytes (read + write each nbytes); returns ms per copy."""
    ms = C.c_double()
    check(lib().bec_time_copy(nbytes, iters, C.byref(ms)), "bec_time_copy")
    return ms.value


def device_count():
    n = C.c_int(0)
    st = lib().bec_device_count(C.byref(n))
    return n.value if st == OK else 0


def ground_states_batched(params, guesses=None, **opt):
    """Batched imaginary-time ground states for a list of CalcParams sharing one lattice."""
    L = lib()
    B = len(params)
    arr = (CalcParams * B)(*params)
    o = gs_options(**opt)
    cells = params[0].cells
    out = np.empty((B, params[0].nvz, params[0].nvy, params[0].nvx), dtype=np.complex128)
    reps = (GsReport * B)()
    g = None if guesses is None else _c128(guesses)
    st = L.bec_ground_states_batched(arr, B, None if g is None else _ptr(g), C.byref(o), None, _ptr(out), reps)
    if st not in (OK, ERR_NOT_CONVERGED):
        check(st, "bec_ground_states_batched")
    assert out.size == B * cells
    return out, [r.as_dict() for r in reps], st


def ground_states_two_component(params, frac_a=None, guesses=None, net=None, spin_weight=1.0, **opt):
    """Two-component ground states (coupled imaginary-time GPEs, gradient flow) for a
    list of CalcParams sharing one lattice.  Returns (out [B][2][nz][ny][nx], reports, status).
    net: a device buffer (an object with .ptr) holding the surrogate network's output
    [B][cells][2] for the start, converted on the GPU (bec_ground_states_two_component_net)."""
    L = lib()
    B = len(params)
    arr = (CalcParams * B)(*params)
    opt.setdefault("method", GS_GRADIENT)
    opt.setdefault("rule", GS_RULE_RESIDUAL)
    opt.setdefault("tau", GS2_DEFAULT_TAU)
    opt.setdefault("beta", GS2_DEFAULT_BETA)
    o = gs_options(**opt)
    p0 = params[0]
    out = np.empty((B, 2, p0.nvz, p0.nvy, p0.nvx), dtype=np.complex128)
    reps = (Gs2Report * B)()
    fr = None if frac_a is None else np.ascontiguousarray(np.broadcast_to(frac_a, (B,)), dtype=np.float64)
    g = None if guesses is None else _c128(guesses)
    if g is not None and g.size != out.size:
        raise ValueError("guesses must hold batch * 2 * cells values")
    if net is not None:
        if g is not None:
            raise ValueError("give guesses or net, not both")
        st = L.bec_ground_states_two_component_net(arr, B, None if fr is None else _ptr(fr), net.ptr,
                                                   float(spin_weight), C.byref(o), None, _ptr(out), reps)
    else:
        st = L.bec_ground_states_two_component(arr, B, None if fr is None else _ptr(fr),
                                               None if g is None else _ptr(g), C.byref(o), None, _ptr(out), reps)
    if st not in (OK, ERR_NOT_CONVERGED):
        check(st, "bec_ground_states_two_component")
    return out, [r.as_dict() for r in reps], st


def params2d(**kw):
    """The 2D program's constants (programs/bec_wigner_2d.cpp/wigner.cpp:106-136) on a 64 x 64 lattice
    (its 128 x 128 does not fit one workgroup); keyword arguments override."""
    d = dict(nvx=64, nvy=64, xmax=20.0, ymax=20.0, g11=6.18, g12=6.03, g22=5.88, E=2.5, mu=26.8, Va=0.2, Vb=0.5,
             dtG=1.0 / 31.0, ntG=32, itmax=3, dtW=0.01, ntW=200, n_meas=1, ne=32, n0=1.0, seed=0x5EED,
             noise_dev=None, gs_rule=0, gs_tol=1e-8, gs_max_iter=100000, n_atoms=0.0, gs_tau=0.5, gs_beta=0.6)
    d.update(kw)
    if d["noise_dev"] is None:
        d["noise_dev"] = 1.0 / np.sqrt((2.0 * d["xmax"] / (d["nvx"] - 1)) * (2.0 * d["ymax"] / (d["nvy"] - 1)))
    return Params2D(**d)


def params2d_from_model(model, **kw):
    """The 2D program's parameters for a two-component cloud given in SI units
    (bec_fill_2d_params): the radial plane of the trap, the axial direction in its
    oscillator ground state; keyword arguments override the result's fields."""
    p = Params2D()
    check(lib().bec_fill_2d_params(C.byref(model), C.byref(p)), "bec_fill_2d_params")
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def run_2d(p, want_avn=True):
    """The 2D program on the GPU: ground state, then ne trajectories (bec_run_2d)."""
    nm = p.n_meas
    m = np.zeros(2 * nm)
    var = np.zeros(2 * nm)
    vis = np.zeros(nm)
    gs = np.empty((p.nvy, p.nvx), dtype=np.complex128)
    avn = np.empty((p.nvy, p.nvx)) if want_avn else None
    rep = Report2D()
    check(lib().bec_run_2d(C.byref(p), _ptr(m), _ptr(var), _ptr(vis), _ptr(gs),
                           _ptr(avn) if want_avn else None, C.byref(rep)), "bec_run_2d")
    return dict(m=m.reshape(2, nm), var=var.reshape(2, nm), vis=vis, gs=gs, avn=avn, report=rep.as_dict())


def params1d(**kw):
    """The 1D program's constants (programs/bec_wigner_2d.cpp/wigner_1d.cpp:93-130); keyword arguments
    override.  tmaxGP / tmaxWig give ntG = 1 + (int)(tmaxGP/dt) and ntW = 1 + (int)(tmaxWig/dt) as the
    program derives them."""
    d = dict(zmax=20.0, dz=1.0, g11=6.18, g12=6.03, g22=5.88, E=2.5, mu=26.8, Va=0.5, Vb=1.0, dt=0.01, ntG=None,
             itmax=3, ntW=None, ne=20, n0=1.0, seed=0x5EED, noise_dev=None, tmaxGP=1.0, tmaxWig=30.0)
    d.update(kw)
    tG, tW = d.pop("tmaxGP"), d.pop("tmaxWig")
    if d["ntG"] is None:
        d["ntG"] = 1 + int(tG / d["dt"])
    if d["ntW"] is None:
        d["ntW"] = 1 + int(tW / d["dt"])
    if d["noise_dev"] is None:
        d["noise_dev"] = np.sqrt(1.0 / d["dz"]) / 2.0
    return Params1D(**d)


def cells_1d(p):
    """Lattice cells of the 1D program's parameters (bec_1d_cells)."""
    n = lib().bec_1d_cells(C.byref(p))
    check(-n if n < 0 else OK, "bec_1d_cells")
    return n


def run_1d(p):
    """The 1D program on the GPU: ground state, then ne trajectories (bec_run_1d)."""
    nv, nt = cells_1d(p), p.ntW
    m = np.zeros(2 * nt)
    var = np.zeros(2 * nt)
    vis = np.zeros(nt)
    gs = np.empty(nv, dtype=np.complex128)
    avn = np.empty(nv)
    ntot = np.empty(p.ntG)
    rep = Report1D()
    check(lib().bec_run_1d(C.byref(p), _ptr(m), _ptr(var), _ptr(vis), _ptr(gs), _ptr(avn), _ptr(ntot),
                           C.byref(rep)), "bec_run_1d")
    return dict(m=m.reshape(2, nt), var=var.reshape(2, nt), vis=vis, gs=gs, avn=avn, ntot=ntot,
                report=rep.as_dict())


class Solver:
    """Host-side mirror of the reference driver (programs/bec_wigner_3d.cuda/wigner.cpp:313-369):
    calculateSteadyState -> initEvolution -> propagate / calculateEvolution -> observables."""

    def __init__(self, calc, precision=FP32, groups=None):
        L = lib()
        self._ctx = C.c_void_p()
        if groups is None:
            groups = [calc]
        self.groups = list(groups)
        arr = (CalcParams * len(self.groups))(*self.groups)
        check(L.bec_create_groups(arr, len(self.groups), precision, None, C.byref(self._ctx)), "bec_create")
        self.precision = precision

    def close(self):
        if self._ctx:
            lib().bec_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def shape(self, g=0):
        c = self.groups[g]
        return (c.ne, c.nvz, c.nvy, c.nvx)

    def calculate_steady_state(self, group=0, guess=None, **opt):
        c = self.groups[group]
        o = gs_options(**opt)
        out = np.empty((c.nvz, c.nvy, c.nvx), dtype=np.complex128)
        rep = GsReport()
        g = None if guess is None else _c128(guess)
        st = lib().bec_calculate_steady_state(self._ctx, group, None if g is None else _ptr(g), C.byref(o),
                                               _ptr(out), C.byref(rep))
        if st not in (OK, ERR_NOT_CONVERGED):
            check(st, "bec_calculate_steady_state")
        return out, rep.as_dict(), st

    def state_residual(self, psi, group=0):
        r, mu = C.c_double(), C.c_double()
        check(lib().bec_state_residual(self._ctx, group, _ptr(_c128(psi)), C.byref(r), C.byref(mu)),
              "bec_state_residual")
        return r.value, mu.value

    def init_evolution(self, gs, seed, group=0, equilibrate=True):
        check(lib().bec_init_evolution(self._ctx, group, _ptr(_c128(gs)), seed, 1 if equilibrate else 0),
              "bec_init_evolution")

    def propagate(self, n_steps, dt=0.0, sample_every=0):
        ntot = sum(c.ne for c in self.groups)
        ns = n_steps // sample_every if sample_every > 0 else 0
        buf = np.empty((max(ns, 1), ntot, 4), dtype=np.float64)
        check(lib().bec_propagate(self._ctx, n_steps, dt, sample_every, _ptr(buf) if ns else None),
              "bec_propagate")
        return buf[:ns]

    def prepare(self, n_steps, sample_every=0, dt=0.0):
        """Capture the step graphs that propagate(n_steps, dt, sample_every) replays (bec_propagate_prepare)."""
        check(lib().bec_propagate_prepare(self._ctx, n_steps, dt, sample_every), "bec_propagate_prepare")

    def propagate_graph(self, n_steps, steps_per_graph, dt=0.0):
        check(lib().bec_propagate_graph(self._ctx, n_steps, dt, steps_per_graph), "bec_propagate_graph")

    def calculate_evolution(self, dt=0.0):
        obs = (Obs * len(self.groups))()
        check(lib().bec_calculate_evolution(self._ctx, dt, obs), "bec_calculate_evolution")
        return [o.as_dict() for o in obs]

    def observables(self, group=0):
        o = Obs()
        check(lib().bec_observables(self._ctx, group, C.byref(o)), "bec_observables")
        return o.as_dict()

    def trajectory_sums(self, group=0):
        out = np.empty((self.groups[group].ne, 4), dtype=np.float64)
        check(lib().bec_trajectory_sums(self._ctx, group, _ptr(out)), "bec_trajectory_sums")
        return out

    def projections(self, alpha=0.0, group=0, static=False):
        """Density projections (bec_projections; static=True: bec_projections_static,
        which allocates no device memory)."""
        c = self.groups[group]
        xa = np.empty((c.nvy, c.nvx)); xb = np.empty_like(xa)
        za = np.empty((c.nvy, c.nvz)); zb = np.empty_like(za)
        aa = np.empty(c.nvz); ab = np.empty_like(aa)
        name = "bec_projections_static" if static else "bec_projections"
        check(getattr(lib(), name)(self._ctx, group, alpha, *[_ptr(v) for v in (xa, xb, za, zb, aa, ab)]), name)
        return dict(xy_a=xa, xy_b=xb, zy_a=za, zy_b=zb, axial_a=aa, axial_b=ab)

    def pulse(self, kind=PULSE_HALF_PI):
        check(lib().bec_pulse(self._ctx, kind), "bec_pulse")

    def pulse_rotation(self, theta, phi=0.0):
        """A resonant pulse of area theta and phase phi on every group (bec_pulse_rotation)."""
        check(lib().bec_pulse_rotation(self._ctx, float(theta), float(phi)), "bec_pulse_rotation")

    def get_state(self, space=SPACE_X, group=0):
        a = np.empty(self.shape(group), dtype=np.complex128)
        b = np.empty_like(a)
        check(lib().bec_get_state(self._ctx, group, space, _ptr(a), _ptr(b)), "bec_get_state")
        return a, b

    def set_state(self, a, b, space=SPACE_X, group=0):
        a, b = _c128(a), _c128(b)
        if a.shape != self.shape(group) or b.shape != self.shape(group):
            raise ValueError(f"state arrays must have shape {self.shape(group)}, got {a.shape} and {b.shape}")
        check(lib().bec_set_state(self._ctx, group, space, _ptr(_c128(a)), _ptr(_c128(b))), "bec_set_state")

    @property
    def t(self):
        return lib().bec_get_time(self._ctx)

    @t.setter
    def t(self, v):
        check(lib().bec_set_time(self._ctx, v), "bec_set_time")

    def time_passes(self, iters=20, group=0):
        ms = np.zeros(3)
        check(lib().bec_time_passes(self._ctx, group, iters, _ptr(ms)), "bec_time_passes")
        return dict(z_pass=ms[0], xy_pass=ms[1], sample_pass=ms[2])

    def schedule(self, group=0):
        sc = Schedule()
        check(lib().bec_get_schedule(self._ctx, group, C.byref(sc)), "bec_get_schedule")
        return sc.as_dict()

    def set_schedule(self, group=0, **kw):
        sc = Schedule()
        check(lib().bec_get_schedule(self._ctx, group, C.byref(sc)), "bec_get_schedule")
        for k, v in kw.items():
            setattr(sc, k, v)
        check(lib().bec_set_schedule(self._ctx, group, C.byref(sc)), "bec_set_schedule")

    def schedule_candidates(self, group=0):
        arr = (Schedule * 64)()
        n = C.c_int(0)
        check(lib().bec_schedule_candidates(self._ctx, group, arr, 64, C.byref(n)), "bec_schedule_candidates")
        return [arr[i].as_dict() for i in range(min(n.value, 64))]

    def autotune(self, steps=20, group=0):
        """Time every candidate launch schedule on this context and keep the fastest."""
        arr = (Schedule * 64)()
        best = Schedule()
        n = C.c_int(0)
        check(lib().bec_autotune(self._ctx, group, steps, arr, 64, C.byref(n), C.byref(best)), "bec_autotune")
        return best.as_dict(), [arr[i].as_dict() for i in range(min(n.value, 64))]

    def set_graphs(self, enable):
        check(lib().bec_set_graphs(self._ctx, 1 if enable else 0), "bec_set_graphs")

    def set_multi_launch(self, enable=True):
        """Step every group in one launch per pass (bec_set_multi_launch)."""
        check(lib().bec_set_multi_launch(self._ctx, 1 if enable else 0), "bec_set_multi_launch")

    def set_loss_noise(self, enable=True, seed=0x1055, first_step=0):
        """Add (or remove) the loss-induced Wigner noise to every nonlinear step (bec_set_loss_noise)."""
        check(lib().bec_set_loss_noise(self._ctx, 1 if enable else 0, int(seed), int(first_step)),
              "bec_set_loss_noise")

    def averages(self, group=0):
        """calculateAverages (wigner.cu:305-337): atoms per trajectory, the scale and the ensemble
        variances of the four quadratures (bec_calculate_averages)."""
        r = AvgReport()
        check(lib().bec_calculate_averages(self._ctx, group, C.byref(r)), "bec_calculate_averages")
        return r.as_dict()

    def loss_noise_step(self, group=0):
        st = C.c_uint(0)
        check(lib().bec_get_loss_noise_step(self._ctx, group, C.byref(st)), "bec_get_loss_noise_step")
        return st.value

    def synchronize(self):
        check(lib().bec_synchronize(self._ctx), "bec_synchronize")


class FFTPlan:
    """batchfftPlan3d / batchfftExecute / batchfftDestroy (programs/bec_wigner_3d.cuda/batchfft.h:83-108)."""

    def __init__(self, nz, ny, nx, batch, precision=FP32):
        self._p = C.c_void_p()
        check(lib().bec_fft_plan3d(C.byref(self._p), nz, ny, nx, precision, batch), "bec_fft_plan3d")
        self.shape = (batch, nz, ny, nx)
        self.precision = precision

    def execute(self, data, sign):
        """Transform a host array through device memory (test helper)."""
        L = lib()
        dt = np.complex64 if self.precision == FP32 else np.complex128
        h = np.ascontiguousarray(data, dtype=dt).reshape(self.shape)
        d = C.c_void_p()
        check(L.bec_device_alloc(C.byref(d), h.nbytes), "bec_device_alloc")
        try:
            check(L.bec_memcpy_h2d(d, _ptr(h), h.nbytes), "bec_memcpy_h2d")
            check(L.bec_fft_execute(self._p, d, d, sign, None), "bec_fft_execute")
            out = np.empty_like(h)
            check(L.bec_memcpy_d2h(_ptr(out), d, h.nbytes), "bec_memcpy_d2h")
        finally:
            L.bec_device_free(d)
        return out

    def close(self):
        if self._p:
            lib().bec_fft_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
