"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests, the smoke check and the benchmark's CPU-baseline leg import this
module; the product library never does.  Arrays are numpy complex128 in the
reference layout (trajectory outermost, then z, y, x with x fastest).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Model(C.Structure):
    """ModelParameters (programs/bec_wigner_3d.cuda/defines.h:25-54)."""
    _fields_ = [("a11", C.c_double), ("a12", C.c_double), ("a22", C.c_double),
                ("m", C.c_double), ("fx", C.c_double), ("fy", C.c_double), ("fz", C.c_double),
                ("N", C.c_int), ("detuning", C.c_double),
                ("gamma111", C.c_double), ("gamma12", C.c_double), ("gamma22", C.c_double),
                ("Va", C.c_double), ("Vb", C.c_double),
                ("nvx", C.c_int), ("nvy", C.c_int), ("nvz", C.c_int), ("itmax", C.c_int),
                ("dt_steady", C.c_double), ("t_equilib", C.c_double), ("dt_evo", C.c_double),
                ("ne", C.c_int)]


class Calc(C.Structure):
    """CalculationParameters (programs/bec_wigner_3d.cuda/defines.h:57-112)."""
    _fields_ = [("N", C.c_int), ("g11", C.c_double), ("g12", C.c_double), ("g22", C.c_double),
                ("mu", C.c_double), ("cells", C.c_int), ("Va", C.c_double), ("Vb", C.c_double),
                ("V", C.c_double), ("lambda_", C.c_double), ("l_rho", C.c_double),
                ("t_rho", C.c_double), ("detuning", C.c_double), ("l111", C.c_double),
                ("l12", C.c_double), ("l22", C.c_double),
                ("nvx", C.c_int), ("nvy", C.c_int), ("nvz", C.c_int),
                ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("dkx", C.c_double), ("dky", C.c_double), ("dkz", C.c_double),
                ("itmax", C.c_int), ("dt_steady", C.c_double), ("t_equilib", C.c_double),
                ("dt_evo", C.c_double), ("ne", C.c_int)]


class GSReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("E_ref", C.c_double),
                ("mu_ref", C.c_double), ("E", C.c_double), ("mu", C.c_double),
                ("N", C.c_double), ("residual", C.c_double)]


class GS2Report(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("E", C.c_double),
                ("mu_a", C.c_double), ("mu_b", C.c_double), ("N_a", C.c_double), ("N_b", C.c_double),
                ("residual", C.c_double)]


class Params2D(C.Structure):
    """File-scope parameters of programs/bec_wigner_2d.cpp/wigner.cpp:106-136."""
    _fields_ = [("nvx", C.c_int), ("nvy", C.c_int), ("xmax", C.c_double), ("ymax", C.c_double),
                ("g11", C.c_double), ("g12", C.c_double), ("g22", C.c_double),
                ("E", C.c_double), ("mu", C.c_double), ("Va", C.c_double), ("Vb", C.c_double),
                ("dtG", C.c_double), ("ntG", C.c_int), ("itmax", C.c_int),
                ("dtW", C.c_double), ("ntW", C.c_int), ("n_meas", C.c_int), ("ne", C.c_int),
                ("n0", C.c_double), ("seed", C.c_uint64), ("noise_dev", C.c_double),
                ("gs_rule", C.c_int), ("gs_tol", C.c_double), ("gs_max_iter", C.c_int),
                ("n_atoms", C.c_double), ("gs_tau", C.c_double), ("gs_beta", C.c_double)]


class Report2D(C.Structure):
    _fields_ = [("N0", C.c_double), ("gs_iterations", C.c_int), ("gs_residual", C.c_double)]


class Params1D(C.Structure):
    """File-scope parameters of programs/bec_wigner_2d.cpp/wigner_1d.cpp:93-130."""
    _fields_ = [("zmax", C.c_double), ("dz", C.c_double),
                ("g11", C.c_double), ("g12", C.c_double), ("g22", C.c_double),
                ("E", C.c_double), ("mu", C.c_double), ("Va", C.c_double), ("Vb", C.c_double),
                ("dt", C.c_double), ("ntG", C.c_int), ("itmax", C.c_int), ("ntW", C.c_int), ("ne", C.c_int),
                ("n0", C.c_double), ("seed", C.c_uint64), ("noise_dev", C.c_double)]


class Report1D(C.Structure):
    _fields_ = [("N0", C.c_double), ("nv", C.c_int), ("gs_residual", C.c_double), ("gs_mu", C.c_double)]


RULE_REFERENCE = 0
RULE_RESIDUAL = 1

_P = C.c_void_p
_D = C.POINTER(C.c_double)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        sig = {
            "orc_derive": (None, [C.POINTER(Model), C.POINTER(Calc)]),
            "orc_ktable": (None, [C.POINTER(Calc), _P]),
            "orc_potential": (C.c_double, [C.POINTER(Calc), C.c_int]),
            "orc_fft1d": (None, [_P, C.c_int, C.c_int]),
            "orc_fft3d": (None, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_fft_axes": (None, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_dft3d_bruteforce": (None, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_ground_state": (C.c_int, [C.POINTER(Calc), _P, C.c_int, C.c_double, C.c_int,
                                           C.c_int, _P, C.POINTER(GSReport)]),
            "orc_tf_state": (None, [C.POINTER(Calc), _P]),
            "orc_state_integral_ref": (C.c_double, [C.POINTER(Calc), _P, C.c_int]),
            "orc_energy_mu": (None, [C.POINTER(Calc), _P, _D, _D, _D]),
            "orc_residual": (C.c_double, [C.POINTER(Calc), _P, _D]),
            "orc_gs_iterate": (None, [C.POINTER(Calc), _P, C.c_int, _D]),
            "orc_apply_h2": (None, [C.POINTER(Calc), _P, _P, _P, _P]),
            "orc_residual2": (C.c_double, [C.POINTER(Calc), _P, _P, _D, _D]),
            "orc_energy2": (C.c_double, [C.POINTER(Calc), _P, _P]),
            "orc_tf_state2": (None, [C.POINTER(Calc), C.c_double, _P, _P]),
            "orc_ground_state2": (C.c_int, [C.POINTER(Calc), C.c_double, _P, _P, C.c_double, C.c_int,
                                            C.c_double, C.c_double, C.c_double, _P, _P,
                                            C.POINTER(GS2Report)]),
            "orc_philox4x32_10": (None, [_P, _P, _P]),
            "orc_noise_cell": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
            "orc_init_ensembles": (None, [C.POINTER(Calc), _P, C.c_uint64, _P, _P]),
            "orc_equilibration_steps": (C.c_int, [C.POINTER(Calc)]),
            "orc_to_kspace": (None, [C.POINTER(Calc), _P, C.c_int]),
            "orc_to_xspace": (None, [C.POINTER(Calc), _P, C.c_int]),
            "orc_init_evolution": (None, [C.POINTER(Calc), _P, C.c_uint64, C.c_int, _P, _P]),
            "orc_propagate": (None, [C.POINTER(Calc), _P, _P, C.c_int, C.c_double, C.c_int]),
            "orc_rotation": (None, [_P, _P, C.c_long, C.c_double, C.c_double]),
            "orc_averages": (None, [C.POINTER(Calc), _P, _P, C.c_int, _P]),
            "orc_loss_noise_cell": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
            "orc_nonlinear_rt_noise": (None, [C.POINTER(Calc), _P, _P, C.c_int, C.c_double, C.c_uint64,
                                              C.c_uint32, C.c_uint32]),
            "orc_propagate_noise": (None, [C.POINTER(Calc), _P, _P, C.c_int, C.c_double, C.c_int, C.c_uint64,
                                           C.c_uint32, C.c_uint32]),
            "orc_kinetic_rt": (None, [C.POINTER(Calc), _P, _P, _P, C.c_int, C.c_double]),
            "orc_nonlinear_rt": (None, [C.POINTER(Calc), _P, _P, C.c_int, C.c_double]),
            "orc_half_pi_pulse": (None, [_P, _P, C.c_long]),
            "orc_pi_pulse": (None, [_P, _P, C.c_long]),
            "orc_rotate_sums": (None, [C.POINTER(Calc), _P, _P, C.c_int, C.c_double, _D, _D]),
            "orc_traj_sums": (None, [C.POINTER(Calc), _P, _P, C.c_int, _P]),
            "orc_cell_density_sums": (None, [C.POINTER(Calc), _P, _P, C.c_int, C.c_double, _P, _P]),
            "orc_visibility_grid": (C.c_double, [C.POINTER(Calc), _P, _P, C.c_int]),
            "orc2d_run": (C.c_int, [C.POINTER(Params2D), _P, _P, _P, _P, _P, C.POINTER(Report2D)]),
            "orc2d_residual": (C.c_double, [C.POINTER(Params2D), _P, _D]),
            "orc_dft1d": (None, [_P, C.c_int, C.c_int]),
            "orc1d_nv": (C.c_int, [C.POINTER(Params1D)]),
            "orc1d_run": (C.c_int, [C.POINTER(Params1D), _P, _P, _P, _P, _P, _P, C.POINTER(Report1D)]),
            "orc1d_residual": (C.c_double, [C.POINTER(Params1D), _P, _D]),
            "orc_num_threads": (C.c_int, []),
            "orc_set_threads": (None, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def model(**kw):
    """Build a Model with the reference defaults of wigner.cpp:43-98 overridden by kw."""
    d = dict(a11=100.4, a12=97.66, a22=95.0, m=1.443160648e-25, fx=97.6, fy=97.6, fz=11.96,
             N=150000, detuning=-41.0, gamma111=0.0, gamma12=0.0, gamma22=0.0, Va=1.0, Vb=1.0,
             nvx=16, nvy=16, nvz=128, itmax=1, dt_steady=1e-5, t_equilib=0.0, dt_evo=4e-5, ne=4)
    d.update(kw)
    return Model(**d)


def derive(m):
    c = Calc()
    lib().orc_derive(C.byref(m), C.byref(c))
    return c


def shape(c, ne=None):
    if ne is None:
        return (c.nvz, c.nvy, c.nvx)
    return (ne, c.nvz, c.nvy, c.nvx)


def ktable(c):
    out = np.empty(c.cells, dtype=np.float64)
    lib().orc_ktable(C.byref(c), _ptr(out))
    return out.reshape(shape(c))


def potential(c):
    L = lib()
    return np.array([L.orc_potential(C.byref(c), i) for i in range(c.cells)]).reshape(shape(c))


def fft3d(x, sign):
    x = _c128(x).copy()
    nz, ny, nx = x.shape[-3:]
    batch = int(np.prod(x.shape[:-3])) if x.ndim > 3 else 1
    lib().orc_fft3d(_ptr(x), nz, ny, nx, batch, sign)
    return x


def fft_axes(x, sign, mask):
    x = _c128(x).copy()
    nz, ny, nx = x.shape[-3:]
    batch = int(np.prod(x.shape[:-3])) if x.ndim > 3 else 1
    lib().orc_fft_axes(_ptr(x), nz, ny, nx, batch, sign, mask)
    return x


def fft1d(x, sign):
    x = _c128(x).copy()
    lib().orc_fft1d(_ptr(x), x.size, sign)
    return x


def dft3d_bruteforce(x, sign):
    x = _c128(x)
    out = np.empty_like(x)
    nz, ny, nx = x.shape
    lib().orc_dft3d_bruteforce(_ptr(x), _ptr(out), nz, ny, nx, sign)
    return out


def tf_state(c):
    out = np.empty(shape(c), dtype=np.complex128)
    lib().orc_tf_state(C.byref(c), _ptr(out))
    return out


def ground_state(c, guess=None, rule=RULE_REFERENCE, tol=1e-6, max_iter=100000, check_every=1):
    out = np.empty(shape(c), dtype=np.complex128)
    rep = GSReport()
    g = None if guess is None else _c128(guess)
    lib().orc_ground_state(C.byref(c), None if g is None else _ptr(g), rule, tol, max_iter,
                           check_every, _ptr(out), C.byref(rep))
    return out, rep


def gs_iterate(c, a_k, n_iter):
    a = _c128(a_k).copy()
    E = C.c_double()
    lib().orc_gs_iterate(C.byref(c), _ptr(a), n_iter, C.byref(E))
    return a, E.value


def state_integral_ref(c, psi_x, energy=True):
    return lib().orc_state_integral_ref(C.byref(c), _ptr(_c128(psi_x)), 1 if energy else 0)


def energy_mu(c, psi_x):
    E, mu, N = C.c_double(), C.c_double(), C.c_double()
    lib().orc_energy_mu(C.byref(c), _ptr(_c128(psi_x)), C.byref(E), C.byref(mu), C.byref(N))
    return E.value, mu.value, N.value


def residual(c, psi_x):
    mu = C.c_double()
    r = lib().orc_residual(C.byref(c), _ptr(_c128(psi_x)), C.byref(mu))
    return r, mu.value


def apply_h2(c, a, b):
    ha = np.empty(shape(c), dtype=np.complex128)
    hb = np.empty(shape(c), dtype=np.complex128)
    lib().orc_apply_h2(C.byref(c), _ptr(_c128(a)), _ptr(_c128(b)), _ptr(ha), _ptr(hb))
    return ha, hb


def residual2(c, a, b):
    ma, mb = C.c_double(), C.c_double()
    r = lib().orc_residual2(C.byref(c), _ptr(_c128(a)), _ptr(_c128(b)), C.byref(ma), C.byref(mb))
    return r, ma.value, mb.value


def energy2(c, a, b):
    return lib().orc_energy2(C.byref(c), _ptr(_c128(a)), _ptr(_c128(b)))


def tf_state2(c, frac_a=0.5):
    a = np.empty(shape(c), dtype=np.complex128)
    b = np.empty(shape(c), dtype=np.complex128)
    lib().orc_tf_state2(C.byref(c), frac_a, _ptr(a), _ptr(b))
    return a, b


def ground_state2(c, frac_a=0.5, guess=None, tol=1e-8, max_iter=5000, alpha=None, tau=1.0, beta=0.85):
    """Two-component ground state (gradient flow, with the library's default step and
    momentum of the coupled flow); returns (a, b, report)."""
    a = np.empty(shape(c), dtype=np.complex128)
    b = np.empty(shape(c), dtype=np.complex128)
    rep = GS2Report()
    ga = gb = None
    if guess is not None:
        ga, gb = _c128(guess[0]), _c128(guess[1])
    lib().orc_ground_state2(C.byref(c), frac_a, None if ga is None else _ptr(ga), None if gb is None else _ptr(gb),
                            tol, max_iter, c.mu if alpha is None else alpha, tau, beta, _ptr(a), _ptr(b),
                            C.byref(rep))
    return a, b, rep


def philox(ctr, key):
    ctr = np.asarray(ctr, dtype=np.uint32)
    key = np.asarray(key, dtype=np.uint32)
    out = np.empty(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(ctr), _ptr(key), _ptr(out))
    return out


def noise_cell(seed, traj, cell, stream=0):
    out = np.empty(4, dtype=np.float64)
    lib().orc_noise_cell(seed, traj, cell, stream, _ptr(out))
    return out


def init_ensembles(c, gs_x, seed):
    a = np.empty(shape(c, c.ne), dtype=np.complex128)
    b = np.empty_like(a)
    lib().orc_init_ensembles(C.byref(c), _ptr(_c128(gs_x)), seed, _ptr(a), _ptr(b))
    return a, b


def equilibration_steps(c):
    return lib().orc_equilibration_steps(C.byref(c))


def to_kspace(c, x):
    x = _c128(x).copy()
    ne = x.size // c.cells
    lib().orc_to_kspace(C.byref(c), _ptr(x), ne)
    return x


def to_xspace(c, k):
    k = _c128(k).copy()
    ne = k.size // c.cells
    lib().orc_to_xspace(C.byref(c), _ptr(k), ne)
    return k


def init_evolution(c, gs_x, seed, equilibrate=True):
    a = np.empty(shape(c, c.ne), dtype=np.complex128)
    b = np.empty_like(a)
    lib().orc_init_evolution(C.byref(c), _ptr(_c128(gs_x)), seed, 1 if equilibrate else 0,
                             _ptr(a), _ptr(b))
    return a, b


def propagate(c, a_k, b_k, dt, n_steps):
    a = _c128(a_k).copy()
    b = _c128(b_k).copy()
    ne = a.size // c.cells
    lib().orc_propagate(C.byref(c), _ptr(a), _ptr(b), ne, dt, n_steps)
    return a, b


def rotation(a, b, theta, phi=0.0):
    """A resonant pulse of area theta and phase phi (orc_rotation)."""
    a, b = _c128(a).copy(), _c128(b).copy()
    lib().orc_rotation(_ptr(a), _ptr(b), a.size, theta, phi)
    return a, b


def averages(c, a_x, b_x):
    """calculateAverages (wigner.cu:305-337) on x-space states."""
    a = _c128(a_x)
    out = np.zeros(6)
    lib().orc_averages(C.byref(c), _ptr(a), _ptr(_c128(b_x)), a.size // c.cells, _ptr(out))
    return dict(zip(("N", "norm", "var_re_a", "var_im_a", "var_re_b", "var_im_b"), out))


def propagate_noise(c, a_k, b_k, dt, n_steps, seed, step0=0, group=0):
    """propagate with the loss-induced noise increments of steps step0 .. step0 + n_steps - 1."""
    a = _c128(a_k).copy()
    b = _c128(b_k).copy()
    ne = a.size // c.cells
    lib().orc_propagate_noise(C.byref(c), _ptr(a), _ptr(b), ne, dt, n_steps, seed, step0, group)
    return a, b


def nonlinear_rt_noise(c, a_x, b_x, dt, seed, step, group=0):
    a = _c128(a_x).copy()
    b = _c128(b_x).copy()
    lib().orc_nonlinear_rt_noise(C.byref(c), _ptr(a), _ptr(b), a.size // c.cells, dt, seed, step, group)
    return a, b


def loss_noise_cell(seed, traj, cell, step, group=0):
    out = np.zeros(4)
    lib().orc_loss_noise_cell(seed, traj, cell, step, group, _ptr(out))
    return out


def nonlinear_rt(c, a_x, b_x, dt):
    a = _c128(a_x).copy()
    b = _c128(b_x).copy()
    lib().orc_nonlinear_rt(C.byref(c), _ptr(a), _ptr(b), a.size // c.cells, dt)
    return a, b


def kinetic_rt(c, a_k, b_k, dt):
    a = _c128(a_k).copy()
    b = _c128(b_k).copy()
    k2 = np.ascontiguousarray(ktable(c))
    lib().orc_kinetic_rt(C.byref(c), _ptr(k2), _ptr(a), _ptr(b), a.size // c.cells, dt)
    return a, b


def half_pi_pulse(a, b):
    a = _c128(a).copy()
    b = _c128(b).copy()
    lib().orc_half_pi_pulse(_ptr(a), _ptr(b), a.size)
    return a, b


def pi_pulse(a, b):
    a = _c128(a).copy()
    b = _c128(b).copy()
    lib().orc_pi_pulse(_ptr(a), _ptr(b), a.size)
    return a, b


def rotate_sums(c, a_x, b_x, alpha):
    n1, n2 = C.c_double(), C.c_double()
    a = _c128(a_x)
    lib().orc_rotate_sums(C.byref(c), _ptr(a), _ptr(_c128(b_x)), a.size // c.cells, alpha,
                          C.byref(n1), C.byref(n2))
    return n1.value, n2.value


def traj_sums(c, a_x, b_x):
    a = _c128(a_x)
    ne = a.size // c.cells
    out = np.empty((ne, 4), dtype=np.float64)
    lib().orc_traj_sums(C.byref(c), _ptr(a), _ptr(_c128(b_x)), ne, _ptr(out))
    return out


def cell_density_sums(c, a_x, b_x, alpha=0.0):
    a = _c128(a_x)
    ne = a.size // c.cells
    da = np.empty(c.cells, dtype=np.float64)
    db = np.empty(c.cells, dtype=np.float64)
    lib().orc_cell_density_sums(C.byref(c), _ptr(a), _ptr(_c128(b_x)), ne, alpha, _ptr(da), _ptr(db))
    return da.reshape(shape(c)), db.reshape(shape(c))


def visibility_grid(c, a_x, b_x):
    a = _c128(a_x)
    return lib().orc_visibility_grid(C.byref(c), _ptr(a), _ptr(_c128(b_x)), a.size // c.cells)


def params2d(**kw):
    """2D program defaults (programs/bec_wigner_2d.cpp/wigner.cpp:106-136)."""
    d = dict(nvx=128, nvy=128, xmax=20.0, ymax=20.0, g11=6.18, g12=6.03, g22=5.88, E=2.5,
             mu=26.8, Va=0.2, Vb=0.5, dtG=1.0 / 31.0, ntG=32, itmax=3, dtW=0.01, ntW=200,
             n_meas=1, ne=32, n0=1.0, seed=0x5EED, noise_dev=None, gs_rule=0, gs_tol=1e-8,
             gs_max_iter=100000, n_atoms=0.0, gs_tau=0.5, gs_beta=0.6)
    d.update(kw)
    if d["noise_dev"] is None:
        dx = 2.0 * d["xmax"] / (d["nvx"] - 1)
        dy = 2.0 * d["ymax"] / (d["nvy"] - 1)
        d["noise_dev"] = 1.0 / np.sqrt(dx * dy)
    return Params2D(**d)


def run2d(p, want_gs=True):
    nm = p.n_meas
    m = np.zeros(2 * nm)
    var = np.zeros(2 * nm)
    vis = np.zeros(nm)
    gs = np.empty((p.nvy, p.nvx), dtype=np.complex128)
    avn = np.empty((p.nvy, p.nvx), dtype=np.float64)
    rep = Report2D()
    lib().orc2d_run(C.byref(p), _ptr(m), _ptr(var), _ptr(vis), _ptr(gs), _ptr(avn), C.byref(rep))
    return dict(m=m.reshape(2, nm), var=var.reshape(2, nm), vis=vis, gs=gs, avn=avn, report=rep)


def residual2d(p, psi):
    mu = C.c_double()
    r = lib().orc2d_residual(C.byref(p), _ptr(_c128(psi)), C.byref(mu))
    return r, mu.value


def dft1d(x, sign):
    """The length-n DFT by its definition, sum_j x_j exp(sign 2 pi i j k / n)."""
    d = _c128(np.array(x, dtype=np.complex128))
    lib().orc_dft1d(_ptr(d), d.size, int(sign))
    return d


def params1d(**kw):
    """1D program defaults (programs/bec_wigner_2d.cpp/wigner_1d.cpp:93-130); tmaxGP / tmaxWig
    give ntG = 1 + (int)(tmaxGP/dt), ntW = 1 + (int)(tmaxWig/dt) as the program does."""
    d = dict(zmax=20.0, dz=1.0, g11=6.18, g12=6.03, g22=5.88, E=2.5, mu=26.8, Va=0.5, Vb=1.0, dt=0.01,
             ntG=None, itmax=3, ntW=None, ne=20, n0=1.0, seed=0x5EED, noise_dev=None, tmaxGP=1.0, tmaxWig=30.0)
    d.update(kw)
    tG, tW = d.pop("tmaxGP"), d.pop("tmaxWig")
    if d["ntG"] is None:
        d["ntG"] = 1 + int(tG / d["dt"])
    if d["ntW"] is None:
        d["ntW"] = 1 + int(tW / d["dt"])
    if d["noise_dev"] is None:
        d["noise_dev"] = np.sqrt(1.0 / d["dz"]) / 2.0
    return Params1D(**d)


def nv1d(p):
    return lib().orc1d_nv(C.byref(p))


def run1d(p):
    nv, nt = nv1d(p), p.ntW
    m = np.zeros(2 * nt)
    var = np.zeros(2 * nt)
    vis = np.zeros(nt)
    gs = np.empty(nv, dtype=np.complex128)
    avn = np.empty(nv)
    ntot = np.empty(p.ntG)
    rep = Report1D()
    lib().orc1d_run(C.byref(p), _ptr(m), _ptr(var), _ptr(vis), _ptr(gs), _ptr(avn), _ptr(ntot), C.byref(rep))
    return dict(m=m.reshape(2, nt), var=var.reshape(2, nt), vis=vis, gs=gs, avn=avn, ntot=ntot, report=rep)


def residual1d(p, psi):
    mu = C.c_double()
    r = lib().orc1d_residual(C.byref(p), _ptr(_c128(psi)), C.byref(mu))
    return r, mu.value


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return lib().orc_num_threads()
