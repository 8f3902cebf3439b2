"""CPU: the C-ABI library loads, exports every entry point include/bec_api.h
declares, and rejects bad arguments before touching a device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


HEADERS = ("bec_api.h", "bec_fno.h")


def _declared(headers=HEADERS):
    names = set()
    for h in headers:
        with open(os.path.join(ROOT, "include", h)) as f:
            text = f.read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(bec_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import bec_amd
    L = bec_amd.lib()
    names = _declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from bec_amd import api, fno
    missing = [n for n in _declared(("bec_api.h",)) if n not in api.EXPORTS]
    missing += [n for n in _declared(("bec_fno.h",)) if n not in fno.EXPORTS]
    assert not missing, missing


def test_oracle_is_not_linked_into_the_product():
    import bec_amd
    L = bec_amd.lib()
    for sym in ("orc_propagate", "orc_fft3d", "orc_ground_state"):
        assert not hasattr(L, sym)
    # and the product sources never import, include or load the oracle
    pat = re.compile(r"(import\s+oracle|from\s+oracle|liboracle|bec_oracle\.h|oracle\.oracle|orc_[a-z])")
    for d in ("bec_amd", "bec_amd/csrc", "include"):
        for fn in os.listdir(os.path.join(ROOT, d)):
            if fn.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                with open(os.path.join(ROOT, d, fn)) as f:
                    assert not pat.search(f.read()), fn


def test_strerror_and_status_codes():
    import bec_amd
    L = bec_amd.lib()
    assert L.bec_strerror(0) == b"ok"
    assert b"shape" in L.bec_strerror(2)
    assert L.bec_strerror(99) == b"unknown status"


def test_invalid_arguments_rejected_on_cpu():
    import bec_amd
    from bec_amd import api
    L = bec_amd.lib()
    # NULL pointers
    assert L.bec_fill_derived_params(None, None) == api.ERR_INVALID
    assert L.bec_create(None, 0, None, None) == api.ERR_INVALID
    # non power-of-two lattice
    m = bec_amd.default_model(nvx=12)
    c = api.CalcParams()
    assert L.bec_fill_derived_params(C.byref(m), C.byref(c)) == api.ERR_SHAPE
    # lattice shapes the kernels are not built for
    for kw in (dict(nvx=16, nvy=32), dict(nvx=64, nvy=64), dict(nvz=512), dict(nvz=32)):
        cc = bec_amd.derive(bec_amd.default_model(**kw))
        ctx = C.c_void_p()
        assert L.bec_create(C.byref(cc), 0, None, C.byref(ctx)) == api.ERR_SHAPE
        assert not ctx.value
    cc = bec_amd.derive(bec_amd.default_model())
    ctx = C.c_void_p()
    assert L.bec_create(C.byref(cc), 7, None, C.byref(ctx)) == api.ERR_INVALID
    # FFT plan shapes
    p = C.c_void_p()
    assert L.bec_fft_plan3d(C.byref(p), 100, 16, 16, 0, 1) == api.ERR_SHAPE
    assert L.bec_fft_plan3d(C.byref(p), 128, 16, 16, 0, 0) == api.ERR_INVALID
    # ground-state options
    o = bec_amd.gs_options(method=1, rule=0)
    reps = (api.GsReport * 1)()
    assert L.bec_ground_states_batched(C.byref(cc), 1, None, C.byref(o), None, None, reps) == api.ERR_INVALID
    o = bec_amd.gs_options(max_iter=0)
    assert L.bec_ground_states_batched(C.byref(cc), 1, None, C.byref(o), None, None, reps) == api.ERR_INVALID
    o = bec_amd.gs_options()
    bad = bec_amd.derive(bec_amd.default_model(nvz=32))
    assert L.bec_ground_states_batched(C.byref(bad), 1, None, C.byref(o), None, None, reps) == api.ERR_SHAPE
    # two-component ground state: only the gradient flow with the residual rule; fractions in (0, 1)
    reps2 = (api.Gs2Report * 1)()
    o = bec_amd.gs_options(method=bec_amd.GS_SPLIT_STEP, rule=bec_amd.GS_RULE_RESIDUAL)
    assert L.bec_ground_states_two_component(C.byref(cc), 1, None, None, C.byref(o), None, None, reps2) == \
        api.ERR_INVALID
    o = bec_amd.gs_options(method=bec_amd.GS_GRADIENT, rule=bec_amd.GS_RULE_RESIDUAL)
    bad_frac = (C.c_double * 1)(1.0)
    assert L.bec_ground_states_two_component(C.byref(cc), 1, bad_frac, None, C.byref(o), None, None, reps2) == \
        api.ERR_INVALID
    # the 2D program: lattices beyond one workgroup's LDS, bad rules
    m2, v2 = (C.c_double * 2)(), (C.c_double * 2)()
    p2 = bec_amd.params2d(nvx=128, nvy=128)
    assert L.bec_run_2d(C.byref(p2), m2, v2, None, None, None, None) == api.ERR_SHAPE
    p2 = bec_amd.params2d(gs_rule=2)
    assert L.bec_run_2d(C.byref(p2), m2, v2, None, None, None, None) == api.ERR_INVALID
    # the 1D program: lattice size from zmax / dz (the program's 41), limits and bad constants
    assert L.bec_1d_cells(C.byref(bec_amd.params1d())) == 41
    assert L.bec_1d_cells(C.byref(bec_amd.params1d(zmax=10.0, dz=0.3))) == 67
    assert L.bec_1d_cells(C.byref(bec_amd.params1d(zmax=300.0, dz=1.0))) == -api.ERR_SHAPE
    assert L.bec_1d_cells(C.byref(bec_amd.params1d(dz=0.0, noise_dev=0.5))) == -api.ERR_INVALID
    assert L.bec_1d_cells(None) == -api.ERR_INVALID
    m1, v1 = (C.c_double * 2)(), (C.c_double * 2)()
    for bad in (dict(zmax=300.0), dict(dt=0.0, ntG=101, ntW=3001), dict(ntG=0), dict(itmax=-1), dict(ntW=0), dict(ne=0),
                dict(ne=100000, ntW=1000)):
        want = api.ERR_SHAPE if ("zmax" in bad or bad.get("ne") == 100000) else api.ERR_INVALID
        assert L.bec_run_1d(C.byref(bec_amd.params1d(**bad)), m1, v1, None, None, None, None, None) == want, bad
    assert L.bec_run_1d(C.byref(bec_amd.params1d()), None, v1, None, None, None, None, None) == api.ERR_INVALID
    # context-less calls
    assert L.bec_set_loss_noise(None, 1, 5, 0) == api.ERR_INVALID
    assert L.bec_pulse_rotation(None, 1.0, 0.0) == api.ERR_INVALID
    step = C.c_uint(0)
    assert L.bec_get_loss_noise_step(None, 0, C.byref(step)) == api.ERR_INVALID
    assert L.bec_propagate(None, 1, 0.0, 0, None) == api.ERR_INVALID
    assert L.bec_observables(None, 0, None) == api.ERR_INVALID
    assert L.bec_autotune(None, 0, 1, None, 0, None, None) == api.ERR_INVALID


def test_default_gs_options_are_the_reference_method():
    import bec_amd
    o = bec_amd.gs_options()
    assert (o.method, o.rule, o.tol, o.check_every) == (bec_amd.GS_SPLIT_STEP, bec_amd.GS_RULE_REFERENCE, 1e-6, 1)


def test_surrogate_entry_points_validate_on_cpu():
    from bec_amd import api, fno
    L = fno.flib()
    d = fno.GemmDesc()
    assert L.bec_gemm(None, None) == api.ERR_INVALID
    assert L.bec_gemm(C.byref(d), None) == api.ERR_INVALID          # NULL operands
    d.A = d.B = d.C = C.c_void_p(16)
    d.M = d.N = d.K = d.batch = 1
    d.ksplit = 4                                                     # split K without a workspace
    assert L.bec_gemm(C.byref(d), None) == api.ERR_INVALID
    h = C.c_void_p()
    good = dict(nz=128, ny=16, nx=16, mz=16, my=16, mx=16, width=32, layers=4, c_in=16, c_out=2, hidden=64,
                max_batch=16)
    for bad in (dict(mz=3), dict(mz=0), dict(mx=32), dict(width=0), dict(max_batch=0)):
        cfg = fno.FnoConfig(**dict(good, **bad))
        assert L.bec_fno_create(C.byref(cfg), 0, C.byref(h)) == api.ERR_INVALID, bad
        assert not h.value
    cfg = fno.FnoConfig(**dict(good, width=24))                      # channel maps need powers of two
    assert L.bec_fno_create(C.byref(cfg), 0, C.byref(h)) == api.ERR_SHAPE
    assert L.bec_fno_forward(None, 1, None, None, None) == api.ERR_INVALID
    assert L.bec_fno_gs2_features(None, None, None, None) == api.ERR_INVALID
