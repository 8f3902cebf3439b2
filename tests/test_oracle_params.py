"""CPU: fillDerivedParameters (programs/bec_wigner_3d.cuda/wigner.cpp:119-176) in the
oracle and in the product library, against independently computed FP64 values of
the Rb-87 configurations (N = 30000 and 70000)."""
import math

import pytest

from tests.helpers import config1_model


# independent FP64 evaluation of the formulas of wigner.cpp:121-176 for the
# 30k-atom cigar trap (128x16x16) and the 70k-atom 256x32x32 lattice
CFG1 = dict(l_rho=1.09161e-6, t_rho=1.63069e-3, lambda_=8.16054, g11=0.061162, g12=0.059492, g22=0.057872,
            mu=4.68247, xmax=3.67226, zmax=29.96763, dx=0.48964, dz=0.47193, l111=5.20438e-9,
            l12=9.77839e-5, l22=1.88046e-5, dt_evo=0.0245296)
CFG70K = dict(mu=6.5715, xmax=4.3504, zmax=35.5016)


def _close(a, b, rel):
    return abs(a - b) <= rel * abs(b)


def test_oracle_config1_values(oracle):
    c = oracle.derive(oracle.model(**config1_model(oracle)))
    for k, v in CFG1.items():
        assert _close(getattr(c, k), v, 2e-5), (k, getattr(c, k), v)
    assert _close(c.dx * c.dy * c.dz, 0.113142, 2e-5)
    assert _close(c.V / (2 * c.dx * c.dy * c.dz), 4.41923, 2e-5)   # vacuum density 1/(2 dV)


def test_oracle_70k_256x32x32(oracle):
    c = oracle.derive(oracle.model(**config1_model(oracle, N=70000, nvx=32, nvy=32, nvz=256)))
    for k, v in CFG70K.items():
        assert _close(getattr(c, k), v, 2e-5), (k, getattr(c, k), v)
    assert _close(c.dx * c.dy * c.dz, 0.0219347, 2e-5)


def test_128x32x32_spacing(oracle):
    c = oracle.derive(oracle.model(**config1_model(oracle, nvx=32, nvy=32)))
    assert _close(c.dx, 0.23692, 2e-5)
    assert _close(c.dx * c.dy * c.dz, 0.0264901, 2e-5)


def test_tf_mu_closed_form(oracle):
    # mu_TF = (15 N g11 / (16 pi lambda sqrt 2))^(2/5)
    c = oracle.derive(oracle.model(N=100000))
    assert _close(c.mu, (15 * 100000 * c.g11 / (16 * math.pi * c.lambda_ * math.sqrt(2))) ** 0.4, 1e-14)
    assert _close(c.dkx, math.pi / c.xmax, 1e-15)      # quirk Q2: pi / xmax


def test_library_matches_oracle(oracle):
    import bec_amd
    kw = config1_model(oracle, detuning=-41.0, N=55000)
    c = bec_amd.derive(bec_amd.default_model(**kw))
    o = oracle.derive(oracle.model(**kw))
    for name, _ in o._fields_:
        assert getattr(c, name) == pytest.approx(getattr(o, name), rel=1e-15, abs=1e-300), name


def test_library_defaults_are_reference_defaults():
    import bec_amd
    m = bec_amd.default_model()
    assert (m.N, m.nvx, m.nvy, m.nvz, m.ne, m.itmax) == (150000, 16, 16, 128, 4, 1)
    assert (m.a11, m.a12, m.a22, m.fx, m.fz, m.detuning) == (100.4, 97.66, 95.0, 97.6, 11.96, -41)
    assert (m.dt_steady, m.dt_evo, m.t_equilib) == (1e-5, 4e-5, 0.0)
