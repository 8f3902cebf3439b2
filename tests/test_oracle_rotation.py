"""CPU: pins the oracle's general resonant pulse U(theta, phi) (orc_rotation):
it equals the reference's pulses at U(pi/2, 0) = applyHalfPiPulse and
U(pi, 0) = applyPiPulse (wigner_kernel.cu:320-341), and its densities at
U(pi/2, -alpha) are halfPiRotate(alpha)'s (wigner_kernel.cu:344-360, restated
independently in orc_rotate_sums); pulses of equal phase add their areas;
every pulse is unitary (the total density is unchanged)."""
import numpy as np
import pytest

from tests.helpers import config1_model, oracle_calc


def _state(n, seed=0):
    r = np.random.default_rng(seed)
    return (r.standard_normal(n) + 1j * r.standard_normal(n), r.standard_normal(n) + 1j * r.standard_normal(n))


def test_reference_pulses_are_special_cases(oracle):
    a, b = _state(1000)
    for theta, ref in ((np.pi / 2, oracle.half_pi_pulse), (np.pi, oracle.pi_pulse)):
        ra, rb = oracle.rotation(a, b, theta, 0.0)
        ea, eb = ref(a, b)
        assert np.allclose(ra, ea, atol=1e-15) and np.allclose(rb, eb, atol=1e-15)


@pytest.mark.parametrize("alpha", [0.0, 0.5, 2.0, 4.5])
def test_measurement_rotation_is_a_half_pi_pulse(oracle, alpha):
    c = oracle_calc(oracle, **config1_model(oracle, nvz=64, ne=2))
    a, b = _state(c.cells * c.ne, 1)
    n1, n2 = oracle.rotate_sums(c, a, b, alpha)
    ra, rb = oracle.rotation(a, b, np.pi / 2, -alpha)
    vac = c.V / (2.0 * c.dx * c.dy * c.dz)
    assert np.sum(np.abs(ra) ** 2 - vac) == pytest.approx(n1, rel=1e-12)
    assert np.sum(np.abs(rb) ** 2 - vac) == pytest.approx(n2, rel=1e-12)


def test_areas_add_and_pulses_are_unitary(oracle):
    a, b = _state(500, 2)
    for phi in (0.0, 0.7, -2.1):
        a1, b1 = oracle.rotation(*oracle.rotation(a, b, 0.4, phi), 1.1, phi)
        a2, b2 = oracle.rotation(a, b, 1.5, phi)
        assert np.allclose(a1, a2, atol=1e-14) and np.allclose(b1, b2, atol=1e-14)
        assert np.allclose(np.abs(a2) ** 2 + np.abs(b2) ** 2, np.abs(a) ** 2 + np.abs(b) ** 2, rtol=1e-14)
    # a full turn (area 2 pi) is -1
    a4, b4 = oracle.rotation(a, b, 2 * np.pi, 0.3)
    assert np.allclose(a4, -a, atol=1e-14) and np.allclose(b4, -b, atol=1e-14)
