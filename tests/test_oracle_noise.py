"""CPU: Philox4x32-10 known-answer vectors.  tests/golden/philox_kat.json was
written by tests/golden/make_philox_vectors.cpp from rocRAND's host engine; the
three vectors below are the published Random123 known answers."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

RANDOM123 = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def test_random123_known_answers(oracle):
    for ctr, key, out in RANDOM123:
        assert list(oracle.philox(ctr, key)) == out


def test_rocrand_vectors(oracle):
    with open(os.path.join(HERE, "golden", "philox_kat.json")) as f:
        cases = json.load(f)
    assert len(cases) >= 8
    for c in cases:
        assert list(oracle.philox(c["ctr"], c["key"])) == c["out"]


def test_noise_cell_box_muller(oracle):
    """Counter {cell, traj, stream, 0}, key = seed; Box-Muller with u = (x + 1) / 2^32
    as fillWithNormalDistribution (programs/bec_wigner_3d.cuda/wigner.cu:15-34)."""
    seed, traj, cell = 0x5EED, 3, 777
    w = oracle.philox([cell, traj, 0, 0], [seed & 0xffffffff, seed >> 32])
    u = (w.astype(np.float64) + 1.0) / 4294967296.0
    r1 = 0.5 * np.sqrt(-2 * np.log(u[0]))
    r2 = 0.5 * np.sqrt(-2 * np.log(u[2]))
    want = [r1 * np.cos(2 * np.pi * u[1]), r1 * np.sin(2 * np.pi * u[1]),
            r2 * np.cos(2 * np.pi * u[3]), r2 * np.sin(2 * np.pi * u[3])]
    assert np.allclose(oracle.noise_cell(seed, traj, cell), want, rtol=1e-15, atol=0)
