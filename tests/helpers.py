"""Shared configuration builders for the parity tests."""
import numpy as np


def config1_model(api, **kw):
    """The 128x16x16 Rb-87 Wigner configuration (30k atoms, losses on)."""
    d = dict(N=30000, nvx=16, nvy=16, nvz=128, ne=4, gamma111=5.4e-30, gamma12=0.78e-13, gamma22=1.5e-14,
             detuning=0.0, dt_evo=4e-5, dt_steady=1e-5, t_equilib=0.0, Va=1.0, Vb=1.0, itmax=1)
    d.update(kw)
    return d


def oracle_calc(O, **kw):
    return O.derive(O.model(**kw))


def gpu_calc(bec, **kw):
    return bec.derive(bec.default_model(**kw))


def rel_l2(a, b, axis=None):
    a = np.asarray(a)
    b = np.asarray(b)
    if axis is None:
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    num = np.sqrt(np.sum(np.abs(a - b) ** 2, axis=axis))
    den = np.sqrt(np.sum(np.abs(b) ** 2, axis=axis))
    return num / den
