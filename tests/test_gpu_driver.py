"""GPU: the YAML driver bec_run (experiment YAML in, the reference's xy-series
YAML out, data/*.yaml schema) runs a small experiment, and its series equal the
same experiment run through the library's Python binding."""
import ast
import os
import subprocess

import numpy as np
import pytest
import yaml

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

EXPERIMENT = """# small driver run
name: driver test
N: 30000
nvx: 16
nvy: 16
nvz: 128
ne: 4
Va: 1
Vb: 1
detuning: 0
gamma111: 5.4e-30
gamma12: 0.78e-13
gamma22: 1.5e-14
dt_evo: 4e-5
dt_steady: 1e-5
precision: fp64
gs_method: gradient
gs_tol: 1e-6
seed: 24301
steps: 500
sample_every: 250
"""


def test_driver_writes_reference_schema_and_matches_api(bec, tmp_path):
    exe = os.path.join(ROOT, "bec_amd", "bec_run")
    assert os.path.exists(exe), "bec_run is built by bec_amd.build"
    cfg = tmp_path / "exp.yaml"
    cfg.write_text(EXPERIMENT)
    prefix = str(tmp_path / "out")
    r = subprocess.run([exe, str(cfg), prefix], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    series = {}
    for tag in ("visibility", "visibility_analytic", "N", "Na", "Nb", "phase"):
        with open(f"{prefix}_{tag}.yaml") as f:
            d = yaml.safe_load(f)
        for key in ("xarray", "yarray", "xname", "yname", "ymin", "ymax", "format", "name", "source"):
            assert key in d, (tag, key)
        assert d["format"] == "xy"
        x, y = np.array(ast.literal_eval(d["xarray"])), np.array(ast.literal_eval(d["yarray"]))
        assert x.shape == y.shape == (3,)
        series[tag] = (x, y, d)
    assert np.allclose(series["N"][0], np.arange(3) * 250 * 4e-5, rtol=1e-6)
    # error bars on the populations
    err = np.array(ast.literal_eval(series["Na"][2]["yerrors"]))
    assert err.shape == (3,) and np.all(err > 0)

    # the same experiment through the binding
    m = dict(N=30000, nvx=16, nvy=16, nvz=128, ne=4, Va=1.0, Vb=1.0, detuning=0.0, gamma111=5.4e-30,
             gamma12=0.78e-13, gamma22=1.5e-14, dt_evo=4e-5, dt_steady=1e-5)
    c = bec.derive(bec.default_model(**m))
    with bec.Solver(c, precision=bec.FP64) as s:
        gs, rep, st = s.calculate_steady_state(method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL, tol=1e-6,
                                               precision=bec.FP64)
        s.init_evolution(gs, 24301, equilibrate=True)
        obs = [s.observables()]
        for _ in range(2):
            s.propagate(250)
            obs.append(s.observables())
    assert np.allclose(series["visibility"][1], [o["visibility_grid"] for o in obs], rtol=1e-10, atol=1e-12)
    assert np.allclose(series["N"][1], [o["N_total"] for o in obs], rtol=1e-10)
    assert np.allclose(series["Na"][1], [o["Na"] / 4 for o in obs], rtol=1e-10)
    assert np.allclose(err, [o["Na_stderr"] for o in obs], rtol=1e-8)


def test_driver_loss_noise_option(bec, tmp_path):
    exe = os.path.join(ROOT, "bec_amd", "bec_run")
    series = {}
    for noise in ("false", "true"):
        cfg = tmp_path / f"exp_{noise}.yaml"
        cfg.write_text(EXPERIMENT.replace("gamma12: 0.78e-13", "gamma12: 0.78e-11") + f"loss_noise: {noise}\n")
        prefix = str(tmp_path / f"out_{noise}")
        r = subprocess.run([exe, str(cfg), prefix], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        with open(f"{prefix}_Na.yaml") as f:
            series[noise] = np.array(ast.literal_eval(yaml.safe_load(f)["yarray"]))
    # the noise is on from the equilibration step on, so every sample carries it
    assert not np.allclose(series["true"], series["false"], rtol=1e-9)
