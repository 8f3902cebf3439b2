"""CPU: the YAML driver rejects a run length or sampling period it cannot honour
before it touches the device (no GPU needed: the checks come first)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "bec_amd", "bec_run")

BASE = """name: args
N: 30000
nvx: 16
nvy: 16
nvz: 64
ne: 2
"""


def _run(tmp_path, extra):
    if not os.path.exists(EXE):
        pytest.fail("bec_run is built by bec_amd.build")
    y = tmp_path / "e.yaml"
    y.write_text(BASE + extra)
    return subprocess.run([EXE, str(y), str(tmp_path / "out")], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=60)


@pytest.mark.parametrize("extra,word", [("sample_every: 0\nsteps: 10\n", "sample_every"),
                                        ("sample_every: -5\nsteps: 10\n", "sample_every"),
                                        ("sample_every: 2.5\nsteps: 10\n", "sample_every"),
                                        ("sample_every: 5\nsteps: -1\n", "steps")])
def test_driver_rejects_bad_period_or_length(tmp_path, extra, word):
    p = _run(tmp_path, extra)
    assert p.returncode == 2, (p.returncode, p.stdout, p.stderr)
    assert word in p.stderr
    assert not (tmp_path / "out_visibility.yaml").exists()
