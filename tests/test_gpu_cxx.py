"""The C++ facade on the GPU: the reference's single_launch driver
(test/single_launch.cc) rebuilt over include/fastsolver_amd.hpp, its saved
solution checked against the Woodbury closed form of the same circulant system."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from fastsolver_amd import gen
from tests.gpu_util import host_system, rel

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fastsolver_amd", "examples", "single_launch")


def _read_regions(path):
    rows = []
    for line in open(path):
        vals = [v for v in line.rstrip("\n").split("\t") if v]
        if vals:
            rows.append([float(v) for v in vals])
    return np.array(rows)


@pytest.mark.parametrize("precision", [0, 1])
def test_single_launch_driver(tmp_path, precision):
    print(f"single_launch precision={precision}", flush=True)
    p = subprocess.run([EXE, str(precision)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0, (p.returncode, p.stderr)
    assert "Residual:" in p.stdout
    # the reference's own error report (compute_L2_error), computed by the facade
    err = float(p.stdout.split("Err:")[1].split()[0])
    assert err <= 1e-12, err
    assert (tmp_path / "soln_ref.txt").exists()
    spec = gen.single_launch_case()
    X, V, K = host_system(spec)
    b = X[:, :spec.nrhs].copy(order="F")
    xw = oracle.woodbury_circulant(spec.n, spec.ranks[0], spec.diag, b)
    x = _read_regions(tmp_path / "global_soln.txt")
    assert x.shape == (spec.n, spec.nrhs)
    assert rel(x, xw) <= 1e-12
