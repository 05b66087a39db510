"""The HIP path (through the C-ABI) against the committed golden fixtures."""
import os
import subprocess

import numpy as np
import pytest

import mat3d_amd as m

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32_polish", 1e-6)])
def test_elastic12_against_fixture(precision, tol):
    f = load("elastic12.npz")
    n = 12
    for tag in ("stress", "strain"):
        s = m.FFT_Analysis3D(n, n, n, "e", precision=precision)
        s.material(f[f"{tag}_mat"], f["ids"])
        s.homogenization(1e-6, 1e-6)
        C = f[f"{tag}_C"]
        assert np.linalg.norm(s.CHomo - C) / np.linalg.norm(C) < tol
        if precision == "fp64":
            assert [i["iterations"] for i in s.info] == list(f[f"{tag}_iters"])
    s = m.FFT_Analysis3D(n, n, n, "e", precision=precision)
    s.material(np.array([[300e9, 0.25, np.inf, 0, 1, 0, 0], [3e9, 0.35, np.inf, 0, 1, 0, 0]]), f["ids"])
    E = np.array([0.0, 0.1, 0.0, 0.0, 0.0, 0.0])
    s.setLoad(E, E)
    info = s.runAnalysis(1e-6, 1e-6, 15)
    eps, sig = s.fields()
    assert np.abs(sig - f["run_sig"]).max() / np.abs(f["run_sig"]).max() < (1e-9 if precision == "fp64" else 1e-5)
    if precision == "fp64":
        assert info["iterations"] == int(f["run_iters"])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32_polish", 1e-6)])
def test_thermal12_against_fixture(precision, tol):
    f = load("thermal12.npz")
    n = 12
    for tag in ("grad", "flux"):
        km, ki = f[f"{tag}_k"]
        s = m.FFT_analysis3d_thermal(n, n, n, precision=precision)
        s.material(km, ki, f["ids"])
        s.homogenization(1e-6, 1e-6)
        K = f[f"{tag}_K"]
        assert np.linalg.norm(s.KHomo - K) / np.linalg.norm(K) < tol
    s = m.FFT_analysis3d_thermal(n, n, n, precision=precision)
    s.material_n(f["n3_k"], f["ids3"])
    s.homogenization(1e-7, 1e-7)
    assert np.linalg.norm(s.KHomo - f["n3_K"]) / np.linalg.norm(f["n3_K"]) < tol


def test_cpp_compat_example_runs():
    """The C++ caller (include/m3d_compat.hpp) drives the same solve."""
    f = load("elastic12.npz")
    exe = os.path.join(ROOT, "mat3d_amd", "examples", "homogenize")
    assert os.path.exists(exe), "build() compiles the example"
    rve = os.path.join(ROOT, "gpurun_out", "rve12.txt")
    os.makedirs(os.path.dirname(rve), exist_ok=True)
    np.savetxt(rve, f["ids"].reshape(-1), fmt="%d")
    mat = f["stress_mat"]
    out = subprocess.run([exe, "12", rve, *(f"{float(v):.17g}" for v in (mat[0, 0], mat[0, 1], mat[1, 0], mat[1, 1])),
                          "1e-6", "1e-6", "H", "fp64"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    C = np.array([[float(v) for v in l.split()] for l in lines[2:8]])
    ref = f["stress_C"]
    assert np.linalg.norm(C - ref) / np.linalg.norm(ref) < 1e-8
