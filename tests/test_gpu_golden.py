"""The HIP path (through the C-ABI) against the committed golden fixtures."""
import os
import subprocess

import numpy as np
import pytest

import mat3d_amd as m
from mat3d_amd import gen

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


def test_cpp_compat_example_runs(tmp_path):
    """The C++ caller (include/m3d_compat.hpp) drives the same solve."""
    f = load("elastic12.npz")
    exe = os.path.join(ROOT, "mat3d_amd", "examples", "homogenize")
    assert os.path.exists(exe), "build() compiles the example"
    rve = str(tmp_path / "rve12.txt")
    np.savetxt(rve, f["ids"].reshape(-1), fmt="%d")
    mat = f["stress_mat"]
    out = subprocess.run([exe, "12", rve, *(f"{float(v):.17g}" for v in (mat[0, 0], mat[0, 1], mat[1, 0], mat[1, 1])),
                          "1e-6", "1e-6", "H", "fp64"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    C = np.array([[float(v) for v in l.split()] for l in lines[2:8]])
    ref = f["stress_C"]
    assert np.linalg.norm(C - ref) / np.linalg.norm(ref) < 1e-8


def _deck(src, tmp, **subs):
    """The reference example deck with some values replaced (key -> value)."""
    toks = open(os.path.join(ROOT, "tests", "golden", "decks", src)).read().split()
    for i in range(0, len(toks) - 1, 2):
        if toks[i] in subs:
            toks[i + 1] = str(subs[toks[i]])
    path = os.path.join(tmp, src)
    with open(path, "w") as f:
        f.write("\n".join(f"{toks[i]}\n{toks[i + 1]}\n" for i in range(0, len(toks) - 1, 2)))
    return path


def test_reference_decks_end_to_end(tmp_path):
    """mat3d_run reads the reference example decks (example/3d, example/3d_thermal),
    writes the homogenization reports (name + ".txt") and the .vti file."""
    exe = os.path.join(ROOT, "mat3d_amd", "examples", "mat3d_run")
    assert os.path.exists(exe), "build() compiles the example"
    n = 16
    ids = gen.spheres(n, 3, 0.2, seed=3)
    np.savetxt(tmp_path / "rve16.txt", ids.reshape(-1), fmt="%d")
    # structural, homogenization, report, vti
    d = _deck("in_struct.txt", tmp_path, **{"Nx=Ny=Nz=": n, "RVE_Name": "rve16.txt",
                                            "SingleAnalysisOrHomogenization=": "H"})
    out = subprocess.run([exe, "--fp64", d], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert out.returncode == 0, out.stderr
    rep = open(tmp_path / "myOutput.txt.txt").read().split("\n")
    j = rep.index("C_Homogenized=")
    C = np.array([[float(x) for x in rep[j + 2 + q].split()] for q in range(6)])
    mat = np.array([[31e9, 0.3, 40e6, 100e6, 1, 8e-2, 8e2], [30e9, 0.2, 250e6, 8006, 0.2, 8e-4, 8e3]])
    s = m.FFT_Analysis3D(n, n, n, "e", precision="fp64")
    s.material(mat, ids)
    s.homogenization(1e-6, 1e-6)
    assert np.allclose(C, s.CHomo, rtol=1e-5, atol=1e-6 * np.abs(s.CHomo).max())
    import xml.etree.ElementTree as ET
    arr = ET.parse(tmp_path / "myvt1i.vti").getroot().find("ImageData/Piece/PointData/DataArray")
    v = np.array(arr.text.split(), dtype=float).reshape(n**3, 14)
    eps, sig = s.fields(5)  # the handle keeps the last load case, as the reference
    assert np.allclose(v[:, :6], eps, rtol=1e-12, atol=1e-12 * np.abs(eps).max())
    assert np.allclose(v[:, 6:12], sig, rtol=1e-12, atol=1e-12 * np.abs(sig).max())
    # thermal, homogenization, report
    d = _deck("in_therm.txt", tmp_path, **{"Nx=Ny=Nz=": n, "RVE_Name": "rve16.txt", "vtiOutput_YorN=": "N"})
    out = subprocess.run([exe, "--fp64", "--thermal", d], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert out.returncode == 0, out.stderr
    rep = open(tmp_path / "myOutputthermal.txt.txt").read().split("\n")
    j = rep.index("KHomogenized=")
    K = np.array([[float(x) for x in rep[j + 2 + q].split()] for q in range(3)])
    t = m.FFT_analysis3d_thermal(n, n, n, precision="fp64")
    t.material(25.0, 30.0, ids)
    t.homogenization(1e-6, 1e-6)
    assert np.allclose(K, t.KHomo, rtol=1e-5)
