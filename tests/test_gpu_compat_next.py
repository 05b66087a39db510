"""The C++ compat classes for the rows beside the 3D solvers
(include/m3d_compat.hpp: percolation, FFT_Analysis2D, recPlan), driven by
mat3d_amd/examples/next_rows the way the reference's drivers call its
classes, give the same public results as the Python mirror on the same
library."""
import os
import subprocess

import numpy as np
import pytest

import mat3d_amd as m
from mat3d_amd import gen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mat3d_amd", "examples", "next_rows")


def run(*args, cwd=None):
    assert os.path.exists(EXE), "build() compiles the example"
    out = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=120, cwd=cwd)
    assert out.returncode == 0, out.stderr
    return out.stdout.strip().splitlines()


def test_percolation_cpp(tmp_path):
    n = 32
    ids = gen.spheres(n, 3, 0.25, seed=4)
    f = tmp_path / "ids.txt"
    np.savetxt(f, ids.reshape(-1), fmt="%d")
    out = tmp_path / "perc_out"
    lines = run("perc", n, f, 1, 0, 0, 1, out)
    p = m.percolation(ids, n, n, n, 1, 0)
    p.runPercolationAnalysis(1)
    c1 = p.C1.reshape(-1).astype(np.int64)
    h = 0
    for v in c1.tolist():
        h = (h * 31 + v) % 1000000007
    assert lines[0] == f"clusters {len(p.all_clusters)}"
    assert float(lines[1].split()[1]) == p.vol_intrstd_phs
    assert lines[2] == f"c1 {int(c1.sum())} {h}"
    assert [int(v) for v in lines[3].split()[1:]] == p.all_clusters
    assert [float(v) for v in lines[4].split()[1:]] == p.clusters_vol_frac
    assert int(lines[5].split()[1]) == p.SurfaceOfTwoPhase(ids, 0, 1)
    # printer / vtiWriter (percolation.cpp:417-473)
    assert [int(v) for v in open(out / "clustersFile.txt").read().split()] == c1.tolist()
    rows = open(out / "clustersdata.txt").read().split("\n")
    assert rows[0] == "           intrested phase :  0     Number of the phases: 1"
    assert rows[2].startswith("connected cluster name        volume fraction")
    for c, fr in enumerate(p.clusters_vol_frac):
        assert rows[3 + c] == "        " + str(2 + c) + " " * 35 + f"{fr:g}"
    vti = open(out / "clusters.vti").read()
    vals = vti.split('format="ascii">\n')[1].split("        </DataArray>")[0].split()
    assert [int(float(v)) for v in vals] == c1.tolist()


def test_plate2d_cpp(tmp_path):
    n = 64
    C = np.array([[float(v) for v in l.split()] for l in run("plate2d", n, "print", cwd=tmp_path)])
    x, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    rve = np.where((x - n // 2) ** 2 + (y - n // 2) ** 2 <= (n / 4.0) ** 2, 0.0, 1.0)
    a = m.FFT_Analysis2D(n, n)
    a.material([68.9e9, 0.35, 400e9, 0.23], rve.reshape(-1))
    a.setLoad([0, 0, 0])
    ref = a.homogenization()
    assert np.array_equal(C, ref)
    # printer (FFT_analysis2D.cpp:328-360): real parts of the last load case's fields
    e, s = a.fields()
    for c in range(3):
        got = np.array(open(tmp_path / f"epsilon{c + 1}.txt").read().split(), dtype=float)
        assert np.allclose(got, e[:, c].real, rtol=1e-5, atol=1e-12 * np.abs(e).max())
        got = np.array(open(tmp_path / f"sigma{c + 1}.txt").read().split(), dtype=float)
        assert np.allclose(got, s[:, c].real, rtol=1e-5, atol=1e-12 * np.abs(s).max())


def test_rec_cpp(tmp_path):
    n = 32
    rve = (gen.spheres(n, 3, 0.25, seed=3) == 0).astype(float)
    tpcf = np.abs(np.fft.fftn(rve)) ** 2 / rve.size
    f = tmp_path / "tpcf.txt"
    np.savetxt(f, tpcf.reshape(-1), fmt="%.17g")
    out = tmp_path / "rec_out"
    lines = run("rec", n, f, 40, 0.8, 5, out)
    g = m.recPlan(np.loadtxt(f).reshape(n, n, n))
    it = g.runRec(40, 0.8, 5)
    assert lines[0] == f"iterations {len(g.trackrealerrorA)}"
    assert float(lines[1].split()[1]) == g.trackrealerrorA[-1]
    assert float(lines[2].split()[1]) == pytest.approx(g.mTilda.real.mean(), rel=1e-12)
    # saveRecImages text files and vtiWriter (reconstruction.cpp:409-477)
    rr = np.array(open(out / "rve_rec.txt").read().split(), dtype=float)
    assert np.allclose(rr, g.mTilda.real.reshape(-1), rtol=1e-5, atol=1e-6)
    err = open(out / "err.txt").read().split("\n")
    assert err[0].startswith("voloum fraction= ") and len([x for x in err[1:] if x]) == min(39, it)
    vti = open(out / "rec.vti").read()
    vals = np.array(vti.split('format="ascii">\n')[1].split("        </DataArray>")[0].split(), dtype=float)
    assert np.allclose(vals.reshape(-1, 2)[:, 1], np.loadtxt(f))
