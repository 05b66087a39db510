"""GPU parity: the HIP solver (libmat3d.so, through the C-ABI) against the FP64 oracle.

Bars: FP64 path -- same iteration counts, C_eff within 1e-9 relative, fields
within 1e-9 relative; FP32+FP64-polish path -- C_eff within 1e-6 relative
voxel stress within 1e-5 relative (max-abs over max-abs).
"""
import numpy as np
import pytest

import mat3d_amd as m
import oracle
from mat3d_amd import gen

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def field_rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def oracle_homog(shape, mat, ids, tol=1e-6, behav="e"):
    o = oracle.Elastic(*shape, behav)
    o.material(mat, ids)
    r = o.homogenization(tol, tol)
    return o, r


@pytest.mark.parametrize("case", ["strain", "stress"])
def test_elastic_homogenization_fp64_matches_oracle(case):
    n = 16
    ids = gen.spheres(n, 3, 0.2, seed=3)
    # strain base: stiff inclusions are not the majority of max-mu -> pick by swapping phases
    mat = gen.MAT_C1 if case == "stress" else gen.MAT_C1[::-1].copy()
    o, r = oracle_homog((n, n, n), mat, ids)
    s = m.FFT_Analysis3D(n, n, n, "e", precision="fp64")
    s.material(mat, ids)
    assert s.strainBase == o.info()["strain_base"]
    assert s.strainBase == (case == "strain")
    s.homogenization(1e-6, 1e-6)
    assert [i["iterations"] for i in s.info] == r["iters"]
    assert rel(s.CHomo, r["C"]) < 1e-9
    assert rel(s.epHomo, r["epHomo"]) < 1e-9
    assert rel(s.sigHomo, r["sigHomo"]) < 1e-9
    eps_o, sig_o = o.fields()  # oracle keeps the last load case (w = 5)
    eps, sig = s.fields(5)
    assert field_rel(sig, sig_o) < 1e-9
    assert field_rel(eps, eps_o) < 1e-9


@pytest.mark.parametrize("shape", [(16, 24, 12), (12, 12, 20), (18, 18, 18), (15, 15, 15), (8, 10, 7)])
def test_elastic_fp64_shapes(shape):
    ids = gen.spheres(max(shape), 3, 0.25, seed=11)[: shape[0], : shape[1], : shape[2]].copy()
    mat = gen.MAT_C2
    o, r = oracle_homog(shape, mat, ids)
    s = m.FFT_Analysis3D(*shape, "e", precision="fp64")
    s.material(mat, ids)
    s.homogenization(1e-6, 1e-6)
    assert [i["iterations"] for i in s.info] == r["iters"]
    assert rel(s.CHomo, r["C"]) < 1e-9


def test_elastic_single_run_fields():
    n = 20
    ids = gen.fibres(n, 2, 0.25, seed=5, length=10)
    mat = gen.MAT_C2
    E = np.array([0.0, 0.1, 0.0, 0.0, 0.0, 0.0])
    o = oracle.Elastic(n, n, n, "e")
    o.material(mat, ids)
    o.set_load(E, E)
    it = o.run(1e-6, 1e-6, 15)
    s = m.FFT_Analysis3D(n, n, n, "e", precision="fp64")
    s.material(mat, ids)
    s.setLoad(E, E)
    info = s.runAnalysis(1e-6, 1e-6, 15)
    assert info["iterations"] == it
    eps_o, sig_o = o.fields()
    eps, sig = s.fields()
    assert field_rel(sig, sig_o) < 1e-9
    assert field_rel(eps, eps_o) < 1e-9


@pytest.mark.parametrize("n", [32, 48])
def test_elastic_fp32_polish_meets_bar(n):
    ids = gen.fibres(n, 3, 0.3, seed=2, length=n // 2)
    mat = gen.MAT_C2
    o, r = oracle_homog((n, n, n), mat, ids)
    s = m.FFT_Analysis3D(n, n, n, "e", precision="fp32_polish")
    s.material(mat, ids)
    s.homogenization(1e-6, 1e-6)
    err = rel(s.CHomo, r["C"])
    its = [i["iterations"] for i in s.info]
    print("fp32+polish", n, "C rel err", err, "iters", its, "oracle", r["iters"],
          "fp64 share", [i["iters_fp64"] for i in s.info])
    assert err < 1e-6
    eps_o, sig_o = o.fields()
    eps, sig = s.fields(5)
    assert field_rel(sig, sig_o) < 1e-5


def test_sphere65_fixture_fp64():
    ids = gen.sphere65()
    mat = np.array([[31e9, 0.3, 40e6, 100e6, 1, 8e-2, 8e2],   # example/3d/main.cpp:110-111 (E_0, v_0)
                    [30e9, 0.2, 250e6, 8006, 0.2, 8e-4, 8e3]])
    o, r = oracle_homog((65, 65, 65), mat, ids)
    s = m.FFT_Analysis3D(65, 65, 65, "e", precision="fp64")
    s.material(mat, ids)
    s.homogenization(1e-6, 1e-6)
    assert [i["iterations"] for i in s.info] == r["iters"]
    assert rel(s.CHomo, r["C"]) < 1e-9


@pytest.mark.parametrize("km,ki", [(25.0, 30.0), (30.0, 25.0), (1.0, 100.0)])
def test_thermal_homogenization_fp64(km, ki):
    n = 16
    ids = gen.spheres(n, 3, 0.2, seed=9)
    o = oracle.Thermal(n, n, n)
    o.material(km, ki, ids)
    r = o.homogenization(1e-6, 1e-6)
    s = m.FFT_analysis3d_thermal(n, n, n, precision="fp64")
    s.material(km, ki, ids)
    s.homogenization(1e-6, 1e-6)
    assert [i["iterations"] for i in s.info] == r["iters"]
    assert rel(s.KHomo, r["K"]) < 1e-9
    g_o, q_o = o.fields()
    g, q = s.fields(2)
    assert field_rel(q, q_o) < 1e-9


def test_thermal_nphase_fp32_polish():
    n = 32
    ids = gen.multi_spheres(n, [(2, 0.15, 2, 4), (0, 0.10, 2, 4)], seed=5)
    k = np.array([0.01, 1.0, 1000.0])
    o = oracle.Thermal(n, n, n)
    o.material_n(k, ids)
    r = o.homogenization(1e-7, 1e-7)
    s = m.FFT_analysis3d_thermal(n, n, n, precision="fp32_polish")
    s.material_n(k, ids)
    s.homogenization(1e-7, 1e-7)
    err = rel(s.KHomo, r["K"])
    print("thermal 3-phase fp32+polish K rel err", err, [i["iterations"] for i in s.info], r["iters"])
    assert err < 1e-6
