"""runAnalysis "ep" / "ed" (basic scheme with mixed control, J2 and damage)
on the GPU against the FP64 oracle, the golden fixture and the uniaxial J2
closed form.  Bar: prop_homogenized and fields within 1e-9 relative, same
iteration counts."""
import os

import numpy as np
import pytest

import j2_model as j2m
import mat3d_amd as m
import oracle
from mat3d_amd import gen

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

MAT_ED = np.array([[400e9, 0.20, np.inf, 0, 1, 1e-3, 2e3],
                   [70e9, 0.30, np.inf, 0, 1, 1e-3, 8e2]])


def run_both(shape, behav, mat, ids, E, X0, nInc, tol=1e-6, precision="fp64"):
    o = oracle.Elastic(*shape, behav)
    o.material(mat, ids)
    o.set_load(E, X0)
    it_o = o.run(tol, tol, nInc)
    P_o = o.prop_homogenized()
    eps_o, sig_o = o.fields()
    s = m.FFT_Analysis3D(*shape, behav, precision=precision)
    s.material(mat, ids)
    s.setLoad(E, X0)
    info = s.runAnalysis(tol, tol, nInc)
    eps, sig = s.fields()
    return (P_o, eps_o, sig_o, it_o), (s.prop_homogenized, eps, sig, info)


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("shape", [(16, 16, 16), (12, 16, 10), (9, 9, 9)])
def test_j2_matches_oracle(shape):
    n = max(shape)
    ids = gen.spheres(n, 3, 0.25, seed=4)[: shape[0], : shape[1], : shape[2]].copy()
    X0 = np.eye(6)[0]
    (P_o, e_o, s_o, it_o), (P, e, s, info) = run_both(shape, "ep", gen.MAT_C4, ids, 0.02 * X0, X0, 8)
    print("j2", shape, "iters oracle", it_o, "gpu", info["iters_fp64"], "last sigma", P[-1, 1])
    assert info["iters_fp64"] == it_o
    assert P.shape == P_o.shape
    assert rel(P, P_o) < 1e-9
    assert rel(s, s_o) < 1e-9 and rel(e, e_o) < 1e-9


@pytest.mark.parametrize("shape,behav", [((32, 32, 32), "ep"), ((32, 64, 48), "ep"), ((32, 32, 32), "ed")])
def test_specialised_kernels_match_oracle(shape, behav):
    """Grids whose line lengths have the compile-time specialised basic-scheme
    kernels (the tests above run the runtime-planned ones)."""
    n = max(shape)
    ids = gen.spheres(n, 4, 0.25, seed=4)[: shape[0], : shape[1], : shape[2]].copy()
    X0 = np.eye(6)[0] if behav == "ep" else np.eye(6)[1]
    mat = gen.MAT_C4 if behav == "ep" else MAT_ED
    E = 0.02 * X0 if behav == "ep" else 0.01 * X0
    (P_o, e_o, s_o, it_o), (P, e, s, info) = run_both(shape, behav, mat, ids, E, X0, 4)
    print(f"[basic fast {shape} {behav}] iters oracle {it_o} gpu {info['iters_fp64']} last row {P[-1, :3]}")
    assert info["iters_fp64"] == it_o
    assert P.shape == P_o.shape
    assert rel(P, P_o) < 1e-9
    # past the damage softening (thousands of iterations) FFT rounding
    # differences grow to ~1e-9 in the fields
    ftol = 1e-9 if behav == "ep" else 1e-7
    assert rel(s, s_o) < ftol and rel(e, e_o) < ftol


@pytest.mark.parametrize("shape,pf,ptol,ftol", [((32, 32, 32), 1000, 2e-7, 1e-5), ((32, 32, 32), 10, 3e-6, 1e-4),
                                                 ((32, 64, 48), 10, 3e-6, 1e-4)])
def test_j2_fp32_phase_against_oracle(shape, pf, ptol, ftol):
    """Opt-in FP32 phase of "ep" (option basic_fp32 on an FP32+polish handle):
    each increment iterates in FP32 (fields, FFTs, return map) to pf x tol,
    then in FP64 under the reference rule from the FP64 recomputation of the
    switch iterate; the history stays FP64.  The early FP32 rounding stays in
    the plastic state, so the fields meet 1e-5 of the FP64 oracle only with
    a small FP32 share (pf 1000); at pf 10 they are within 1e-4.  The FP32
    and FP64 iterations add up to the oracle's count."""
    n = max(shape)
    ids = gen.spheres(n, 4, 0.25, seed=4)[: shape[0], : shape[1], : shape[2]].copy()
    X0 = np.eye(6)[0]
    o = oracle.Elastic(*shape, "ep")
    o.material(gen.MAT_C4, ids)
    o.set_load(0.02 * X0, X0)
    it_o = o.run(1e-6, 1e-6, 4)
    P_o = o.prop_homogenized()
    e_o, s_o = o.fields()
    s = m.FFT_Analysis3D(*shape, "ep", precision="fp32_polish")
    s.set_option("basic_fp32", 1)
    s.set_option("polish_factor", pf)
    s.material(gen.MAT_C4, ids)
    s.setLoad(0.02 * X0, X0)
    info = s.runAnalysis(1e-6, 1e-6, 4)
    e, sg = s.fields()
    P = s.prop_homogenized
    print(f"[j2 fp32 phase {shape} pf {pf}] iters oracle {it_o} gpu fp32 {info['iters_fp32']} fp64 {info['iters_fp64']} "
          f"prop {rel(P, P_o):.2e} sigma {rel(sg, s_o):.2e} eps {rel(e, e_o):.2e}")
    assert info["iters_fp32"] > 0
    assert abs(info["iters_fp32"] + info["iters_fp64"] - it_o) <= 4
    assert P.shape == P_o.shape
    assert rel(P, P_o) < ptol
    assert rel(sg, s_o) < ftol and rel(e, e_o) < ftol


def test_basic_fp32_handle_defaults_to_fp64():
    """Without the opt-in an FP32+polish handle runs "ep"/"ed" in FP64: the
    same iterations and results as an FP64 handle."""
    n = 32
    ids = gen.spheres(n, 4, 0.25, seed=4)
    X0 = np.eye(6)[1]
    out = []
    for prec in ("fp64", "fp32_polish"):
        s = m.FFT_Analysis3D(n, n, n, "ed", precision=prec)
        s.material(MAT_ED, ids)
        s.setLoad(0.01 * X0, X0)
        out.append((s.runAnalysis(1e-6, 1e-6, 2), s.prop_homogenized.copy()))
    assert out[1][0]["iters_fp32"] == 0 and out[1][0]["iters_fp64"] == out[0][0]["iters_fp64"]
    assert np.array_equal(out[0][1], out[1][1])


def test_damage_matches_oracle():
    n = 16
    ids = gen.spheres(n, 3, 0.25, seed=6)
    X0 = np.eye(6)[1]
    (P_o, e_o, s_o, it_o), (P, e, s, info) = run_both((n, n, n), "ed", MAT_ED, ids, 0.01 * X0, X0, 6)
    print("damage iters", it_o, info["iters_fp64"], "d mean", P[:, 7])
    assert info["iters_fp64"] == it_o
    assert P.shape == P_o.shape
    assert rel(P, P_o) < 1e-9
    assert rel(s, s_o) < 1e-9


def test_j2_golden_fixture():
    f = np.load(os.path.join(G, "j2_8.npz"), allow_pickle=False)
    s = m.FFT_Analysis3D(8, 8, 8, "ep", precision="fp64")
    s.material(f["mat"], f["ids"])
    s.setLoad(f["E"], f["X0"])
    s.runAnalysis(1e-6, 1e-6, int(f["nInc"]))
    assert rel(s.prop_homogenized, f["prop"]) < 1e-9


def test_j2_uniaxial_closed_form_gpu():
    n = 32
    ids = np.ones((n, n, n), dtype=np.int32)
    ids[0, 0, 0] = 0
    E, nu, sy, H = 70e9, 0.3, 100e6, 1e9
    mat = np.array([[E, nu, sy, H, 1, 0, 0], [E, nu, sy, H, 1, 0, 0]])
    s = m.FFT_Analysis3D(n, n, n, "ep", precision="fp64")
    s.material(mat, ids)
    X0 = np.eye(6)[0]
    s.setLoad(0.004 * X0, X0)
    s.runAnalysis(1e-10, 1e-10, 20)
    for eps, s0 in s.prop_homogenized[:, :2]:
        ref = E * eps if s0 <= sy else (eps + sy / H) / (1 / E + 1 / H)
        assert abs(s0 - ref) < 1e-6 * max(abs(ref), sy)


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 32, 32)])
def test_j2_power_law_matches_oracle(shape):
    """The reference CLI's material table (example/3d/main.cpp:107-108):
    matrix with power-law hardening n = 0.2, inclusions with linear
    hardening, both yielding.  16^3 runs the runtime-planned basic kernels,
    32^3 the specialised ones; both run the Newton branch of the return map."""
    n = shape[0]
    ids = gen.spheres(n, 3, 0.2, seed=11)
    X0 = np.eye(6)[1]
    (P_o, e_o, s_o, it_o), (P, e, s, info) = run_both(shape, "ep", j2m.MAT_REF_TABLE, ids, 0.02 * X0, X0, 6)
    print(f"[j2 n=0.2 {shape}] iters oracle {it_o} gpu {info['iters_fp64']} last row {P[-1, :3]}")
    assert it_o > 0
    assert info["iters_fp64"] == it_o
    assert P.shape == P_o.shape
    # axial stress (load along tensor axis 1: column 2): the macroscopic tangent has dropped, both phases yield
    assert P[-1, 2] - P[-2, 2] < 0.5 * (P[1, 2] - P[0, 2])
    assert rel(P, P_o) < 1e-9
    assert rel(s, s_o) < 1e-9 and rel(e, e_o) < 1e-9


@pytest.mark.parametrize("n,case", [(32, 0), (32, 1), (192, 0), (192, 1)])
def test_j2_uniaxial_power_law_gpu(n, case):
    """Homogeneous "ep" RVE, hardening exponent n = 0.2, uniaxial stress
    control: every prop_homogenized row against the scalar root of the flow
    rule (tests/j2_model.py).  192^3 is the elastoplastic workload's grid,
    on its specialised basic-scheme kernels."""
    E, nu, sy, H, nexp, emax = j2m.POWER_LAW_CASES[case]
    ids = np.ones((n, n, n), dtype=np.int32)
    ids[0, 0, 0] = 0
    mat = np.array([[E, nu, sy, H, nexp, 0, 0]] * 2)
    s = m.FFT_Analysis3D(n, n, n, "ep", precision="fp64")
    s.material(mat, ids)
    X0 = np.eye(6)[0]
    s.setLoad(emax * X0, X0)
    s.runAnalysis(1e-10, 1e-10, 20)
    assert s.prop_homogenized.shape == (20, 9)
    assert j2m.check_uniaxial_rows(s.prop_homogenized, E, sy, H, nexp) >= 10


def test_j2_uniaxial_linear_192():
    """Linear hardening (n = 1) at 192^3, the elastoplastic workload's grid:
    sigma = E eps below yield, (eps + sy/H) / (1/E + 1/H) above."""
    n = 192
    ids = np.ones((n, n, n), dtype=np.int32)
    ids[0, 0, 0] = 0
    E, nu, sy, H = 70e9, 0.3, 100e6, 1e9
    mat = np.array([[E, nu, sy, H, 1, 0, 0], [E, nu, sy, H, 1, 0, 0]])
    s = m.FFT_Analysis3D(n, n, n, "ep", precision="fp64")
    s.material(mat, ids)
    X0 = np.eye(6)[0]
    s.setLoad(0.004 * X0, X0)
    s.runAnalysis(1e-10, 1e-10, 20)
    assert j2m.check_uniaxial_rows(s.prop_homogenized, E, sy, H, 1.0) >= 10


def test_j2_c4_microstructure_matches_oracle():
    """The elastoplastic workload's first ten increments (spheres of radius 8
    at 25%, E = 0.02 e_1 over 50 increments: 0.004 e_1 over 10) on a 48^3
    grid against the oracle: same iterations, prop_homogenized and fields
    within 1e-9."""
    n = 48
    ids = gen.spheres(n, 8, 0.25, seed=4)
    X0 = np.eye(6)[0]
    (P_o, e_o, s_o, it_o), (P, e, s, info) = run_both((n, n, n), "ep", gen.MAT_C4, ids, 0.004 * X0, X0, 10)
    print(f"[j2 C4 48^3] iters oracle {it_o} gpu {info['iters_fp64']} last row {P[-1, :3]}")
    assert info["iters_fp64"] == it_o
    assert P[-1, 1] - P[-2, 1] < 0.5 * (P[1, 1] - P[0, 1])  # the macroscopic tangent has dropped: yielding
    assert rel(P, P_o) < 1e-9
    assert rel(s, s_o) < 1e-9 and rel(e, e_o) < 1e-9


def test_damage_uniaxial_closed_form_gpu():
    """Homogeneous "ed" RVE, uniaxial stress control: d = 1 - exp(-P (sqrt(E) eps - Y0)),
    sigma = (1 - d) E eps."""
    n = 32
    ids = np.ones((n, n, n), dtype=np.int32)
    ids[0, 0, 0] = 0
    E, nu, Pd, Y0 = 70e9, 0.3, 5e-4, 500.0
    mat = np.array([[E, nu, np.inf, 0, 1, Pd, Y0], [E, nu, np.inf, 0, 1, Pd, Y0]])
    s = m.FFT_Analysis3D(n, n, n, "ed", precision="fp64")
    s.material(mat, ids)
    X0 = np.eye(6)[0]
    s.setLoad(0.004 * X0, X0)
    s.runAnalysis(1e-10, 1e-10, 10)
    P = s.prop_homogenized
    assert P.shape == (10, 9)
    for eps, s0, d in P[:, [0, 1, 7]]:
        Y = np.sqrt(E) * eps
        dref = 1 - np.exp(-Pd * (Y - Y0)) if Y > Y0 else 0.0
        assert abs(d - dref) < 1e-8
        assert abs(s0 - (1 - dref) * E * eps) < 1e-6 * abs((1 - dref) * E * eps)
