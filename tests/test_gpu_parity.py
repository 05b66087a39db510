"""GPU parity: the HIP solver (libmat3d.so, through the C-ABI) against the FP64 oracle.

Bars: FP64 path -- same iteration counts, C_eff within 1e-9 relative, fields
within 1e-9 relative; FP32+FP64-polish path -- C_eff within 1e-6 relative
(the bar of the FP32 path), voxel stress within 1e-5 relative (max-abs over max-abs).
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
    assert its == r["iters"]
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


@pytest.mark.parametrize("case", ["elastic32", "elastic_generic", "thermal32"])
def test_anchored_polish_matches_fp64_polish(case):
    """FP32 + polish with the anchored FP32 correction iterations (refine=1,
    the default) against plain FP64 polish iterations (refine=0) and the
    oracle: same iteration counts, same C_eff / K_eff to FP32-switch level."""
    if case == "thermal32":
        n = 32
        shape = (n, n, n)
        ids = gen.multi_spheres(n, [(2, 0.15, 2, 4), (0, 0.10, 2, 4)], seed=5)
        k = np.array([0.01, 1.0, 1000.0])
        o = oracle.Thermal(*shape)
        o.material_n(k, ids)
        r = o.homogenization(1e-7, 1e-7)
        ref, iters_o = r["K"], r["iters"]
        out = {}
        for refine in (1, 0):
            s = m.FFT_analysis3d_thermal(*shape, precision="fp32_polish")
            s.set_option("refine", refine)
            s.material_n(k, ids)
            s.homogenization(1e-7, 1e-7)
            out[refine] = (s.KHomo.copy(), [i["iterations"] for i in s.info], s.fields(1)[1])
            s.close()
        q_o = None
    else:
        shape = (32, 32, 32) if case == "elastic32" else (12, 16, 10)
        ids = gen.fibres(32, 3, 0.3, seed=2, length=16)[: shape[0], : shape[1], : shape[2]].copy()
        mat = gen.MAT_C2
        o, r = oracle_homog(shape, mat, ids)
        ref, iters_o = r["C"], r["iters"]
        out = {}
        for refine in (1, 0):
            s = m.FFT_Analysis3D(*shape, "e", precision="fp32_polish")
            s.set_option("refine", refine)
            s.material(mat, ids)
            s.homogenization(1e-6, 1e-6)
            out[refine] = (s.CHomo.copy(), [i["iterations"] for i in s.info], s.fields(3)[1])
            s.close()
    for refine in (1, 0):
        C, its, _ = out[refine]
        print(f"[{case}] refine={refine} rel err {rel(C, ref):.3e} iters {its} oracle {iters_o}")
        assert rel(C, ref) < 1e-6
    assert out[1][1] == out[0][1]
    assert rel(out[1][0], out[0][0]) < 1e-6
    assert field_rel(out[1][2], out[0][2]) < 1e-5


@pytest.mark.parametrize("shape,precision", [((64, 64, 64), "fp64"), ((12, 16, 10), "fp64"), ((32, 32, 32), "fp32_polish")])
def test_basic_scheme_fixed_load_matches_oracle(shape, precision):
    """Basic Moulinec-Suquet scheme with the unit loads held fixed ("ep"
    reference medium: arithmetic means, strain base), scheme option 1."""
    n = max(shape)
    ids = gen.spheres(n, 6 if n >= 64 else 3, 0.2, seed=1)[: shape[0], : shape[1], : shape[2]].copy()
    mat = gen.MAT_C1
    o = oracle.Elastic(*shape, "ep")
    o.set_scheme(1)
    o.material(mat, ids)
    r = o.homogenization(1e-6, 1e-6)
    s = m.FFT_Analysis3D(*shape, "ep", precision=precision)
    s.set_option("scheme", 1)
    s.material(mat, ids)
    s.homogenization(1e-6, 1e-6)
    its = [i["iterations"] for i in s.info]
    err = rel(s.CHomo, r["C"])
    print(f"[basic {shape} {precision}] iters {its} oracle {r['iters']} C rel err {err:.3e}")
    if precision == "fp64":
        assert its == r["iters"]
        assert err < 1e-9
    else:
        assert err < 1e-6
    _, sig_o = o.fields()
    _, sig = s.fields(5)
    assert field_rel(sig, sig_o) < (1e-9 if precision == "fp64" else 1e-5)


# ---------------------------------------------------------------------------
# compile-time specialised kernels: fixed iteration counts against the oracle
def _fixed_iters_elastic(shape, ids, mat, w, k, precision):
    o = oracle.Elastic(*shape, "e")
    o.material(mat, ids)
    o.run_fixed(w, k)
    eps_o, sig_o = o.fields()
    sb = o.info()["strain_base"]
    o.close()
    s = m.FFT_Analysis3D(*shape, "e", precision=precision)
    s.material(mat, ids)
    s.set_option("cap", k - 1)
    E = np.zeros(6)
    E[w] = 0.0005 if sb else 1e6
    s.setLoad(E, E)
    info = s.runAnalysis(-1.0, -1.0)
    assert info["iterations"] == k and info["status"] == 2
    eps, sig = s.fields()
    return field_rel(eps, eps_o), field_rel(sig, sig_o)


@pytest.mark.parametrize("shape,w,k", [((64, 64, 64), 0, 6), ((96, 96, 96), 3, 4), ((128, 128, 128), 5, 3),
                                       ((192, 192, 192), 1, 2), ((256, 256, 256), 2, 2), ((32, 64, 128), 4, 5),
                                       ((48, 48, 48), 0, 6)])
def test_fast_kernels_fp64_fixed_iterations(shape, w, k):
    n = max(shape)
    ids = gen.fibres(n, 4, 0.3, seed=2, length=min(64, n // 2))[: shape[0], : shape[1], : shape[2]].copy()
    e_eps, e_sig = _fixed_iters_elastic(shape, ids, gen.MAT_C2, w, k, "fp64")
    assert e_eps < 1e-9 and e_sig < 1e-9, (e_eps, e_sig)


@pytest.mark.parametrize("shape,w,k", [((64, 64, 64), 0, 6), ((128, 128, 128), 5, 3), ((256, 256, 256), 2, 2)])
def test_fast_kernels_fp32_fixed_iterations(shape, w, k):
    """FP32 fields/FFTs (no polish: the cap ends the run in the FP32 phase):
    bounded by FP32 round-off of the transforms."""
    n = max(shape)
    ids = gen.fibres(n, 4, 0.3, seed=2, length=min(64, n // 2))
    e_eps, e_sig = _fixed_iters_elastic(shape, ids, gen.MAT_C2, w, k, "fp32_polish")
    assert e_eps < 2e-5 and e_sig < 2e-5, (e_eps, e_sig)


@pytest.mark.parametrize("n,precision", [(64, "fp64"), (128, "fp64"), (128, "fp32_polish")])
def test_fast_kernels_thermal_fixed_iterations(n, precision):
    ids = gen.multi_spheres(n, [(2, 0.15, 3, 6), (0, 0.10, 3, 6)], seed=5)
    k3 = np.array([0.01, 1.0, 1000.0])
    k = 3
    o = oracle.Thermal(n, n, n)
    o.material_n(k3, ids)
    o.set_cap(k - 1)
    o.set_load(np.array([0.0, 1.0, 0.0]))
    o.run(-1.0, -1.0)
    g_o, q_o = o.fields()
    o.close()
    s = m.FFT_analysis3d_thermal(n, n, n, precision=precision)
    s.material_n(k3, ids)
    s.set_option("cap", k - 1)
    s.setLoad([0.0, 1.0, 0.0])
    info = s.runAnalysis(-1.0, -1.0)
    assert info["iterations"] == k
    g, q = s.fields()
    tol = 1e-9 if precision == "fp64" else 2e-5
    assert field_rel(g, g_o) < tol and field_rel(q, q_o) < tol


# ---------------------------------------------------------------------------
# size-independent properties at full size: exact laminate solutions
@pytest.mark.parametrize("n,precision", [(512, "fp32_polish"), (256, "fp64")])
def test_thermal_laminate_exact_full_size(n, precision):
    ids = np.ones((n, n, n), dtype=np.int32)
    ids[: n // 4] = 0  # layers normal to x (tensor axis 2)
    km, ki = 1.0, 20.0
    s = m.FFT_analysis3d_thermal(n, n, n, precision=precision)
    s.material(km, ki, ids)
    s.homogenization(1e-7, 1e-7)
    print("laminate", n, precision, [(i["iterations"], i["status"]) for i in s.info], s.KHomo.diagonal())
    assert all(i["status"] == 1 for i in s.info)
    f = 0.25
    par, ser = f * ki + (1 - f) * km, 1 / (f / ki + (1 - f) / km)
    K = s.KHomo
    assert abs(K[0, 0] - par) < 1e-6 * par and abs(K[1, 1] - par) < 1e-6 * par
    assert abs(K[2, 2] - ser) < 1e-6 * ser


def test_elastic_laminate_exact_256_fp32_polish():
    from test_oracle import laminate_exact

    n = 256
    ids = np.ones((n, n, n), dtype=np.int32)
    ids[:, :, : 3 * n // 8] = 0  # layers normal to z (tensor axis 0)
    mat = gen.MAT_C2
    s = m.FFT_Analysis3D(n, n, n, "e", precision="fp32_polish")
    s.material(mat, ids)
    # the stop rule bounds the change per iteration, not the distance to the
    # fixed point: a tight tolerance makes the converged iterate exact to 1e-6
    s.homogenization(1e-10, 1e-10)
    assert all(i["status"] == 1 for i in s.info)
    print("elastic laminate 256", [(i["iterations"], i["status"]) for i in s.info])
    f0 = 3 / 8
    sb = s.strainBase
    ep = np.zeros((6, 6)); sg = np.zeros((6, 6))
    for w in range(6):
        if sb:
            E = np.eye(6)[w] * 0.0005
        else:
            Cl = np.stack([laminate_exact(mat[:, :2], [f0, 1 - f0], 0, b, sb)[1] for b in np.eye(6)], 1)
            E = np.linalg.solve(Cl, np.eye(6)[w] * 1e6)
        e, sgm = laminate_exact(mat[:, :2], [f0, 1 - f0], 0, E, sb)
        ep[:, w] = e
        sg[w, :] = sgm
    Cref = np.linalg.inv(ep) @ sg if sb else np.linalg.inv(np.linalg.inv(sg) @ ep)
    assert rel(s.CHomo, Cref) < 1e-6


def test_thermal_three_phase_laminate_512_material_n():
    """The 512^3 conduction workload's phases through material_n: pores
    k = 0.01 (phase 0, 10%), matrix k = 1 (phase 1, 75%), inclusions
    k = 1000 (phase 2, 15%), here as layers normal to x.  The laminate model
    (tests/laminate_model.py) predicts the base, the stop iteration of every
    load case and the exact series/parallel K_eff; FP32 + polish must stop
    where the FP64 rule does."""
    import laminate_model as lm

    n = 512
    pat = np.ones(n, dtype=np.int32)
    pat[:77] = 2
    pat[200:251] = 0
    L = lm.Layers(pat, "x", (n, n, n))
    k = np.array([0.01, 1.0, 1000.0])
    tol = 1e-7
    model = lm.solve_thermal(L, k, tol, tol)
    assert all(lm.margin(r, tol, tol) > 1e-3 for r in model["runs"])
    s = m.FFT_analysis3d_thermal(n, n, n, precision="fp32_polish")
    s.material_n(k, L.rve())
    s.homogenization(tol, tol)
    its = [i["iterations"] for i in s.info]
    print("3-phase laminate 512", its, [r["iters"] for r in model["runs"]], s.KHomo.diagonal())
    assert its == [r["iters"] for r in model["runs"]]
    assert all(i["status"] == 1 for i in s.info)
    f = L.frac
    K = s.KHomo
    # the model's own K at its stop iterate: <q>/<grad T> of each load case
    # (the stop rule bounds the change per iteration, not the distance to
    # the fixed point: at contrast 1e5 the in-plane iterate stops 1.6e-4 from
    # the exact parallel value <k>)
    for w in range(3):
        r = model["runs"][w]
        q = np.array([r["q"][p][w] for p in range(3)])
        g = np.array([r["u"][p][w] if model["grad_base"] else r["q"][p][w] / k[p] for p in range(3)])
        Km = float(f @ q) / float(f @ g)
        assert abs(K[w, w] - Km) < 1e-6 * Km, (w, K[w, w], Km)
    # the load along the normal (tensor index 2 is x) is exact after 2 iterations: the series value
    ser = float(1 / (f @ (1 / k)))
    assert abs(K[2, 2] - ser) < 1e-6 * ser


@pytest.mark.parametrize("kind", ["elastic", "thermal"])
def test_fp32_phase_reaching_the_cap_returns_the_fp64_iterate(kind):
    """A solve whose FP32 phase reaches the iteration cap is finished in FP64:
    its last iterations are redone as polish iterates, so the result is the
    FP64 iterate at the cap that the reference returns
    (FFT_analysis3D.cpp:785-792, FFT_thermal3D.cpp:474-480)."""
    n = 32
    if kind == "elastic":
        ids = gen.fibres(n, 3, 0.3, seed=2, length=n // 2)
        cap, tol = 12, 1e-6
        o = oracle.Elastic(n, n, n, "e")
        o.set_cap(cap)
        o.material(gen.MAT_C2, ids)
        r = o.homogenization(tol, tol)
        ref, oi = r["C"], r["iters"]
        s = m.FFT_Analysis3D(n, n, n, "e", precision="fp32_polish")
        s.set_option("cap", cap)
        s.material(gen.MAT_C2, ids)
        s.homogenization(tol, tol)
        got = s.CHomo
    else:
        ids = gen.multi_spheres(n, [(2, 0.15, 2, 4), (0, 0.10, 2, 4)], seed=5)
        k = np.array([0.01, 1.0, 1000.0])
        cap, tol = 40, 1e-7
        o = oracle.Thermal(n, n, n)
        o.set_cap(cap)
        o.material_n(k, ids)
        r = o.homogenization(tol, tol)
        ref, oi = r["K"], r["iters"]
        s = m.FFT_analysis3d_thermal(n, n, n, precision="fp32_polish")
        s.set_option("cap", cap)
        s.material_n(k, ids)
        s.homogenization(tol, tol)
        got = s.KHomo
    its = [i["iterations"] for i in s.info]
    print(kind, "capped", its, oi, [i["iters_fp64"] for i in s.info], rel(got, ref))
    assert oi == [cap + 1] * len(oi)  # every load case of the oracle reaches the cap
    assert its == oi
    assert all(i["status"] == 2 and i["iters_fp64"] > 0 for i in s.info)
    assert rel(got, ref) < 1e-6
