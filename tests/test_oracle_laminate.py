"""Stop decision of the oracle pinned by the laminate model (tests/laminate_model.py).

On layered RVEs the reference's accelerated iteration is a recursion on one
tensor per phase with closed-form Green operators.  The model predicts, per
load case and per iteration, the equilibrium measure eqVar, the relative
change of the field sum (Delta-sum) and the stop iteration from the
reference's own rules (medium, base, normalisation, stop rule).  The oracle,
which runs the full 3D FFT iteration, must reproduce every value and the
stop iteration of every load case.  A wrong reference medium (arithmetic for
geometric means), a wrong phase pair in the Delta-sum denominator (three-phase
cases, where phases 0 and 1 are not the extreme ones), a wrong base or a
wrong eqVar weight changes these sequences.
"""
import numpy as np
import pytest

import oracle
import laminate_model as lm

# (pattern along the axis, axis, grid shape).  Thermal flux-base grids keep
# Nx = Ny: the reference's flux-base Green table maps x through -(Ny - x)
# (FFT_thermal3D.cpp:150-151), which gives a zero frequency, and 0/0, at
# x = Ny when Ny < Nx.
X12 = ([1, 1, 1, 0, 0, 1, 1, 1, 1, 0, 1, 1], "x", (12, 3, 4))
Z10 = ([2, 2, 0, 1, 1, 1, 0, 2, 1, 1], "z", (4, 4, 10))
Z12 = ([1, 1, 0, 0, 0, 1, 1, 1, 1, 0, 1, 1], "z", (3, 3, 12))
Y9 = ([0, 2, 2, 1, 1, 1, 0, 1, 2], "y", (2, 9, 4))


def _split(hist, iters):
    out, i = [], 0
    for n in iters:
        out.append(hist[i:i + n])
        i += n
    assert i == len(hist)
    return out


def _compare(runs, hist, iters):
    """eqVar and Delta-sum are dimensionless: rounding of the 3D FFT shows at
    about 1e-16 of the fields, so 1e-12 is an absolute floor far below every
    tolerance used here"""
    assert [r["iters"] for r in runs] == list(iters)
    for r, h in zip(runs, _split(hist, iters)):
        assert np.allclose(h[:, 2], r["dsum"], rtol=1e-7, atol=1e-12), "Delta-sum sequence"
        assert np.allclose(h[:, 1], r["eqvar"], rtol=1e-7, atol=1e-12), "eqVar sequence"
        # the first iteration's Delta-eq compares with the previous load case's last eqVar
        assert np.allclose(h[1:, 3], r["deq"][1:], rtol=1e-6, atol=1e-12), "Delta-eq sequence"


def _robust(runs, tf, te):
    for r in runs:
        assert lm.margin(r, tf, te) > 1e-6, "a stop quantity sits within rounding of its tolerance"


@pytest.mark.parametrize("km,ki", [(1.0, 50.0), (50.0, 1.0), (25.0, 30.0)])
def test_thermal_two_phase_laminate(km, ki):
    pat, axis, shape = X12
    L = lm.Layers(pat, axis, shape)
    tf, te = 1e-6, 1e-6
    model = lm.solve_thermal(L, [ki, km], tf, te, two_phase=True)
    _robust(model["runs"], tf, te)
    o = oracle.Thermal(*shape, nthreads=2)
    o.material(km, ki, L.rve())
    inf = o.info()
    assert bool(inf["grad_base"]) == model["grad_base"]
    assert inf["k0"] == pytest.approx(model["k0"], rel=1e-15)
    assert model["grad_base"] == (km < ki)
    r = o.homogenization(tf, te)
    _compare(model["runs"], o.history(), r["iters"])


def test_thermal_three_phase_laminate():
    """N-phase extension: k0 = sqrt(kmin kmax) from phases 1 and 2, Delta-sum
    normalised by phases 0 and 1."""
    pat, axis, shape = Z10
    L = lm.Layers(pat, axis, shape)
    k = [5.0, 1.0, 40.0]
    tf, te = 1e-7, 1e-7
    model = lm.solve_thermal(L, k, tf, te)
    _robust(model["runs"], tf, te)
    o = oracle.Thermal(*shape, nthreads=2)
    o.material_n(k, L.rve())
    inf = o.info()
    assert bool(inf["grad_base"]) == model["grad_base"]
    assert inf["k0"] == pytest.approx(np.sqrt(40.0), rel=1e-15)
    # the normaliser 1/2 (k[0] + k[1]) = 3 differs from that of the extreme
    # phases (20.5): a wrong pair scales every Delta-sum by 6.8
    r = o.homogenization(tf, te)
    assert max(r["iters"]) > 10
    _compare(model["runs"], o.history(), r["iters"])


@pytest.mark.parametrize("case", ["strain", "stress"])
def test_elastic_two_phase_laminate(case):
    pat, axis, shape = Z12
    L = lm.Layers(pat, axis, shape)
    # phase 0 stiff: the minority (stress base) or, with the ids swapped, the majority (strain base)
    if case == "strain":
        L = lm.Layers([1 - p for p in pat], axis, shape)
    E, nu = [300e9, 3e9], [0.25, 0.35]
    tf, te = 1e-6, 1e-6
    model = lm.solve_elastic(L, E, nu, tf, te)
    assert model["strain_base"] == (case == "strain")
    _robust(model["runs"], tf, te)
    mat = np.array([[E[0], nu[0], np.inf, 0, 1, 0, 0], [E[1], nu[1], np.inf, 0, 1, 0, 0]])
    o = oracle.Elastic(*shape, "e", nthreads=2)
    o.material(mat, L.rve())
    inf = o.info()
    assert inf["strain_base"] == model["strain_base"]
    assert inf["la0"] == pytest.approx(model["la0"], rel=1e-14)
    assert inf["mu0"] == pytest.approx(model["mu0"], rel=1e-14)
    o.clear_history()
    r = o.homogenization(tf, te)
    assert max(r["iters"]) > 10
    _compare(model["runs"], o.history(), r["iters"])


@pytest.mark.parametrize("geom", ["z", "y"])
def test_elastic_three_phase_laminate(geom):
    """Three phases, phases 0 and 1 not the extreme ones; along y the
    mechanical check's kz >= 1 range sees no layer mode (eqVar is 0)."""
    pat, axis, shape = Z10 if geom == "z" else Y9
    L = lm.Layers(pat, axis, shape)
    E, nu = [70e9, 120e9, 400e9], [0.3, 0.22, 0.2]
    tf, te = 1e-7, 1e-7
    model = lm.solve_elastic(L, E, nu, tf, te)
    _robust(model["runs"], tf, te)
    mat = np.array([[E[p], nu[p], np.inf, 0, 1, 0, 0] for p in range(3)])
    o = oracle.Elastic(*shape, "e", nthreads=2)
    o.material(mat, L.rve())
    inf = o.info()
    assert inf["strain_base"] == model["strain_base"]
    assert inf["mu0"] == pytest.approx(model["mu0"], rel=1e-14)
    o.clear_history()
    r = o.homogenization(tf, te)
    if geom == "y":
        assert all(max(x["eqvar"]) == 0 for x in model["runs"])
    scale = max(max(max(x["eqvar"]) for x in model["runs"]), 1e-300)
    _compare(model["runs"], o.history(), r["iters"])


def test_model_fixed_point_is_the_exact_laminate():
    """The model's fixed point is the exact laminate: a gradient along the
    normal carries the series flux 1 / <1/k> in every layer, one across it
    the parallel flux <k> (both bases)."""
    pat, axis, shape = Z12
    L = lm.Layers(pat, axis, shape)
    f = L.frac
    for k in ([50.0, 1.0], [1.0, 50.0]):
        m = lm.solve_thermal(L, k, 1e-13, 1e-13)
        series = 1 / (f[0] / k[0] + f[1] / k[1])
        parallel = f[0] * k[0] + f[1] * k[1]
        for w, want in ((0, series), (1, parallel), (2, parallel)):
            r = m["runs"][w]
            assert r["iters"] < 500
            q = [r["q"][p][w] for p in range(2)]
            g = [r["u"][p][w] if m["grad_base"] else r["q"][p][w] / k[p] for p in range(2)]
            mean_q = f[0] * q[0] + f[1] * q[1]
            mean_g = f[0] * g[0] + f[1] * g[1]
            assert mean_q / mean_g == pytest.approx(want, rel=1e-9)
