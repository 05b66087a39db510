"""The oracle against the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py)."""
import os

import numpy as np

import oracle

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_green21_fixture():
    f = load("green21.npz")
    for z, gs, gd in zip(f["zets"], f["strain"], f["stress"]):
        assert np.allclose(oracle.green21(z, float(f["la0"]), float(f["mu0"]), True), gs, rtol=1e-14, atol=0)
        assert np.allclose(oracle.green21(z, float(f["la0"]), float(f["mu0"]), False), gd, rtol=1e-14, atol=1e-30)


def test_rfft3_fixture():
    f = load("rfft3.npz")
    X = oracle.rfft3(f["x"], 2)
    assert np.abs(X - f["X"]).max() < 1e-12 * np.abs(f["X"]).max()


def test_elastic12_fixture():
    f = load("elastic12.npz")
    n = 12
    for tag in ("stress", "strain"):
        o = oracle.Elastic(n, n, n, "e", nthreads=3)
        o.material(f[f"{tag}_mat"], f["ids"])
        r = o.homogenization(1e-6, 1e-6)
        assert r["iters"] == list(f[f"{tag}_iters"])
        assert np.abs(r["C"] - f[f"{tag}_C"]).max() < 1e-10 * np.abs(f[f"{tag}_C"]).max()


def test_thermal12_fixture():
    f = load("thermal12.npz")
    n = 12
    for tag in ("grad", "flux"):
        km, ki = f[f"{tag}_k"]
        t = oracle.Thermal(n, n, n, nthreads=3)
        t.material(km, ki, f["ids"])
        r = t.homogenization(1e-6, 1e-6)
        assert r["iters"] == list(f[f"{tag}_iters"])
        assert np.abs(r["K"] - f[f"{tag}_K"]).max() < 1e-10 * np.abs(f[f"{tag}_K"]).max()


def test_j2_fixture():
    f = load("j2_8.npz")
    o = oracle.Elastic(8, 8, 8, "ep", nthreads=2)
    o.material(f["mat"], f["ids"])
    o.set_load(f["E"], f["X0"])
    o.run(1e-6, 1e-6, int(f["nInc"]))
    P = o.prop_homogenized()
    assert np.abs(P - f["prop"]).max() <= 1e-9 * np.abs(f["prop"]).max()
