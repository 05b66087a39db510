"""FFT_Analysis2D on the GPU (m3d_2d_*) against the FP64 oracle: fields after
runAnalysis and CHomo after homogenization on square and non-square grids
(the latter exercises the reference's x-frequency map, which uses Ny), the
exact laminate stiffness at 128^2, and the reference's example call
sequence (material, setLoad, runAnalysis)."""
import numpy as np
import pytest

import mat3d_amd as m
import oracle

pytestmark = pytest.mark.gpu

Q = [68.9e9, 0.35, 400e9, 0.23]   # example/2d/main.cpp:33-37


def _rve(Nx, Ny, seed):
    rng = np.random.default_rng(seed)
    x, y = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")
    r = np.ones((Nx, Ny))
    for _ in range(8):
        cx, cy, rad = rng.uniform(0, Nx), rng.uniform(0, Ny), rng.uniform(3, 8)
        dx = np.minimum(np.abs(x - cx), Nx - np.abs(x - cx))
        dy = np.minimum(np.abs(y - cy), Ny - np.abs(y - cy))
        r[dx * dx + dy * dy <= rad * rad] = 0
    return r.reshape(-1)


@pytest.mark.parametrize("Nx,Ny", [(64, 64), (32, 96), (128, 48)])
def test_run_and_homogenization_match_oracle(Nx, Ny):
    rve = _rve(Nx, Ny, Nx + Ny)
    E = np.array([0, 0, 0.005], dtype=complex)   # example/2d/main.cpp:28-29
    o = oracle.Elastic2D(Nx, Ny)
    o.material(Q, rve)
    o.set_load(E)
    o.run()
    eo, so = o.fields()
    g = m.FFT_Analysis2D(Nx, Ny)
    g.material(Q, rve)
    g.setLoad(E)
    g.runAnalysis()
    eg, sg = g.fields()
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()
    print(f"[2d {Nx}x{Ny}] eps {rel(eg, eo):.2e} sigma {rel(sg, so):.2e}")
    assert g.itNo == 25
    assert rel(eg, eo) < 1e-11 and rel(sg, so) < 1e-11
    Co, _, _ = o.homogenization()
    Cg = g.homogenization()
    print(f"[2d {Nx}x{Ny}] CHomo {rel(Cg, Co):.2e}")
    assert rel(Cg, Co) < 1e-11


def test_laminate_exact_at_128():
    from test_oracle_2d import C2, lame, laminate2

    n = 128
    rve = np.ones((n, n))
    rve[:, :48] = 0          # layers normal to y
    q = [70e9, 0.3, 140e9, 0.25]
    g = m.FFT_Analysis2D(n, n)
    g.material(q, rve)
    g.setLoad([0, 0, 0])
    C = g.homogenization()
    f0 = 48 / n
    ref = laminate2([C2(*lame(140e9, 0.25)), C2(*lame(70e9, 0.3))], [f0, 1 - f0], 1)
    assert np.abs(C - ref).max() / np.abs(ref).max() < 1e-9
