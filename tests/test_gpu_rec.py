"""recPlan / runRec on the GPU (m3d_rec_*) against the FP64 oracle: the
fixed point, the iteration through the relaxed projection (j < 100), the
re-randomisations and the clamp phase (j >= 100), and the volume-fraction
search.

The iteration is chaotic: the projection is discontinuous at 0 and 1, and a
rounding difference (the device's atan2 / cos / sin and FFT summation order
against the host's) grows about tenfold per iteration after the first ten
(measured: 1e-15 at iteration 10, 1e-3 by iteration 25).  The reference's own
output would move the same way under another FFT library or libm.  So the
trajectories are compared to 1e-9 over the first 12 iterations, and long runs
through their statistics: final error, volume fraction, and status."""
import numpy as np
import pytest

import mat3d_amd as m
import oracle
from mat3d_amd import gen

pytestmark = pytest.mark.gpu


def _rve(shape, seed=4):
    n = max(shape)
    return (gen.spheres(n, 4, 0.25, seed=seed)[: shape[0], : shape[1], : shape[2]] == 0).astype(float)


def _tpcf(rve):
    return np.abs(np.fft.fftn(rve)) ** 2 / rve.size


def _both(rve, n, beta, seed):
    t = _tpcf(rve)
    o = oracle.Rec(t)
    it_o = o.run(n, beta, seed)
    m_o, e_o = o.result()
    g = m.recPlan(t)
    it_g = g.runRec(n, beta, seed)
    _both.search = (o.search_steps, g.search_steps)
    return (it_o, o.status, m_o, e_o), (it_g, g.status, g.mTilda, g.trackrealerrorA)


def test_true_structure_is_a_fixed_point():
    rve = _rve((32, 32, 32))
    g = m.recPlan(_tpcf(rve))
    g.set_initial(rve)
    assert g.runRec(30, 0.5, 3) == 1 and g.status == 1
    assert g.trackrealerrorA[0] < 1e-10
    assert np.array_equal(g.mTilda.real, rve) and g.search_steps == 0


@pytest.mark.parametrize("shape,beta", [((32, 32, 32), 0.8), ((32, 64, 48), 0.5)])
def test_relaxed_projection_matches_oracle(shape, beta):
    rve = _rve(shape)
    (it_o, st_o, m_o, e_o), (it_g, st_g, m_g, e_g) = _both(rve, 12, beta, 7)
    dev = np.abs(e_g - e_o) / e_o
    agree = float((m_g.real == m_o.real).mean())
    print(f"[rec {shape}] iterations {it_o}/{it_g} err {e_o[-1]:.9e} / {e_g[-1]:.9e} max rel dev {dev.max():.2e} "
          f"voxels agreeing {agree:.6f}")
    assert (it_g, st_g) == (it_o, st_o) == (12, 2)
    assert dev.max() < 1e-9
    assert agree > 0.9999


def test_long_run_statistics_match_oracle():
    """320 iterations: re-randomisations at iterations 150 and 300, the
    clamp phase from j = 100 and the reset of j after 200.  The trajectories
    have decorrelated by then, so the GPU runs of three seeds are held to the
    oracle's seed-to-seed spread: the mean error of the last 50 iterations
    within the oracle's range widened by 15%, and every result binary with
    the volume fraction of the structure (within the search's 1e-4, unless
    the 0.001 threshold steps jump over that band until the 201-step limit;
    then within 2e-3)."""
    rve = _rve((32, 32, 32), seed=5)
    t = _tpcf(rve)
    tails_o, tails_g = [], []
    for seed in (9, 10, 11):
        (it_o, st_o, m_o, e_o), (it_g, st_g, m_g, e_g) = _both(rve, 320, 0.8, seed)
        assert (it_g, st_g) == (it_o, st_o) == (320, 2)
        assert abs(e_g[:12] / e_o[:12] - 1).max() < 1e-9
        tails_o.append(e_o[-50:].mean())
        tails_g.append(e_g[-50:].mean())
        for mm, steps in zip((m_o, m_g), _both.search):
            assert set(np.unique(mm.real)) <= {0.0, 1.0}
            d = abs(mm.real.mean() - rve.mean())
            assert d < 1e-4 or (steps == 201 and d < 2e-3)
    print(f"[rec long] mean error of the last 50 iterations, oracle {np.round(tails_o, 4)} gpu {np.round(tails_g, 4)}")
    lo, hi = min(tails_o) / 1.15, max(tails_o) * 1.15
    assert all(lo <= v <= hi for v in tails_g)
