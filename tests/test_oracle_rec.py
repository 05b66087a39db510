"""The oracle's restatement of recPlan / runRec (src/rec/reconstruction.cpp)
pinned on the CPU: its 3D complex FFT against numpy, its random stream's
mixing function against the published splitmix64 output, the fixed point of
the iteration, a hand-computed one-voxel run through both branches of the
relaxed projection, and the volume-fraction search."""
import numpy as np

import oracle
from mat3d_amd import gen


def test_c2c3_matches_numpy():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((6, 8, 10)) + 1j * rng.standard_normal((6, 8, 10))
    assert np.abs(oracle.c2c3(x, -1) - np.fft.fftn(x)).max() < 1e-12
    assert np.abs(oracle.c2c3(x, 1) - np.fft.ifftn(x) * x.size).max() < 1e-12


def test_random_stream_mixing_is_splitmix64():
    # splitmix64 seeded with 1234567: its first three outputs as published
    # with the reference implementation (state advanced by the golden gamma)
    g = 0x9E3779B97F4A7C15
    out = [oracle.rec_mix((1234567 + k * g) % 2**64) for k in range(3)]
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423]
    u = [oracle.rec_uniform(5, k, i) for k in range(3) for i in range(1000)]
    assert min(u) >= -1.0 and max(u) < 1.0 and abs(np.mean(u)) < 0.05


def _tpcf(rve):
    return np.abs(np.fft.fftn(rve)) ** 2 / rve.size


def test_true_structure_is_a_fixed_point():
    rve = (gen.spheres(16, 3, 0.25, seed=4) == 0).astype(float)
    r = oracle.Rec(_tpcf(rve))
    r.set_initial(rve)
    assert r.run(50, 0.5, 7) == 1 and r.status == 1
    m, err = r.result()
    assert err[0] < 1e-12
    assert np.array_equal(m.real, rve) and r.search_steps == 0


def test_one_voxel_projection_branches():
    """N = 1: the FFT is the identity and normC = 1.5.  From 0.3 with
    beta 0.5: 1.5 > 1 -> 0.3 - 0.5 (1.5 - 1) = 0.05; again -> -0.2; then
    arg = pi gives -1.5 < 0 -> -0.2 + 0.5 * 1.5 = 0.55.  Errors |m| - 1.5."""
    r = oracle.Rec(np.full((1, 1, 1), 2.25))
    r.set_initial(np.array([0.3]))
    assert r.run(3, 0.5, 1) == 3 and r.status == 2
    m, err = r.result()
    assert np.allclose(err, [1.2, 1.45, 1.3], rtol=0, atol=1e-15)
    # volume fraction 1.5 cannot be met by one voxel: the search runs out
    # after 201 steps with the voxel (0.55) above every threshold tried
    assert r.search_steps == 201 and m.real[0, 0, 0] == 1.0


def test_random_start_keeps_the_volume_fraction():
    rve = (gen.spheres(16, 3, 0.25, seed=4) == 0).astype(float)
    r = oracle.Rec(_tpcf(rve))
    it = r.run(60, 0.8, 11)
    m, err = r.result()
    assert it == 60 and r.status == 2 and err[-1] < err[0]
    assert set(np.unique(m.real)) <= {0.0, 1.0}
    # the search stops within 1e-4 of the target fraction, or after 201
    # steps; at 16^3 one voxel is 2.4e-4 of the volume
    d = abs(m.real.mean() - rve.mean())
    assert d < 1e-4 or r.search_steps == 201
    assert d <= 2.0 / rve.size
