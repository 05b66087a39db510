"""The generators' voxel work on the GPU (m3d_gen_*) against the numpy
generators of mat3d_amd/gen.py on the same seeds: balls and fibres bit for
bit (integer tests; the fibre distances in numpy's operation order without
fused multiply-adds); Voronoi cells against scipy's periodic k-d tree, equal
except where two seeds are at exactly the same distance."""
import numpy as np
import pytest

from mat3d_amd import gen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,r,frac,seed", [(64, 6, 0.20, 1), (48, 8, 0.25, 4), (40, 3, 0.15, 9)])
def test_spheres_device_equals_host(n, r, frac, seed):
    a = gen.spheres(n, r, frac, seed)
    b = gen.spheres(n, r, frac, seed, device=True)
    assert np.array_equal(a, b)


def test_multi_spheres_device_equals_host():
    groups = [(2, 0.15, 3, 10), (0, 0.10, 3, 10)]
    a = gen.multi_spheres(64, groups, seed=5)
    b = gen.multi_spheres(64, groups, seed=5, device=True)
    assert np.array_equal(a, b) and set(np.unique(b)) == {0, 1, 2}


@pytest.mark.parametrize("n,r,frac,seed,length", [(64, 4, 0.30, 2, 32), (96, 3, 0.20, 7, 48)])
def test_fibres_device_equals_host(n, r, frac, seed, length):
    a = gen.fibres(n, r, frac, seed, length=length)
    b = gen.fibres(n, r, frac, seed, length=length, device=True)
    print(f"[gen fibres {n}] fraction {(b == 0).mean():.4f}, equal voxels {(a == b).mean():.8f}")
    assert np.array_equal(a, b)


@pytest.mark.parametrize("shape,grains,seed", [(32, 20, 1000), ((48, 32, 40), 120, 7), (64, 200, 1003)])
def test_voronoi_device_matches_kdtree(shape, grains, seed):
    a = gen.voronoi(shape, grains, seed)
    b = gen.voronoi(shape, grains, seed, device=True)
    diff = np.argwhere(a != b)
    print(f"[gen voronoi {shape} {grains}] differing voxels {len(diff)}")
    if len(diff):
        # only exact ties may differ: both seeds at the same minimum-image distance
        box = np.array(a.shape, dtype=float)
        pts = np.random.default_rng(seed).uniform(0, 1, size=(grains, 3)) * box
        for v in diff:
            d = [np.abs(v - pts[k]) for k in (a[tuple(v)], b[tuple(v)])]
            d = [np.minimum(x, box - x) for x in d]
            assert np.isclose((d[0] ** 2).sum(), (d[1] ** 2).sum(), rtol=1e-12, atol=0)
    assert len(diff) <= 1e-4 * a.size
