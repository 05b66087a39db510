"""percolation (src/percolation/percolation.cpp) on the GPU through the
C-ABI (m3d_perc_*, Python mirror mat3d_amd.percolation) against the oracle's
restatement of the reference's sweeps: the relabelled field C1, all_clusters,
clusters_vol_frac and vol_intrstd_phs bit for bit wherever the oracle
reports that the reference's sweeps reached their fixed point; the hand-made
cases of the oracle's own tests; SurfaceOfTwoPhase counts.  At the full
256^3 fibre structure the outputs are checked through properties (sizes add
up to the phase volume, descending fractions, relabelled counts agree with
the fractions) and, at 96^3, against a sparse-graph search of the rule."""
import numpy as np
import pytest

import mat3d_amd as m3d
import oracle
from mat3d_amd import gen
from perc_rule import components, expected, snake, surface

pytestmark = pytest.mark.gpu


def both(ids, nph, phase):
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    N = ids.shape
    p = m3d.percolation(ids, *N, nph, phase)
    p.runPercolationAnalysis(1)
    o = oracle.Percolation(ids, nph, phase).run()
    return p, o


def same(p, o):
    c1, cl, fr, vol, fp = o
    assert fp, "the reference's sweeps stopped short of their fixed point on this field"
    assert p.vol_intrstd_phs == vol
    assert p.all_clusters == cl.tolist()
    assert p.clusters_vol_frac == fr.tolist()
    assert np.array_equal(p.C1, c1)


@pytest.mark.parametrize("N,d,seed", [(8, 0.3, 1), (16, 0.35, 3), (24, 0.12, 4), (32, 0.3, 5), (40, 0.2, 6),
                                      (48, 0.1, 7), (33, 0.25, 8)])
def test_random_fields_against_oracle(N, d, seed):
    rng = np.random.default_rng(seed)
    ids = (rng.random((N, N, N)) < d).astype(np.int32) + 1
    p, o = both(ids, 2, 2)
    print(f"[perc random {N} {d}] clusters {len(p.all_clusters)}")
    same(p, o)


@pytest.mark.parametrize("phase", [0, 1])
def test_fibres_64_against_oracle(phase):
    ids = gen.c1_rve(64)
    p, o = both(ids, int(ids.max()), phase)
    print(f"[perc fibres64 phase {phase}] clusters {len(p.all_clusters)}")
    same(p, o)


def test_spheres_multi_phase_against_oracle():
    ids = gen.multi_spheres(48, [(2, 0.15, 2, 6), (0, 0.10, 2, 6)], seed=5)
    for ph in (0, 1, 2):
        p, o = both(ids, 2, ph)
        same(p, o)


def test_hand_cases():
    ids = np.full((6, 6, 6), 2, dtype=np.int32)
    same(*both(ids, 2, 2))
    iso = np.ones((8, 8, 8), dtype=np.int32)
    iso[::2, ::2, ::2] = 2
    same(*both(iso, 2, 2))
    for trig in (False, True):
        d = np.zeros((6, 6, 6), dtype=np.int32)
        d[3, 2, 0] = d[2, 3, 0] = 1
        if trig:
            d[0, 3, 3] = 1
        p, o = both(d, 1, 1)
        same(p, o)
        assert len(p.all_clusters) == 2


def test_no_interest_voxels():
    ids = np.ones((7, 7, 7), dtype=np.int32)
    p, o = both(ids, 2, 2)
    same(p, o)
    assert p.all_clusters == [] and p.vol_intrstd_phs == 0


def test_winding_path_is_one_cluster():
    # the reference's fixed sweep count leaves this one path split (the
    # oracle shows it); the GPU gives the fixed point of the same rule
    f = snake(16)
    ids = np.where(f, 2, 1).astype(np.int32)
    p, o = both(ids, 2, 2)
    assert not o[4] and len(o[1]) > 1
    e = expected(ids, 2, 2)
    assert p.all_clusters == e[1].tolist() == [3]
    assert np.array_equal(p.C1, e[0])
    # the reference's own sweeps, with rounds enough to reach their fixed
    # point, give the GPU's result bit for bit
    rounds = 10
    while not (o := oracle.Percolation(ids, 2, 2).run(rounds))[4]:
        rounds *= 2
        assert rounds <= 320
    same(p, o)


def test_run_twice_and_rerun_after_surface():
    rng = np.random.default_rng(11)
    ids = (rng.random((20, 20, 20)) < 0.3).astype(np.int32) + 1
    p = m3d.percolation(ids, 20, 20, 20, 2, 2)
    p.runPercolationAnalysis(1)
    a = (p.C1.copy(), list(p.all_clusters), list(p.clusters_vol_frac))
    p.runPercolationAnalysis(8)
    assert np.array_equal(a[0], p.C1) and a[1] == p.all_clusters and a[2] == p.clusters_vol_frac
    other = (rng.random((20, 20, 20)) < 0.2).astype(np.int32) + 1
    p.SurfaceOfTwoPhase(other, 2, 1)
    assert p.commonSurface == surface(other, 2, 1)
    p.runPercolationAnalysis(1)
    same(p, oracle.Percolation(other, 2, 2).run())


def test_errors():
    ids = np.ones((8, 8, 6), dtype=np.int32)
    p = m3d.percolation(ids, 8, 8, 6, 2, 2)
    with pytest.raises(m3d.M3DError):
        p.runPercolationAnalysis(1)  # Nx = Ny = Nz needed
    bad = np.ones((6, 6, 6), dtype=np.int32)
    bad[1, 2, 3] = 5  # reads as a cluster label when noOfPhases = 2
    q = m3d.percolation(bad, 6, 6, 6, 2, 2)
    with pytest.raises(m3d.M3DError):
        q.runPercolationAnalysis(1)


@pytest.mark.parametrize("shape", [(9, 9, 9), (12, 7, 5), (32, 32, 32)])
def test_surface_against_oracle(shape):
    rng = np.random.default_rng(sum(shape))
    ids = rng.integers(0, 3, size=shape).astype(np.int32)
    p = m3d.percolation(ids, *shape, 2, 1)
    o = oracle.Percolation(ids, 2, 1)
    for a, b in ((0, 1), (1, 2), (2, 0), (1, 1), (1, -1)):
        got = p.SurfaceOfTwoPhase(ids, a, b)
        assert got == o.surface(ids, a, b) == surface(ids, a, b)


def test_rule_graph_96():
    ids = gen.fibres(96, 3, 0.2, 7, length=48)
    for ph in (0, 1):
        p = m3d.percolation(ids, 96, 96, 96, 1, ph)
        p.runPercolationAnalysis(1)
        e = expected(ids, 1, ph)
        assert np.array_equal(p.C1, e[0])
        assert p.all_clusters == e[1].tolist()
        assert p.clusters_vol_frac == e[2].tolist()


def test_fibres_256_properties():
    ids = gen.c2_rve(256, device=True)
    for ph in (0, 1):
        p = m3d.percolation(ids, 256, 256, 256, 1, ph)
        p.runPercolationAnalysis(1)
        vol = float(np.sum(ids == ph))
        K = len(p.all_clusters)
        print(f"[perc fibres256 phase {ph}] clusters {K} volume {vol:.0f}")
        assert p.vol_intrstd_phs == vol
        fr = np.array(p.clusters_vol_frac)
        assert np.all(np.diff(fr) <= 0)
        sizes = np.rint(fr * vol).astype(np.int64)
        assert np.array_equal(sizes / vol, fr) and sizes.sum() == vol
        # labels: strictly increasing in first-appearance order, within range
        cl = np.array(p.all_clusters)
        assert np.all(np.diff(cl) > 0) and cl[0] == 2 and cl[-1] < 2 + vol
        # relabelled counts: label 2 + r holds every cluster of the r-th size
        lab = p.C1[ids == ph]
        counts = np.bincount(lab - 2)
        for r in np.flatnonzero(counts):
            assert r == 0 or sizes[r - 1] > sizes[r]
            assert counts[r] == sizes[sizes == sizes[r]].sum()
        assert np.array_equal(p.C1[ids != ph], ids[ids != ph])


@pytest.mark.parametrize("N", [1, 2, 3, 65])
def test_small_and_odd_boxes(N):
    rng = np.random.default_rng(N)
    for d in (0.0, 0.5, 1.0):
        ids = (rng.random((N, N, N)) < d).astype(np.int32) + 1
        p, o = both(ids, 2, 2)
        same(p, o)
        assert p.SurfaceOfTwoPhase(ids, 2, 1) == surface(ids, 2, 1)
