"""The oracle's restatement of src/percolation/percolation.cpp pinned on the
CPU.  Known answers worked out by hand from the reference's text: one
cluster for a full box, one cluster per voxel for isolated voxels, the
diagonal pair that only joins when a permutation of the cube's corner is an
interest voxel (the trigger tests :127-145), the reference's stop after its
fixed 5 x 8 x 6 sweeps on a long winding path, and closed-form face-contact
counts.  On random fields the oracle's fixed point equals the connected sets
of the rule graph found by a sparse-graph search (tests/perc_rule.py), a
different route from the sweeps."""
import numpy as np
import pytest

import oracle
from perc_rule import components, expected, snake, surface


def run(ids, nph, phase):
    return oracle.Percolation(np.asarray(ids, dtype=np.int32), nph, phase).run()


def test_full_box_is_one_cluster():
    ids = np.full((6, 6, 6), 2, dtype=np.int32)
    c1, cl, fr, vol, fp = run(ids, 2, 2)
    assert fp and vol == 216
    assert cl.tolist() == [3] and fr.tolist() == [1.0]
    assert np.all(c1 == 3)


def test_isolated_voxels():
    N = 8
    ids = np.ones((N, N, N), dtype=np.int32)
    ids[::2, ::2, ::2] = 2
    c1, cl, fr, vol, fp = run(ids, 2, 2)
    K = (N // 2) ** 3
    assert fp and vol == K
    # labels phasePlus + raster rank of the voxel, in raster order
    assert cl.tolist() == list(range(3, 3 + K))
    assert np.all(fr == 1.0 / K)
    # equal fractions: every cluster takes the first position of its value
    assert np.all(c1[::2, ::2, ::2] == 3) and np.all(c1[ids == 1] == 1)


def diagonal_pair(trigger):
    ids = np.zeros((6, 6, 6), dtype=np.int32)
    ids[3, 2, 0] = ids[2, 3, 0] = 1
    if trigger:
        ids[0, 3, 3] = 1  # a permutation of the cube corner (3, 3, 0)
    return ids


def test_diagonal_pair_alone_stays_apart():
    c1, cl, fr, vol, fp = run(diagonal_pair(False), 1, 1)
    assert fp and vol == 2
    assert cl.tolist() == [2, 3] and fr.tolist() == [0.5, 0.5]
    assert c1[3, 2, 0] == 2 and c1[2, 3, 0] == 2


def test_diagonal_pair_joined_by_permuted_trigger():
    c1, cl, fr, vol, fp = run(diagonal_pair(True), 1, 1)
    assert fp and vol == 3
    # raster ranks: (0,3,3) -> 0, (2,3,0) -> 1, (3,2,0) -> 2
    assert cl.tolist() == [2, 3]
    assert fr.tolist() == [2 / 3, 1 / 3]
    assert c1[2, 3, 0] == 2 and c1[3, 2, 0] == 2 and c1[0, 3, 3] == 3


def test_shared_min_corner_joins_a_diagonal_pair():
    ids = np.zeros((4, 4, 4), dtype=np.int32)
    ids[1, 0, 0] = ids[0, 1, 0] = ids[0, 0, 0] = 1
    c1, cl, fr, vol, fp = run(ids, 1, 1)
    assert cl.tolist() == [2] and fr.tolist() == [1.0]


def test_sweeps_stop_short_on_a_winding_path():
    # one connected path; the reference's fixed number of sweeps does not
    # carry the first label along all of it
    f = snake(16)
    assert components(f)[0] == 1
    c1, cl, fr, vol, fp = run(np.where(f, 2, 1), 2, 2)
    assert not fp and len(cl) > 1
    f8 = snake(8)
    assert components(f8)[0] == 1
    c1, cl, fr, vol, fp = run(np.where(f8, 2, 1), 2, 2)
    assert fp and cl.tolist() == [3]


@pytest.mark.parametrize("N,d,seed", [(8, 0.3, 1), (12, 0.2, 2), (16, 0.35, 3), (24, 0.12, 4), (32, 0.3, 5)])
def test_fixed_point_matches_the_rule_graph(N, d, seed):
    rng = np.random.default_rng(seed)
    ids = (rng.random((N, N, N)) < d).astype(np.int32) + 1
    c1, cl, fr, vol, fp = run(ids, 2, 2)
    assert fp
    e = expected(ids, 2, 2)
    assert np.array_equal(c1, e[0])
    assert np.array_equal(cl, e[1])
    assert np.array_equal(fr, e[2])
    assert vol == e[3]


def test_no_interest_voxels():
    c1, cl, fr, vol, fp = run(np.ones((5, 5, 5)), 2, 2)
    assert len(cl) == 0 and vol == 0 and np.all(c1 == 1)


def test_surface_closed_forms():
    N = 6
    half = np.ones((N, N, N), dtype=np.int32)
    half[N // 2:] = 2
    p = oracle.Percolation(half, 2, 1)
    # phase-2 voxels on the plane x = N/2 see phase 1 at -x
    assert p.surface(half, 2, 1) == N * N
    assert p.surface(half, 1, 2) == 0
    # -1 is the margin: every voxel on the x = 0, y = 0, z = 0 faces
    one = np.ones((N, N, N), dtype=np.int32)
    assert p.surface(one, 1, -1) == 3 * N * N
    # checkerboard: every phase-0 voxel sees phase 1 at each in-box -face
    x, y, z = np.indices((N, N, N))
    cb = ((x + y + z) % 2).astype(np.int32)
    want = int(np.sum(((x > 0).astype(int) + (y > 0) + (z > 0))[cb == 0]))
    assert p.surface(cb, 0, 1) == want


def test_surface_random_against_numpy():
    rng = np.random.default_rng(7)
    ids = rng.integers(0, 3, size=(9, 9, 9)).astype(np.int32)
    p = oracle.Percolation(ids, 2, 1)
    for a, b in ((0, 1), (1, 2), (2, 0), (1, 1)):
        assert p.surface(ids, a, b) == surface(ids, a, b)


def test_more_rounds_reach_the_rule_graph_on_the_winding_path():
    # the reference's own sweeps, given rounds enough, end at the connected
    # sets of the rule graph (one cluster) on the path where 5 rounds stop short
    f = snake(16)
    ids = np.where(f, 2, 1).astype(np.int32)
    e = expected(ids, 2, 2)
    rounds = 5
    while True:
        c1, cl, fr, vol, fp = oracle.Percolation(ids, 2, 2).run(rounds)
        if fp:
            break
        rounds *= 2
        assert rounds <= 320
    assert rounds > 5
    assert np.array_equal(c1, e[0]) and cl.tolist() == e[1].tolist() == [3]
