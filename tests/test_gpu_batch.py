"""GPU parity of the batch API (m3d_batch_*): many RVEs of mixed grid shapes.

Every batch result must equal what the loop over RVEs gives: the FP64 oracle
per RVE (C_eff within 1e-9 relative, same iteration counts, FP64 path; 1e-6
for FP32+FP64 polish) and, bit for bit, the single-RVE handle (solves are
independent and their reductions do not depend on the batch size).
"""
import numpy as np
import pytest

import mat3d_amd as m
import oracle
from mat3d_amd import gen

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def grain_mat(g, seed):
    r = np.random.default_rng(seed)
    mat = np.zeros((g, 7))
    mat[:, 0] = r.uniform(50e9, 300e9, size=g)
    mat[:, 1] = r.uniform(0.2, 0.35, size=g)
    mat[:, 2], mat[:, 4] = np.inf, 1
    return mat


def problems():
    """Mixed shapes (cubic, odd, non-cubic), polycrystals and two-phase
    inclusions in both the strain and the stress base."""
    out = []
    for i, (shape, g) in enumerate([((16, 16, 16), 5), ((12, 12, 12), 4), ((16, 16, 16), 7), ((8, 12, 10), 3),
                                    ((16, 16, 16), 2), ((9, 9, 9), 6)]):
        out.append((shape, gen.voronoi(shape, g, 50 + i), grain_mat(g, 70 + i)))
    sph = gen.spheres(16, 3, 0.2, seed=3)
    out.append(((16, 16, 16), sph, gen.MAT_C1.copy()))              # stiff minority: stress base
    out.append(((16, 16, 16), sph, gen.MAT_C1[::-1].copy()))        # stiff majority: strain base
    return out


def run_batch(probs, precision, **opts):
    b = m.Batch(precision)
    for k, v in opts.items():
        b.set_option(k, v)
    for i, (shape, ids, mat) in enumerate(probs):
        assert b.add(mat, ids, shape) == i
    b.homogenization(1e-6, 1e-6)
    res = [b.result(i) for i in range(len(probs))]
    st = b.stats()
    b.close()
    return res, st


@pytest.fixture(scope="module")
def probs():
    return problems()


@pytest.fixture(scope="module")
def oracle_results(probs):
    out = []
    for shape, ids, mat in probs:
        o = oracle.Elastic(*shape, "e")
        o.material(mat, ids)
        r = o.homogenization(1e-6, 1e-6)
        r["strain_base"] = o.info()["strain_base"]
        out.append(r)
    return out


def test_batch_fp64_matches_oracle_per_rve(probs, oracle_results):
    res, st = run_batch(probs, "fp64")
    bases = set()
    for i, (r, o) in enumerate(zip(res, oracle_results)):
        print(f"[batch fp64] rve {i} shape {probs[i][0]} base {'strain' if o['strain_base'] else 'stress'} "
              f"iters {[x['iterations'] for x in r['info']]} vs {o['iters']} relC {rel(r['CHomo'], o['C']):.2e}")
        assert [x["iterations"] for x in r["info"]] == o["iters"]
        assert rel(r["CHomo"], o["C"]) < 1e-9
        assert rel(r["SHomo"], o["S"]) < 1e-9
        assert rel(r["epHomo"], o["epHomo"]) < 1e-9
        assert rel(r["sigHomo"], o["sigHomo"]) < 1e-9
        assert abs(r["ret"] - o["ret"]) <= 1e-9 * abs(o["ret"])
        bases.add(o["strain_base"])
    assert bases == {0, 1} or bases == {False, True}
    assert st["count"] == len(probs)
    expect = sum(np.prod(p[0]) * sum(o["iters"]) for p, o in zip(probs, oracle_results))
    assert st["voxel_iterations"] == pytest.approx(expect)


def test_batch_fp32_polish_matches_oracle_per_rve(probs, oracle_results):
    res, _ = run_batch(probs, "fp32_polish")
    for i, (r, o) in enumerate(zip(res, oracle_results)):
        e = rel(r["CHomo"], o["C"])
        print(f"[batch fp32+polish] rve {i} relC {e:.2e}")
        assert e < 1e-6
        assert all(x["status"] == 1 for x in r["info"])


def test_batch_equals_single_handle_bitwise(probs):
    res, _ = run_batch(probs, "fp64")
    for i, (shape, ids, mat) in enumerate(probs):
        s = m.FFT_Analysis3D(*shape, "e", precision="fp64")
        s.material(mat, ids)
        ret = s.homogenization(1e-6, 1e-6)
        assert np.array_equal(res[i]["CHomo"], s.CHomo), f"rve {i}"
        assert np.array_equal(res[i]["epHomo"], s.epHomo), f"rve {i}"
        assert res[i]["ret"] == ret
        s.close()


def test_batch_split_into_device_batches_gives_same_results(probs):
    full, _ = run_batch(probs, "fp32_polish")
    split, _ = run_batch(probs, "fp32_polish", max_solves=6)  # one RVE per device batch
    for a, b in zip(full, split):
        assert np.array_equal(a["CHomo"], b["CHomo"])
        assert [x["iterations"] for x in a["info"]] == [x["iterations"] for x in b["info"]]


# ------------------------------------------- batches on the specialised kernels
def problems_fast():
    """Voronoi polycrystals of the 64-RVE workload's kind (20-200 grains,
    E in [50, 300] GPa, nu in [0.2, 0.35]) on grids whose line lengths have
    the compile-time specialised kernels (32, 48 = 16 x 3, 64), one of them
    made strain-based (its largest grain is the stiffest), so one launch
    mixes bases, RVE coefficient tables and solves that stop at different
    iterations."""
    out = []
    for i, (n, g) in enumerate([(32, 20), (48, 60), (32, 200), (48, 120), (64, 40), (32, 35)]):
        ids = gen.voronoi(n, g, 1100 + i)
        mat = grain_mat(g, 1200 + i)
        if i == 5:
            big = np.bincount(ids.ravel(), minlength=g).argmax()
            mat[big, 0] = 350e9
            mat[big, 1] = 0.2
        out.append(((n, n, n), ids, mat))
    return out


@pytest.fixture(scope="module")
def probs_fast():
    return problems_fast()


@pytest.fixture(scope="module")
def oracle_fast(probs_fast):
    out = []
    for shape, ids, mat in probs_fast:
        o = oracle.Elastic(*shape, "e")
        o.material(mat, ids)
        r = o.homogenization(1e-6, 1e-6)
        r["strain_base"] = o.info()["strain_base"]
        out.append(r)
    return out


def test_specialised_batch_fp64_matches_oracle(probs_fast, oracle_fast):
    res, _ = run_batch(probs_fast, "fp64")
    bases, iters = set(), set()
    for i, (r, o) in enumerate(zip(res, oracle_fast)):
        its = [x["iterations"] for x in r["info"]]
        print(f"[fast batch fp64] rve {i} n {probs_fast[i][0][0]} grains {len(probs_fast[i][2])} "
              f"base {'strain' if o['strain_base'] else 'stress'} iters {its} vs {o['iters']} "
              f"relC {rel(r['CHomo'], o['C']):.2e}")
        assert its == o["iters"]
        assert rel(r["CHomo"], o["C"]) < 1e-9
        assert rel(r["epHomo"], o["epHomo"]) < 1e-9
        assert rel(r["sigHomo"], o["sigHomo"]) < 1e-9
        bases.add(bool(o["strain_base"]))
        iters.update(its)
    assert bases == {False, True}
    assert len(iters) > 1


def test_specialised_batch_fp32_polish_matches_oracle(probs_fast, oracle_fast):
    res, _ = run_batch(probs_fast, "fp32_polish")
    for i, (r, o) in enumerate(zip(res, oracle_fast)):
        e = rel(r["CHomo"], o["C"])
        print(f"[fast batch fp32+polish] rve {i} relC {e:.2e} iters {[x['iterations'] for x in r['info']]} "
              f"vs {o['iters']}")
        assert e < 1e-6
        assert all(x["status"] == 1 for x in r["info"])
        # the FP32 phase's switch and rollback leave the reference's stop iteration unchanged
        assert [x["iterations"] for x in r["info"]] == o["iters"]


def test_full_voronoi_workload_uniform_moduli():
    """The 64-RVE workload at its sizes ({32, 48, 64, 96, 128}^3, 20-200
    grains, 384 solves) with every grain given the same moduli: the RVEs
    are homogeneous, so C_eff equals the isotropic C exactly and every solve
    stops after the reference's minimum of 2 iterations -- a property of the
    whole batch path at full size, independent of the oracle."""
    probs = gen.c3_batch(64, device=True)
    E, nu = 120e9, 0.3
    la, mu = nu * E / (1 + nu) / (1 - 2 * nu), E / (1 + nu) / 2
    C = np.zeros((6, 6))
    C[:3, :3] = la
    C[np.arange(6), np.arange(6)] += np.r_[2 * mu * np.ones(3), 2 * mu * np.ones(3)]
    for prec in ("fp64", "fp32_polish"):
        b = m.Batch(prec)
        for n, ids, mat in probs:
            mat = mat.copy()
            mat[:, 0], mat[:, 1] = E, nu
            b.add(mat, ids)
        b.homogenization(1e-6, 1e-6)
        worst = 0.0
        for i in range(len(probs)):
            r = b.result(i)
            worst = max(worst, rel(r["CHomo"], C))
            assert [x["iterations"] for x in r["info"]] == [2] * 6, f"rve {i} {prec}"
        print(f"[uniform voronoi batch {prec}] worst relC {worst:.2e}")
        assert worst < (1e-12 if prec == "fp64" else 1e-6)
        assert b.stats()["count"] == 64
        b.close()
