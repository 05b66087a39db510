"""Write the golden fixtures under tests/golden/ from the FP64 oracle.

Small inputs only (<= 12^3).  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from mat3d_amd import gen  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(2024)
    # Green operator tables (21 components, FFT_analysis3D.cpp:155-233)
    zets = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 2, -3], [-4, 1, 5], [3, -3, 2], [-32, 5, 7], [0, -2, 7]],
                    dtype=float)
    la0, mu0 = 7.3e9, 4.1e9
    gs = np.array([oracle.green21(z, la0, mu0, True) for z in zets])
    gd = np.array([oracle.green21(z, la0, mu0, False) for z in zets])
    np.savez_compressed(os.path.join(OUT, "green21.npz"), zets=zets, la0=la0, mu0=mu0, strain=gs, stress=gd)

    # rank-3 r2c of a seeded field (6 components, odd/even axes)
    x = rng.standard_normal((5, 6, 8, 6))
    np.savez_compressed(os.path.join(OUT, "rfft3.npz"), x=x, X=oracle.rfft3(x, 1))

    # elastic homogenization at 12^3, both bases (phase order swapped)
    n = 12
    ids = gen.spheres(n, 2, 0.2, seed=3)
    out = {"ids": ids}
    for tag, mat in (("stress", gen.MAT_C1), ("strain", gen.MAT_C1[::-1].copy())):
        o = oracle.Elastic(n, n, n, "e", nthreads=1)
        o.material(mat, ids)
        r = o.homogenization(1e-6, 1e-6)
        assert o.info()["strain_base"] == (tag == "strain")
        out[f"{tag}_mat"] = mat
        for k in ("C", "S", "epHomo", "sigHomo"):
            out[f"{tag}_{k}"] = r[k]
        out[f"{tag}_iters"] = np.array(r["iters"])
        o.close()
    # single analysis with the example deck's load (example/3d/in_struct.txt)
    E = np.array([0.0, 0.1, 0.0, 0.0, 0.0, 0.0])
    o = oracle.Elastic(n, n, n, "e", nthreads=1)
    o.material(gen.MAT_C2, ids)
    o.set_load(E, E)
    out["run_iters"] = np.array(o.run(1e-6, 1e-6, 15))
    out["run_eps"], out["run_sig"] = o.fields()
    o.close()
    np.savez_compressed(os.path.join(OUT, "elastic12.npz"), **out)

    # thermal homogenization at 12^3: 2-phase reference API (both bases) and 3 phases
    ids3 = gen.multi_spheres(n, [(2, 0.15, 1, 2), (0, 0.10, 1, 2)], seed=5)
    out = {"ids": ids, "ids3": ids3}
    for tag, (km, ki) in (("grad", (25.0, 30.0)), ("flux", (30.0, 25.0))):
        t = oracle.Thermal(n, n, n, nthreads=1)
        t.material(km, ki, ids)
        r = t.homogenization(1e-6, 1e-6)
        assert t.info()["grad_base"] == (tag == "grad")
        out[f"{tag}_k"] = np.array([km, ki])
        out[f"{tag}_K"] = r["K"]
        out[f"{tag}_iters"] = np.array(r["iters"])
        t.close()
    t = oracle.Thermal(n, n, n, nthreads=1)
    k3 = np.array([0.01, 1.0, 1000.0])
    t.material_n(k3, ids3)
    r = t.homogenization(1e-7, 1e-7)
    out["n3_k"], out["n3_K"], out["n3_iters"] = k3, r["K"], np.array(r["iters"])
    t.close()
    np.savez_compressed(os.path.join(OUT, "thermal12.npz"), **out)

    # J2 elastoplastic incremental basic scheme at 8^3 ("ep", FFT_analysis3D.cpp:796-853)
    n = 8
    ids = gen.spheres(n, 2, 0.2, seed=4)
    o = oracle.Elastic(n, n, n, "ep", nthreads=1)
    o.material(gen.MAT_C4, ids)
    X0 = np.eye(6)[0]
    o.set_load(0.01 * X0, X0)
    o.run(1e-6, 1e-6, 10)
    np.savez_compressed(os.path.join(OUT, "j2_8.npz"), ids=ids, mat=gen.MAT_C4, E=0.01 * X0, X0=X0, nInc=10,
                        prop=o.prop_homogenized())
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
