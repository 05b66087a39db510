"""The oracle's restatement of FFT_Analysis2D (src/2d/FFT_analysis2D.cpp)
pinned on the CPU: a homogeneous plate returns C exactly; two-phase
laminates along x and along y return the exact laminate stiffness (tensor
shear Voigt [11, 22, 12]; contrast 2, where 25 basic-scheme iterations are
converged); the material tables are the reference's closed forms."""
import numpy as np
import pytest

import oracle


def lame(E, v):
    return v * E / (1 + v) / (1 - 2 * v), E / (1 + v) / 2


def C2(la, mu):
    return np.array([[la + 2 * mu, la, 0], [la, la + 2 * mu, 0], [0, 0, 2 * mu]])


VO = [(0, 0), (1, 1), (0, 1)]


def laminate2(Cs, fracs, k):
    """Exact 2D laminate stiffness, layers normal to tensor axis k: phase
    strains E + sym(a_p n), equal tractions, sum f_p a_p = 0."""
    n = np.zeros(2)
    n[k] = 1

    def sym(a):
        T = (np.outer(a, n) + np.outer(n, a)) / 2
        return np.array([T[i, j] for i, j in VO])

    def trac(s):
        T = np.array([[s[0], s[2]], [s[2], s[1]]])
        return T @ n

    out = np.zeros((3, 3))
    for w in range(3):
        E = np.zeros(3)
        E[w] = 1.0
        P = len(Cs)
        A = np.zeros((2 * P, 2 * P))
        b = np.zeros(2 * P)
        basis = [sym(e) for e in np.eye(2)]
        for p in range(1, P):
            for c in range(2):
                for d in range(2):
                    A[2 * (p - 1) + c, 2 * p + d] += trac(Cs[p] @ basis[d])[c]
                    A[2 * (p - 1) + c, d] -= trac(Cs[0] @ basis[d])[c]
                b[2 * (p - 1) + c] = trac(Cs[0] @ E)[c] - trac(Cs[p] @ E)[c]
        for c in range(2):
            for p in range(P):
                A[2 * (P - 1) + c, 2 * p + c] = fracs[p]
        a = np.linalg.solve(A, b).reshape(P, 2)
        sig = sum(f * (C @ (E + sym(ap))) for f, C, ap in zip(fracs, Cs, a))
        out[:, w] = sig
    return out


def test_homogeneous_plate():
    q = [70e9, 0.3, 70e9, 0.3]
    o = oracle.Elastic2D(16, 16)
    o.material(q, np.ones(256))
    o.set_load([0, 0, 0])
    C, ep, sg = o.homogenization()
    assert np.allclose(C, C2(*lame(70e9, 0.3)), rtol=1e-12, atol=1e-3)
    assert np.allclose(ep, 0.005 * np.eye(3), rtol=1e-12, atol=1e-18)


@pytest.mark.parametrize("axis", [0, 1])
def test_laminate_exact(axis):
    Nx, Ny = 16, 16
    rve = np.ones((Nx, Ny))
    if axis == 0:
        rve[:6, :] = 0   # layers normal to x (tensor index 0 is x)
    else:
        rve[:, :6] = 0
    q = [70e9, 0.3, 140e9, 0.25]
    o = oracle.Elastic2D(Nx, Ny)
    o.material(q, rve.reshape(-1))
    o.set_load([0, 0, 0])
    C, ep, sg = o.homogenization()
    f0 = 6 / 16
    ref = laminate2([C2(*lame(140e9, 0.25)), C2(*lame(70e9, 0.3))], [f0, 1 - f0], axis)
    assert np.abs(C - ref).max() / np.abs(ref).max() < 1e-9
