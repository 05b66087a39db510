"""Pins of the FP64 oracle (CPU).  Each check is chosen so a plausible slip in
the restatement (dropped term, wrong sign or index, transposed operand) fails:

* transforms       -- numpy.fft (pocketfft) and a dense DFT on tiny grids;
* Green operators  -- Gamma0 C0 Gamma0 = Gamma0, Gamma0 C0 on compatible
                      strains is the identity, Delta0 = C0 - C0 Gamma0 C0;
* material         -- S = C^-1 and the spectral closed form of C_acc, S_acc;
* eqVar            -- closed forms on single Fourier modes;
* full solver      -- homogeneous RVE (C_eff = C, 2 iterations), exact
                      laminate solutions (elastic and thermal), the
                      reference's own Hashin-Shtrikman bounds, the uniaxial
                      J2 curve with linear hardening and with power-law
                      hardening (n = 0.2, a scalar root by bisection).
"""
import numpy as np
import pytest

import j2_model as j2m
import oracle
from mat3d_amd import gen

M = np.array([1, 1, 1, 2, 2, 2], dtype=float)  # shear doubling of every contraction
VO = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]  # Voigt order [11,22,33,23,13,12]
WMAP = [[0, 1, 2, 3, 4, 5], [1, 6, 7, 8, 9, 10], [2, 7, 11, 12, 13, 14], [3, 8, 12, 15, 16, 17],
        [4, 9, 13, 16, 18, 19], [5, 10, 14, 17, 19, 20]]


def iso_C(la, mu):
    C = np.zeros((6, 6))
    C[:3, :3] = la
    C[np.arange(3), np.arange(3)] += 2 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = 2 * mu
    return C


def lame(E, nu):
    return nu * E / (1 + nu) / (1 - 2 * nu), E / (1 + nu) / 2


# ------------------------------------------------------------- transforms
@pytest.mark.parametrize("shape", [(4, 6, 8, 6), (5, 7, 9, 3), (8, 8, 8, 1), (6, 10, 12, 6), (3, 5, 65, 2),
                                   (16, 12, 2, 2), (9, 9, 26, 3)])
def test_rfft3_irfft3_match_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    X = oracle.rfft3(x)
    Xn = np.fft.rfftn(x, axes=(0, 1, 2))
    assert np.abs(X - Xn).max() <= 1e-12 * np.abs(Xn).max()
    # non-Hermitian input: irfftn semantics (imaginary part of kz=0 / Nyquist bins ignored)
    Y = rng.standard_normal(X.shape) + 1j * rng.standard_normal(X.shape)
    y = oracle.irfft3(Y, shape[2])
    yn = np.fft.irfftn(Y, s=shape[:3], axes=(0, 1, 2)) * np.prod(shape[:3])
    assert np.abs(y - yn).max() <= 1e-12 * np.abs(yn).max()
    # round trip
    assert np.abs(oracle.irfft3(X, shape[2]) / np.prod(shape[:3]) - x).max() < 1e-13


def test_rfft3_dense_dft():
    Nx, Ny, Nz = 3, 4, 5
    rng = np.random.default_rng(1)
    x = rng.standard_normal((Nx, Ny, Nz, 1))
    X = oracle.rfft3(x)[..., 0]
    ref = np.zeros((Nx, Ny, Nz // 2 + 1), dtype=complex)
    for kx in range(Nx):
        for ky in range(Ny):
            for kz in range(Nz // 2 + 1):
                for a in range(Nx):
                    for b in range(Ny):
                        for c in range(Nz):
                            ref[kx, ky, kz] += x[a, b, c, 0] * np.exp(-2j * np.pi * (kx * a / Nx + ky * b / Ny + kz * c / Nz))
    assert np.abs(X - ref).max() < 1e-12


# ---------------------------------------------------------- Green operators
def op(g21):
    """6x6 operator acting on tensor-shear Voigt vectors: (G:t)_a = sum_b G_ab t_b m_b."""
    G = np.array([[g21[WMAP[a][b]] for b in range(6)] for a in range(6)])
    return G * M[None, :]


@pytest.mark.parametrize("zet", [(1.0, 0.0, 0.0), (1.0, 2.0, -3.0), (-4.0, 1.0, 5.0), (0.0, -2.0, 7.0)])
def test_green_identities(zet):
    la0, mu0 = 7.3e9, 4.1e9
    C0 = iso_C(la0, mu0)
    G = op(oracle.green21(zet, la0, mu0, True))
    D = op(oracle.green21(zet, la0, mu0, False))
    s = np.abs(G).max()
    assert np.abs(G @ C0 @ G - G).max() < 1e-12 * s                      # projector
    assert np.abs(D - (C0 - C0 @ G @ C0)).max() < 1e-12 * np.abs(D).max()  # Delta0 = C0 - C0 Gamma0 C0
    # Gamma0 C0 is the identity on compatible strains eps = sym(xi (x) u)
    xi = np.array(zet)
    for u in np.eye(3):
        e = np.array([(xi[i] * u[j] + xi[j] * u[i]) / 2 for i, j in VO])
        assert np.abs(G @ C0 @ e - e).max() < 1e-12 * np.abs(e).max()
    # and Delta0 kills equilibrated-compatible directions: D applied to C0 e is 0
    for u in np.eye(3):
        e = np.array([(xi[i] * u[j] + xi[j] * u[i]) / 2 for i, j in VO])
        assert np.abs(D @ e).max() < 1e-9 * np.abs(D).max() * np.abs(e).max()


# ----------------------------------------------------------------- material
def test_material_tables():
    n = 8
    ids = np.zeros((n, n, n), dtype=np.int32)
    ids[:, :, :3] = 1
    mat = np.array([[200e9, 0.3, np.inf, 0, 1, 0, 0], [5e9, 0.4, np.inf, 0, 1, 0, 0]])
    o = oracle.Elastic(n, n, n, "e", nthreads=1)
    o.material(mat, ids)
    t = o.matrices(2)
    info = o.info()
    la0, mu0 = info["la0"], info["mu0"]
    K0 = la0 + 2 * mu0 / 3
    for p, (E, nu) in enumerate(mat[:, :2]):
        la, mu = lame(E, nu)
        K = la + 2 * mu / 3
        C = iso_C(la, mu)
        assert np.allclose(t["C"][p], C, rtol=1e-14, atol=0)
        S = t["S"][p]
        assert abs(S[0, 0] - 1 / E) < 1e-12 / E and abs(S[0, 1] + nu / E) < 1e-12 / E
        assert abs(S[3, 3] - 1 / (2 * mu)) < 1e-12 / mu
        # spectral closed forms: C_acc = 2K0/(K+K0) J + 2mu0/(mu+mu0) Kdev
        J = np.zeros((6, 6)); J[:3, :3] = 1 / 3
        Kd = np.diag([1, 1, 1, 1, 1, 1.0]) - J
        assert np.abs(t["Cacc"][p] - (2 * K0 / (K + K0) * J + 2 * mu0 / (mu + mu0) * Kd)).max() < 1e-13
        assert np.abs(t["Sacc"][p] - (2 * K / (K + K0) * J + 2 * mu / (mu + mu0) * Kd)).max() < 1e-13
    assert np.allclose(t["vol"], [5 / 8, 3 / 8])


# -------------------------------------------------------------------- eqVar
def test_eqvar_single_modes():
    Nx, Ny, Nz = 4, 6, 8
    N = Nx * Ny * Nz
    o = oracle.Elastic(Nx, Ny, Nz, "e", nthreads=1)
    x, y, z = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")
    sig = np.zeros((Nx, Ny, Nz, 6))
    sig[..., 0] = 1 + np.cos(2 * np.pi * z / Nz)          # kz = 1, weight xi0^2 = 1
    eq, cc, dd = o.eqvar(sig.reshape(-1, 6))
    assert abs(dd - N) < 1e-9 and abs(cc - N / 2) < 1e-9 and abs(eq - 1 / (2 * N)) < 1e-15
    sig[..., 3] = np.cos(2 * np.pi * (z / Nz - y / Ny))    # (kz=1, ky=-1): linear shear weight 2*xi1 = -2
    eq2, cc2, dd2 = o.eqvar(sig.reshape(-1, 6))
    assert abs(dd2 - N) < 1e-9
    assert abs(cc2 - (N / 2 + (N / 2) * np.sqrt(2))) < 1e-9  # the two modes sit in different bins
    sig[..., 0] = 1.0
    eq3, cc3, _ = o.eqvar(sig.reshape(-1, 6))
    assert abs(cc3 - (N / 2) * np.sqrt(2)) < 1e-9


# ------------------------------------------------------------ full solver
def test_homogeneous_rve():
    n = 8
    ids = np.ones((n, n, n), dtype=np.int32)
    ids[0, 0, 0] = 0
    E, nu = 70e9, 0.3
    mat = np.array([[E, nu, np.inf, 0, 1, 0, 0], [E, nu, np.inf, 0, 1, 0, 0]])
    o = oracle.Elastic(n, n, n, "e", nthreads=2)
    o.material(mat, ids)
    r = o.homogenization(1e-6, 1e-6)
    la, mu = lame(E, nu)
    C = iso_C(la, mu)
    assert np.abs(r["C"] - C).max() < 1e-9 * np.abs(C).max()
    assert r["iters"] == [2] * 6


def laminate_exact(mats, fracs, k, E, strain_base):
    """Mean fields of a laminate with normal along tensor axis k under the
    mean strain E (tensor-shear Voigt): phase strains eps_p = E + sym(a_p n),
    traction sigma_p n equal in all phases, sum f_p a_p = 0."""
    n = np.zeros(3); n[k] = 1
    P = len(mats)
    Cs = [iso_C(*lame(*m)) for m in mats]

    def sym(a):
        T = (np.outer(a, n) + np.outer(n, a)) / 2
        return np.array([T[i, j] for i, j in VO])

    def trac(s):
        T = np.zeros((3, 3))
        for q, (i, j) in enumerate(VO):
            T[i, j] = T[j, i] = s[q]
        return T @ n

    # unknowns a_p (3P); equations: traction_p = traction_0 (3(P-1)), sum f a = 0 (3)
    A = np.zeros((3 * P, 3 * P)); b = np.zeros(3 * P)
    basis = [sym(e) for e in np.eye(3)]
    for p in range(1, P):
        for c in range(3):
            for d in range(3):
                A[3 * (p - 1) + c, 3 * p + d] += trac(Cs[p] @ basis[d])[c]
                A[3 * (p - 1) + c, d] -= trac(Cs[0] @ basis[d])[c]
            b[3 * (p - 1) + c] = trac(Cs[0] @ E)[c] - trac(Cs[p] @ E)[c]
    for c in range(3):
        for p in range(P):
            A[3 * (P - 1) + c, 3 * p + c] = fracs[p]
    a = np.linalg.solve(A, b).reshape(P, 3)
    eps = sum(f * (E + sym(ap)) for f, ap in zip(fracs, a))
    sig = sum(f * Cp @ (E + sym(ap)) for f, Cp, ap in zip(fracs, Cs, a))
    return eps, sig


@pytest.mark.parametrize("axis,layers", [(0, 5), (1, 5), (2, 5), (0, 7), (2, 7)])
def test_elastic_laminate_exact(axis, layers):
    n = 12
    ids = np.ones((n, n, n), dtype=np.int32)
    sl = [slice(None)] * 3
    sl[2 - axis] = slice(0, layers)  # tensor axis k is array axis 2-k (index 0 = z)
    ids[tuple(sl)] = 0
    mat = np.array([[300e9, 0.25, np.inf, 0, 1, 0, 0], [30e9, 0.35, np.inf, 0, 1, 0, 0]])
    o = oracle.Elastic(n, n, n, "e", nthreads=4)
    o.material(mat, ids)
    sb = o.info()["strain_base"]
    r = o.homogenization(1e-11, 1e-11)
    f0 = layers / n
    assert sb == (layers > n / 2)
    ep = np.zeros((6, 6)); sg = np.zeros((6, 6))
    for w in range(6):
        E = np.zeros(6)
        E[w] = 0.0005 if sb else 1.0
        if sb:
            e, s = laminate_exact(mat[:, :2], [f0, 1 - f0], axis, E, sb)
        else:  # stress control: find the mean strain giving <sigma> = e_w * 1e6 (linear)
            Cl = np.stack([laminate_exact(mat[:, :2], [f0, 1 - f0], axis, b, sb)[1] for b in np.eye(6)], 1)
            Ebar = np.linalg.solve(Cl, np.eye(6)[w] * 1e6)
            e, s = laminate_exact(mat[:, :2], [f0, 1 - f0], axis, Ebar, sb)
        ep[:, w] = e
        sg[w, :] = s
    Cref = np.linalg.inv(ep) @ sg if sb else np.linalg.inv(np.linalg.inv(sg) @ ep)
    assert np.abs(r["C"] - Cref).max() < 1e-6 * np.abs(Cref).max()


def test_hashin_shtrikman_bounds_contain_ceff():
    n = 24
    ids = gen.spheres(n, 3, 0.2, seed=4)
    o = oracle.Elastic(n, n, n, "e", nthreads=4)
    o.material(gen.MAT_C1, ids)
    r = o.homogenization(1e-6, 1e-6)
    Emin, Kmin, Gmin, Emax, Kmax, Gmax = r["hs"]
    S = r["S"]
    for i in range(3):
        assert Emin <= 1 / S[i, i] <= Emax
    assert Emin < Emax


@pytest.mark.parametrize("km,ki", [(25.0, 30.0), (1.0, 100.0)])
def test_thermal_laminate_series_parallel(km, ki):
    n = 12
    ids = np.ones((n, n, n), dtype=np.int32)
    ids[:4] = 0  # layers normal to x (array axis 0 = tensor index 2)
    o = oracle.Thermal(n, n, n, nthreads=4)
    o.material(km, ki, ids)
    r = o.homogenization(1e-11, 1e-11)
    f = 4 / n
    par = f * ki + (1 - f) * km
    ser = 1 / (f / ki + (1 - f) / km)
    K = r["K"]
    assert abs(K[0, 0] - par) < 1e-7 * par and abs(K[1, 1] - par) < 1e-7 * par
    assert abs(K[2, 2] - ser) < 1e-7 * ser
    assert np.abs(K - np.diag(np.diag(K))).max() < 1e-9 * par


def test_thermal_homogeneous_and_nphase_consistent():
    n = 8
    ids = gen.spheres(n, 2, 0.2, seed=1)
    o = oracle.Thermal(n, n, n, nthreads=2)
    o.material(5.0, 5.0, ids)
    r = o.homogenization(1e-8, 1e-8)
    assert np.abs(r["K"] - 5 * np.eye(3)).max() < 1e-10
    # the N-phase table with 2 phases picks its base by volume (the 2-phase
    # reference rule effectively by conductivity); both schemes converge to the
    # same K at tight tolerance on an odd grid (on even grids the two bases
    # use opposite Nyquist signs, FFT_thermal3D.cpp:113-120 vs :150-157)
    n = 9
    ids = gen.spheres(n, 2, 0.2, seed=1)
    a = oracle.Thermal(n, n, n, nthreads=2)
    a.material(25.0, 30.0, ids)
    ra = a.homogenization(1e-11, 1e-11)
    b = oracle.Thermal(n, n, n, nthreads=2)
    b.material_n(np.array([30.0, 25.0]), ids)
    rb = b.homogenization(1e-11, 1e-11)
    assert a.info()["grad_base"] != b.info()["grad_base"]
    assert np.abs(ra["K"] - rb["K"]).max() < 1e-7 * 30


def test_j2_uniaxial_linear_hardening():
    """Homogeneous J2 RVE under mixed (uniaxial stress) control along tensor
    axis 0: sigma = E eps below yield, eps = sigma/E + (sigma - sy)/H above."""
    n = 4
    ids = np.ones((n, n, n), dtype=np.int32)
    ids[0, 0, 0] = 0
    E, nu, sy, H = 70e9, 0.3, 100e6, 1e9
    mat = np.array([[E, nu, sy, H, 1, 0, 0], [E, nu, sy, H, 1, 0, 0]])
    o = oracle.Elastic(n, n, n, "ep", nthreads=1)
    o.material(mat, ids)
    X0 = np.eye(6)[0]
    o.set_load(0.004 * X0, X0)
    o.run(1e-10, 1e-10, 20)
    P = o.prop_homogenized()
    for row in P:
        eps, s0 = row[0], row[1]
        if s0 <= sy:
            ref = E * eps
        else:
            ref = (eps + sy / H) / (1 / E + 1 / H)
        assert abs(s0 - ref) < 1e-6 * max(abs(ref), sy), (eps, s0, ref)
        assert np.abs(row[2:7]).max() < 1e-6 * max(abs(s0), 1.0)  # uniaxial: other components vanish
    assert P[-1, 1] > sy  # the path went plastic


@pytest.mark.parametrize("E,nu,sy,H,n_exp,emax", j2m.POWER_LAW_CASES)
def test_j2_uniaxial_power_law(E, nu, sy, H, n_exp, emax):
    """Homogeneous "ep" RVE with hardening exponent n = 0.2 under mixed
    (uniaxial stress) control along tensor axis 0 against the scalar root of
    the power-law flow rule (tests/j2_model.py).  An initial plastic strain
    of 0 instead of EP0 moves the yield point by H EP0^0.2 (1.7% of H); a
    Newton slip (wrong derivative, exponent n instead of n-1) moves every
    plastic row."""
    n = 4
    ids = np.ones((n, n, n), dtype=np.int32)
    ids[0, 0, 0] = 0
    mat = np.array([[E, nu, sy, H, n_exp, 0, 0]] * 2)
    o = oracle.Elastic(n, n, n, "ep", nthreads=1)
    o.material(mat, ids)
    X0 = np.eye(6)[0]
    o.set_load(emax * X0, X0)
    assert o.run(1e-10, 1e-10, 20) > 0  # -1: the return map failed (NaN or no convergence)
    P = o.prop_homogenized()
    assert P.shape == (20, 9)
    assert j2m.check_uniaxial_rows(P, E, sy, H, n_exp) >= 10


def test_damage_uniaxial_closed_form():
    """Homogeneous "ed" RVE under mixed (uniaxial stress) control along tensor
    axis 0: Y = sqrt(eps:C eps) = sqrt(E) eps, d = 1 - exp(-P (Y - Y0)) once
    Y > Y0 (monotone loading), sigma = (1 - d) E eps (FFT_analysis3D.cpp:619-638)."""
    n = 4
    ids = np.ones((n, n, n), dtype=np.int32)
    ids[0, 0, 0] = 0
    E, nu, Pd, Y0 = 70e9, 0.3, 5e-4, 500.0
    mat = np.array([[E, nu, np.inf, 0, 1, Pd, Y0], [E, nu, np.inf, 0, 1, Pd, Y0]])
    o = oracle.Elastic(n, n, n, "ed", nthreads=1)
    o.material(mat, ids)
    X0 = np.eye(6)[0]
    o.set_load(0.004 * X0, X0)
    o.run(1e-10, 1e-10, 10)
    P = o.prop_homogenized()
    assert P.shape == (10, 9)
    for eps, s0, d in P[:, [0, 1, 7]]:
        Y = np.sqrt(E) * eps
        dref = 1 - np.exp(-Pd * (Y - Y0)) if Y > Y0 else 0.0
        assert abs(d - dref) < 1e-8, (eps, d, dref)
        sref = (1 - dref) * E * eps
        assert abs(s0 - sref) < 1e-6 * abs(sref), (eps, s0, sref)
    assert P[-1, 7] > 0.1  # the path went well into damage


@pytest.mark.parametrize("axis", [0, 2])
def test_basic_scheme_laminate_exact(axis):
    """Basic scheme with the load held fixed ("ep" medium: arithmetic means,
    strain base) converges to the exact laminate solution."""
    n = 12
    layers = 5
    ids = np.ones((n, n, n), dtype=np.int32)
    sl = [slice(None)] * 3
    sl[2 - axis] = slice(0, layers)
    ids[tuple(sl)] = 0
    mat = np.array([[300e9, 0.25, np.inf, 0, 1, 0, 0], [30e9, 0.35, np.inf, 0, 1, 0, 0]])
    o = oracle.Elastic(n, n, n, "ep", nthreads=4)
    o.set_scheme(1)
    o.material(mat, ids)
    assert o.info()["strain_base"]
    r = o.homogenization(1e-12, 1e-12)
    f0 = layers / n
    ep = np.zeros((6, 6)); sg = np.zeros((6, 6))
    for w in range(6):
        E = np.zeros(6); E[w] = 0.0005
        e, s = laminate_exact(mat[:, :2], [f0, 1 - f0], axis, E, True)
        ep[:, w] = e
        sg[w, :] = s
    Cref = np.linalg.inv(ep) @ sg
    assert np.abs(r["C"] - Cref).max() < 1e-6 * np.abs(Cref).max()


def test_basic_and_accelerated_schemes_share_the_fixed_point():
    """Same RVE, same reference medium: the basic scheme (C_acc = I) needs more
    iterations but converges to the accelerated scheme's C_eff."""
    n = 16
    ids = gen.spheres(n, 3, 0.2, seed=3)
    out = {}
    for basic in (0, 1):
        o = oracle.Elastic(n, n, n, "ep", nthreads=4)
        o.set_scheme(basic)
        o.material(gen.MAT_C1, ids)
        out[basic] = o.homogenization(1e-10, 1e-10)
    rel = np.linalg.norm(out[1]["C"] - out[0]["C"]) / np.linalg.norm(out[0]["C"])
    assert rel < 1e-8
    assert min(out[1]["iters"]) > max(out[0]["iters"])
