"""Closed-form model of the accelerated (Eyre-Milton) fixed-point iteration on
a laminate, used to pin the stop decision of the oracle and of the GPU path.

On an RVE made of flat layers normal to one grid axis, with a uniform load,
every field of the reference iteration is constant within each phase: the
initial field is uniform, the constitutive maps act per voxel, and the Green
operator of a field that varies along one axis n only acts as one fixed
tensor (that of the direction n) on its fluctuation.  The whole iteration is
then a recursion on one tensor per phase, with the Green operators written
from continuum mechanics, not from the reference's component tables:

* elastic strain base: eps_c = E - Gamma(n) (tau - <tau>), tau = sigma - C0 eps,
  Gamma(n) t = sym(n (x) A^-1 t n), A = mu0 I + (lambda0 + mu0) n n the
  acoustic tensor of C0 along n;
* elastic stress base: sigma_c = E - Delta(n) (tau - <tau>), tau = eps - S0 sigma,
  Delta = C0 - C0 Gamma C0;
* thermal gradient base: g_c = E + (n n / k0) (tau - <tau>), tau = k0 g - q;
* thermal flux base: q_c = E + k0 (I - n n) (tau - <tau>), tau = q / k0 - g;

followed by the relaxation u <- u + A_acc (u_c - u) with A_acc = 2 (M + M0)^-1 M0
(spectral form of an isotropic map) and the dual field.  The spectrum of a
phase-wise constant field along n is the field's per-phase value times the
1D DFT of the phase indicator, so the equilibrium measure of the check and
the change of the field sum are evaluated in closed form per iteration.  The
rules that decide where a solve stops are those of the reference's text:
reference medium (FFT_analysis3D.cpp:496-501 geometric K0 and mu0 for "e",
FFT_thermal3D.cpp:415 k0 = sqrt(ki km)), base choice (:503-512,
FFT_thermal3D.cpp:401-410), Delta-sum normalisation with phases 0 and 1
(:399,408; FFT_thermal3D.cpp:346,356), eqVar weights and frequency ranges
(:351-392; FFT_thermal3D.cpp:303-337), stop rule and caps (:785-792;
FFT_thermal3D.cpp:474-480).
"""
from __future__ import annotations

import numpy as np

VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
I3 = np.eye(3)


def to_tensor(v):
    t = np.zeros((3, 3))
    for c, (i, j) in enumerate(VOIGT):
        t[i, j] = t[j, i] = v[c]
    return t


def to_voigt(t):
    return np.array([t[i, j] for i, j in VOIGT])


def lame(E, nu):
    mu = E / (2 * (1 + nu))
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    return lam, mu


def iso(lam, mu, t):
    """lam tr(t) I + 2 mu t"""
    return lam * np.trace(t) * I3 + 2 * mu * t


def iso_spectral(a_sph, a_dev, t):
    """the isotropic map a_sph on the spherical part, a_dev on the deviator"""
    s = np.trace(t) / 3 * I3
    return a_sph * s + a_dev * (t - s)


def axis_vector(axis):
    """unit vector of grid axis 'z' | 'y' | 'x' in tensor indices (0 = z)"""
    n = np.zeros(3)
    n[{"z": 0, "y": 1, "x": 2}[axis]] = 1.0
    return n


def gamma_strain(n, lam0, mu0, t):
    Ainv = (I3 - np.outer(n, n)) / mu0 + np.outer(n, n) / (lam0 + 2 * mu0)
    v = Ainv @ (t @ n)
    w = np.outer(n, v)
    return 0.5 * (w + w.T)


def fmap(k, N, thermal):
    """signed frequency of DFT index k: the mechanical sites map k > (N-1)/2,
    the thermal ones k > N/2, to -(N - k)"""
    lim = N // 2 if thermal else (N - 1) // 2
    return np.where(k <= lim, k, -(N - k)).astype(float)


class Layers:
    """pattern[s]: phase of the layer at position s along axis (length Na);
    shape: (Nx, Ny, Nz) of the grid."""

    def __init__(self, pattern, axis, shape):
        self.pattern = np.asarray(pattern, dtype=np.int32)
        self.axis = axis
        self.shape = tuple(shape)
        self.N = int(np.prod(shape))
        Na = {"x": shape[0], "y": shape[1], "z": shape[2]}[axis]
        assert self.pattern.size == Na
        self.Na = Na
        self.Nperp = self.N // Na
        self.nph = int(self.pattern.max()) + 1
        self.frac = np.array([np.mean(self.pattern == p) for p in range(self.nph)])
        self.chi = np.array([np.fft.fft((self.pattern == p).astype(float)) for p in range(self.nph)])

    def rve(self):
        """phase ids in the reference's voxel order (z fastest)"""
        Nx, Ny, Nz = self.shape
        ids = np.empty((Nx, Ny, Nz), dtype=np.int32)
        if self.axis == "x":
            ids[:] = self.pattern[:, None, None]
        elif self.axis == "y":
            ids[:] = self.pattern[None, :, None]
        else:
            ids[:] = self.pattern[None, None, :]
        return ids

    def check_modes(self, thermal):
        """DFT indices k along the axis that the equilibrium check sums, and
        their signed frequencies: the mechanical check sums kz >= 1 of the half
        spectrum, the thermal one kz >= 0 (FFT_analysis3D.cpp:368,
        FFT_thermal3D.cpp:317); a layer normal along x or y has its spectrum
        in the kz = 0 plane."""
        if self.axis == "z":
            k = np.arange(0 if thermal else 1, self.Na // 2 + 1)
        else:
            k = np.arange(self.Na) if thermal else np.arange(0)
        return k, fmap(k, self.Na, thermal)


def _eqvar_elastic(L, sig):
    """cc / dd / N of FFT_analysis3D.cpp:351-392 for phase-wise stresses sig[p] (Voigt)"""
    N = L.N
    mean = sum(L.frac[p] * sig[p] for p in range(L.nph))
    w = np.array([1, 1, 1, 2, 2, 2.0])
    dd = np.sqrt(np.sum(w * (N * mean) ** 2))
    k, xi_a = L.check_modes(False)
    cc = 0.0
    ia = {"z": 0, "y": 1, "x": 2}[L.axis]
    for kk, xa in zip(k, xi_a):
        sh = L.Nperp * sum(sig[p] * L.chi[p][kk] for p in range(L.nph))
        xi = np.zeros(3)
        xi[ia] = xa
        W = np.array([xi[0] ** 2, xi[1] ** 2, xi[2] ** 2, 2 * xi[1] + 2 * xi[2], 2 * xi[0] + 2 * xi[2],
                      2 * xi[0] + 2 * xi[1]])
        cc += np.sqrt(abs(np.sum(np.abs(sh) ** 2 * W)))
    return cc / dd / N


def _eqvar_thermal(L, q):
    N = L.N
    mean = sum(L.frac[p] * q[p] for p in range(L.nph))
    dd = np.sqrt(np.sum((N * mean) ** 2))
    k, xi_a = L.check_modes(True)
    ia = {"z": 0, "y": 1, "x": 2}[L.axis]
    cc = 0.0
    for kk, xa in zip(k, xi_a):
        sh = L.Nperp * sum(q[p] * L.chi[p][kk] for p in range(L.nph))
        cc += np.sqrt(np.abs(sh[ia]) ** 2 * xa * xa)
    return cc / dd / N


def elastic_medium(lam, mu):
    """reference medium of behaviour "e" (FFT_analysis3D.cpp:496-501)"""
    K = lam + 2 * mu / 3
    K0 = np.sqrt(K.min() * K.max())
    mu0 = np.sqrt(mu.min() * mu.max())
    return K0 - 2 * mu0 / 3, mu0


def elastic_strain_base(mu, frac):
    """the first phase of largest mu must hold the largest volume fraction
    (FFT_analysis3D.cpp:503-512)"""
    return frac[int(np.argmax(mu))] == frac.max()


def solve_elastic(L, E_list, nu, tol_field, tol_eq, cap=2000):
    """Run the homogenization's load cases on the laminate.  E_list: Young's
    moduli per phase, nu: Poisson ratios.  Returns dict(strain_base, la0, mu0,
    runs=[dict(iters, dsum=[...], eqvar=[...], deq=[...])] per load case)."""
    lam, mu = lame(np.asarray(E_list, float), np.asarray(nu, float))
    K = lam + 2 * mu / 3
    la0, mu0 = elastic_medium(lam, mu)
    K0 = la0 + 2 * mu0 / 3
    sb = elastic_strain_base(mu, L.frac)
    n = axis_vector(L.axis)
    C00 = lam + 2 * mu          # C[p](0,0)
    S00 = 1.0 / np.asarray(E_list, float)  # S[p](0,0)
    P = range(L.nph)
    p1 = 1 if L.nph > 1 else 0
    C0 = lambda t: iso(la0, mu0, t)
    S0 = lambda t: iso_spectral(1 / (3 * K0), 1 / (2 * mu0), t)
    runs = []
    for w in range(6):
        Ev = np.zeros(6)
        Ev[w] = 0.0005 if sb else 1e6     # FFT_analysis3D.cpp:872-875
        Et = to_tensor(Ev)
        u = [Et.copy() for _ in P]      # the unknown: eps (strain base) or sigma
        if sb:
            dual = [iso(lam[p], mu[p], u[p]) for p in P]
        else:
            dual = [iso_spectral(1 / (3 * K[p]), 1 / (2 * mu[p]), u[p]) for p in P]
        eq_old = 0.0
        it = 0
        rec = dict(dsum=[], eqvar=[], deq=[])
        while True:
            sum_old = L.N * sum(L.frac[p] * to_voigt(dual[p]).sum() for p in P)
            if sb:
                tau = [dual[p] - C0(u[p]) for p in P]
            else:
                tau = [dual[p] - S0(u[p]) for p in P]
            tbar = sum(L.frac[p] * tau[p] for p in P)
            if sb:
                uc = [Et - gamma_strain(n, la0, mu0, tau[p] - tbar) for p in P]
            else:
                uc = []
                for p in P:
                    t = tau[p] - tbar
                    ct = C0(t)
                    uc.append(Et - (ct - C0(gamma_strain(n, la0, mu0, ct))))
            for p in P:
                if sb:
                    a = iso_spectral(2 * K0 / (K[p] + K0), 2 * mu0 / (mu[p] + mu0), uc[p] - u[p])
                else:
                    a = iso_spectral(2 * K[p] / (K[p] + K0), 2 * mu[p] / (mu[p] + mu0), uc[p] - u[p])
                u[p] = u[p] + a
            if sb:
                dual = [iso(lam[p], mu[p], u[p]) for p in P]
                sig = [to_voigt(dual[p]) for p in P]
                norm = 0.5 * C00[0] + 0.5 * C00[p1]
            else:
                dual = [iso_spectral(1 / (3 * K[p]), 1 / (2 * mu[p]), u[p]) for p in P]
                sig = [to_voigt(u[p]) for p in P]
                norm = 0.5 * S00[0] + 0.5 * S00[p1]
            it += 1
            eqv = _eqvar_elastic(L, sig)
            tot = L.N * sum(L.frac[p] * to_voigt(dual[p]).sum() for p in P)
            dsum = abs((tot - sum_old) / (L.N * norm * Ev.sum()))
            deq = abs(eqv - eq_old)
            eq_old = eqv
            rec["dsum"].append(dsum)
            rec["eqvar"].append(eqv)
            rec["deq"].append(deq)
            if it > 1 and deq < tol_eq and dsum < tol_field:
                break
            if it > cap:
                break
        rec["iters"] = it
        runs.append(rec)
    return dict(strain_base=bool(sb), la0=la0, mu0=mu0, runs=runs)


def thermal_setup(k, L, two_phase):
    """(gradient base, k0).  two_phase: the reference's material(km, ki) with
    k = [ki, km] (FFT_thermal3D.cpp:373-422: vol_1 = RVE.sum()/xyz is an integer
    division, 0 unless every voxel is matrix); otherwise the N-phase rule of
    the library (k0 = sqrt(kmin kmax), gradient base iff the first largest-k
    phase holds the largest volume fraction)."""
    k = np.asarray(k, float)
    if two_phase:
        ki, km = k[0], k[1]
        vol_1 = float(np.count_nonzero(L.pattern) * L.Nperp // L.N)
        grad = (vol_1 >= 0.5 and km > ki) or (vol_1 < 0.5 and km < ki)
        return bool(grad), float(np.sqrt(ki * km))
    return bool(L.frac[int(np.argmax(k))] == L.frac.max()), float(np.sqrt(k.min() * k.max()))


def solve_thermal(L, k, tol_field, tol_eq, two_phase=False, cap=500):
    """Run the three load cases E = e_w on the laminate (FFT_thermal3D.cpp:489-548)."""
    k = np.asarray(k, float)
    grad, k0 = thermal_setup(k, L, two_phase)
    n = axis_vector(L.axis)
    nn = np.outer(n, n)
    P = range(L.nph)
    p1 = 1 if L.nph > 1 else 0
    runs = []
    for w in range(3):
        E = np.zeros(3)
        E[w] = 1.0
        u = [E.copy() for _ in P]       # gradient (gradient base) or flux
        dual = [k[p] * u[p] if grad else u[p] / k[p] for p in P]
        eq_old, it = 0.0, 0
        rec = dict(dsum=[], eqvar=[], deq=[])
        while True:
            sum_old = L.N * sum(L.frac[p] * dual[p].sum() for p in P)
            tau = [k0 * u[p] - dual[p] if grad else u[p] / k0 - dual[p] for p in P]
            tbar = sum(L.frac[p] * tau[p] for p in P)
            if grad:
                uc = [E + nn @ (tau[p] - tbar) / k0 for p in P]
                u = [u[p] + 2 * k0 / (k[p] + k0) * (uc[p] - u[p]) for p in P]
                dual = [k[p] * u[p] for p in P]
                q = dual
                norm = 0.5 * k[0] + 0.5 * k[p1]
            else:
                uc = [E + k0 * (I3 - nn) @ (tau[p] - tbar) for p in P]
                u = [u[p] + 2 * k[p] / (k[p] + k0) * (uc[p] - u[p]) for p in P]
                dual = [u[p] / k[p] for p in P]
                q = u
                norm = 0.5 / k[0] + 0.5 / k[p1]
            it += 1
            eqv = _eqvar_thermal(L, q)
            tot = L.N * sum(L.frac[p] * dual[p].sum() for p in P)
            dsum = abs((tot - sum_old) / (L.N * norm * E.sum()))
            deq = abs(eqv - eq_old)
            eq_old = eqv
            rec["dsum"].append(dsum)
            rec["eqvar"].append(eqv)
            rec["deq"].append(deq)
            if it > 1 and deq < tol_eq and dsum < tol_field:
                break
            if it > cap:
                break
        rec["iters"] = it
        rec["u"], rec["q"] = u, q
        runs.append(rec)
    return dict(grad_base=grad, k0=k0, runs=runs)


def margin(run, tol_field, tol_eq):
    """smallest relative distance of any stop-rule quantity from its
    tolerance over the run: the stop iteration is a rounding-robust
    prediction only if this is well above rounding"""
    m = np.inf
    for ds, de in zip(run["dsum"], run["deq"]):
        m = min(m, abs(ds / tol_field - 1), abs(de / tol_eq - 1))
    return m
