"""FP64 CPU oracle for the 3D FFT homogenization loop -- TEST INFRASTRUCTURE.

This package is the parity checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it.
It wraps ``m3d_oracle.c`` (a double-precision C restatement of
src/3d/FFT_analysis3D.cpp and src/3dTherm/FFT_thermal3D.cpp) through ctypes.

Array conventions follow the reference (and the product's C-ABI):
  * real fields are voxel-major ``[N, C]`` (Eigen ``C x N`` column-major);
    voxel index ``i = z + Nz*(y + Ny*x)`` (z fastest);
  * half spectra are ``[Nx, Ny, Nz//2+1, C]`` complex;
  * material tables are ``nph x 7`` ``[E, nu, sigma_y, H, n, P, Y0]``.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_SRC = os.path.join(_HERE, "m3d_oracle.c")

BEH = {"e": 0, "ep": 1, "ed": 2}


def build(force: bool = False) -> str:
    """Compile m3d_oracle.c into oracle/liboracle.so (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(_SRC):
        cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-o", _LIB_PATH + ".tmp", _SRC, "-lm"]
        subprocess.check_call(cmd)
        os.replace(_LIB_PATH + ".tmp", _LIB_PATH)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        vp, ip, dp = ct.c_void_p, ct.c_int, ct.POINTER(ct.c_double)
        i32p = ct.POINTER(ct.c_int32)
        ipp = ct.POINTER(ct.c_int)
        sig = {
            "m3o_rfft3": (None, [ip, ip, ip, ip, dp, dp, ip]),
            "m3o_irfft3": (None, [ip, ip, ip, ip, dp, dp, ip]),
            "m3o_el_create": (vp, [ip, ip, ip, ip, ip]),
            "m3o_el_destroy": (None, [vp]),
            "m3o_el_set_cap": (None, [vp, ip]),
            "m3o_el_set_threads": (None, [vp, ip, ip]),
            "m3o_el_set_scheme": (None, [vp, ip]),
            "m3o_el_material": (ip, [vp, dp, i32p, ip]),
            "m3o_el_set_load": (None, [vp, dp, dp]),
            "m3o_el_run": (ip, [vp, ct.c_double, ct.c_double, ip]),
            "m3o_el_homogenize": (ct.c_double, [vp, ct.c_double, ct.c_double, dp, dp, dp, dp, dp, ipp]),
            "m3o_el_fields": (None, [vp, dp, dp]),
            "m3o_el_info": (None, [vp, dp]),
            "m3o_el_matrices": (None, [vp, dp, dp, dp, dp, dp, dp, dp]),
            "m3o_el_history": (ip, [vp, dp, ip]),
            "m3o_el_clear_history": (None, [vp]),
            "m3o_el_prop": (None, [vp, dp, ipp]),
            "m3o_el_j2_state": (None, [vp, dp, dp]),
            "m3o_el_damage_state": (None, [vp, dp]),
            "m3o_green21": (None, [dp, ct.c_double, ct.c_double, ip, dp]),
            "m3o_el_apply_green": (None, [vp, dp, dp]),
            "m3o_el_eqvar": (None, [vp, dp, dp]),
            "m3o_el_run_fixed": (ip, [vp, ip, ip]),
            "m3o_th_create": (vp, [ip, ip, ip, ip]),
            "m3o_th_destroy": (None, [vp]),
            "m3o_th_set_cap": (None, [vp, ip]),
            "m3o_th_material": (ip, [vp, dp, ip, i32p, ip]),
            "m3o_th_set_load": (None, [vp, dp]),
            "m3o_th_run": (ip, [vp, ct.c_double, ct.c_double]),
            "m3o_th_homogenize": (None, [vp, ct.c_double, ct.c_double, dp, dp, dp, dp, dp, ipp]),
            "m3o_th_fields": (None, [vp, dp, dp]),
            "m3o_th_info": (None, [vp, dp]),
            "m3o_th_history": (ip, [vp, dp, ip]),
            "m3o_th_clear_history": (None, [vp]),
            "m3o_rec_uniform": (ct.c_double, [ct.c_uint64, ct.c_uint64, ct.c_uint64]),
            "m3o_rec_mix": (ct.c_uint64, [ct.c_uint64]),
            "m3o_c2c3": (None, [ip, ip, ip, dp, ip, ip]),
            "m3o_rec_create": (vp, [dp, ip, ip, ip, ip]),
            "m3o_rec_destroy": (None, [vp]),
            "m3o_rec_set_initial": (None, [vp, dp]),
            "m3o_rec_run": (ip, [vp, ip, ct.c_double, ct.c_uint64, ipp, ipp]),
            "m3o_rec_result": (ip, [vp, dp, dp, ip]),
            "m3o_2d_create": (vp, [ip, ip, ip]),
            "m3o_2d_destroy": (None, [vp]),
            "m3o_2d_material": (None, [vp, dp, dp]),
            "m3o_2d_set_load": (None, [vp, dp]),
            "m3o_2d_run": (None, [vp]),
            "m3o_2d_homogenization": (None, [vp, dp, dp, dp]),
            "m3o_2d_fields": (None, [vp, dp, dp]),
            "m3o_perc_create": (vp, [i32p, ip, ip, ip, ip, ip]),
            "m3o_perc_destroy": (None, [vp]),
            "m3o_perc_run": (None, [vp]),
            "m3o_perc_run_rounds": (None, [vp, ip]),
            "m3o_perc_result": (ip, [vp, i32p, ipp, dp, dp, ip]),
            "m3o_perc_surface": (ip, [vp, i32p, ip, ip]),
            "m3o_perc_fixed_point": (ip, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _i32p(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_int32))


def _threads(n):
    return int(n) if n else int(os.environ.get("M3D_ORACLE_THREADS", os.cpu_count() or 1))


def rfft3(x: np.ndarray, nthreads: int = 0) -> np.ndarray:
    """Unnormalised rank-3 r2c of ``x[Nx,Ny,Nz,C]`` -> ``[Nx,Ny,Nz//2+1,C]`` complex."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    Nx, Ny, Nz, C = x.shape
    out = np.empty((Nx, Ny, Nz // 2 + 1, C), dtype=np.complex128)
    lib().m3o_rfft3(Nx, Ny, Nz, C, _dp(x), _dp(out.view(np.float64)), _threads(nthreads))
    return out


def irfft3(X: np.ndarray, Nz: int, nthreads: int = 0) -> np.ndarray:
    """Unnormalised rank-3 c2r of ``X[Nx,Ny,Nz//2+1,C]`` -> ``[Nx,Ny,Nz,C]``."""
    X = np.ascontiguousarray(X, dtype=np.complex128)
    Nx, Ny, Nzh, C = X.shape
    assert Nzh == Nz // 2 + 1
    out = np.empty((Nx, Ny, Nz, C), dtype=np.float64)
    lib().m3o_irfft3(Nx, Ny, Nz, C, _dp(X.view(np.float64)), _dp(out), _threads(nthreads))
    return out


def green21(zet, la0, mu0, strain_base: bool) -> np.ndarray:
    z = np.ascontiguousarray(zet, dtype=np.float64)
    g = np.empty(21)
    lib().m3o_green21(_dp(z), la0, mu0, 1 if strain_base else 0, _dp(g))
    return g


class Elastic:
    """Oracle mirror of ``FFT_Analysis3D`` (include/FFT_analysis3D.hpp:36-215)."""

    def __init__(self, Nx, Ny, Nz, behav="e", nthreads=0):
        self.shape = (Nx, Ny, Nz)
        self.N = Nx * Ny * Nz
        self._h = lib().m3o_el_create(Nx, Ny, Nz, BEH[behav], _threads(nthreads))
        if not self._h:
            raise MemoryError("oracle create failed")

    def close(self):
        if self._h:
            lib().m3o_el_destroy(self._h)
            self._h = None

    __del__ = close

    def set_cap(self, cap):
        lib().m3o_el_set_cap(self._h, int(cap))

    def set_threads(self, fft_threads, loop_threads):
        """threads of the FFTs and of the per-voxel loops (the reference: 3 and 1)"""
        lib().m3o_el_set_threads(self._h, int(fft_threads), int(loop_threads))

    def set_scheme(self, basic):
        """1: basic scheme with the load held fixed (update eps <- eps_c)."""
        lib().m3o_el_set_scheme(self._h, int(basic))

    def material(self, mat, rve, nph=None):
        mat = np.asarray(mat, dtype=np.float64)
        nph = mat.shape[0] if nph is None else nph
        matf = np.ascontiguousarray(mat.T)  # column-major nph x 7
        rve = np.ascontiguousarray(np.asarray(rve).reshape(-1), dtype=np.int32)
        assert rve.size == self.N
        rc = lib().m3o_el_material(self._h, _dp(matf), _i32p(rve), int(nph))
        if rc != 0:
            raise ValueError(f"oracle material() failed: {rc}")

    def set_load(self, E, X0=None):
        E = np.ascontiguousarray(E, dtype=np.float64)
        X0 = E if X0 is None else np.ascontiguousarray(X0, dtype=np.float64)
        lib().m3o_el_set_load(self._h, _dp(E), _dp(X0))

    def run(self, tol_field, tol_eq, nInc=1):
        return lib().m3o_el_run(self._h, tol_field, tol_eq, int(nInc))

    def run_fixed(self, w, k):
        return lib().m3o_el_run_fixed(self._h, int(w), int(k))

    def homogenization(self, tol_field, tol_eq):
        C, S, ep, sg, hs = (np.zeros(36), np.zeros(36), np.zeros(36), np.zeros(36), np.zeros(6))
        iters = (ct.c_int * 6)()
        ret = lib().m3o_el_homogenize(self._h, tol_field, tol_eq, _dp(C), _dp(S), _dp(ep), _dp(sg), _dp(hs), iters)
        f = lambda a: a.reshape(6, 6).T.copy()  # column-major -> numpy
        return dict(C=f(C), S=f(S), epHomo=f(ep), sigHomo=f(sg), hs=hs, ret=ret, iters=list(iters))

    def fields(self):
        eps = np.empty((self.N, 6))
        sig = np.empty((self.N, 6))
        lib().m3o_el_fields(self._h, _dp(eps), _dp(sig))
        return eps, sig

    def info(self):
        o = np.zeros(8)
        lib().m3o_el_info(self._h, _dp(o))
        return dict(strain_base=bool(o[0]), la0=o[1], mu0=o[2], itNo=int(o[3]), eqVar=o[4], diff_sum=o[5],
                    diff_eq=o[6], status=int(o[7]))

    def matrices(self, nph):
        C, S, Ca, Sa = (np.zeros(36 * nph) for _ in range(4))
        C0, S0, vol = np.zeros(36), np.zeros(36), np.zeros(nph)
        lib().m3o_el_matrices(self._h, _dp(C), _dp(S), _dp(Ca), _dp(Sa), _dp(C0), _dp(S0), _dp(vol))
        f = lambda a: a.reshape(-1, 6, 6).transpose(0, 2, 1).copy()
        return dict(C=f(C), S=f(S), Cacc=f(Ca), Sacc=f(Sa), C0=f(C0)[0], S0=f(S0)[0], vol=vol)

    def history(self):
        buf = np.zeros(4 * 4096)
        n = lib().m3o_el_history(self._h, _dp(buf), 4096)
        return buf[: 4 * n].reshape(n, 4)

    def clear_history(self):
        lib().m3o_el_clear_history(self._h)

    def prop_homogenized(self):
        rows = ct.c_int(0)
        lib().m3o_el_prop(self._h, None, ct.byref(rows))
        out = np.zeros(9 * max(rows.value, 1))
        lib().m3o_el_prop(self._h, _dp(out), ct.byref(rows))
        return out[: 9 * rows.value].reshape(9, rows.value).T.copy()

    def j2_state(self):
        a, b = np.zeros(self.N), np.zeros(self.N)
        lib().m3o_el_j2_state(self._h, _dp(a), _dp(b))
        return a, b

    def damage_state(self):
        d = np.zeros(self.N)
        lib().m3o_el_damage_state(self._h, _dp(d))
        return d

    def apply_green(self, t2):
        t2 = np.ascontiguousarray(t2, dtype=np.complex128)
        t3 = np.empty_like(t2)
        lib().m3o_el_apply_green(self._h, _dp(t2.view(np.float64)), _dp(t3.view(np.float64)))
        return t3

    def eqvar(self, sig):
        sig = np.ascontiguousarray(sig, dtype=np.float64)
        o = np.zeros(3)
        lib().m3o_el_eqvar(self._h, _dp(sig), _dp(o))
        return o


class Thermal:
    """Oracle mirror of ``FFT_analysis3d_thermal`` (include/FFT_thermal3D.hpp:37-154)."""

    def __init__(self, Nx, Ny, Nz, nthreads=0):
        self.shape = (Nx, Ny, Nz)
        self.N = Nx * Ny * Nz
        self._h = lib().m3o_th_create(Nx, Ny, Nz, _threads(nthreads))

    def close(self):
        if self._h:
            lib().m3o_th_destroy(self._h)
            self._h = None

    __del__ = close

    def set_cap(self, cap):
        lib().m3o_th_set_cap(self._h, int(cap))

    def material(self, km, ki, rve):
        """Reference 2-phase API: RVE != 0 -> km, RVE == 0 -> ki."""
        k = np.array([ki, km], dtype=np.float64)
        rve = np.ascontiguousarray(np.asarray(rve).reshape(-1), dtype=np.int32)
        rc = lib().m3o_th_material(self._h, _dp(k), 2, _i32p(rve), 1)
        if rc:
            raise ValueError(rc)

    def material_n(self, k, rve):
        """N-phase extension: conductivity table indexed by phase id."""
        k = np.ascontiguousarray(k, dtype=np.float64)
        rve = np.ascontiguousarray(np.asarray(rve).reshape(-1), dtype=np.int32)
        rc = lib().m3o_th_material(self._h, _dp(k), int(k.size), _i32p(rve), 0)
        if rc:
            raise ValueError(rc)

    def set_load(self, E):
        E = np.ascontiguousarray(E, dtype=np.float64)
        lib().m3o_th_set_load(self._h, _dp(E))

    def run(self, tol_field, tol_eq):
        return lib().m3o_th_run(self._h, tol_field, tol_eq)

    def homogenization(self, tol_field, tol_eq):
        K, R, G, Q, est = np.zeros(9), np.zeros(9), np.zeros(9), np.zeros(9), np.zeros(6)
        iters = (ct.c_int * 3)()
        lib().m3o_th_homogenize(self._h, tol_field, tol_eq, _dp(K), _dp(R), _dp(G), _dp(Q), _dp(est), iters)
        f = lambda a: a.reshape(3, 3).T.copy()
        return dict(K=f(K), R=f(R), tempGradHomo=f(G), heatFluxHomo=f(Q), est=est, iters=list(iters))

    def fields(self):
        tg = np.empty((self.N, 3))
        q = np.empty((self.N, 3))
        lib().m3o_th_fields(self._h, _dp(tg), _dp(q))
        return tg, q

    def info(self):
        o = np.zeros(6)
        lib().m3o_th_info(self._h, _dp(o))
        return dict(grad_base=bool(o[0]), k0=o[1], itNo=int(o[2]), eqVar=o[3], diff_sum=o[4], diff_eq=o[5])

    def history(self):
        buf = np.zeros(4 * 4096)
        n = lib().m3o_th_history(self._h, _dp(buf), 4096)
        return buf[: 4 * n].reshape(n, 4)


def c2c3(x: np.ndarray, sign: int, nthreads: int = 0) -> np.ndarray:
    """Unnormalised complex 3D FFT of x[Nx, Ny, Nz] (sign -1 forward, +1 backward)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    Nx, Ny, Nz = a.shape
    lib().m3o_c2c3(Nx, Ny, Nz, a.view(np.float64).ctypes.data_as(ct.POINTER(ct.c_double)), int(sign), _threads(nthreads))
    return a


def rec_uniform(seed: int, stream: int, i: int) -> float:
    return lib().m3o_rec_uniform(seed, stream, i)


def rec_mix(z: int) -> int:
    return lib().m3o_rec_mix(z)


class Rec:
    """recPlan(double *TPCF, Nx, Ny, Nz) + runRec(n, beta)
    (src/rec/reconstruction.cpp:211-243, 285-406)."""

    def __init__(self, tpcf, nthreads=0):
        t = np.ascontiguousarray(tpcf, dtype=np.float64)
        self.shape = t.shape
        self._tpcf = t
        self._h = lib().m3o_rec_create(_dp(t), *t.shape, _threads(nthreads))
        self._init = None

    def close(self):
        if self._h:
            lib().m3o_rec_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_initial(self, m0):
        self._init = np.ascontiguousarray(m0, dtype=np.float64).reshape(-1)
        lib().m3o_rec_set_initial(self._h, _dp(self._init))

    def run(self, n, beta, seed):
        st, ctr = ct.c_int(0), ct.c_int(0)
        it = lib().m3o_rec_run(self._h, int(n), float(beta), int(seed), ct.byref(st), ct.byref(ctr))
        self.status, self.search_steps = st.value, ctr.value
        return it

    def result(self):
        """(mTilda as complex [Nx, Ny, Nz], trackrealerrorA)"""
        m = np.zeros(self.shape, dtype=np.complex128)
        err = np.zeros(max(1, 100000), dtype=np.float64)
        k = lib().m3o_rec_result(self._h, m.view(np.float64).ctypes.data_as(ct.POINTER(ct.c_double)), _dp(err),
                                 err.size)
        return m, err[:k].copy()


class Elastic2D:
    """FFT_Analysis2D (src/2d/FFT_analysis2D.cpp): basic scheme on complex 2D
    fields [E11, E22, E12], 25 fixed iterations per runAnalysis."""

    def __init__(self, Nx, Ny, nthreads=0):
        self.shape = (Nx, Ny)
        self._h = lib().m3o_2d_create(Nx, Ny, _threads(nthreads))

    def close(self):
        if self._h:
            lib().m3o_2d_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def material(self, props, rve):
        self._q = np.ascontiguousarray(props, dtype=np.float64)
        self._r = np.ascontiguousarray(rve, dtype=np.float64).reshape(-1)
        lib().m3o_2d_material(self._h, _dp(self._q), _dp(self._r))

    def set_load(self, E):
        e = np.ascontiguousarray(np.asarray(E, dtype=np.complex128).reshape(3))
        lib().m3o_2d_set_load(self._h, _dp(e.view(np.float64)))

    def run(self):
        lib().m3o_2d_run(self._h)

    def homogenization(self):
        C, ep, sg = np.zeros(9), np.zeros(9), np.zeros(9)
        lib().m3o_2d_homogenization(self._h, _dp(C), _dp(ep), _dp(sg))
        return C.reshape(3, 3, order="F"), ep.reshape(3, 3, order="F"), sg.reshape(3, 3, order="F")

    def fields(self):
        n = self.shape[0] * self.shape[1]
        e = np.zeros((n, 3), dtype=np.complex128)
        s = np.zeros((n, 3), dtype=np.complex128)
        lib().m3o_2d_fields(self._h, _dp(e.view(np.float64)), _dp(s.view(np.float64)))
        return e, s


class Percolation:
    """percolation (src/percolation/percolation.cpp): the reference's
    min-label sweeps (5 rounds x 8 orders x 6 trigger tests), cluster list,
    volume fractions and relabelling, and SurfaceOfTwoPhase."""

    def __init__(self, rve, nph, phase):
        r = np.ascontiguousarray(rve, dtype=np.int32)
        self.shape = r.shape
        self._r = r.reshape(-1)
        self._h = lib().m3o_perc_create(_i32p(self._r), *self.shape, int(nph), int(phase))

    def close(self):
        if self._h:
            lib().m3o_perc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def run(self, rounds=5):
        """-> (C1, all_clusters, clusters_vol_frac, vol_intrstd_phs, fixed point
        reached); rounds: rounds of the 8 x 6 sweeps (the reference runs 5)"""
        if len(set(self.shape)) != 1:
            raise ValueError("the reference's sweeps need Nx = Ny = Nz")
        lib().m3o_perc_run_rounds(self._h, int(rounds))
        n = int(np.prod(self.shape))
        c1 = np.empty(n, dtype=np.int32)
        cl = np.zeros(n + 1, dtype=np.int32)
        fr = np.zeros(n + 1)
        vol = ct.c_double(0)
        k = lib().m3o_perc_result(self._h, _i32p(c1), cl.ctypes.data_as(ct.POINTER(ct.c_int)), _dp(fr),
                                  ct.byref(vol), n + 1)
        return (c1.reshape(self.shape), cl[:k].copy(), fr[:k].copy(), vol.value,
                bool(lib().m3o_perc_fixed_point(self._h)))

    def surface(self, rve, p1, p2):
        r = np.ascontiguousarray(rve, dtype=np.int32).reshape(-1)
        return int(lib().m3o_perc_surface(self._h, _i32p(r), int(p1), int(p2)))
