"""mat3d_amd -- MI355X-native 3D FFT homogenization (Moulinec-Suquet / Eyre-Milton).

Python mirror of the reference's solver classes over the C-ABI of
``libmat3d.so`` (declared in include/m3d.h):

  * ``FFT_Analysis3D``         -- include/FFT_analysis3D.hpp:36-215
  * ``FFT_analysis3d_thermal`` -- include/FFT_thermal3D.hpp:37-154

Method names, argument order and meaning follow the reference.  Array
conventions are the reference's: real fields are ``(N, C)`` voxel-major
(Eigen ``C x N`` column-major), voxel index ``i = z + Nz*(y + Ny*x)``;
material tables are ``nph x 7`` ``[E, nu, sigma_y, H, n, P, Y0]``.

The compute path is the HIP library only: importing a solver without a
built ``libmat3d.so`` raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# M3D_LIB selects another build of the library (e.g. one with other tuning macros)
LIB_PATH = os.environ.get("M3D_LIB") or os.path.join(_HERE, "libmat3d.so")

M3D_OK = 0
ERRORS = {-1: "M3D_ERR_ARG", -2: "M3D_ERR_UNSUPPORTED", -3: "M3D_ERR_OOM", -4: "M3D_ERR_HIP",
          -5: "M3D_ERR_STATE", -6: "M3D_ERR_NUMERIC"}
PRECISION = {"fp64": 0, "fp32_polish": 1}

# every symbol include/m3d.h declares
EXPORTS = [
    "m3d_version",
    "m3d_elastic_create", "m3d_elastic_destroy", "m3d_elastic_material", "m3d_elastic_set_load",
    "m3d_elastic_run", "m3d_elastic_homogenization", "m3d_elastic_get_fields", "m3d_elastic_info",
    "m3d_elastic_set_option", "m3d_elastic_last_error", "m3d_elastic_profile", "m3d_elastic_get_prop", "m3d_elastic_get_state",
    "m3d_elastic_get_tuning", "m3d_tuning_report", "m3d_tuning_file", "m3d_device_copy",
    "m3d_thermal_create", "m3d_thermal_destroy", "m3d_thermal_material", "m3d_thermal_material_n",
    "m3d_thermal_set_load", "m3d_thermal_run", "m3d_thermal_homogenization", "m3d_thermal_get_fields",
    "m3d_thermal_set_option", "m3d_thermal_last_error",
    "m3d_batch_create", "m3d_batch_destroy", "m3d_batch_add", "m3d_batch_homogenization", "m3d_batch_results",
    "m3d_batch_stats", "m3d_batch_set_option", "m3d_batch_last_error",
    "m3d_rec_create", "m3d_rec_destroy", "m3d_rec_last_error", "m3d_rec_set_initial", "m3d_rec_run", "m3d_rec_result",
    "m3d_gen_last_error", "m3d_gen_balls", "m3d_gen_cylinders", "m3d_gen_cylinders_noalloc", "m3d_gen_voronoi",
    "m3d_2d_create", "m3d_2d_destroy", "m3d_2d_last_error", "m3d_2d_material", "m3d_2d_set_load", "m3d_2d_run",
    "m3d_2d_homogenization", "m3d_2d_get_fields",
    "m3d_perc_create", "m3d_perc_destroy", "m3d_perc_last_error", "m3d_perc_run", "m3d_perc_result",
    "m3d_perc_surface", "m3d_perc_set_timing", "m3d_perc_timings",
]


def tuning_report():
    """Every autotuner choice this process timed or read from the tuning file
    (one line each) and the tuning file's path."""
    need = ct.c_int(0)
    lib().m3d_tuning_report(None, 0, ct.byref(need))
    buf = ct.create_string_buffer(max(1, need.value))
    lib().m3d_tuning_report(buf, len(buf), ct.byref(need))
    return {"file": lib().m3d_tuning_file().decode(), "choices": [l for l in buf.value.decode().splitlines() if l]}


class RunInfo(ct.Structure):
    _fields_ = [("iterations", ct.c_int), ("status", ct.c_int), ("iters_fp32", ct.c_int), ("iters_fp64", ct.c_int),
                ("eqVar", ct.c_double), ("diff_sum", ct.c_double), ("diff_eq", ct.c_double),
                ("time_ms", ct.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class M3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libmat3d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m mat3d_amd.build` "
                               "(there is no CPU fallback)")
        L = ct.CDLL(LIB_PATH)
        vp, ip, dp, cp = ct.c_void_p, ct.c_int, ct.POINTER(ct.c_double), ct.c_char_p
        i32p = ct.POINTER(ct.c_int32)
        rip = ct.POINTER(RunInfo)
        ipp = ct.POINTER(ct.c_int)
        sig = {
            "m3d_version": (cp, []),
            "m3d_elastic_create": (ip, [ip, ip, ip, cp, ip, ct.POINTER(vp)]),
            "m3d_elastic_destroy": (ip, [vp]),
            "m3d_elastic_material": (ip, [vp, dp, i32p, ip]),
            "m3d_elastic_set_load": (ip, [vp, dp, dp]),
            "m3d_elastic_run": (ip, [vp, ct.c_double, ct.c_double, ip, rip]),
            "m3d_elastic_homogenization": (ip, [vp, ct.c_double, ct.c_double, dp, dp, dp, dp, dp, rip]),
            "m3d_elastic_get_fields": (ip, [vp, ip, dp, dp]),
            "m3d_elastic_info": (ip, [vp, dp, ipp, dp, dp]),
            "m3d_elastic_set_option": (ip, [vp, cp, ct.c_double]),
            "m3d_elastic_last_error": (cp, [vp]),
            "m3d_elastic_profile": (ip, [vp, ip, ip, dp, ipp]),
            "m3d_elastic_get_prop": (ip, [vp, dp, ipp]),
            "m3d_elastic_get_state": (ip, [vp, dp, dp]),
            "m3d_elastic_get_tuning": (ip, [vp, ipp, ipp, ipp, ipp, ipp, ipp]),
            "m3d_tuning_report": (ip, [ct.c_char_p, ip, ipp]),
            "m3d_tuning_file": (cp, []),
            "m3d_device_copy": (ip, [ct.c_double, ip, dp]),
            "m3d_rec_create": (ip, [dp, ip, ip, ip, ct.POINTER(vp)]),
            "m3d_rec_destroy": (ip, [vp]),
            "m3d_rec_last_error": (cp, [vp]),
            "m3d_rec_set_initial": (ip, [vp, dp]),
            "m3d_rec_run": (ip, [vp, ip, ct.c_double, ct.c_uint64, ipp, ipp, ipp]),
            "m3d_rec_result": (ip, [vp, dp, dp, ip, ipp]),
            "m3d_gen_last_error": (cp, []),
            "m3d_gen_balls": (ip, [ip, ip, ip, i32p, ip, ct.c_int32, i32p]),
            "m3d_gen_cylinders": (ip, [ip, ip, ip, dp, ip, ct.c_double, ct.c_double, ct.c_int32, ip, ct.c_longlong,
                                       i32p, ipp, ct.POINTER(ct.c_longlong)]),
            "m3d_gen_cylinders_noalloc": (ip, [ip, ip, ip, dp, ip, ct.c_double, ct.c_double, ct.c_int32, ip,
                                               ct.c_longlong, vp, ipp, ct.POINTER(ct.c_longlong)]),
            "m3d_gen_voronoi": (ip, [ip, ip, ip, dp, ip, i32p]),
            "m3d_2d_create": (ip, [ip, ip, ct.POINTER(vp)]),
            "m3d_2d_destroy": (ip, [vp]),
            "m3d_2d_last_error": (cp, [vp]),
            "m3d_2d_material": (ip, [vp, dp, dp]),
            "m3d_2d_set_load": (ip, [vp, dp]),
            "m3d_2d_run": (ip, [vp, ipp]),
            "m3d_2d_homogenization": (ip, [vp, dp, dp, dp]),
            "m3d_2d_get_fields": (ip, [vp, dp, dp]),
            "m3d_perc_create": (ip, [i32p, ip, ip, ip, ip, ip, ct.POINTER(vp)]),
            "m3d_perc_destroy": (ip, [vp]),
            "m3d_perc_last_error": (cp, [vp]),
            "m3d_perc_run": (ip, [vp, ip, ipp, dp]),
            "m3d_perc_result": (ip, [vp, i32p, i32p, dp, ip]),
            "m3d_perc_surface": (ip, [vp, i32p, ip, ip, ct.POINTER(ct.c_longlong)]),
            "m3d_perc_set_timing": (ip, [vp, ip]),
            "m3d_perc_timings": (ip, [vp, dp, ip]),
            "m3d_thermal_create": (ip, [ip, ip, ip, ip, ct.POINTER(vp)]),
            "m3d_thermal_destroy": (ip, [vp]),
            "m3d_thermal_material": (ip, [vp, ct.c_double, ct.c_double, i32p]),
            "m3d_thermal_material_n": (ip, [vp, dp, ip, i32p]),
            "m3d_thermal_set_load": (ip, [vp, dp]),
            "m3d_thermal_run": (ip, [vp, ct.c_double, ct.c_double, rip]),
            "m3d_thermal_homogenization": (ip, [vp, ct.c_double, ct.c_double, dp, dp, dp, dp, rip]),
            "m3d_thermal_get_fields": (ip, [vp, ip, dp, dp]),
            "m3d_thermal_set_option": (ip, [vp, cp, ct.c_double]),
            "m3d_thermal_last_error": (cp, [vp]),
            "m3d_batch_create": (ip, [ip, ct.POINTER(vp)]),
            "m3d_batch_destroy": (ip, [vp]),
            "m3d_batch_add": (ip, [vp, ip, ip, ip, dp, i32p, ip, ipp]),
            "m3d_batch_homogenization": (ip, [vp, ct.c_double, ct.c_double]),
            "m3d_batch_results": (ip, [vp, ip, dp, dp, dp, dp, dp, rip]),
            "m3d_batch_stats": (ip, [vp, ipp, dp, dp]),
            "m3d_batch_set_option": (ip, [vp, cp, ct.c_double]),
            "m3d_batch_last_error": (cp, [vp]),
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


def _colmajor(m):
    return np.ascontiguousarray(np.asarray(m, dtype=np.float64).T).reshape(-1)


def device_copy(gib=2.0, reps=5):
    """Average ms of one device-to-device copy of gib GiB (the runtime's copy
    kernel) after a fill and an untimed copy; for measurement tools."""
    ms = ct.c_double(0)
    rc = lib().m3d_device_copy(float(gib), int(reps), ct.byref(ms))
    if rc != 0:
        raise M3DError(rc, "m3d_device_copy failed")
    return ms.value


class FFT_Analysis3D:
    """Mirror of ``FFT_Analysis3D`` (include/FFT_analysis3D.hpp:36-215) on the HIP solver."""

    def __init__(self, Nx1, Ny1, Nz1, mater_behav="e", precision="fp64"):
        self.Nx, self.Ny, self.Nz = int(Nx1), int(Ny1), int(Nz1)
        self.xyz = self.Nx * self.Ny * self.Nz
        self.mater_behav = mater_behav
        h = ct.c_void_p()
        rc = lib().m3d_elastic_create(self.Nx, self.Ny, self.Nz, mater_behav.encode(), PRECISION[precision],
                                      ct.byref(h))
        if rc != M3D_OK:
            raise M3DError(rc, "m3d_elastic_create failed")
        self._h = h
        self.last_info = None
        self.info = []

    def _chk(self, rc):
        if rc != M3D_OK:
            raise M3DError(rc, lib().m3d_elastic_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().m3d_elastic_destroy(self._h)
            self._h = None

    __del__ = close

    def set_option(self, key, value):
        self._chk(lib().m3d_elastic_set_option(self._h, key.encode(), float(value)))

    def material(self, mat, RVE1, noPhases1=None):
        mat = np.asarray(mat, dtype=np.float64)
        nph = mat.shape[0] if noPhases1 is None else int(noPhases1)
        rve = np.ascontiguousarray(np.asarray(RVE1).reshape(-1), dtype=np.int32)
        if rve.size != self.xyz:
            raise ValueError("RVE size does not match the grid")
        m = _colmajor(mat[:nph])
        self._chk(lib().m3d_elastic_material(self._h, _dp(m), _i32p(rve), nph))
        self.noPhases = nph
        vol = np.zeros(nph)
        sb = ct.c_int(0)
        la0, mu0 = ct.c_double(0), ct.c_double(0)
        self._chk(lib().m3d_elastic_info(self._h, _dp(vol), ct.byref(sb), ct.byref(la0), ct.byref(mu0)))
        self.vol_fractions = vol
        self.strainBase = bool(sb.value)
        self.stressBase = not self.strainBase
        self.la0, self.mu0 = la0.value, mu0.value

    def setLoad(self, E1, X01=None):
        E = np.ascontiguousarray(np.asarray(E1, dtype=np.float64).reshape(6))
        X0 = E if X01 is None else np.ascontiguousarray(np.asarray(X01, dtype=np.float64).reshape(6))
        self._chk(lib().m3d_elastic_set_load(self._h, _dp(E), _dp(X0)))
        self.E, self.X0 = E.copy(), X0.copy()

    def runAnalysis(self, tol_field, tol_eq, nInc1=1):
        info = RunInfo()
        self._chk(lib().m3d_elastic_run(self._h, float(tol_field), float(tol_eq), int(nInc1), ct.byref(info)))
        self.last_info = info.as_dict()
        self.info = [self.last_info]
        return self.last_info

    def homogenization(self, tol_field, tol_eq, outputName=None):
        C, S, ep, sg = (np.zeros(36) for _ in range(4))
        ret = ct.c_double(0)
        infos = (RunInfo * 6)()
        self._chk(lib().m3d_elastic_homogenization(self._h, float(tol_field), float(tol_eq), _dp(C), _dp(S), _dp(ep),
                                                   _dp(sg), ct.byref(ret), infos))
        f = lambda a: a.reshape(6, 6).T.copy()
        self.CHomo, self.SHomo, self.epHomo, self.sigHomo = f(C), f(S), f(ep), f(sg)
        self.E11, self.E22, self.E33 = 1 / self.SHomo[0, 0], 1 / self.SHomo[1, 1], 1 / self.SHomo[2, 2]
        self.info = [i.as_dict() for i in infos]
        return ret.value

    def fields(self, lc=0):
        eps = np.empty((self.xyz, 6))
        sig = np.empty((self.xyz, 6))
        self._chk(lib().m3d_elastic_get_fields(self._h, int(lc), _dp(eps), _dp(sig)))
        return eps, sig

    @property
    def prop_homogenized(self):
        """rows x 9 [eps_t, sigma_1..6, sigma_eff|d, eps_p|0] of the last "ep"/"ed" run."""
        rows = ct.c_int(0)
        self._chk(lib().m3d_elastic_get_prop(self._h, None, ct.byref(rows)))
        out = np.zeros(9 * max(rows.value, 1))
        self._chk(lib().m3d_elastic_get_prop(self._h, _dp(out), ct.byref(rows)))
        return out[: 9 * rows.value].reshape(9, rows.value).T.copy()

    def history(self):
        """(epsilon_p_eff or d, sigma_eff) per voxel after an "ep"/"ed" runAnalysis."""
        a = np.empty(self.xyz)
        b = np.empty(self.xyz)
        self._chk(lib().m3d_elastic_get_state(self._h, _dp(a), _dp(b)))
        return a, b

    def tuning(self):
        """Autotuned variants of the FP32 x/y passes for the last solve's shape."""
        v = [ct.c_int(0) for _ in range(6)]
        self._chk(lib().m3d_elastic_get_tuning(self._h, *(ct.byref(x) for x in v)))
        return {"xgreen_threads": v[0].value, "xgreen_wide_tile": v[1].value, "ypass_threads": v[2].value,
                "xgreen_radix_plan": v[3].value, "ypass_radix_plan": v[4].value, "from_tuning_file": v[5].value}

    def profile_kernels(self, precision="fp32_polish", reps=10):
        """Average ms per launch of the per-iteration kernels (HIP events on the
        solver's stream) and the geometry used for byte accounting."""
        ms = np.zeros(6)
        geom = (ct.c_int * 9)()
        self._chk(lib().m3d_elastic_profile(self._h, PRECISION[precision], int(reps), _dp(ms), geom))
        names = ["zpass_update", "ypass_fwd", "xpass_green", "ypass_inv", "zpass_check", "xpass_check"]
        g = list(geom)
        return dict(zip(names, ms)), dict(zip(["Nx", "Ny", "Nz", "Nzh", "C", "solves", "TY", "TKy", "TKx"], g))

    @property
    def epsilon(self):
        return self.fields(0)[0]

    @property
    def sigma(self):
        return self.fields(0)[1]


class FFT_analysis3d_thermal:
    """Mirror of ``FFT_analysis3d_thermal`` (include/FFT_thermal3D.hpp:37-154) on the HIP solver."""

    def __init__(self, Nx1, Ny1, Nz1, precision="fp64"):
        self.Nx, self.Ny, self.Nz = int(Nx1), int(Ny1), int(Nz1)
        self.xyz = self.Nx * self.Ny * self.Nz
        h = ct.c_void_p()
        rc = lib().m3d_thermal_create(self.Nx, self.Ny, self.Nz, PRECISION[precision], ct.byref(h))
        if rc != M3D_OK:
            raise M3DError(rc, "m3d_thermal_create failed")
        self._h = h
        self._E = None
        self.info = []

    def _chk(self, rc):
        if rc != M3D_OK:
            raise M3DError(rc, lib().m3d_thermal_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().m3d_thermal_destroy(self._h)
            self._h = None

    __del__ = close

    def set_option(self, key, value):
        self._chk(lib().m3d_thermal_set_option(self._h, key.encode(), float(value)))

    def material(self, km, ki, RVE1):
        rve = np.ascontiguousarray(np.asarray(RVE1).reshape(-1), dtype=np.int32)
        self._chk(lib().m3d_thermal_material(self._h, float(km), float(ki), _i32p(rve)))

    def material_n(self, k, RVE1):
        k = np.ascontiguousarray(k, dtype=np.float64)
        rve = np.ascontiguousarray(np.asarray(RVE1).reshape(-1), dtype=np.int32)
        self._chk(lib().m3d_thermal_material_n(self._h, _dp(k), int(k.size), _i32p(rve)))

    def setLoad(self, E1):
        E = np.ascontiguousarray(np.asarray(E1, dtype=np.float64).reshape(3))
        self._chk(lib().m3d_thermal_set_load(self._h, _dp(E)))

    def runAnalysis(self, tol_field, tol_eq):
        info = RunInfo()
        self._chk(lib().m3d_thermal_run(self._h, float(tol_field), float(tol_eq), ct.byref(info)))
        self.info = [info.as_dict()]
        return self.info[0]

    def homogenization(self, tol_field, tol_eq, filename=None):
        K, R, G, Q = (np.zeros(9) for _ in range(4))
        infos = (RunInfo * 3)()
        self._chk(lib().m3d_thermal_homogenization(self._h, float(tol_field), float(tol_eq), _dp(K), _dp(R), _dp(G),
                                                   _dp(Q), infos))
        f = lambda a: a.reshape(3, 3).T.copy()
        self.KHomo, self.RHomo, self.tempGradHomo, self.heatFluxHomo = f(K), f(R), f(G), f(Q)
        self.K11, self.K22, self.K33 = self.KHomo[0, 0], self.KHomo[1, 1], self.KHomo[2, 2]
        self.info = [i.as_dict() for i in infos]

    def fields(self, lc=0):
        g = np.empty((self.xyz, 3))
        q = np.empty((self.xyz, 3))
        self._chk(lib().m3d_thermal_get_fields(self._h, int(lc), _dp(g), _dp(q)))
        return g, q

    @property
    def tempGradient(self):
        return self.fields(0)[0]

    @property
    def heatFlux(self):
        return self.fields(0)[1]


class Batch:
    """Many independent elastic homogenizations in one call.

    The reference's callers construct one ``FFT_Analysis3D`` per RVE and call
    ``material`` and ``homogenization`` on each (example/3d/main.cpp:117-128);
    ``Batch`` returns the results of that loop, with RVEs of one grid shape
    solved together on the device."""

    def __init__(self, precision="fp32_polish"):
        h = ct.c_void_p()
        rc = lib().m3d_batch_create(PRECISION[precision], ct.byref(h))
        if rc != M3D_OK:
            raise M3DError(rc, "m3d_batch_create failed")
        self._h = h
        self.shapes = []

    def _chk(self, rc):
        if rc != M3D_OK:
            raise M3DError(rc, lib().m3d_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().m3d_batch_destroy(self._h)
            self._h = None

    __del__ = close

    def set_option(self, key, value):
        self._chk(lib().m3d_batch_set_option(self._h, key.encode(), float(value)))

    def add(self, mat, rve, shape=None):
        """Adds one RVE (``rve``: (Nx, Ny, Nz) phase ids, or flat with ``shape``);
        ``mat``: (nph, 7) [E, nu, sigma_y, H, n, P, Y0].  Returns its index."""
        rve = np.asarray(rve)
        shape = tuple(rve.shape) if shape is None else tuple(shape)
        if len(shape) != 3:
            raise ValueError("shape must be (Nx, Ny, Nz)")
        flat = np.ascontiguousarray(rve.reshape(-1), dtype=np.int32)
        if flat.size != shape[0] * shape[1] * shape[2]:
            raise ValueError("RVE size does not match its shape")
        mat = np.asarray(mat, dtype=np.float64)
        m = _colmajor(mat)
        idx = ct.c_int(-1)
        self._chk(lib().m3d_batch_add(self._h, shape[0], shape[1], shape[2], _dp(m), _i32p(flat), mat.shape[0],
                                      ct.byref(idx)))
        self.shapes.append(shape)
        return idx.value

    def homogenization(self, tol_field, tol_eq):
        self._chk(lib().m3d_batch_homogenization(self._h, float(tol_field), float(tol_eq)))

    def result(self, i):
        C, S, ep, sg = (np.zeros(36) for _ in range(4))
        ret = ct.c_double(0)
        infos = (RunInfo * 6)()
        self._chk(lib().m3d_batch_results(self._h, int(i), _dp(C), _dp(S), _dp(ep), _dp(sg), ct.byref(ret), infos))
        f = lambda a: a.reshape(6, 6).T.copy()
        return {"CHomo": f(C), "SHomo": f(S), "epHomo": f(ep), "sigHomo": f(sg), "ret": ret.value,
                "info": [x.as_dict() for x in infos]}

    def stats(self):
        n, ms, vi = ct.c_int(0), ct.c_double(0), ct.c_double(0)
        self._chk(lib().m3d_batch_stats(self._h, ct.byref(n), ct.byref(ms), ct.byref(vi)))
        return {"count": n.value, "ms": ms.value, "voxel_iterations": vi.value}


class recPlan:
    """recPlan(double *TPCF, Nx, Ny, Nz) and runRec(n, b)
    (include/reconstruction.hpp:55-74, src/rec/reconstruction.cpp:211-406):
    phase retrieval of a two-phase microstructure from its two-point
    correlation function, on the GPU in FP64.  mTilda and trackrealerrorA
    are the reference's public results."""

    def __init__(self, TPCF, Nx1=None, Ny1=None, Nz1=None):
        t = np.ascontiguousarray(TPCF, dtype=np.float64)
        shape = t.shape if Nx1 is None else (Nx1, Ny1, Nz1)
        self.shape = tuple(int(v) for v in shape)
        h = ct.c_void_p()
        rc = lib().m3d_rec_create(_dp(t.reshape(-1)), *self.shape, ct.byref(h))
        if rc != 0:
            raise M3DError(rc, "m3d_rec_create failed")
        self._h = h
        self.mTilda = None
        self.trackrealerrorA = None

    def _chk(self, rc):
        if rc != 0:
            raise M3DError(rc, lib().m3d_rec_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().m3d_rec_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_initial(self, m0):
        self._m0 = np.ascontiguousarray(m0, dtype=np.float64).reshape(-1)
        self._chk(lib().m3d_rec_set_initial(self._h, _dp(self._m0)))

    def runRec(self, n, b, seed=1):
        it, st, ss = ct.c_int(0), ct.c_int(0), ct.c_int(0)
        self._chk(lib().m3d_rec_run(self._h, int(n), float(b), int(seed), ct.byref(it), ct.byref(st), ct.byref(ss)))
        self.status, self.search_steps = st.value, ss.value
        m = np.zeros(self.shape, dtype=np.complex128)
        err = np.zeros(max(1, it.value), dtype=np.float64)
        k = ct.c_int(0)
        self._chk(lib().m3d_rec_result(self._h, m.view(np.float64).ctypes.data_as(ct.POINTER(ct.c_double)), _dp(err),
                                       err.size, ct.byref(k)))
        self.mTilda = m
        self.trackrealerrorA = err[: k.value].copy()
        return it.value


class FFT_Analysis2D:
    """FFT_Analysis2D(Nx1, Ny1) (include/FFT_analysis2D.hpp:82-98,
    src/2d/FFT_analysis2D.cpp): the 2D basic scheme on complex fields
    [E11, E22, E12], 25 fixed iterations per runAnalysis, on the GPU (FP64)."""

    def __init__(self, Nx1, Ny1):
        self.Nx, self.Ny = int(Nx1), int(Ny1)
        self.xy = self.Nx * self.Ny
        h = ct.c_void_p()
        rc = lib().m3d_2d_create(self.Nx, self.Ny, ct.byref(h))
        if rc != 0:
            raise M3DError(rc, "m3d_2d_create failed")
        self._h = h
        self.itNo = 0

    def _chk(self, rc):
        if rc != 0:
            raise M3DError(rc, lib().m3d_2d_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().m3d_2d_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def material(self, properties, RVE1):
        q = np.ascontiguousarray(properties, dtype=np.float64)
        r = np.ascontiguousarray(RVE1, dtype=np.float64).reshape(-1)
        self._chk(lib().m3d_2d_material(self._h, _dp(q), _dp(r)))

    def setLoad(self, E1):
        e = np.ascontiguousarray(np.asarray(E1, dtype=np.complex128).reshape(3))
        self._chk(lib().m3d_2d_set_load(self._h, _dp(e.view(np.float64))))

    def runAnalysis(self):
        it = ct.c_int(0)
        self._chk(lib().m3d_2d_run(self._h, ct.byref(it)))
        self.itNo = it.value

    def homogenization(self):
        C, ep, sg = np.zeros(9), np.zeros(9), np.zeros(9)
        self._chk(lib().m3d_2d_homogenization(self._h, _dp(C), _dp(ep), _dp(sg)))
        self.CHomo = C.reshape(3, 3, order="F")
        self.epHomo = ep.reshape(3, 3, order="F")
        self.sigHomo = sg.reshape(3, 3, order="F")
        return self.CHomo

    def fields(self):
        """(epsilon, sigma) as complex arrays (Nx*Ny, 3)"""
        e = np.zeros((self.xy, 3), dtype=np.complex128)
        s = np.zeros((self.xy, 3), dtype=np.complex128)
        self._chk(lib().m3d_2d_get_fields(self._h, _dp(e.view(np.float64)), _dp(s.view(np.float64))))
        return e, s


def _phase_ids(RVE, n):
    """RVE values as int32 phase ids, truncated toward zero as the
    reference's int conversion C1[i] = RVE(i) does."""
    r = np.asarray(RVE).reshape(-1)
    if r.size != n:
        raise ValueError(f"RVE has {r.size} values, expected {n}")
    return np.ascontiguousarray(r.astype(np.int32))


class percolation:
    """percolation(RVE, Nx1, Ny1, Nz1, noOfPhases, intrestPhase)
    (include/percolation.hpp:36-71, src/percolation/percolation.cpp): the
    clusters of one phase and the face contacts of two phases, on the GPU.
    RVE holds Nx*Ny*Nz phase ids, index (x*Ny + y)*Nz + z.  Public results
    as in the reference: all_clusters, clusters_vol_frac (descending),
    vol_intrstd_phs, commonSurface, and C1 (the relabelled field, shape
    (Nx, Ny, Nz))."""

    def __init__(self, RVE, Nx1, Ny1, Nz1, noOfPhases, intrestPhase):
        self.NX, self.NY, self.NZ = int(Nx1), int(Ny1), int(Nz1)
        self.noOfPhases, self.intrestPhase = int(noOfPhases), int(intrestPhase)
        ids = _phase_ids(RVE, self.NX * self.NY * self.NZ)
        h = ct.c_void_p()
        rc = lib().m3d_perc_create(_i32p(ids), self.NX, self.NY, self.NZ, self.noOfPhases, self.intrestPhase,
                                   ct.byref(h))
        if rc != 0:
            raise M3DError(rc, "m3d_perc_create failed")
        self._h = h
        self.all_clusters = []
        self.clusters_vol_frac = []
        self.vol_intrstd_phs = 0.0
        self.commonSurface = 0
        self.C1 = ids.reshape(self.NX, self.NY, self.NZ)

    def _chk(self, rc):
        if rc != 0:
            raise M3DError(rc, lib().m3d_perc_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().m3d_perc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def runPercolationAnalysis(self, total_iter=1):
        """percolation.cpp:117-382 (total_iter is unused, as there)."""
        k, vol = ct.c_int(0), ct.c_double(0)
        self._chk(lib().m3d_perc_run(self._h, int(total_iter), ct.byref(k), ct.byref(vol)))
        c1 = np.empty((self.NX, self.NY, self.NZ), dtype=np.int32)
        cl = np.zeros(max(1, k.value), dtype=np.int32)
        fr = np.zeros(max(1, k.value), dtype=np.float64)
        self._chk(lib().m3d_perc_result(self._h, _i32p(c1), _i32p(cl), _dp(fr), k.value))
        self.C1 = c1
        self.all_clusters = [int(v) for v in cl[: k.value]]
        self.clusters_vol_frac = [float(v) for v in fr[: k.value]]
        self.vol_intrstd_phs = vol.value

    def SurfaceOfTwoPhase(self, RVE, phase1, phase2):
        """percolation.cpp:384-414: sets commonSurface (and C1 to RVE)."""
        ids = _phase_ids(RVE, self.NX * self.NY * self.NZ)
        c = ct.c_longlong(0)
        self._chk(lib().m3d_perc_surface(self._h, _i32p(ids), int(phase1), int(phase2), ct.byref(c)))
        self.commonSurface = c.value
        self.C1 = ids.reshape(self.NX, self.NY, self.NZ)
        return c.value

    def timings(self, on=None):
        """Stage durations (ms) of the last run; timings(True) turns the
        events on for the next runs."""
        if on is not None:
            self._chk(lib().m3d_perc_set_timing(self._h, int(bool(on))))
            return None
        ms = np.zeros(6)
        self._chk(lib().m3d_perc_timings(self._h, _dp(ms), 6))
        return dict(zip(("init", "link", "enumerate", "sizes", "sort", "relabel"), ms.tolist()))
