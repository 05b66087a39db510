"""Benchmark: 3D FFT homogenization throughput on one MI355X.

Workload (default): the 256^3 two-phase elastic fibre RVE -- straight
cylinders of radius 4 voxels and length 64, random orientation, 30% fibre
fraction (seed 2), fibre E=300 GPa nu=0.25, matrix E=3 GPa nu=0.35 --
homogenized with the accelerated (Eyre-Milton) scheme over its 6 load cases
to tol 1e-6 / 1e-6, FP32 fields and FFTs with FP64 reductions and FP64 polish.

A step is one full homogenization (6 batched load cases to the stop rule).
value = sum over load cases of (voxels x iterations) / wall time of the timed
steps, with the microstructure already resident in HBM.

Also reported: the plain FP64 GPU path on the same input (voxel-iters/s and
the C_eff relative difference), the per-kernel roofline of the dominant
kernel (HIP-event timings on the solver's stream), and the FP64 CPU oracle
timed on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
# HBM bytes per launch of the percolation link kernel on the 256^3 fibre
# structure: FETCH_SIZE x 2 + WRITE_SIZE x 1 (copy-calibrated), rocprofv3 --pmc
MEASURED_TRAFFIC_PERC_LINK = 256.1e6

METRIC = "voxel-iters/s at 256³ (FP32+FP64 polish vs FP64 GPU), C_eff rel err"

# HBM bytes per full-batch launch (6 load cases) of the FP32 kernels on the
# 256^3 workload, from rocprofv3 FETCH_SIZE (x2.000) + WRITE_SIZE (x1.000)
# passes, each scale factor taken from a copy kernel of known size
# (profiles/elastic256_fibres_hbm_traffic.md)
MEASURED_TRAFFIC_256 = {
    "zpass_update": 9.805e9,
    "ypass_fwd": 5.128e9,
    "xpass_green": 4.999e9,
    "ypass_inv": 5.086e9,
    "zpass_check": 4.952e9,
    "xpass_check": 2.797e9,
}


def kernel_bytes(geom, r):
    """Algorithmic HBM bytes per launch of each per-iteration kernel: every
    kernel reads its inputs and writes its outputs once (r = bytes per real)."""
    N = geom["Nx"] * geom["Ny"] * geom["Nz"]
    H = geom["Nx"] * geom["Ny"] * geom["Nzh"]
    C, S = geom["C"], geom["solves"]
    spec = H * C * 2 * r
    real = N * C * r
    return {
        "zpass_update": S * (2 * spec + 2 * real + N),
        "ypass_fwd": S * 2 * spec,
        "xpass_green": S * 2 * spec,
        "ypass_inv": S * 2 * spec,
        "zpass_check": S * (real + N + spec),
        "xpass_check": S * spec,
    }


def iteration_bytes_per_voxel(Nz, C, r):
    """Algorithmic bytes per voxel-iteration of the reference dataflow: three
    C-component 3D FFTs (read real, write half spectrum), the update of the
    unknown field (read + write), one phase byte."""
    return 3 * C * (r + 2 * r * (Nz // 2 + 1) / Nz) + 2 * C * r + 1


def job_roofline(bytes_per_s, basis):
    """whole-job roofline: algorithmic bytes per voxel-iteration x voxel-iterations/s
    against the HBM peak (no per-kernel traffic measured for this line)"""
    gbs = bytes_per_s / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "basis": basis}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ids, mat, n, threads, iters=3, fft_threads=None):
    """FP64 oracle iterations of load case 0.  fft_threads given: FFTs on that
    many threads and the per-voxel loops on `threads` (the reference's own
    setting is 3 FFTW threads and serial loops, FFT_analysis3D.cpp:101-102)."""
    import oracle

    o = oracle.Elastic(n, n, n, "e", nthreads=threads)
    if fft_threads is not None:
        o.set_threads(fft_threads, threads)
    o.material(mat, ids)
    o.run_fixed(0, 1)  # warm-up (page-in, plans)
    t0 = time.perf_counter()
    done = o.run_fixed(0, iters)
    dt = time.perf_counter() - t0
    o.close()
    return done * n**3 / dt, dt, done


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 GPU comparison")
    ap.add_argument("--profile-reps", type=int, default=10)
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="solver option for the FP32+polish handle (m3d_elastic_set_option), repeatable")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run a fill and three device copies of GIB GiB (torch kernels), the known-size "
                         "reference for HBM byte counters, then exit")
    ap.add_argument("--workload", default="fibres256",
                    choices=["fibres256", "voronoi_batch", "thermal512", "j2_192", "basic64", "rec128", "plate512",
                             "perc256"],
                    help="fibres256 is the headline line; the others measure the batch, thermal, J2, "
                         "basic-scheme, reconstruction, 2D and percolation paths on their own workloads")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "sharding the load cases over several GPUs is not launched by this build.", flush=True)
        sys.exit(2)

    if args.copy_calibration > 0:
        return copy_calibration(args.copy_calibration)
    if args.workload != "fibres256":
        return OTHER[args.workload](args)

    import mat3d_amd as m
    from mat3d_amd import gen

    n = args.size
    t0 = time.perf_counter()
    ids = gen.fibres(n, 4, 0.30, seed=2, length=64, device=True)
    t_gen = time.perf_counter() - t0
    mat = gen.MAT_C2
    N = n**3
    print(f"[bench] generated {n}^3 fibre RVE in {t_gen:.1f}s, fibre fraction {(ids == 0).mean():.4f}", flush=True)

    s = m.FFT_Analysis3D(n, n, n, "e", precision="fp32_polish")
    for kv in args.set:
        k, v = kv.split("=", 1)
        s.set_option(k, float(v))
    s.material(mat, ids)
    for i in range(args.warmup):
        s.homogenization(args.tol, args.tol)
        print(f"[bench] warmup {i} iterations {[d['iterations'] for d in s.info]}", flush=True)
    vox_it = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        s.homogenization(args.tol, args.tol)
        vox_it += N * sum(d["iterations"] for d in s.info)
    dt = time.perf_counter() - t0
    value = vox_it / dt
    info32 = s.info
    C32 = s.CHomo.copy()
    print(f"[bench] fp32+polish: {value:.4e} voxel-it/s, {dt / args.steps * 1e3:.1f} ms/step, "
          f"iters {[d['iterations'] for d in info32]} fp64 share {[d['iters_fp64'] for d in info32]}", flush=True)

    fp64 = None
    c_err = None
    if not args.no_fp64:
        d = m.FFT_Analysis3D(n, n, n, "e", precision="fp64")
        d.material(mat, ids)
        d.homogenization(args.tol, args.tol)  # warm-up
        t0 = time.perf_counter()
        d.homogenization(args.tol, args.tol)
        dt64 = time.perf_counter() - t0
        v64 = N * sum(x["iterations"] for x in d.info) / dt64
        c_err = float(np.linalg.norm(C32 - d.CHomo) / np.linalg.norm(d.CHomo))
        fp64 = {"value": v64, "ms_per_step": dt64 * 1e3, "iterations": [x["iterations"] for x in d.info]}
        print(f"[bench] fp64: {v64:.4e} voxel-it/s, C_eff rel diff fp32+polish vs fp64 = {c_err:.3e}", flush=True)
        d.close()

    # per-kernel timings (HIP events on the solver stream) and roofline
    tuning = s.tuning()
    ms, geom = s.profile_kernels("fp32_polish", reps=args.profile_reps)
    kb = kernel_bytes(geom, 4)
    kernels = {}
    for k in ms:
        gbs = kb[k] / (ms[k] * 1e-3) / 1e9
        kernels[k] = {"ms": round(ms[k], 4), "bytes": kb[k], "GB/s": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)}
    per_iter = ["zpass_update", "ypass_fwd", "xpass_green", "ypass_inv"]
    dom = max(per_iter, key=lambda k: ms[k])
    it_ms = sum(ms[k] for k in per_iter)
    B = iteration_bytes_per_voxel(n, 6, 4)
    it_gbs = B * N * geom["solves"] / (it_ms * 1e-3) / 1e9
    print(f"[bench] kernels: {json.dumps(kernels)}", flush=True)
    s.close()
    copy_gbs = copy_peak()
    print(f"[bench] device copy (read + write): {copy_gbs:.1f} GB/s", flush=True)

    cpu = None
    if not args.no_cpu:
        # the box's CPU share for one GPU is 16 threads, whatever nproc shows
        threads = min(16, os.cpu_count() or 1)
        # samples of about 10 s each: 15 iterations on every thread, 3 with the reference's threading
        v, tcpu, done = cpu_baseline(ids, mat, n, threads, iters=15)
        print(f"[bench] cpu oracle: {v:.4e} voxel-it/s on {threads} threads", flush=True)
        vr, tr, dr = cpu_baseline(ids, mat, n, 1, iters=3, fft_threads=min(3, threads))
        print(f"[bench] cpu oracle, reference threading (3 FFT threads, serial voxel loops): {vr:.4e} voxel-it/s",
              flush=True)
        cpu = {"value": v, "unit": "voxel-iters/s", "cores": threads, "kind": "port",
               "sample": f"{n}^3 fibre RVE, load case 0, {done} accelerated FP64 iterations after 1 warm-up "
                         f"({tcpu:.1f}s; oracle's own FP64 FFT, FFTW absent; FFTs and voxel loops on all "
                         f"{threads} threads)",
               "nproc": os.cpu_count(), "cpu_model": cpu_model(),
               "reference_threading": {"value": vr, "unit": "voxel-iters/s", "cores": min(3, threads),
                                       "fft_threads": min(3, threads), "loop_threads": 1,
                                       "sample": f"{n}^3 fibre RVE, load case 0, {dr} accelerated FP64 iterations "
                                                 f"after 1 warm-up ({tr:.1f}s)"}}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64polish",
        "data": "synthetic: seeded fibre RVE (radius 4, length 64, 30% fibres, seed 2), generated in-process",
        "config": {"workload": f"{n}^3 two-phase elastic fibres (E 300/3 GPa), accelerated scheme, "
                               f"6 load cases batched, tol {args.tol:g}/{args.tol:g}",
                   "grid": [n, n, n], "load_cases": 6, "scheme": "accelerated", "base": "stress"},
        "iterations": [d["iterations"] for d in info32],
        "iters_fp64_polish": [d["iters_fp64"] for d in info32],
        "fp64_gpu": fp64,
        "speedup_vs_fp64_gpu": (value / fp64["value"]) if fp64 else None,
        "c_eff_rel_err_vs_fp64_gpu": c_err,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(kernels[dom]["GB/s"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kernels[dom]["frac"],
                     "traffic": (MEASURED_TRAFFIC_256.get(dom) if (n == 256 and geom["solves"] == 6) else None),
                     "copy_measured": round(copy_gbs, 1), "frac_of_copy": round(kernels[dom]["GB/s"] / copy_gbs, 4)},
        "roofline_iteration": {"bytes_per_voxel_iter": round(B, 2), "ms_per_iter_batch": round(it_ms, 4),
                               "achieved": round(it_gbs, 1), "frac": round(it_gbs / HBM_PEAK_GBS, 4)},
        "kernels": kernels,
        "autotune": dict(tuning, **m.tuning_report()),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def copy_peak(gib=2.0, reps=5):
    """Read+write rate of the runtime's device copy kernel over gib GiB: the
    practical streaming ceiling beside the 8 TB/s vendor peak."""
    import mat3d_amd as m

    ms = m.device_copy(gib, reps)
    return 2.0 * gib * 2**30 / (ms * 1e-3) / 1e9


def copy_calibration(gib):
    """A fill (writes B bytes) and four copies (read B, write B each) of
    B = gib GiB, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import mat3d_amd as m

    ms = m.device_copy(gib, 3)
    print(f"[bench] copy calibration: fill + 4 copies of {int(gib * 2**30)} bytes, {ms:.3f} ms per copy", flush=True)


def _line(args, metric, value, unit, ms, dtype, workload, extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic, seeded, generated in-process", "config": {"workload": workload}}
    out.update(extra)
    import mat3d_amd as m

    out.setdefault("autotune", m.tuning_report())
    print(json.dumps(out), flush=True)


def bench_voronoi_batch(args):
    """64 periodic Voronoi polycrystals, grids drawn from {32,48,64,96,128}^3,
    20-200 grains with E in [50,300] GPa and nu in [0.2,0.35], 6 load cases
    each (384 solves), FP32 + FP64 polish to 1e-6, through the batch API."""
    import mat3d_amd as m
    from mat3d_amd import gen

    t0 = time.perf_counter()
    probs = gen.c3_batch(64, device=True)
    print(f"[bench] generated 64 Voronoi RVEs in {time.perf_counter() - t0:.1f}s, "
          f"sizes {sorted(set(p[0] for p in probs))}", flush=True)
    b = m.Batch("fp32_polish")
    for n, ids, mat in probs:
        b.add(mat, ids)
    for i in range(args.warmup):
        b.homogenization(args.tol, args.tol)
    vi = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        b.homogenization(args.tol, args.tol)
        vi += b.stats()["voxel_iterations"]
    dt = time.perf_counter() - t0
    st = b.stats()
    its = [x["iterations"] for i in range(len(probs)) for x in b.result(i)["info"]]
    # algorithmic bytes of one homogenization of every RVE (FP32 accounting per grid)
    byts = sum(n**3 * iteration_bytes_per_voxel(n, 6, 4) * sum(x["iterations"] for x in b.result(i)["info"])
               for i, (n, _, _) in enumerate(probs))
    print(f"[bench] voronoi batch: {vi / dt:.4e} voxel-it/s, {dt / args.steps * 1e3:.1f} ms/step, "
          f"iterations min/mean/max {min(its)}/{np.mean(its):.1f}/{max(its)}, device ms {st['ms']:.1f}", flush=True)
    _line(args, "voxel-iters/s, 64-RVE Voronoi batch (384 solves)", vi / dt, "voxel-iters/s", dt / args.steps * 1e3,
          "f32+f64polish", "64 Voronoi RVEs {32..128}^3 x 6 load cases, accelerated scheme, tol 1e-6",
          {"iterations_mean": float(np.mean(its)), "solves": len(its),
           "roofline": job_roofline(byts * args.steps / dt, "FP32 algorithmic bytes per voxel-iteration of each grid "
                                                            "(194.1-197.5 B) x its voxel-iterations, per second")})


def bench_thermal512(args):
    """512^3 three-phase conduction: matrix k=1, inclusions k=1000 (15%),
    pores k=0.01 (10%), radii 3-10, 3 load cases, FP32 + FP64 polish to 1e-7."""
    import mat3d_amd as m
    from mat3d_amd import gen

    n = 512
    t0 = time.perf_counter()
    ids = gen.c5_rve(n, device=True)
    print(f"[bench] generated {n}^3 three-phase RVE in {time.perf_counter() - t0:.1f}s", flush=True)
    k = np.array([0.01, 1.0, 1000.0])  # phase 0 pores, 1 matrix, 2 inclusions
    s = m.FFT_analysis3d_thermal(n, n, n, precision="fp32_polish")
    s.material_n(k, ids)
    for i in range(args.warmup):
        s.homogenization(1e-7, 1e-7)
    vi = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        s.homogenization(1e-7, 1e-7)
        vi += n**3 * sum(d["iterations"] for d in s.info)
    dt = time.perf_counter() - t0
    print(f"[bench] thermal512: {vi / dt:.4e} voxel-it/s, iterations {[d['iterations'] for d in s.info]} "
          f"status {[d['status'] for d in s.info]}, K diag {np.diag(s.KHomo)}", flush=True)
    B = iteration_bytes_per_voxel(n, 3, 4)
    _line(args, "voxel-iters/s, 512^3 three-phase thermal", vi / dt, "voxel-iters/s", dt / args.steps * 1e3,
          "f32+f64polish", "512^3 thermal, 3 phases (k 0.01/1/1000), 3 load cases, tol 1e-7, cap 500",
          {"iterations": [d["iterations"] for d in s.info], "status": [d["status"] for d in s.info],
           "iters_fp64": [d["iters_fp64"] for d in s.info], "K_eff_diag": list(np.diag(s.KHomo)),
           "roofline": job_roofline(B * vi / dt, f"FP32 algorithmic bytes per voxel-iteration ({B:.1f} B, C = 3) "
                                                 "x voxel-iterations/s")})


def bench_j2_192(args):
    """192^3 J2 matrix (E=70 GPa, nu=0.3, sigma_y=100 MPa, H=1 GPa) with 25%
    elastic spheres (E=400 GPa, nu=0.2), uniaxial strain to 2% in 50
    increments, basic scheme with mixed control in FP64 (the value), beside
    the opt-in FP32 phase per increment (FP32 fields, FFTs and return map to
    10 x tol, then FP64) with its speedup and its difference from FP64."""
    import mat3d_amd as m
    from mat3d_amd import gen

    n = 192
    ids = gen.c4_rve(n, device=True)
    X0 = np.eye(6)[0]

    def run(precision, steps, **opt):
        s = m.FFT_Analysis3D(n, n, n, "ep", precision=precision)
        for k, v in opt.items():
            s.set_option(k, v)
        s.material(gen.MAT_C4, ids)
        vi = 0
        t0 = time.perf_counter()
        for i in range(steps):
            s.setLoad(0.02 * X0, X0)
            info = s.runAnalysis(1e-6, 1e-6, 50)
            vi += n**3 * (info["iters_fp32"] + info["iters_fp64"])
        dt = time.perf_counter() - t0
        eps, sig = s.fields()
        return vi / dt, dt / steps, info, s.prop_homogenized.copy(), sig

    v64, t64, i64, P64, sig64 = run("fp64", args.steps)
    print(f"[bench] j2_192 fp64: {v64:.4e} voxel-it/s, {i64['iters_fp64']} iterations over 50 increments, "
          f"final sigma_11 {P64[-1, 1]:.4e}", flush=True)
    extra = {"iterations": i64["iters_fp64"], "final_sigma11": P64[-1, 1]}
    if not args.no_fp64:
        v32, t32, i32, P32, sig32 = run("fp32_polish", 1, basic_fp32=1)
        d_prop = float(np.abs(P32 - P64).max() / np.abs(P64).max())
        d_sig = float(np.abs(sig32 - sig64).max() / np.abs(sig64).max())
        print(f"[bench] j2_192 fp32 phase (opt-in): {v32:.4e} voxel-it/s, iterations fp32 {i32['iters_fp32']} "
              f"fp64 {i32['iters_fp64']}; prop diff {d_prop:.2e}, sigma diff {d_sig:.2e}", flush=True)
        extra["fp32_phase_opt_in"] = {"value": v32, "speedup_vs_fp64": v32 / v64, "iterations_fp32": i32["iters_fp32"],
                                      "iterations_fp64": i32["iters_fp64"], "prop_rel_diff": d_prop,
                                      "sigma_rel_diff": d_sig}
    # FP64 algorithmic bytes per voxel-iteration with the J2 history read
    # every iteration (sigma_old, eps_old: 48 B each, eps_p: 8 B)
    B = iteration_bytes_per_voxel(n, 6, 8) + 104
    extra["roofline"] = job_roofline(B * v64, f"FP64 algorithmic bytes per voxel-iteration with the J2 history "
                                              f"({B:.1f} B) x voxel-iterations/s")
    _line(args, "voxel-iters/s, 192^3 J2 50 increments", v64, "voxel-iters/s", t64 * 1e3, "f64",
          "192^3 J2 matrix + elastic spheres, uniaxial 2% in 50 increments, basic scheme", extra)


def bench_basic64(args):
    """64^3 two-phase elastic RVE (matrix E=70 GPa nu=0.30, spheres r=6 at 20%,
    E=400 GPa nu=0.20, seed 1), basic Moulinec-Suquet scheme with the 6 unit
    strain loads held fixed, arithmetic-mean reference medium, FP64, tol 1e-6."""
    import mat3d_amd as m
    from mat3d_amd import gen

    n = 64
    ids = gen.c1_rve(n, device=True)
    s = m.FFT_Analysis3D(n, n, n, "ep", precision="fp64")
    s.set_option("scheme", 1)
    s.material(gen.MAT_C1, ids)
    for i in range(args.warmup):
        s.homogenization(1e-6, 1e-6)
    vi = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        s.homogenization(1e-6, 1e-6)
        vi += n**3 * sum(d["iterations"] for d in s.info)
    dt = time.perf_counter() - t0
    its = [d["iterations"] for d in s.info]
    print(f"[bench] basic64: {vi / dt:.4e} voxel-it/s, {dt / args.steps * 1e3:.1f} ms/step, iterations {its}, "
          f"C11 {s.CHomo[0, 0]:.6e}", flush=True)
    _line(args, "voxel-iters/s, 64^3 basic scheme FP64 (6 strain loads)", vi / dt, "voxel-iters/s",
          dt / args.steps * 1e3, "f64", "64^3 spheres r=6 20% (E 400/70 GPa), basic scheme, fixed strain loads, tol 1e-6",
          {"iterations": its, "C_eff": s.CHomo.tolist()})


def bench_rec128(args):
    """Phase-retrieval reconstruction (recPlan / runRec) of a 128^3 two-phase
    structure (spheres of radius 6 at 25%, seed 3) from its two-point
    correlation function: 600 iterations (beta 0.8, random stream seed 1;
    re-randomisations at 150, 300, 450) in FP64 on the GPU, beside the
    oracle on the host's cores for a few iterations."""
    import mat3d_amd as m
    from mat3d_amd import gen

    n = 128
    N = n**3
    rve = (gen.spheres(n, 6, 0.25, seed=3, device=True) == 0).astype(float)
    tpcf = np.abs(np.fft.fftn(rve)) ** 2 / N
    g = m.recPlan(tpcf)
    g.runRec(20, 0.8, 1)  # warm-up: graph capture, first launches
    iters = 600
    vi = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        it = g.runRec(iters, 0.8, 1)
        vi += N * it
    dt = time.perf_counter() - t0
    value = vi / dt
    # per iteration: forward and inverse 3D FFT (read + write 16 B each),
    # phase (read MT, normC; write MH), projection (read MH, mT; write mT)
    B = 32 + 40 + 32 + 48
    it_gbs = B * value / 1e9
    print(f"[bench] rec128: {value:.4e} voxel-it/s, {it} iterations in {dt / args.steps:.3f} s, final error "
          f"{g.trackrealerrorA[-1]:.4f}, volume fraction {g.mTilda.real.mean():.5f} (target {rve.mean():.5f})",
          flush=True)
    cpu = None
    if not args.no_cpu:
        import oracle

        threads = min(16, os.cpu_count() or 1)
        o = oracle.Rec(tpcf, nthreads=threads)
        o.run(1, 0.8, 1)
        k = 3
        t1 = time.perf_counter()
        o.run(k, 0.8, 1)
        tc = time.perf_counter() - t1
        cpu = {"value": N * k / tc, "unit": "voxel-iters/s", "cores": threads, "kind": "port",
               "sample": f"128^3, {k} iterations of runRec after one warm-up ({tc:.1f}s; oracle's own FP64 FFT)"}
        print(f"[bench] cpu oracle: {N * k / tc:.4e} voxel-it/s on {threads} threads", flush=True)
    _line(args, "voxel-iters/s, 128^3 phase-retrieval reconstruction", value, "voxel-iters/s",
          dt / args.steps * 1e3, "f64", "128^3 recPlan/runRec, 600 iterations, beta 0.8",
          {"iterations": it, "final_error": float(g.trackrealerrorA[-1]),
           "roofline_iteration": {"bytes_per_voxel_iter": B, "achieved": round(it_gbs, 1), "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": round(it_gbs / HBM_PEAK_GBS, 4)},
           "cpu_baseline": cpu})


def bench_plate512(args):
    """FFT_Analysis2D homogenization of a 512^2 plate with 40 random discs
    (the example's moduli: matrix 68.9 GPa / 0.35, inclusion 400 GPa / 0.23):
    3 load cases x 25 iterations in FP64 on the GPU, beside the oracle."""
    import mat3d_amd as m

    n = 512
    rng = np.random.default_rng(3)
    x, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    rve = np.ones((n, n))
    for _ in range(40):
        cx, cy, rad = rng.uniform(0, n), rng.uniform(0, n), rng.uniform(10, 30)
        dx = np.minimum(np.abs(x - cx), n - np.abs(x - cx))
        dy = np.minimum(np.abs(y - cy), n - np.abs(y - cy))
        rve[dx * dx + dy * dy <= rad * rad] = 0
    rve = rve.reshape(-1)
    q = [68.9e9, 0.35, 400e9, 0.23]
    g = m.FFT_Analysis2D(n, n)
    g.material(q, rve)
    g.setLoad([0, 0, 0])
    g.homogenization()  # warm-up
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.setLoad([0, 0, 0])
        C = g.homogenization()
    dt = time.perf_counter() - t0
    vi = n * n * 75 * args.steps
    print(f"[bench] plate512: {vi / dt:.4e} point-iterations/s, {dt / args.steps * 1e3:.1f} ms per homogenization, "
          f"C11 {C[0, 0]:.4e}", flush=True)
    cpu = None
    if not args.no_cpu:
        import oracle

        threads = min(16, os.cpu_count() or 1)
        o = oracle.Elastic2D(n, n, nthreads=threads)
        o.material(q, rve)
        o.set_load([0, 0, 0])
        t1 = time.perf_counter()
        Co, _, _ = o.homogenization()
        tc = time.perf_counter() - t1
        cpu = {"value": n * n * 75 / tc, "unit": "point-iterations/s", "cores": threads, "kind": "port",
               "sample": f"one 512^2 homogenization (3 x 25 iterations, {tc:.1f}s; oracle's own FP64 FFT)"}
        print(f"[bench] cpu oracle: {n * n * 75 / tc:.4e} point-iterations/s on {threads} threads, "
              f"CHomo rel diff {np.abs(C - Co).max() / np.abs(Co).max():.2e}", flush=True)
    _line(args, "point-iterations/s, 512^2 FFT_Analysis2D homogenization", vi / dt, "point-iterations/s",
          dt / args.steps * 1e3, "f64", "512^2 plate, 40 discs, 3 load cases x 25 iterations",
          {"CHomo": C.tolist(), "cpu_baseline": cpu})


def bench_perc256(args):
    """percolation (runPercolationAnalysis) of the fibre phase of the 256^3
    fibre structure of the headline line (radius 4, 30%, seed 2): cluster
    labelling, sizes, ranks and the relabelled field on the GPU per step,
    inputs resident; the oracle's sweeps on the host beside it on one
    128^3 octant."""
    import ctypes as ct

    import mat3d_amd as m
    from mat3d_amd import gen

    n = 256
    N = n**3
    ids = gen.fibres(n, 4, 0.30, seed=2, length=64, device=True)
    p = m.percolation(ids, n, n, n, 1, 0)
    L = m.lib()
    k, vol = ct.c_int(0), ct.c_double(0)

    def run():
        rc = L.m3d_perc_run(p._h, 1, ct.byref(k), ct.byref(vol))
        if rc != 0:
            raise m.M3DError(rc, L.m3d_perc_last_error(p._h).decode())

    for _ in range(max(1, args.warmup)):
        run()
    steps = max(args.steps, 10)
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    dt = time.perf_counter() - t0
    value = N * steps / dt
    # stage durations from events on the handle's stream (separate runs)
    p.timings(True)
    st = {}
    for _ in range(5):
        run()
        for key, v in p.timings().items():
            st[key] = st.get(key, 0.0) + v / 5
    p.timings(False)
    # k_link moves at least the flag byte of every cube corner voxel and the
    # parent (4 B) of every interest voxel
    link_bytes = N * 1 + 4 * vol.value
    link_gbs = link_bytes / (st["link"] * 1e-3) / 1e9
    # the whole analysis: phase ids in (4 B), relabelled field out (4 B)
    pipe_gbs = 8 * N / (dt / steps) / 1e9
    print(f"[bench] perc256: {value:.4e} voxels/s, {dt / steps * 1e3:.3f} ms per analysis, {k.value} clusters, "
          f"volume {vol.value:.0f}; stages (ms) " + ", ".join(f"{a} {b:.3f}" for a, b in st.items()), flush=True)
    p.runPercolationAnalysis(1)
    largest = p.clusters_vol_frac[0]
    cpu = None
    if not args.no_cpu:
        import oracle

        h = n // 2
        sub = np.ascontiguousarray(ids[:h, :h, :h])
        t1 = time.perf_counter()
        o = oracle.Percolation(sub, 1, 0).run()
        tc = time.perf_counter() - t1
        cpu = {"value": h**3 / tc, "unit": "voxels/s", "cores": 1, "kind": "port",
               "sample": f"the oracle's 5 x 8 x 6 sweeps on the 128^3 octant (0..127)^3 of the same structure "
                         f"({tc:.1f}s, {len(o[1])} clusters, fixed point {o[4]})"}
        print(f"[bench] cpu oracle: {h**3 / tc:.4e} voxels/s on 1 thread", flush=True)
    _line(args, "voxels/s, 256^3 percolation analysis of the fibre phase", value, "voxels/s", dt / steps * 1e3,
          "int32", "256^3 fibres (radius 4, 30%, seed 2), runPercolationAnalysis of phase 0",
          {"clusters": k.value, "largest_cluster_fraction": largest, "stages_ms": {a: round(b, 4) for a, b in st.items()},
           "roofline": {"bound": "hbm", "kernel": "k_link", "achieved": round(link_gbs, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(link_gbs / HBM_PEAK_GBS, 4), "traffic": MEASURED_TRAFFIC_PERC_LINK,
                        "bytes_per_launch": link_bytes},
           "roofline_analysis": {"bytes_per_voxel": 8, "achieved": round(pipe_gbs, 1), "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": round(pipe_gbs / HBM_PEAK_GBS, 4)},
           "cpu_baseline": cpu})


OTHER = {"voronoi_batch": bench_voronoi_batch, "thermal512": bench_thermal512, "j2_192": bench_j2_192,
         "basic64": bench_basic64, "rec128": bench_rec128, "plate512": bench_plate512, "perc256": bench_perc256}


if __name__ == "__main__":
    main()
