"""Build libmat3d.so (HIP, gfx950) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", f) for f in ("m3d_elastic.hip", "m3d_thermal.hip", "m3d_basic.hip", "m3d_rec.hip",
                                                   "m3d_gen.hip", "m3d_2d.hip", "m3d_perc.hip")]
LIB = os.path.join(HERE, "libmat3d.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wno-unused-value",
    "-Wno-unused-result",
    # packed FP32 VALU forms are kept out of generated code on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False, out: str | None = None, defines: list[str] | None = None) -> str:
    """Compile the three units in parallel, then link libmat3d.so (or
    ``out``); ``defines`` are extra -D macros (tuning parameters)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    lib = out or LIB
    tag = "" if out is None else "." + os.path.basename(out).replace(".so", "")
    dflags = [f"-D{d}" for d in (defines or [])]
    objs, procs = [], []
    for src in SRCS:
        obj = os.path.join(HERE, "csrc", os.path.basename(src).replace(".hip", tag + ".o"))
        cmd = [hipcc, *[f for f in HIPCC_FLAGS if f != "-shared"], *dflags, "-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
        objs.append(obj)
    failed = False
    for pr in procs:
        txt, _ = pr.communicate()
        msgs = [l for l in txt.splitlines() if "packed-fp32-ops" not in l]
        if pr.returncode != 0:
            failed = True
            sys.stderr.write("\n".join(msgs) + "\n")
    if failed:
        raise RuntimeError("hipcc failed building libmat3d.so")
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib + ".tmp", *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libmat3d.so failed")
    os.replace(lib + ".tmp", lib)
    for o in objs:
        os.remove(o)
    return lib


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", default=None, help="library path (default: mat3d_amd/libmat3d.so)")
    ap.add_argument("-D", dest="defines", action="append", default=[], help="extra macro, e.g. -D M3D_KZ_WAVES=4")
    a = ap.parse_args()
    print(build(verbose=True, out=a.out, defines=a.defines))
