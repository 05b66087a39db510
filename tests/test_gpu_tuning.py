"""The autotuner's tuning file: choices timed in one process are taken by the
next without timing, and every variant the file can select (radix plan,
threads, kz tile of the x pass with the Green step and of the y pass) gives
results within the FP32 + polish bar of the FP64 oracle.  Each case runs in
a child process so that it starts with an empty in-process cache."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
sys.path.insert(0, {root!r})
import numpy as np
import mat3d_amd as m
from mat3d_amd import gen
n = {n}
ids = gen.fibres(n, 3, 0.3, seed=2, length=n // 2)
s = m.FFT_Analysis3D(n, n, n, "e", precision="fp32_polish")
s.material(gen.MAT_C2, ids)
s.homogenization(1e-6, 1e-6)
print("RESULT " + json.dumps({{"tuning": s.tuning(), "C": s.CHomo.tolist(), "report": m.tuning_report()}}), flush=True)
"""


def child(n, path):
    env = dict(os.environ, M3D_TUNE_FILE=str(path))
    r = subprocess.run([sys.executable, "-u", "-c", CHILD.format(root=ROOT, n=n)], env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")][-1]
    return json.loads(line[len("RESULT "):])


def test_tuning_file_is_written_and_read(tmp_path):
    path = tmp_path / "tune.txt"
    a = child(32, path)
    assert a["tuning"]["from_tuning_file"] == 0
    lines = path.read_text().split("\n")
    assert len([l for l in lines if l.strip()]) == 1
    assert a["report"]["file"] == str(path)
    b = child(32, path)
    print("tuning", a["tuning"], b["tuning"], a["report"]["choices"])
    assert b["tuning"]["from_tuning_file"] == 1
    for k in ("xgreen_threads", "xgreen_wide_tile", "ypass_threads", "xgreen_radix_plan", "ypass_radix_plan"):
        assert a["tuning"][k] == b["tuning"][k]
    assert "tuning file" in b["report"]["choices"][0]


@pytest.mark.parametrize("variant", [(256, 0, 0, 256, 0), (512, 1, 1, 512, 1), (256, 1, 1, 256, 1), (512, 0, 1, 256, 0)])
def test_every_tuning_variant_meets_the_bar(tmp_path, variant):
    import oracle
    from mat3d_amd import gen

    n = 64  # both radix plans differ at 64 ([16, 4] and [8, 8]); at 32 they coincide
    path = tmp_path / "tune.txt"
    first = child(n, path)
    dev = path.read_text().split()[0]
    xg_nt, xg_wide, xg_rp, y_nt, y_rp = variant
    path.write_text(f"{dev} {n} {n} {n} 6 6 {xg_nt} {xg_wide} {xg_rp} {y_nt} {y_rp}\n")
    r = child(n, path)
    t = r["tuning"]
    assert t["from_tuning_file"] == 1
    assert (t["xgreen_threads"], t["xgreen_wide_tile"], t["xgreen_radix_plan"], t["ypass_threads"],
            t["ypass_radix_plan"]) == variant
    ids = gen.fibres(n, 3, 0.3, seed=2, length=n // 2)
    o = oracle.Elastic(n, n, n, "e")
    o.material(gen.MAT_C2, ids)
    ref = o.homogenization(1e-6, 1e-6)["C"]
    C = np.array(r["C"])
    err = np.linalg.norm(C - ref) / np.linalg.norm(ref)
    print("variant", variant, "C rel err", err, "timed choice", first["tuning"])
    assert err < 1e-6
