"""On-disk formats (include/m3d_io.hpp): input decks, reports, printer and
.vti writer, driven by the C++ probe tests/io_probe.cpp (CPU only).

The reports are checked line by line against the reference's layout
(src/3d/FFT_analysis3D.cpp:944-996, src/3dTherm/FFT_thermal3D.cpp:597-620)
and their numbers against the oracle (Hashin-Shtrikman bounds, Mori-Tanaka
and HS conductivity estimates)."""
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle
from mat3d_amd import gen

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DECKS = os.path.join(HERE, "golden", "decks")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("probe") / "io_probe")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(ROOT, "include"), "-o", exe,
                           os.path.join(HERE, "io_probe.cpp")])
    return exe


def run(probe, *args, stdin=None):
    return subprocess.run([probe, *args], input=stdin, capture_output=True, text=True, check=True).stdout


def test_matrix_in_eigen_stream_format(probe):
    # Eigen's default IOFormat: right-aligned to the widest entry, one space apart
    assert run(probe, "matrix") == ("        1.5        0.25 1.23457e+08\n"
                                    "         -2           1       0.001\n"
                                    "      3e+06           0          -7\n")


def test_structural_deck_of_the_reference_example(probe):
    out = run(probe, "deck", os.path.join(DECKS, "in_struct.txt")).split("\n")
    assert out[0] == "65 3.1e+10 0.3 3e+10 0.2 1e-06 1e-06"
    assert out[1] == "rve65.txt 1 myvt1i.vti 0 myOutput.txt"
    assert out[2].split() == ["0", "0.1", "0", "0", "0", "0"]


def test_thermal_deck_of_the_reference_example(probe):
    out = run(probe, "tdeck", os.path.join(DECKS, "in_therm.txt")).split("\n")
    assert out[0] == "65 25 30 1e-06 1e-06"
    assert out[1] == "rve65.txt 1 myvt1i5.vti 1 myOutputthermal.txt"


def _mat_text(a):
    return " ".join(repr(float(x)) for x in np.asarray(a).T.reshape(-1))  # column-major


def test_elastic_report_layout_and_bounds(probe):
    n = 12
    ids = gen.spheres(n, 2, 0.2, seed=3)
    mat = gen.MAT_C1
    o = oracle.Elastic(n, n, n, "e", nthreads=2)
    o.material(mat, ids)
    r = o.homogenization(1e-6, 1e-6)
    sb = o.info()["strain_base"]
    vol = o.matrices(2)["vol"]
    head = [mat[0, 0], mat[1, 0], mat[0, 1], mat[1, 1], vol[0], vol[1]]
    spec = (f"{n} {n} {n} {int(sb)} 2 " + " ".join(repr(float(x)) for x in head) + " " +
            " ".join(_mat_text(r[k]) for k in ("C", "S", "epHomo", "sigHomo")))
    txt = run(probe, "report_e", stdin=spec)
    lines = txt.split("\n")
    assert lines[0] == "*" * 73 and lines[1] == ""
    assert lines[2] == "        F F T     H O M O G E N I Z A T I O N    B Y   M A T 3 D"
    assert lines[4].strip() == ("Stran-Based Algorithm Is Used" if sb else "Stress-Based Algorithm Is Used")
    assert lines[5] == f"                           Nx={n}   Ny={n}   Nz={n}"
    hs = {}
    for ln in lines:
        for tok in ln.split():
            if "=" in tok and tok.split("=")[0] in ("E_min", "K_min", "G_min", "E_max", "K_max", "G_max"):
                k, v = tok.split("=")
                hs[k] = float(v)
    ref = dict(zip(["E_min", "K_min", "G_min", "E_max", "K_max", "G_max"], r["hs"]))
    for k in ref:
        assert hs[k] == pytest.approx(ref[k], rel=1e-5), k
    i = lines.index("Homogenized Elastic Modules:")
    e11 = float(lines[i + 1].split()[0].split("=")[1])
    assert e11 == pytest.approx(1 / r["S"][0, 0], rel=1e-5)
    j = lines.index("C_Homogenized=")
    C = np.array([[float(x) for x in lines[j + 2 + q].split()] for q in range(6)])
    assert np.allclose(C, r["C"], rtol=1e-5, atol=1e-6 * np.abs(r["C"]).max())
    assert "C1=" in lines and "S1=" in lines


def test_thermal_report_layout_and_estimates(probe):
    n = 10
    ids = gen.spheres(n, 2, 0.2, seed=9)
    o = oracle.Thermal(n, n, n)
    o.material(25.0, 30.0, ids)
    r = o.homogenization(1e-6, 1e-6)
    spec = (f"{n} {n} {n} 25.0 30.0 {int(ids.sum())} " +
            " ".join(_mat_text(r[k]) for k in ("K", "R", "tempGradHomo", "heatFluxHomo")))
    lines = run(probe, "report_t", stdin=spec).split("\n")
    # RVE.sum()/xyz is an integer division in the reference: 0 here
    assert lines[0] == "Volume Fraction of the White Phase (or 1)=0"
    assert lines[lines.index("K phase matrix  ( or 1)=") + 1].split() == ["25", "0", "0"]
    assert lines[lines.index("Nx=10   Ny=10   10")] == "Nx=10   Ny=10   10"
    mt = float(lines[lines.index("KHomogenized Mori-Tanaka=") + 2])
    low = float(lines[lines.index("Lower bound HS=") + 2])
    up = float(lines[lines.index("Upper bound LS=") + 2])
    assert (mt, low, up) == pytest.approx(tuple(r["est"][:3]), rel=1e-5)
    j = lines.index("KHomogenized=")
    K = np.array([[float(x) for x in lines[j + 2 + q].split()] for q in range(3)])
    assert np.allclose(K, r["K"], rtol=1e-5)


def _bounds(probe, c, K, G, fm, km, fi, ki):
    spec = f"{len(c)} " + " ".join(repr(float(x)) for x in [*c, *K, *G, fm, km, fi, ki])
    return [float(x) for x in run(probe, "bounds", stdin=spec).split()]


def test_hs_bounds_against_the_oracle_many_phases(probe):
    """The report's Hashin-Shtrikman bounds (comparison-medium form) equal the
    oracle's restatement of FFT_analysis3D.cpp:911-942 on a 6-grain RVE."""
    n = 8
    ids = gen.voronoi(n, 6, seed=11)
    rng = np.random.default_rng(4)
    mat = np.zeros((6, 7))
    mat[:, 0] = rng.uniform(50e9, 300e9, 6)
    mat[:, 1] = rng.uniform(0.2, 0.35, 6)
    mat[:, 2] = np.inf
    mat[:, 4] = 1
    o = oracle.Elastic(n, n, n, "e", nthreads=2)
    o.material(mat, ids)
    r = o.homogenization(1e-4, 1e-4)
    vol = o.matrices(6)["vol"]
    E, v = mat[:, 0], mat[:, 1]
    G = E / (2 * (1 + v))
    K = E / (3 * (1 - 2 * v))
    b = _bounds(probe, vol, K, G, 0.5, 1.0, 0.5, 2.0)
    hs = r["hs"]  # E_min K_min G_min E_max K_max G_max
    got = [b[2], b[0], b[1], b[5], b[3], b[4]]
    assert np.allclose(got, hs, rtol=1e-12, atol=0)
    # the bounds bracket the computed effective bulk modulus (isotropic part)
    C = r["C"]
    Keff = C[:3, :3].sum() / 9
    assert hs[1] <= Keff <= hs[4]


def test_hs_bounds_two_phase_textbook_form(probe):
    """Two phases: the classic two-phase bounds, K1 + c2 / (1/(K2-K1) +
    3c1/(3K1+4G1)) and G1 + c2 / (1/(G2-G1) + 6c1(K1+2G1)/(5G1(3K1+4G1))),
    with phase 1 the softer one; equal fractions of one phase give it back."""
    K1, G1, K2, G2, c1 = 58.3e9, 26.9e9, 222.2e9, 166.7e9, 0.7
    c2 = 1 - c1
    b = _bounds(probe, [c1, c2], [K1, K2], [G1, G2], 0.5, 1.0, 0.5, 2.0)
    klo = K1 + c2 / (1 / (K2 - K1) + 3 * c1 / (3 * K1 + 4 * G1))
    khi = K2 + c1 / (1 / (K1 - K2) + 3 * c2 / (3 * K2 + 4 * G2))
    glo = G1 + c2 / (1 / (G2 - G1) + 6 * c1 * (K1 + 2 * G1) / (5 * G1 * (3 * K1 + 4 * G1)))
    ghi = G2 + c1 / (1 / (G1 - G2) + 6 * c2 * (K2 + 2 * G2) / (5 * G2 * (3 * K2 + 4 * G2)))
    assert np.allclose(b[:2], [klo, glo], rtol=1e-13)
    assert np.allclose(b[3:5], [khi, ghi], rtol=1e-13)
    one = _bounds(probe, [1.0, 0.0], [K1, K2], [G1, G2], 1.0, 3.0, 0.0, 5.0)
    assert np.allclose(one[:6], [K1, G1, 9 * K1 * G1 / (3 * K1 + G1)] * 2, rtol=1e-13)
    assert one[6] == pytest.approx(3.0, rel=1e-14) and one[7] == pytest.approx(3.0, rel=1e-14)


def test_conductivity_estimates_against_oracle_and_maxwell(probe):
    """Mori-Tanaka / HS conductivity estimates at the true matrix fraction
    equal the oracle's restatement of FFT_thermal3D.cpp:583-592 (its est[3:6]);
    the matrix-referenced value is Maxwell's km (ki + 2km + 2f(ki-km)) /
    (ki + 2km - f(ki-km)) for inclusion fraction f."""
    n = 10
    ids = gen.spheres(n, 2, 0.25, seed=9)
    for km, ki in ((25.0, 30.0), (1.0, 1000.0), (40.0, 0.5)):
        o = oracle.Thermal(n, n, n)
        o.material(km, ki, ids)
        r = o.homogenization(1e-4, 1e-4)
        fm = float(np.count_nonzero(ids)) / ids.size
        fi = 1 - fm
        b = _bounds(probe, [1.0], [1.0], [1.0], fm, km, fi, ki)
        mt, lo, up = r["est"][3:6]
        assert b[6] == pytest.approx(mt, rel=1e-13) and b[6] == pytest.approx(lo, rel=1e-13)
        assert b[7] == pytest.approx(up, rel=1e-13)
        maxwell = km * (ki + 2 * km + 2 * fi * (ki - km)) / (ki + 2 * km - fi * (ki - km))
        assert b[6] == pytest.approx(maxwell, rel=1e-13)
        K = r["K"]
        assert min(b[6], b[7]) <= np.trace(K) / 3 <= max(b[6], b[7])


def test_vti_is_valid_image_data(probe, tmp_path):
    f = str(tmp_path / "a.vti")
    run(probe, "vti", f)
    root = ET.parse(f).getroot()
    assert root.attrib["type"] == "ImageData"
    img = root.find("ImageData")
    assert img.attrib["WholeExtent"] == "0 1 0 2 0 3"
    arr = img.find("Piece/PointData/DataArray")
    assert arr.attrib["NumberOfComponents"] == "4" and arr.attrib["type"] == "Float64"
    vals = np.array(arr.text.split(), dtype=float).reshape(24, 4)
    i = np.arange(24)
    assert np.array_equal(vals[:, 0], i) and np.array_equal(vals[:, 1], -i)
    assert np.array_equal(vals[:, 2], 0.5 * i) and np.array_equal(vals[:, 3], 0 * i)


def test_percolation_printer_layout(probe, tmp_path):
    """percolation::printer (src/percolation/percolation.cpp:417-441)."""
    d = tmp_path / "clusters"
    run(probe, "perc", str(d))
    assert open(d / "clustersFile.txt").read() == "".join(f"{i}\n" for i in range(24))
    assert open(d / "clustersdata.txt").read() == (
        "           intrested phase :  1     Number of the phases: 2\n"
        "-------------------------------------------------------------------------\n"
        "connected cluster name        volume fraction custer/volume fraction intrested phase\n"
        "        3                                   0.5\n"
        "        4                                   0.25\n"
        "        5                                   0.333333\n")


def test_2d_printer_layout(probe, tmp_path):
    """FFT_Analysis2D::printer (src/2d/FFT_analysis2D.cpp:328-360), real parts
    one per line; the sigma files hold sigma (the reference writes epsilon)."""
    subprocess.run([probe, "p2d"], cwd=tmp_path, check=True)
    assert (tmp_path / "Results").is_dir()
    for c in range(3):
        assert open(tmp_path / f"epsilon{c + 1}.txt").read() == "".join(f"{10 * c + i + 0.5:g}\n" for i in range(4))
        assert open(tmp_path / f"sigma{c + 1}.txt").read() == "".join(f"{-(10 * c + i + 0.5):g}\n" for i in range(4))


def test_rec_text_layout(probe, tmp_path):
    """recPlan::saveRecImages text files (src/rec/reconstruction.cpp:422-436)."""
    d = tmp_path / "rec"
    run(probe, "rec", str(d))
    assert open(d / "rve_rec.txt").read() == "0\n1\n0\n1\n"
    assert open(d / "err.txt").read() == "voloum fraction= 0.5\n3\n2\n"
