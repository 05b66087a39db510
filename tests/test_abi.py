"""CPU checks of the C-ABI library: it loads and exports every entry point
include/m3d.h declares (no compute calls: there is no GPU here)."""
import ctypes as ct
import os
import re

import pytest

import mat3d_amd as m

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "m3d.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(m3d_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_python_export_list():
    assert sorted(m.EXPORTS) == declared_symbols()


def test_library_loads_and_exports_every_declared_symbol():
    L = ct.CDLL(m.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert m.lib().m3d_version().startswith(b"mat3d_amd")


def test_null_handle_is_an_error_not_a_crash():
    L = m.lib()
    assert L.m3d_elastic_destroy(None) == -1
    assert L.m3d_thermal_destroy(None) == -1
    assert L.m3d_elastic_set_load(None, None, None) == -1
    assert L.m3d_elastic_last_error(None) == b"null handle"
    out = ct.c_void_p()
    # argument validation happens before any device call
    assert L.m3d_elastic_create(1, 8, 8, b"e", 0, ct.byref(out)) == -1
    assert L.m3d_elastic_create(8, 8, 8, b"xx", 0, ct.byref(out)) == -1
    assert L.m3d_elastic_create(8, 8, 8, b"e", 7, ct.byref(out)) == -1
    assert L.m3d_thermal_create(8, 8, 0, 0, ct.byref(out)) == -1


def test_batch_host_side_validation():
    """m3d_batch_create/add only touch the host: argument errors come back as codes."""
    import numpy as np

    from mat3d_amd import gen

    b = m.Batch("fp64")
    with pytest.raises(m.M3DError) as e:
        b.homogenization(1e-6, 1e-6)  # nothing added
    assert e.value.code == -5
    with pytest.raises(m.M3DError) as e:
        b.add(gen.MAT_C1, np.full((8, 8, 8), 5, dtype=np.int32))  # phase id out of range
    assert e.value.code == -1
    with pytest.raises(m.M3DError) as e:
        b.add(gen.MAT_C1, np.zeros((17, 17, 17), dtype=np.int32))  # prime factor 17
    assert e.value.code == -2
    with pytest.raises(m.M3DError):
        b.result(0)
    assert b.add(gen.MAT_C1, gen.spheres(8, 2, 0.1, seed=1)) == 0
    assert b.add(gen.MAT_C1, np.ones((6, 8, 10), dtype=np.int32)) == 1
    assert b.stats()["count"] == 2
    with pytest.raises(m.M3DError):
        b.set_option("no_such_option", 1)
    b.close()
    L = m.lib()
    assert L.m3d_batch_destroy(None) == -1
    assert L.m3d_batch_create(5, ct.byref(ct.c_void_p())) == -1


def test_null_handles_of_the_other_rows():
    """reconstruction, 2D solver, percolation and generators: null handles and
    bad sizes come back as codes before any device call."""
    L = m.lib()
    out = ct.c_void_p()
    assert L.m3d_rec_destroy(None) == -1
    assert L.m3d_rec_last_error(None) is not None
    assert L.m3d_2d_destroy(None) == -1
    assert L.m3d_2d_create(0, 8, ct.byref(out)) == -1
    assert L.m3d_perc_destroy(None) == -1
    assert L.m3d_perc_last_error(None) == b"null handle"
    assert L.m3d_perc_run(None, 1, None, None) == -1
    assert L.m3d_perc_result(None, None, None, None, 0) == -1
    assert L.m3d_perc_set_timing(None, 1) == -1
    assert L.m3d_perc_create(None, 8, 8, 8, 2, 1, ct.byref(out)) == -1
    assert L.m3d_perc_create(None, 0, 8, 8, 2, 1, ct.byref(out)) == -1
    assert L.m3d_gen_balls(0, 8, 8, None, 0, 1, None) == -1
    assert L.m3d_gen_voronoi(8, 8, 8, None, 0, None) == -1


def test_tuning_file_path_follows_the_environment(monkeypatch):
    L = m.lib()
    monkeypatch.setenv("M3D_TUNE_FILE", "/tmp/m3d_tune_probe.txt")
    assert L.m3d_tuning_file() == b"/tmp/m3d_tune_probe.txt"
    monkeypatch.setenv("M3D_TUNE_FILE", "none")
    assert L.m3d_tuning_file() == b""
    monkeypatch.delenv("M3D_TUNE_FILE")
    monkeypatch.setenv("XDG_CACHE_HOME", "/tmp/xdg")
    assert L.m3d_tuning_file() == b"/tmp/xdg/mat3d_amd/tune.txt"
    # nothing timed or read in this process (no GPU): an empty report
    rep = m.tuning_report()
    assert rep["choices"] == []
