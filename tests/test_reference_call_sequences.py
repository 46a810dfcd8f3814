"""The call sequences of the reference's own ctypes scripts, replayed against this library:
python/verify_c_interface.py:147-247 (array initialisation) and python/test_mesh_from_files.py
:120-245 (FESOM2 mesh directory, zero velocity, five steps with statistics).  The host part runs
on the CPU; the steps that use the device are in the GPU test at the end."""
from pathlib import Path

import ctypes
import numpy as np
import pytest

from fesom2_tracer_dwarf_amd import abi, mesh_io

GOLDEN = Path(__file__).resolve().parent / "golden"


def _lib():
    lib = abi.load()
    lib.tracer_cleanup()
    return lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def test_verify_c_interface_sequence():
    lib = _lib()
    nl, nod2D, edge2D, elem2D = 10, 5, 8, 4
    assert lib.tracer_init_mpi(0) == 0
    assert lib.tracer_set_mesh_dims(nl, nod2D, edge2D, elem2D) == 0
    assert lib.tracer_set_partition(nod2D, 0) == 0
    assert lib.tracer_allocate_tracers(2, nl, nod2D, 2) == 0
    temperature = np.zeros((nl - 1, nod2D), order="F")
    for k in range(nl - 1):
        temperature[k, :] = 20.0 - 15.0 * (k / (nl - 2))
    assert lib.tracer_set_values(1, _fp(temperature), nl, nod2D) == 0
    salinity = np.full((nl - 1, nod2D), 35.0, order="F")
    assert lib.tracer_set_values(2, _fp(salinity), nl, nod2D) == 0
    assert lib.tracer_set_velocity_zero(nl, nod2D, edge2D) == 0
    assert lib.tracer_finalize_mpi() == 0
    lib.tracer_cleanup()


def _mesh_dir(tmp_path):
    with np.load(GOLDEN / "pi_mesh.npz") as d:
        raw = {k: d[k] for k in d.files}
    mesh_io.write_mesh_files(raw, tmp_path / "pi")
    return str(tmp_path / "pi").encode()


def _mesh_from_files_setup(lib, path):
    """Steps 1-6b of python/test_mesh_from_files.py; returns (nl, nod2D, temperature)."""
    assert lib.tracer_init_mpi(0) == 0
    assert lib.tracer_load_mesh_from_files(path) == 0
    nz, nn = ctypes.c_int(), ctypes.c_int()
    assert lib.tracer_get_size(ctypes.byref(nz), ctypes.byref(nn)) == 0
    nl, nod2D = nz.value + 1, nn.value
    assert lib.tracer_allocate_tracers(2, nl, nod2D, 2) == 0
    temperature = np.zeros((nz.value, nn.value), order="F")
    for k in range(nz.value):
        temperature[k, :] = 20.0 - 15.0 * (k / max(1, nz.value - 1))
    assert lib.tracer_set_values(1, _fp(temperature), nl, nod2D) == 0
    salinity = np.full((nz.value, nn.value), 35.0, order="F")
    assert lib.tracer_set_values(2, _fp(salinity), nl, nod2D) == 0
    assert lib.tracer_set_velocity_zero(nl, nod2D, 3 * nod2D) == 0  # the script's estimate of edge2D
    assert lib.tracer_init_ale_arrays() == 0
    return nl, nod2D, temperature


def test_mesh_from_files_sequence_host_part(tmp_path):
    lib = _lib()
    nl, nod2D, _ = _mesh_from_files_setup(lib, _mesh_dir(tmp_path))
    assert (nl, nod2D) == (48, 3140)
    lib.tracer_cleanup()


@pytest.mark.gpu
def test_mesh_from_files_sequence_runs(tmp_path):
    """Steps 6c-8: advection set-up, initialisation complete, five steps at rest; the statistics
    of tracer 1 (over all levels, as the reference's tracer_get_stats) stay those of the field set."""
    lib = _lib()
    print("case reference call sequence test_mesh_from_files (pi mesh, 5 steps)", flush=True)
    nl, nod2D, temperature = _mesh_from_files_setup(lib, _mesh_dir(tmp_path))
    assert lib.tracer_init_for_advection() == 0
    assert lib.tracer_init_complete() == 0
    for step in range(1, 6):
        assert lib.tracer_advect_step(1.0e-3) == 0
        tmin, tmax, tsum = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        assert lib.tracer_get_stats(ctypes.byref(tmin), ctypes.byref(tmax), ctypes.byref(tsum)) == 0
        assert (tmin.value, tmax.value) == (5.0, 20.0), step
        assert abs(tsum.value - temperature.sum()) <= 1e-12 * abs(temperature.sum())
    assert lib.tracer_finalize_mpi() == 0
    lib.tracer_cleanup()


@pytest.mark.gpu
def test_demo_tracer_restart_workflow(tmp_path):
    """python/demo_tracer.py:210-260 through the Python wrapper's reference-named methods: restart
    files in, ten steps with statistics, restart files out (a reference-written restart)."""
    import shutil
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    print("case demo_tracer restart workflow (10 steps)", flush=True)
    path = tmp_path / "fesom_bin_restart"
    shutil.copytree(GOLDEN / "restart_ref" / "rst2", path)
    before = (path / "np1" / "t_tracer.0").read_bytes()
    t = FESOMTracerAdvection()
    t.lib.tracer_cleanup()
    assert t.initialize_mpi(0) == 0
    assert t.initialize(str(path)) == 0
    sums = []
    for _ in range(10):
        assert t.advect_step(1.0e-3) == 0
        st = t.get_statistics()
        assert np.isfinite(st["min"]) and st["min"] <= st["max"]
        sums.append(st["sum"])
    assert t.finalize(str(path)) == 0
    assert t.finalize_mpi() == 0
    assert (path / "np1" / "t_tracer.0").read_bytes() != before  # the advanced tracers were written
    t.cleanup()
