"""The C-ABI library loads, exports every entry point of include/fesomtra.h, and rejects bad
calls the way the reference's interface does (status -1, message) -- without touching a GPU."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from fesom2_tracer_dwarf_amd import abi

HEADER = Path(__file__).resolve().parents[1] / "include" / "fesomtra.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(tracer_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    names = declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    assert sorted(abi.exported_symbols()) == names


def test_reference_entry_points_all_present():
    """The 21 bind(C) names of lib/tracer_c_interface.F90 keep their names."""
    ref = ["tracer_init_mpi", "tracer_finalize_mpi", "tracer_init", "tracer_advect_step",
           "tracer_get_stats", "tracer_get_size", "tracer_get_values", "tracer_finalize",
           "tracer_run_workflow", "tracer_cleanup", "tracer_set_mesh_dims", "tracer_set_levels",
           "tracer_set_thickness", "tracer_set_partition", "tracer_allocate_tracers",
           "tracer_set_values", "tracer_set_velocity_zero", "tracer_set_velocity",
           "tracer_init_complete", "tracer_load_mesh_partition", "tracer_load_mesh_from_files",
           "tracer_init_ale_arrays", "tracer_init_for_advection"]
    lib = abi.load()
    assert all(hasattr(lib, n) for n in ref)


def test_error_paths_without_gpu():
    lib = abi.load()
    lib.tracer_cleanup()
    assert lib.tracer_init(b"/no/such/dir") == -1
    assert b"restart" in lib.tracer_last_error()
    assert lib.tracer_advect_step(1.0) == -1                      # before init
    a, b, c = ctypes.c_double(1), ctypes.c_double(1), ctypes.c_double(1)
    assert lib.tracer_get_stats(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == -1
    assert (a.value, b.value, c.value) == (0.0, 0.0, 0.0)
    assert lib.tracer_set_mesh_dims(2, 10, 30, 20) == -1          # nl < 3
    assert lib.tracer_set_mesh_dims(10, 4, 12, 8) == 0
    assert lib.tracer_set_partition(4, 3) == -1                   # halo not supported
    assert lib.tracer_allocate_tracers(2, 10, 4, 2) == 0
    v = np.zeros((4, 9))
    assert lib.tracer_set_values(3, abi.dptr(v), 10, 4) == -1      # invalid tracer_id
    assert lib.tracer_set_values(0, abi.dptr(v), 10, 4) == -1
    assert lib.tracer_set_values(1, abi.dptr(v), 11, 4) == -1      # wrong nl
    assert lib.tracer_set_scheme(1, b"MFCT", b"QR4C", b"FCT", 1.0, 1.0) == -1  # no MUSCL support
    assert lib.tracer_set_scheme(1, b"UPWX", b"QR4C", b"FCT", 1.0, 1.0) == -1
    assert b"Unknown horizontal" in lib.tracer_last_error()
    assert lib.tracer_set_scheme(1, b"UPW1", b"QR4C", b"FCT", 1.0, 1.0) == 0
    assert lib.tracer_init_complete() == -1                       # no connectivity yet
    uv = np.ones((9, 12, 2))
    w = np.zeros((4, 10))
    assert lib.tracer_set_velocity(abi.dptr(uv), abi.dptr(w), 10, 4, 12) == -1
    lib.tracer_cleanup()


def test_setup_validation_without_gpu():
    """Mesh checks made at tracer_init_complete (before any device work): inconsistent cavity tops,
    node levels below element levels, missing edge_tri(1); option setters reject bad values."""
    from fesom2_tracer_dwarf_amd import configs
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    lib = abi.load()
    case = configs.bathymetry_case("v", 6, 8, seed=3, nsteps=1, lev_lo=3)
    m = case.mesh
    t = FESOMTracerAdvection()
    # a node column starting below the tops of its elements (ulevels_nod2D > ulevels)
    t.setup_case(case, upload=False)
    ul = m.ulevels_nod2D.astype(np.int32).copy()
    ul[3] = 2
    assert lib.tracer_set_levels(abi.iptr(m.nlevels_nod2D.astype(np.int32)), abi.iptr(ul), m.nod2D) == 0
    assert lib.tracer_init_complete() == -1
    assert b"starts deeper" in lib.tracer_last_error()
    # an element whose top leaves it no layer
    t.setup_case(case, upload=False)
    eu = m.ulevels.astype(np.int32).copy()
    eu[2] = m.nlevels[2]
    assert lib.tracer_set_elem_levels(abi.iptr(m.nlevels.astype(np.int32)), abi.iptr(eu), m.elem2D) == 0
    assert lib.tracer_init_complete() == -1
    assert b"no layer" in lib.tracer_last_error()
    # a node shallower than one of its elements
    t.setup_case(case, upload=False)
    nl = m.nlevels_nod2D.astype(np.int32).copy()
    el = m.elem2D_nodes[0] - 1
    nl[el[0]] = m.nlevels[0] - 1
    assert lib.tracer_set_levels(abi.iptr(nl), abi.iptr(m.ulevels_nod2D.astype(np.int32)), m.nod2D) == 0
    assert lib.tracer_init_complete() == -1
    assert b"fewer levels" in lib.tracer_last_error()
    # edge_tri(1) missing
    t.setup_case(case, upload=False)
    et = m.edge_tri.astype(np.int32).copy()
    et[5, 0] = 0
    assert lib.tracer_set_connectivity(abi.iptr(m.elem2D_nodes.astype(np.int32)), abi.iptr(m.edges.astype(np.int32)),
                                       abi.iptr(et), abi.iptr(m.nod_in_elem2D.astype(np.int32)),
                                       abi.iptr(m.nod_in_elem2D_num.astype(np.int32)), m.max_adj, m.nod2D,
                                       m.elem2D, m.edge2D) == 0
    assert lib.tracer_init_complete() == -1
    assert b"edge_tri(1)" in lib.tracer_last_error()
    # option setters
    assert lib.tracer_set_tracers_per_pass(3) == -1
    assert lib.tracer_set_tracers_per_pass(128) == -1
    assert lib.tracer_set_tracers_per_pass(64) == 0
    assert lib.tracer_set_tracers_per_pass(0) == 0
    assert lib.tracer_set_precision(7) == -1
    assert lib.tracer_set_launch_config(0, 16, 16, 16) == -1
    assert lib.tracer_set_scheme(1, b"UPW1", b"UPW1", b"FCX", 1.0, 1.0) == -1
    assert lib.tracer_set_scheme(1, b"UPW1", b"QR4X", b"FCT", 1.0, 1.0) == -1
    assert b"Unknown vertical" in lib.tracer_last_error()
    assert lib.tracer_set_scheme(9, b"UPW1", b"UPW1", b"FCT", 1.0, 1.0) == -1
    lib.tracer_cleanup()
