"""FESOM2 file meshes: the reader and the spherical geometry (mesh_io), on the 'pi' mesh of the
reference's own test data (tests/golden/pi_mesh.npz, written by tests/golden/make_pi_mesh.py).

Pinned by facts of the mesh files and by closed forms: the reference's level rule
(nlevels_nod2D = max over the node's elements), the boundary-edge count against edgenum.out, the
left-element convention of edge_tri, clockwise triangles after the orientation test, the P1
gradients of longitude and latitude, the surface area sum, and, through the CPU restatement,
conservation, uniform-field preservation and FCT boundedness on the mesh.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from fesom2_tracer_dwarf_amd import configs, mesh_io
from fesom2_tracer_dwarf_amd.mesh import R_EARTH

GOLDEN = Path(__file__).resolve().parent / "golden"
REF_PI = Path("/root/reference/tests/data/pi")


@pytest.fixture(scope="module")
def raw():
    with np.load(GOLDEN / "pi_mesh.npz") as d:
        return {k: d[k] for k in d.files}


@pytest.fixture(scope="module")
def pi(raw):
    return mesh_io.sphere_mesh(raw)


@pytest.mark.skipif(not REF_PI.is_dir(), reason="reference test data not present")
def test_reader_reproduces_fixture(raw):
    fresh = mesh_io.read_mesh_files(REF_PI)
    assert sorted(fresh) == sorted(raw)
    for k in raw:
        np.testing.assert_array_equal(fresh[k], raw[k])


def test_counts_and_levels(raw, pi):
    assert (pi.nod2D, pi.elem2D, pi.edge2D, pi.nl) == (3140, 5839, 8986, 48)
    boundary = int(np.sum(pi.edge_tri[:, 1] == 0))
    assert boundary == 455
    assert raw["edge_num"][0] == pi.edge2D and raw["edge_num"][1] == pi.edge2D - boundary
    # nlevels_nod2D = max over the node's elements (the rule the reference's mesh files obey)
    lev = np.array([pi.nlevels[pi.nod_in_elem2D[n, :pi.nod_in_elem2D_num[n]] - 1].max() for n in range(pi.nod2D)])
    np.testing.assert_array_equal(lev, pi.nlevels_nod2D)
    assert np.all(pi.nlevels_nod2D_min <= pi.nlevels_nod2D)
    np.testing.assert_allclose(pi.zbar[:3], [0.0, -5.0, -10.0])


def test_orientation_conventions(pi):
    en = pi.elem2D_nodes - 1
    xy = pi.coord_nod2D
    b = xy[en[:, 1]] - xy[en[:, 0]]
    c = xy[en[:, 2]] - xy[en[:, 0]]
    cross = mesh_io.trim_lon(b[:, 0]) * c[:, 1] - b[:, 1] * mesh_io.trim_lon(c[:, 0])
    assert np.all(cross < 0)  # clockwise in (lon, lat)
    # edge_tri(1) is the element left of edges(1) -> edges(2)
    ed = pi.edges - 1
    ce = np.cos(0.5 * (xy[ed[:, 0], 1] + xy[ed[:, 1], 1]))
    side = pi.edge_dxdy[:, 0] * ce * pi.edge_cross_dxdy[:, 1] - pi.edge_dxdy[:, 1] * pi.edge_cross_dxdy[:, 0]
    assert np.all(side > 0)
    # up/down-wind triangles contain the edge's end node
    ud = pi.edge_up_dn_tri
    for s_, col in ((0, 0), (1, 1)):
        has = ud[:, s_] > 0
        assert np.all(np.any(pi.elem2D_nodes[ud[has, s_] - 1] == pi.edges[has, col][:, None], axis=1))


def test_gradients_and_areas_closed_forms(pi):
    en = pi.elem2D_nodes - 1
    lon, lat = pi.coord_nod2D[en, 0], pi.coord_nod2D[en, 1]
    inside = (lon.max(axis=1) - lon.min(axis=1)) < math.pi  # elements not across the longitude wrap
    g = pi.gradient_sca
    dlon = np.sum(g[:, :3] * lon, axis=1)
    dlat = np.sum(g[:, 3:] * lat, axis=1)
    np.testing.assert_allclose(dlon[inside] * R_EARTH * pi.elem_cos[inside], 1.0, rtol=1e-9)
    np.testing.assert_allclose(dlat * R_EARTH, 1.0, rtol=1e-9)
    np.testing.assert_allclose(np.sum(g[:, :3], axis=1), 0.0, atol=1e-20)
    # every element is wet in the surface layer: the surface areas add up to the element areas
    np.testing.assert_allclose(pi.area[:, 0].sum(), pi.elem_area.sum(), rtol=1e-12)
    assert 3.0e14 < pi.elem_area.sum() < 3.7e14  # the world ocean


def test_oracle_on_pi_mesh(oracle_mod, pi):
    wet = configs.wet_mask(pi)
    vol = (pi.areasvol[:, :pi.nl - 1] * pi.hnode)[wet]
    # FCT keeps both tracers within their initial range (divergence-free layers, w = 0)
    case = configs.sphere_case("pi", pi, seed=3, ntracers=2, nsteps=5)
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[5])
    for i in range(2):
        v0, v = case.tracers[i][wet], ref[5][i][wet]
        assert v.min() >= v0.min() - 1e-9 * abs(v0.min()) and v.max() <= v0.max() * (1 + 1e-9)
        assert abs(np.sum(v * vol) - np.sum(v0 * vol)) <= 1e-13 * abs(np.sum(v0 * vol))
    # a uniform tracer stays uniform
    case.tracers[0] = np.where(wet, 7.0, 0.0)
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[5])
    assert np.max(np.abs(ref[5][0][wet] - 7.0)) <= 1e-12 * 7.0
    # with an interior vertical velocity the volume sum is still conserved
    case = configs.sphere_case("pi_w", pi, seed=4, ntracers=2, nsteps=3, w_cfl=0.3)
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    for i in range(2):
        v0, v = case.tracers[i][wet], ref[3][i][wet]
        assert abs(np.sum(v * vol) - np.sum(v0 * vol)) <= 1e-13 * abs(np.sum(v0 * vol))


def test_write_read_round_trip(raw, tmp_path):
    mesh_io.write_mesh_files(raw, tmp_path / "pi")
    back = mesh_io.read_mesh_files(tmp_path / "pi")
    for k in raw:
        np.testing.assert_array_equal(back[k], raw[k])


def test_library_reads_mesh_directory_without_gpu(raw, tmp_path):
    """tracer_load_mesh_from_files parses the files and fills the mesh state on the host (the
    device is touched only at tracer_init_complete); a missing directory is an error."""
    from fesom2_tracer_dwarf_amd import abi
    lib = abi.load()
    lib.tracer_cleanup()
    assert lib.tracer_load_mesh_from_files(str(tmp_path / "none").encode()) == -1
    assert b"cannot open" in lib.tracer_last_error()
    mesh_io.write_mesh_files(raw, tmp_path / "pi")
    assert lib.tracer_load_mesh_from_files(str(tmp_path / "pi").encode()) == 0
    assert lib.tracer_allocate_tracers(2, 48, 3140, 2) == 0
    assert lib.tracer_allocate_tracers(2, 47, 3140, 2) == -1          # nl of the loaded mesh is 48
    assert lib.tracer_set_scheme(1, b"MFCT", b"QR4C", b"FCT", 1.0, 1.0) == 0  # MUSCL support derived
    lib.tracer_cleanup()
