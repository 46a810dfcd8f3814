"""FESOM2 mesh files: reader and the geometry of a mesh on the sphere.

A FESOM2 mesh directory holds (read by lib/oce_mesh.F90 ``read_mesh``, ``load_edges`` and
``find_levels``):

==============  =========================================================================
``nod2d.out``   N, then N lines ``index lon lat coast`` (degrees; coast 1 on the coastline)
``elem2d.out``  T, then T lines with the three node indices of an element
``aux3d.out``   nl, then the nl standard depths ``zbar`` (then per-node depths, unused here)
``nlvls.out``   N lines: ``nlevels_nod2D``
``elvls.out``   T lines: ``nlevels`` of each element
``edgenum.out`` ``edge2D`` and ``edge2D_in``
``edges.out``   E lines: the two nodes of an edge
``edge_tri.out`` E lines: the elements left and right of an edge (0: none, a boundary edge)
==============  =========================================================================

:func:`read_mesh_files` returns these as arrays; :func:`sphere_mesh` derives every array the
advection step needs, restating the reference's spherical (non-Cartesian) mesh set-up:
coordinates rotated to the model grid (``g2r``, lib/gen_modules_rotate_grid.F90:92-119, with the
default Euler angles of lib/gen_modules_config.F90:86-88), triangle orientation (``test_tri``,
lib/oce_mesh.F90:1565-1607), areas (``mesh_areas``, :2082-2195) and ``elem_cos``, edge vectors,
cross distances and P1 gradients (``mesh_auxiliary_arrays``, :2385-2534).  Longitude
differences are brought into [-pi, pi] (the reference's ``trim_cyclic`` compares radian
differences with ``cyclic_length`` = 360 and therefore never wraps; the intent, a minimum-image
difference, is applied here).
"""

from __future__ import annotations

import math
from pathlib import Path
from typing import Dict, Tuple

import numpy as np

from .mesh import R_EARTH, Mesh, _nboundary_lay, _ragged_lists, _up_down_triangles

TWO_PI = 2.0 * math.pi


def _ints(path: Path) -> np.ndarray:
    return np.array(path.read_text().split(), dtype=np.int64)


def read_mesh_files(path) -> Dict[str, np.ndarray]:
    """Raw contents of a FESOM2 mesh directory (indices 1-based as in the files)."""
    p = Path(path)
    nod = np.array(p.joinpath("nod2d.out").read_text().split()[1:], dtype=np.float64).reshape(-1, 4)
    el = _ints(p / "elem2d.out")
    elem = el[1:].reshape(int(el[0]), 3)
    aux = p.joinpath("aux3d.out").read_text().split()
    nl = int(aux[0])
    zbar = np.array(aux[1:1 + nl], dtype=np.float64)
    edge_num = _ints(p / "edgenum.out")
    return {
        "nod_index": nod[:, 0].astype(np.int64), "lon": nod[:, 1], "lat": nod[:, 2],
        "coast": nod[:, 3].astype(np.int64),
        "elem2D_nodes": elem, "zbar": zbar,
        "nlevels_nod2D": _ints(p / "nlvls.out"), "nlevels": _ints(p / "elvls.out"),
        "edges": _ints(p / "edges.out").reshape(-1, 2), "edge_tri": _ints(p / "edge_tri.out").reshape(-1, 2),
        "edge_num": edge_num,
    }


def write_mesh_files(raw: Dict[str, np.ndarray], path) -> None:
    """Writes arrays in the layout of :func:`read_mesh_files` as a FESOM2 mesh directory (full
    double precision, so that reading it back gives the same numbers)."""
    p = Path(path)
    p.mkdir(parents=True, exist_ok=True)
    N = len(raw["lon"])
    lines = [str(N)] + [f"{int(i)} {float(x)!r} {float(y)!r} {int(c)}" for i, x, y, c in
                        zip(raw["nod_index"], raw["lon"], raw["lat"], raw["coast"])]
    p.joinpath("nod2d.out").write_text("\n".join(lines) + "\n")
    el = np.asarray(raw["elem2D_nodes"])
    p.joinpath("elem2d.out").write_text("\n".join([str(len(el))] + [" ".join(map(str, r)) for r in el.tolist()]) + "\n")
    z = np.asarray(raw["zbar"])
    p.joinpath("aux3d.out").write_text("\n".join([str(len(z))] + [repr(float(v)) for v in z]) + "\n")
    p.joinpath("nlvls.out").write_text("\n".join(map(str, np.asarray(raw["nlevels_nod2D"]).tolist())) + "\n")
    p.joinpath("elvls.out").write_text("\n".join(map(str, np.asarray(raw["nlevels"]).tolist())) + "\n")
    p.joinpath("edgenum.out").write_text("\n".join(map(str, np.asarray(raw["edge_num"]).tolist())) + "\n")
    for name, key in (("edges.out", "edges"), ("edge_tri.out", "edge_tri")):
        p.joinpath(name).write_text("\n".join(" ".join(map(str, r)) for r in np.asarray(raw[key]).tolist()) + "\n")


def euler_matrix(alpha: float = 50.0, beta: float = 15.0, gamma: float = -90.0) -> np.ndarray:
    """Rotation from geographic to model coordinates for Euler angles in degrees: first alpha about
    z, then beta about the new x, then gamma about the new z (set_mesh_transform_matrix,
    lib/gen_modules_rotate_grid.F90:33-56)."""
    al, be, ga = (math.radians(a) for a in (alpha, beta, gamma))
    return np.array([
        [math.cos(ga) * math.cos(al) - math.sin(ga) * math.cos(be) * math.sin(al),
         math.cos(ga) * math.sin(al) + math.sin(ga) * math.cos(be) * math.cos(al),
         math.sin(ga) * math.sin(be)],
        [-math.sin(ga) * math.cos(al) - math.cos(ga) * math.cos(be) * math.sin(al),
         -math.sin(ga) * math.sin(al) + math.cos(ga) * math.cos(be) * math.cos(al),
         math.cos(ga) * math.sin(be)],
        [math.sin(be) * math.sin(al), -math.sin(be) * math.cos(al), math.cos(be)]])


def geo_to_model(lon: np.ndarray, lat: np.ndarray, rot: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """g2r: geographic (radians) to rotated model coordinates (radians)."""
    g = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])
    r = rot @ g
    rlat = np.arcsin(np.clip(r[2], -1.0, 1.0))
    rlon = np.where((r[0] == 0.0) & (r[1] == 0.0), 0.0, np.arctan2(r[1], r[0]))
    return rlon, rlat


def trim_lon(d: np.ndarray) -> np.ndarray:
    """Longitude difference brought into [-pi, pi]."""
    d = np.where(d > math.pi, d - TWO_PI, d)
    return np.where(d < -math.pi, d + TWO_PI, d)


def sphere_mesh(raw: Dict[str, np.ndarray], rotate: bool = True,
                euler: Tuple[float, float, float] = (50.0, 15.0, -90.0)) -> Mesh:
    """Every array of :class:`Mesh` for a FESOM2 file mesh (see the module docstring)."""
    lon = np.radians(np.asarray(raw["lon"], dtype=np.float64))
    lat = np.radians(np.asarray(raw["lat"], dtype=np.float64))
    if rotate:
        lon, lat = geo_to_model(lon, lat, euler_matrix(*euler))
    xy = np.stack([lon, lat], axis=1)
    N = xy.shape[0]
    enodes = np.asarray(raw["elem2D_nodes"], dtype=np.int64) - 1
    T = enodes.shape[0]
    ednodes = np.asarray(raw["edges"], dtype=np.int64) - 1
    edtri = np.asarray(raw["edge_tri"], dtype=np.int64) - 1      # -1: no element
    edtri = np.where(edtri < 0, -1, edtri)
    E = ednodes.shape[0]
    zbar = np.asarray(raw["zbar"], dtype=np.float64)
    if zbar.size > 1 and zbar[1] > 0:
        zbar = -zbar                                               # zbar is negative downwards
    nl = zbar.size

    # test_tri: clockwise in (lon, lat)
    b = xy[enodes[:, 1]] - xy[enodes[:, 0]]
    c = xy[enodes[:, 2]] - xy[enodes[:, 0]]
    r = trim_lon(b[:, 0]) * c[:, 1] - b[:, 1] * trim_lon(c[:, 0])
    flip = r > 0
    enodes[flip] = enodes[flip][:, [0, 2, 1]]

    # nod_in_elem2D in ascending element order; levels from the level files (no cavities)
    nie, nie_num = _ragged_lists(enodes.ravel(), np.repeat(np.arange(T, dtype=np.int64), 3), N)
    max_adj = nie.shape[1]
    nie_valid = np.arange(max_adj)[None, :] < nie_num[:, None]
    nlevels = np.asarray(raw["nlevels"], dtype=np.int32)
    ulevels = np.ones(T, dtype=np.int32)
    nlevels_nod2D = np.asarray(raw["nlevels_nod2D"], dtype=np.int32)
    nlevels_nod2D_min = np.where(nie_valid, nlevels[nie], np.iinfo(np.int32).max).min(axis=1).astype(np.int32)
    ulevels_nod2D = np.ones(N, dtype=np.int32)

    # areas (mesh_areas): cos of the mean latitude of the three nodes scales longitude
    p1, p2, p3 = xy[enodes[:, 0]], xy[enodes[:, 1]], xy[enodes[:, 2]]
    ay = np.cos((p1[:, 1] + p2[:, 1] + p3[:, 1]) / 3.0)
    a = np.stack([trim_lon(p2[:, 0] - p1[:, 0]) * ay, p2[:, 1] - p1[:, 1]], 1)
    bb = np.stack([trim_lon(p3[:, 0] - p1[:, 0]) * ay, p3[:, 1] - p1[:, 1]], 1)
    elem_area = 0.5 * np.abs(a[:, 0] * bb[:, 1] - bb[:, 0] * a[:, 1])
    area = np.zeros((N, nl), dtype=np.float64)
    levels = np.arange(1, nl + 1)[None, :]
    for k in range(max_adj):
        el = nie[:, k]
        wet = (levels <= (nlevels[el] - 1)[:, None]) & (levels >= ulevels[el][:, None]) & nie_valid[:, k][:, None]
        area += np.where(wet, (elem_area[el] / 3.0)[:, None], 0.0)
    node_wet = (levels >= ulevels_nod2D[:, None]) & (levels <= (nlevels_nod2D - 1)[:, None])
    areasvol = np.where(node_wet, area, 0.0)
    elem_area = elem_area * R_EARTH * R_EARTH
    area = area * R_EARTH * R_EARTH
    areasvol = areasvol * R_EARTH * R_EARTH

    # element centres (elem_center: longitudes unwrapped around the smallest) and elem_cos
    ax = lon[enodes]
    amin = ax.min(axis=1, keepdims=True)
    ax = np.where(ax - amin >= math.pi, ax - TWO_PI, ax)
    ax = np.where(ax - amin < -math.pi, ax + TWO_PI, ax)
    cx = ax.sum(axis=1) / 3.0
    cy = lat[enodes].sum(axis=1) / 3.0
    elem_cos = np.cos(cy)

    # edge vectors in radians, cross distances from the edge centre to the element centres (m)
    edge_dxdy = np.stack([trim_lon(lon[ednodes[:, 1]] - lon[ednodes[:, 0]]),
                          lat[ednodes[:, 1]] - lat[ednodes[:, 0]]], 1)
    ea, eb = xy[ednodes[:, 0]].copy(), xy[ednodes[:, 1]].copy()
    d = ea[:, 0] - eb[:, 0]
    ea[:, 0] = np.where(d > math.pi, ea[:, 0] - TWO_PI, ea[:, 0])
    eb[:, 0] = np.where(d < -math.pi, eb[:, 0] - TWO_PI, eb[:, 0])
    mid = 0.5 * (ea + eb)
    ecd = np.zeros((E, 4), dtype=np.float64)
    for s_ in range(2):
        el = edtri[:, s_]
        has = el >= 0
        elc = np.where(has, el, 0)
        v = np.stack([trim_lon(cx[elc] - mid[:, 0]) * elem_cos[elc], cy[elc] - mid[:, 1]], 1) * R_EARTH
        ecd[:, 2 * s_:2 * s_ + 2] = np.where(has[:, None], v, 0.0)

    # P1 gradients (gradient_sca)
    dX31 = trim_lon(p3[:, 0] - p1[:, 0]) * elem_cos
    dX21 = trim_lon(p2[:, 0] - p1[:, 0]) * elem_cos
    dY31 = p3[:, 1] - p1[:, 1]
    dY21 = p2[:, 1] - p1[:, 1]
    dfac = -0.5 * R_EARTH / elem_area
    gradient_sca = np.stack([(-dY31 + dY21) * dfac, dY31 * dfac, -dY21 * dfac,
                             (dX31 - dX21) * dfac, -dX31 * dfac, dX21 * dfac], axis=1)

    # vertical grid (full cells; no ALE motion)
    layer_h = zbar[:-1] - zbar[1:]
    Z = 0.5 * (zbar[:-1] + zbar[1:])
    hnode = np.broadcast_to(layer_h, (N, nl - 1)).copy()
    helem = np.broadcast_to(layer_h, (T, nl - 1)).copy()

    up_dn = _up_down_triangles(xy, enodes, ednodes, nie, nie_num, TWO_PI, math.inf, require_all=False)
    nbl = _nboundary_lay(nl, N, ednodes, edtri, nlevels)
    return Mesh(
        nl=nl, nod2D=N, elem2D=T, edge2D=E, max_adj=max_adj, coord_nod2D=xy,
        elem2D_nodes=(enodes + 1).astype(np.int32), edges=(ednodes + 1).astype(np.int32),
        edge_tri=(edtri + 1).astype(np.int32),
        nod_in_elem2D=np.where(nie_valid, nie + 1, 0).astype(np.int32), nod_in_elem2D_num=nie_num.astype(np.int32),
        nlevels=nlevels, ulevels=ulevels, nlevels_nod2D=nlevels_nod2D, ulevels_nod2D=ulevels_nod2D,
        nlevels_nod2D_min=nlevels_nod2D_min, ulevels_nod2D_max=np.ones(N, dtype=np.int32),
        elem_area=elem_area, area=area, areasvol=areasvol, edge_dxdy=edge_dxdy, edge_cross_dxdy=ecd,
        elem_cos=elem_cos, gradient_sca=gradient_sca, zbar=zbar, Z=Z, hnode=hnode, hnode_new=hnode.copy(),
        helem=helem, zbar_3d_n=np.broadcast_to(zbar, (N, nl)).copy(), Z_3d_n=np.broadcast_to(Z, (N, nl - 1)).copy(),
        edge_up_dn_tri=(up_dn + 1).astype(np.int32), nboundary_lay=nbl)


def edge_lengths(mesh: Mesh) -> np.ndarray:
    """Edge lengths in metres (local tangent plane at the edge's mean latitude)."""
    en = mesh.edges - 1
    lat = mesh.coord_nod2D[:, 1]
    c = np.cos(0.5 * (lat[en[:, 0]] + lat[en[:, 1]]))
    return R_EARTH * np.hypot(mesh.edge_dxdy[:, 0] * c, mesh.edge_dxdy[:, 1])
