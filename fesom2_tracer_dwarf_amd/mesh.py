"""Doubly-periodic triangular FESOM meshes with their geometry and vertical structure.

Host-side preprocessing (runs once per mesh, NumPy, deterministic from a seed).  Every array is
produced in the reference's conventions so it can be handed unchanged to the reference Fortran,
to the C-ABI library and to the CPU restatement:

* 1-based indices, 0 meaning "no element" (``edge_tri(2, e) == 0`` marks a boundary edge);
* Fortran column-major layout: an array the reference declares ``a(k, n)`` is stored here as a
  C-ordered NumPy array of shape ``(n, k)`` (same bytes);
* triangles listed clockwise, as ``test_tri`` enforces (lib/oce_mesh.F90:1565-1615);
* ``edge_tri(1, e)`` is the element to the left of ``edges(1,e) -> edges(2,e)``
  (lib/MOD_MESH.F90:35-37);
* ``nod_in_elem2D(:, n)`` lists every element touching ``n`` in ascending element order, the
  order in which lib/analytic_mesh.F90:675-694 fills it, but complete across the periodic wrap;
* ``edge_dxdy`` in radians, ``edge_cross_dxdy`` / areas in metres, ``elem_cos == 1`` (Cartesian),
  all with minimum-image differences across the wrap (lib/oce_mesh.F90:2313-2480 restated).

The reference's own analytic generator (lib/analytic_mesh.F90) stores the wrong element as
``edge_tri(1)`` for horizontal/vertical edges and leaves ``nod_in_elem2D`` incomplete at the wrap,
so it is replaced here rather than reproduced.
"""

from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

R_EARTH = 6367500.0  # lib/oce_modules.F90:24


@dataclasses.dataclass
class Mesh:
    """All mesh arrays of one single-partition FESOM mesh (see module docstring for layout)."""

    nl: int
    nod2D: int
    elem2D: int
    edge2D: int
    max_adj: int
    coord_nod2D: np.ndarray          # (N, 2) float64, radians
    elem2D_nodes: np.ndarray         # (T, 3) int32
    edges: np.ndarray                # (E, 2) int32
    edge_tri: np.ndarray             # (E, 2) int32
    nod_in_elem2D: np.ndarray        # (N, max_adj) int32 (0-padded)
    nod_in_elem2D_num: np.ndarray    # (N,) int32
    nlevels: np.ndarray              # (T,) int32
    ulevels: np.ndarray              # (T,) int32
    nlevels_nod2D: np.ndarray        # (N,) int32
    ulevels_nod2D: np.ndarray        # (N,) int32
    nlevels_nod2D_min: np.ndarray    # (N,) int32
    ulevels_nod2D_max: np.ndarray    # (N,) int32
    elem_area: np.ndarray            # (T,) m^2
    area: np.ndarray                 # (N, nl) m^2
    areasvol: np.ndarray             # (N, nl) m^2
    edge_dxdy: np.ndarray            # (E, 2) radians
    edge_cross_dxdy: np.ndarray      # (E, 4) metres
    elem_cos: np.ndarray             # (T,)
    gradient_sca: np.ndarray         # (T, 6) 1/m
    zbar: np.ndarray                 # (nl,) interface depths (<= 0)
    Z: np.ndarray                    # (nl-1,) mid-layer depths
    hnode: np.ndarray                # (N, nl-1) m
    hnode_new: np.ndarray            # (N, nl-1) m
    helem: np.ndarray                # (T, nl-1) m
    zbar_3d_n: np.ndarray            # (N, nl)
    Z_3d_n: np.ndarray               # (N, nl-1)
    edge_up_dn_tri: np.ndarray       # (E, 2) int32
    nboundary_lay: np.ndarray        # (N,) int32
    Lx: float = 0.0
    Ly: float = 0.0
    nx: int = 0
    ny: int = 0
    # positions in metres on the integer lattice used for exact predicates
    xy_m: Optional[np.ndarray] = None

    @property
    def n_node_levels(self) -> int:
        """N_L = sum over nodes of wet layers."""
        return int(np.sum(self.nlevels_nod2D - self.ulevels_nod2D))

    @property
    def n_elem_levels(self) -> int:
        return int(np.sum(self.nlevels - self.ulevels))


def stretched_layers(n_layers: int, h_top: float, h_bottom: float) -> np.ndarray:
    """Layer thicknesses growing geometrically from ``h_top`` to ``h_bottom``."""
    if n_layers == 1:
        return np.array([h_top], dtype=np.float64)
    r = (h_bottom / h_top) ** (1.0 / (n_layers - 1))
    return h_top * r ** np.arange(n_layers, dtype=np.float64)


def _min_image(d: np.ndarray, period: float) -> np.ndarray:
    """Minimum-image difference on a periodic axis (exact for lattice coordinates); an infinite
    period leaves ``d`` unchanged."""
    if not np.isfinite(period):
        return d
    d = np.where(d > 0.5 * period, d - period, d)
    return np.where(d < -0.5 * period, d + period, d)


def _ragged_lists(owner: np.ndarray, member: np.ndarray, n_owner: int):
    """Group ``member`` by ``owner`` keeping ascending member order; returns (table, count)."""
    order = np.lexsort((member, owner))
    own = owner[order]
    mem = member[order]
    count = np.bincount(own, minlength=n_owner).astype(np.int32)
    width = int(count.max()) if count.size else 0
    start = np.zeros(n_owner + 1, dtype=np.int64)
    np.cumsum(count, out=start[1:])
    pos = np.arange(own.size, dtype=np.int64) - start[own]
    table = np.zeros((n_owner, width), dtype=np.int32)
    table[own, pos] = mem
    return table, count


def periodic_mesh(nx: int, ny: int, Lx: float, Ly: float, layer_h: np.ndarray, *,
                  elem_nlevels: Optional[np.ndarray] = None, elem_ulevels: Optional[np.ndarray] = None,
                  seed: Optional[int] = None, permute: bool = False) -> Mesh:
    """Build a doubly periodic ``nx`` x ``ny`` cell mesh, two triangles per cell.

    ``layer_h`` gives the nl-1 nominal layer thicknesses (surface first).  ``elem_nlevels`` gives
    per-element level counts (in canonical element order) for bathymetry; default: all full depth.
    ``elem_ulevels`` gives each element's first wet layer (an ice-shelf cavity where > 1; at least
    two wet layers must remain); default: 1 everywhere.
    With ``permute`` the node, element and edge numbering, the edge directions and the starting
    vertex of every triangle are shuffled from ``seed`` (the geometry is unchanged).
    """
    if nx < 3 or ny < 3:
        raise ValueError("periodic mesh needs at least 3x3 cells")
    layer_h = np.asarray(layer_h, dtype=np.float64)
    nl = layer_h.size + 1
    N = nx * ny
    T = 2 * N
    E = 3 * N
    dx = Lx / nx
    dy = Ly / ny

    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ii = ii.ravel()
    jj = jj.ravel()

    def nid(i, j):
        return (np.mod(j, ny) * nx + np.mod(i, nx)).astype(np.int64)

    a = nid(ii, jj)
    b = nid(ii + 1, jj)
    c = nid(ii + 1, jj + 1)
    d = nid(ii, jj + 1)
    # Elements, clockwise: lower-right (a, c, b) and upper-left (a, d, c) of each cell.
    enodes = np.empty((T, 3), dtype=np.int64)
    enodes[0::2] = np.stack([a, c, b], axis=1)
    enodes[1::2] = np.stack([a, d, c], axis=1)

    def lower(i, j):  # element (a, c, b) of cell (i, j)
        return 2 * nid(i, j)

    def upper(i, j):  # element (a, d, c) of cell (i, j)
        return 2 * nid(i, j) + 1

    # Edges: +x, +y and the +x+y diagonal from each node; edge_tri(1) on the left.
    ednodes = np.empty((E, 2), dtype=np.int64)
    edtri = np.empty((E, 2), dtype=np.int64)
    ednodes[0::3] = np.stack([a, b], axis=1)
    edtri[0::3] = np.stack([lower(ii, jj), upper(ii, jj - 1)], axis=1)
    ednodes[1::3] = np.stack([a, d], axis=1)
    edtri[1::3] = np.stack([lower(ii - 1, jj), upper(ii, jj)], axis=1)
    ednodes[2::3] = np.stack([a, c], axis=1)
    edtri[2::3] = np.stack([upper(ii, jj), lower(ii, jj)], axis=1)

    if elem_nlevels is None:
        elem_nlev = np.full(T, nl, dtype=np.int64)
    else:
        elem_nlev = np.asarray(elem_nlevels, dtype=np.int64).copy()
        if elem_nlev.shape != (T,):
            raise ValueError("elem_nlevels must have one entry per element")
        if elem_nlev.min() < 3 or elem_nlev.max() > nl:
            raise ValueError("element level counts must lie in [3, nl]")
    elem_ulev = np.ones(T, dtype=np.int64) if elem_ulevels is None else np.asarray(elem_ulevels, dtype=np.int64).copy()
    if elem_ulev.shape != (T,) or elem_ulev.min() < 1 or np.any(elem_nlev - elem_ulev < 2):
        raise ValueError("elem_ulevels: one entry >= 1 per element, with nlevels - ulevels >= 2")

    # Lattice coordinates: metres are integers when dx, dy are, which keeps the up/down-wind
    # predicates exact.
    xy = np.stack([ii * dx, jj * dy], axis=1).astype(np.float64)

    if permute:
        rng = np.random.default_rng(seed)
        pn = rng.permutation(N)          # new id of canonical node
        pt = rng.permutation(T)
        pe = rng.permutation(E)
        flip = rng.random(E) < 0.5
        rot = rng.integers(0, 3, size=T)
        inv_n = np.empty(N, dtype=np.int64)
        inv_n[pn] = np.arange(N)
        inv_t = np.empty(T, dtype=np.int64)
        inv_t[pt] = np.arange(T)
        inv_e = np.empty(E, dtype=np.int64)
        inv_e[pe] = np.arange(E)
        xy = xy[inv_n]
        enodes = pn[enodes][inv_t]
        rot = rot[inv_t]
        idx = (np.arange(3)[None, :] + rot[:, None]) % 3
        enodes = np.take_along_axis(enodes, idx, axis=1)
        elem_nlev = elem_nlev[inv_t]
        elem_ulev = elem_ulev[inv_t]
        ednodes = pn[ednodes]
        edtri = pt[edtri]
        ednodes[flip] = ednodes[flip][:, ::-1]
        edtri[flip] = edtri[flip][:, ::-1]
        ednodes = ednodes[inv_e]
        edtri = edtri[inv_e]

    return _finish_mesh(nl, xy, enodes, ednodes, edtri, elem_nlev, layer_h, Lx, Ly, nx, ny, elem_ulev)


def _finish_mesh(nl, xy, enodes, ednodes, edtri, elem_nlev, layer_h, Lx, Ly, nx, ny, elem_ulev=None) -> Mesh:
    """Derive adjacency, levels, geometry and MUSCL support arrays (0-based inputs)."""
    N = xy.shape[0]
    T = enodes.shape[0]
    E = ednodes.shape[0]

    # nod_in_elem2D: ascending element order per node.
    owner = enodes.ravel()
    member = np.repeat(np.arange(T, dtype=np.int64), 3)
    nie, nie_num = _ragged_lists(owner, member, N)
    max_adj = nie.shape[1]
    nie_valid = np.arange(max_adj)[None, :] < nie_num[:, None]

    # Levels: a node is wet from the shallowest top to the deepest bottom of its elements
    # (ulevels == 1 everywhere unless a cavity is given).
    nlevels = elem_nlev.astype(np.int32)
    ulevels = (np.ones(T) if elem_ulev is None else elem_ulev).astype(np.int32)
    lev_tab = np.where(nie_valid, nlevels[nie], 0)
    nlevels_nod2D = lev_tab.max(axis=1).astype(np.int32)       # node = max over its elements
    nlevels_nod2D_min = np.where(nie_valid, nlevels[nie], np.iinfo(np.int32).max).min(axis=1)
    nlevels_nod2D_min = nlevels_nod2D_min.astype(np.int32)     # lib/oce_mesh.F90:1522-1560
    ulevels_nod2D = np.where(nie_valid, ulevels[nie], np.iinfo(np.int32).max).min(axis=1).astype(np.int32)
    ulevels_nod2D_max = np.where(nie_valid, ulevels[nie], 0).max(axis=1).astype(np.int32)

    # ---------------------------------------------------------------- geometry (metres)
    p1 = xy[enodes[:, 0]]
    p2 = xy[enodes[:, 1]]
    p3 = xy[enodes[:, 2]]
    d21 = np.stack([_min_image(p2[:, 0] - p1[:, 0], Lx), _min_image(p2[:, 1] - p1[:, 1], Ly)], 1)
    d31 = np.stack([_min_image(p3[:, 0] - p1[:, 0], Lx), _min_image(p3[:, 1] - p1[:, 1], Ly)], 1)
    cross = d21[:, 0] * d31[:, 1] - d31[:, 0] * d21[:, 1]
    if np.any(cross >= 0):
        raise ValueError("triangles must be listed clockwise")
    elem_area = 0.5 * np.abs(cross)                           # lib/oce_mesh.F90:2082-2095
    centroid = p1 + (d21 + d31) / 3.0                         # unwrapped around vertex 1

    # area(nz, n): a third of every wet element touching n (lib/oce_mesh.F90:2134-2160).
    area = np.zeros((N, nl), dtype=np.float64)
    levels = np.arange(1, nl + 1)[None, :]
    for k in range(max_adj):
        el = nie[:, k]
        ok = nie_valid[:, k]
        wet = (levels <= (nlevels[el] - 1)[:, None]) & (levels >= ulevels[el][:, None]) & ok[:, None]
        area += np.where(wet, (elem_area[el] / 3.0)[:, None], 0.0)
    node_wet = (levels >= ulevels_nod2D[:, None]) & (levels <= (nlevels_nod2D - 1)[:, None])
    areasvol = np.where(node_wet, area, 0.0)                  # lib/oce_mesh.F90:2185-2193

    # Edge vectors (radians, lib/oce_mesh.F90:2417-2424) and cross distances (metres, :2431-2452).
    e1 = xy[ednodes[:, 0]]
    e2 = xy[ednodes[:, 1]]
    ev = np.stack([_min_image(e2[:, 0] - e1[:, 0], Lx), _min_image(e2[:, 1] - e1[:, 1], Ly)], 1)
    edge_dxdy = ev / R_EARTH
    mid = e1 + 0.5 * ev
    ecd = np.zeros((E, 4), dtype=np.float64)
    for s in range(2):
        el = edtri[:, s]
        has = el >= 0
        elc = np.where(has, el, 0)
        v = centroid[elc] - mid
        v = np.stack([_min_image(v[:, 0], Lx), _min_image(v[:, 1], Ly)], 1)
        ecd[:, 2 * s:2 * s + 2] = np.where(has[:, None], v, 0.0)
    # Left-element convention check: the left centroid lies to the left of the directed edge.
    side = ev[:, 0] * ecd[:, 1] - ev[:, 1] * ecd[:, 0]
    if np.any(side <= 0):
        raise ValueError("edge_tri(1) must be the element to the left of the edge")

    elem_cos = np.ones(T, dtype=np.float64)
    # P1 gradient coefficients (lib/oce_mesh.F90:2512-2534) for clockwise triangles.
    dfac = -0.5 / elem_area
    dX31, dX21 = d31[:, 0] * elem_cos, d21[:, 0] * elem_cos
    dY31, dY21 = d31[:, 1], d21[:, 1]
    gradient_sca = np.stack([(-dY31 + dY21) * dfac, dY31 * dfac, -dY21 * dfac,
                             (dX31 - dX21) * dfac, -dX31 * dfac, dX21 * dfac], axis=1)

    # Vertical grid (no ALE motion: hnode_new == hnode).
    zbar = np.concatenate([[0.0], -np.cumsum(layer_h)])
    Z = 0.5 * (zbar[:-1] + zbar[1:])
    # layers above a cavity's top have no thickness
    lay = np.arange(1, nl)[None, :]
    hnode = np.where(lay >= ulevels_nod2D[:, None], layer_h[None, :], 0.0)
    hnode_new = hnode.copy()
    helem = np.where(lay >= ulevels[:, None], layer_h[None, :], 0.0)
    zbar_3d_n = np.broadcast_to(zbar, (N, nl)).copy()
    Z_3d_n = np.broadcast_to(Z, (N, nl - 1)).copy()

    up_dn = _up_down_triangles(xy, enodes, ednodes, nie, nie_num, Lx, Ly)
    nbl = _nboundary_lay(nl, N, ednodes, edtri, nlevels)

    return Mesh(
        nl=nl, nod2D=N, elem2D=T, edge2D=E, max_adj=max_adj,
        coord_nod2D=xy / R_EARTH,
        elem2D_nodes=(enodes + 1).astype(np.int32),
        edges=(ednodes + 1).astype(np.int32),
        edge_tri=(edtri + 1).astype(np.int32),
        nod_in_elem2D=np.where(nie_valid, nie + 1, 0).astype(np.int32),
        nod_in_elem2D_num=nie_num.astype(np.int32),
        nlevels=nlevels, ulevels=ulevels,
        nlevels_nod2D=nlevels_nod2D, ulevels_nod2D=ulevels_nod2D,
        nlevels_nod2D_min=nlevels_nod2D_min, ulevels_nod2D_max=ulevels_nod2D_max,
        elem_area=elem_area, area=area, areasvol=areasvol,
        edge_dxdy=edge_dxdy, edge_cross_dxdy=ecd, elem_cos=elem_cos, gradient_sca=gradient_sca,
        zbar=zbar, Z=Z, hnode=hnode, hnode_new=hnode_new, helem=helem,
        zbar_3d_n=zbar_3d_n, Z_3d_n=Z_3d_n,
        edge_up_dn_tri=(up_dn + 1).astype(np.int32), nboundary_lay=nbl,
        Lx=Lx, Ly=Ly, nx=nx, ny=ny, xy_m=xy,
    )


def _up_down_triangles(xy, enodes, ednodes, nie, nie_num, Lx, Ly, require_all: bool = True) -> np.ndarray:
    """Up- and down-wind triangle of every edge (restates lib/oce_muscl_adv.F90:245-338).

    For edge n1->n2 with x = X(n2) - X(n1), the upwind triangle is the element around n1 whose
    corner at n1 contains -x, the downwind one the element around n2 containing +x.  Every
    element satisfying one of the three tests is assigned in ``nod_in_elem2D`` order (the
    reference's ``cycle`` continues the search), so the last match wins; 0-based result, -1 if
    none matches (allowed with ``require_all=False``: edges at a coast, as in the reference).
    Coordinates wrap with periods ``Lx``, ``Ly`` (``inf``: no wrap).
    """
    E = ednodes.shape[0]
    out = np.full((E, 2), -1, dtype=np.int64)
    xvec = xy[ednodes[:, 1]] - xy[ednodes[:, 0]]
    xvec = np.stack([_min_image(xvec[:, 0], Lx), _min_image(xvec[:, 1], Ly)], 1)
    for side, (node_col, sgn) in enumerate([(0, -1.0), (1, 1.0)]):
        node = ednodes[:, node_col]
        x = sgn * xvec
        for k in range(nie.shape[1]):
            ok = k < nie_num[node]
            elem = nie[node, k]
            en = enodes[elem]
            pos = np.argmax(en == node[:, None], axis=1)
            # b, c: the two other vertices seen from `node`, in the reference's case order.
            ib = np.where(pos == 0, 1, 0)
            ic = np.where(pos == 2, 1, 2)
            pn = xy[node]
            pb = xy[en[np.arange(E), ib]]
            pc = xy[en[np.arange(E), ic]]
            bvec = np.stack([_min_image(pb[:, 0] - pn[:, 0], Lx), _min_image(pb[:, 1] - pn[:, 1], Ly)], 1)
            cvec = np.stack([_min_image(pc[:, 0] - pn[:, 0], Lx), _min_image(pc[:, 1] - pn[:, 1], Ly)], 1)
            cr = np.sum(cvec * cvec, axis=1)
            bx = np.sum(bvec * cvec, axis=1) / cr
            by = (-bvec[:, 0] * cvec[:, 1] + bvec[:, 1] * cvec[:, 0]) / cr
            xx = np.sum(x * cvec, axis=1) / cr
            xy_ = (-x[:, 0] * cvec[:, 1] + x[:, 1] * cvec[:, 0]) / cr
            ab = np.arctan2(by, bx)
            ax = np.arctan2(xy_, xx)
            hit = ((ab > 0) & (ax > 0) & (ax < ab)) | ((ab < 0) & (ax < 0) & (ax > ab)) | \
                  (ab == ax) | (ax == 0)
            hit &= ok
            out[hit, side] = elem[hit]
    if require_all and np.any(out < 0):
        raise ValueError("an edge has no up- or down-wind triangle")
    return out


def _nboundary_lay(nl, N, ednodes, edtri, nlevels) -> np.ndarray:
    """Layer below which a node stops having two wet triangles on every edge
    (restates lib/oce_muscl_adv.F90:96-159; boundary edges force 0)."""
    nbl = np.full(N, nl - 1, dtype=np.int64)
    boundary = np.any(edtri < 0, axis=1)
    lev = np.where(boundary, 0, np.min(np.where(edtri >= 0, nlevels[np.maximum(edtri, 0)], nl), axis=1) - 1)
    for col in range(2):
        np.minimum.at(nbl, ednodes[:, col], lev)
    return nbl.astype(np.int32)


def cavity_ulevels(mesh_nx: int, mesh_ny: int, elem_nlev: np.ndarray, depth: int, radius: float = 0.3) -> np.ndarray:
    """First wet layer of every element (canonical order) under an ice shelf over a disc of the
    given radius (a fraction of the domain) at the domain's centre: 1 + ceil(depth (1 - r/R))
    inside, 1 outside, and at most nlevels - 2 so that two wet layers remain."""
    ii, jj = np.meshgrid(np.arange(mesh_nx), np.arange(mesh_ny), indexing="xy")
    ii = ii.ravel().astype(np.float64)
    jj = jj.ravel().astype(np.float64)
    cx = np.empty(2 * ii.size)
    cy = np.empty(2 * ii.size)
    cx[0::2] = (ii + 2.0 / 3.0) / mesh_nx
    cy[0::2] = (jj + 1.0 / 3.0) / mesh_ny
    cx[1::2] = (ii + 1.0 / 3.0) / mesh_nx
    cy[1::2] = (jj + 2.0 / 3.0) / mesh_ny
    r = np.hypot(cx - 0.5, cy - 0.5) / radius
    ul = 1 + np.where(r < 1.0, np.ceil(depth * (1.0 - r)), 0.0).astype(np.int64)
    return np.minimum(ul, np.asarray(elem_nlev, dtype=np.int64) - 2)


def smooth_bathymetry_levels(mesh_nx: int, mesh_ny: int, nl: int, lo: int, hi: int,
                             seed: int, n_modes: int = 6) -> np.ndarray:
    """Element level counts in [lo, hi] from a seeded sum of periodic Fourier modes, evaluated
    at element centroids in canonical element order (cell-major, lower triangle first)."""
    rng = np.random.default_rng(seed)
    ii, jj = np.meshgrid(np.arange(mesh_nx), np.arange(mesh_ny), indexing="xy")
    ii = ii.ravel().astype(np.float64)
    jj = jj.ravel().astype(np.float64)
    cx = np.empty(2 * ii.size)
    cy = np.empty(2 * ii.size)
    cx[0::2] = (ii + 2.0 / 3.0) / mesh_nx
    cy[0::2] = (jj + 1.0 / 3.0) / mesh_ny
    cx[1::2] = (ii + 1.0 / 3.0) / mesh_nx
    cy[1::2] = (jj + 2.0 / 3.0) / mesh_ny
    f = np.zeros_like(cx)
    for _ in range(n_modes):
        kx = int(rng.integers(-3, 4))
        ky = int(rng.integers(1, 4))
        amp = rng.uniform(0.3, 1.0)
        ph = rng.uniform(0, 2 * math.pi)
        f += amp * np.cos(2 * math.pi * (kx * cx + ky * cy) + ph)
    f = (f - f.min()) / (f.max() - f.min())
    lev = lo + np.floor(f * (hi - lo + 1)).astype(np.int64)
    return np.clip(lev, lo, hi)


def _pad_levels(a: np.ndarray, width: int) -> np.ndarray:
    """Pads the level axis (last) to ``width`` by repeating the deepest value; the padding lies
    below every column's bottom and is never read."""
    if a.shape[-1] == width:
        return a
    pad = [(0, 0)] * (a.ndim - 1) + [(0, width - a.shape[-1])]
    return np.pad(a, pad, mode="edge")


def union_mesh(meshes) -> Mesh:
    """Disjoint union of independent meshes as one mesh (no edge or element joins two of them).

    Node, element and edge numbers of mesh i are offset by the counts of meshes 0..i-1; the level
    axis is padded to the largest nl.  The coordinates place mesh i in cell (i mod G, i div G) of
    a G x G grid of equal cells (G = ceil(sqrt(count)), a cell 1.5 times the largest member's
    extent), so the union's bounding box is about square: the library orders nodes along a
    Hilbert curve scaled to that box per axis, and the 256 / width consecutive nodes of one tile
    then form compact patches whose halos fit the tiled kernels (a strip of 32 meshes along x,
    squeezed 32:1 by the scaling, gave long thin tiles whose halos overflowed nearly every tile at
    width 4, and the whole union ran untiled).  A step of the union is a step of every member mesh:
    the stencils never cross between components.
    """
    meshes = list(meshes)
    nl = max(m.nl for m in meshes)
    max_adj = max(m.max_adj for m in meshes)
    no = eo = to = 0
    ext = np.array([np.ptp(m.coord_nod2D, axis=0) if m.nod2D > 1 else np.zeros(2) for m in meshes])
    cell = ext.max(axis=0) * 1.5 + 1e-3
    cols = int(np.ceil(np.sqrt(len(meshes))))
    parts = {k: [] for k in ("coord", "xy", "enodes", "edges", "etri", "nie", "nie_num", "nlev",
                             "ulev", "nlev_n", "ulev_n", "nlev_min", "ulev_max", "elem_area",
                             "area", "areasvol", "edxdy", "ecd", "ecos", "gsca", "hnode",
                             "hnode_new", "helem", "zbar3", "Z3", "updn", "nbl")}

    def shift(a, off):
        a = np.asarray(a, dtype=np.int64)
        return np.where(a > 0, a + off, 0).astype(np.int32)

    for i, m in enumerate(meshes):
        c = m.coord_nod2D - m.coord_nod2D.min(axis=0)
        c[:, 0] += (i % cols) * cell[0]
        c[:, 1] += (i // cols) * cell[1]
        parts["coord"].append(c)
        if m.xy_m is not None:
            parts["xy"].append(m.xy_m)
        parts["enodes"].append(shift(m.elem2D_nodes, no))
        parts["edges"].append(shift(m.edges, no))
        parts["etri"].append(shift(m.edge_tri, to))
        nie = np.zeros((m.nod2D, max_adj), dtype=np.int32)
        nie[:, :m.max_adj] = shift(m.nod_in_elem2D, to)
        parts["nie"].append(nie)
        parts["nie_num"].append(m.nod_in_elem2D_num)
        parts["nlev"].append(m.nlevels)
        parts["ulev"].append(m.ulevels)
        parts["nlev_n"].append(m.nlevels_nod2D)
        parts["ulev_n"].append(m.ulevels_nod2D)
        parts["nlev_min"].append(m.nlevels_nod2D_min)
        parts["ulev_max"].append(m.ulevels_nod2D_max)
        parts["elem_area"].append(m.elem_area)
        parts["area"].append(_pad_levels(m.area, nl))
        parts["areasvol"].append(_pad_levels(m.areasvol, nl))
        parts["edxdy"].append(m.edge_dxdy)
        parts["ecd"].append(m.edge_cross_dxdy)
        parts["ecos"].append(m.elem_cos)
        parts["gsca"].append(m.gradient_sca)
        parts["hnode"].append(_pad_levels(m.hnode, nl - 1))
        parts["hnode_new"].append(_pad_levels(m.hnode_new, nl - 1))
        parts["helem"].append(_pad_levels(m.helem, nl - 1))
        parts["zbar3"].append(_pad_levels(m.zbar_3d_n, nl))
        parts["Z3"].append(_pad_levels(m.Z_3d_n, nl - 1))
        parts["updn"].append(shift(m.edge_up_dn_tri, to))
        parts["nbl"].append(m.nboundary_lay)
        no += m.nod2D
        eo += m.edge2D
        to += m.elem2D
    deepest = max(meshes, key=lambda m: m.nl)
    cat = {k: np.concatenate(v) for k, v in parts.items() if v}
    return Mesh(nl=nl, nod2D=no, elem2D=to, edge2D=eo, max_adj=max_adj,
                coord_nod2D=cat["coord"], elem2D_nodes=cat["enodes"], edges=cat["edges"],
                edge_tri=cat["etri"], nod_in_elem2D=cat["nie"], nod_in_elem2D_num=cat["nie_num"],
                nlevels=cat["nlev"], ulevels=cat["ulev"], nlevels_nod2D=cat["nlev_n"],
                ulevels_nod2D=cat["ulev_n"], nlevels_nod2D_min=cat["nlev_min"],
                ulevels_nod2D_max=cat["ulev_max"], elem_area=cat["elem_area"], area=cat["area"],
                areasvol=cat["areasvol"], edge_dxdy=cat["edxdy"], edge_cross_dxdy=cat["ecd"],
                elem_cos=cat["ecos"], gradient_sca=cat["gsca"], zbar=deepest.zbar.copy(),
                Z=deepest.Z.copy(), hnode=cat["hnode"], hnode_new=cat["hnode_new"],
                helem=cat["helem"], zbar_3d_n=cat["zbar3"], Z_3d_n=cat["Z3"],
                edge_up_dn_tri=cat["updn"], nboundary_lay=cat["nbl"],
                xy_m=cat.get("xy"))
