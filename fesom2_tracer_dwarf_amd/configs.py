"""Synthetic workloads: meshes, tracers, velocities and time steps for the benchmark configurations.

A *case* bundles one mesh, its element velocities ``uv(2, nl-1, elem)`` (stored ``(T, nl-1, 2)``),
vertical velocities ``w, w_i, w_e (nl, nod2D)`` (stored ``(N, nl)``), the tracers
``values(nl-1, nod2D)`` (stored ``(ntr, N, nl-1)``), the per-tracer scheme selection and ``dt``.

Workloads (sizes fixed here; seeds make every array reproducible bit for bit):

* ``analytic20``  20x20 periodic cells, 2000 km, 10 levels of 10 m, T Gaussian 0-20, S 34-36,
  psi = A sin sin with max|u| = 0.1 m/s, w = 0, MFCT/QR4C/FCT, CFL 0.5, 10 steps.
* ``mesh1000``    1000x1000 periodic, 48 levels stretched 5-250 m, smooth random bathymetry with
  10-48 levels per element, randomly permuted numbering, velocity decaying as exp(-|z|/500 m),
  2 tracers, 100 steps (FP64).
* ``mesh500x64``  500x500 periodic, 48 levels, bathymetry as above, 64 tracers (T, S and 62 passive
  tracers made of 4 random Gaussian blobs with amplitudes in [0, 1]), 200 steps.
* ``multimesh32`` 32 independent periodic meshes, nx = ny log-uniform in [64, 512], nl uniform in
  [10, 80], psi amplitude U[0.05, 0.5] m/s, own bathymetry, 4 tracers each, 100 steps.
  ``union_case`` advances them together as one disconnected mesh, each member with its own dt
  (a time-step factor per node, stepped with dt = 1).
"""

from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

from .mesh import Mesh, cavity_ulevels, periodic_mesh, smooth_bathymetry_levels, stretched_layers, union_mesh

# Scheme names as the reference spells them (lib/oce_adv_tra_driver.F90:244-282,286).
HOR_SCHEMES = ("UPW1", "MUSCL", "MFCT")
VER_SCHEMES = ("UPW1", "QR4C", "CDIFF", "PPM")
LIM_SCHEMES = ("NONE", "FCT")


@dataclasses.dataclass
class Scheme:
    hor: str = "MFCT"
    ver: str = "QR4C"
    lim: str = "FCT"
    ph: float = 1.0
    pv: float = 1.0

    def __post_init__(self):
        if self.hor not in HOR_SCHEMES:
            raise ValueError(f"unknown horizontal advection type {self.hor!r}")
        if self.ver not in VER_SCHEMES:
            raise ValueError(f"unknown vertical advection type {self.ver!r}")
        if self.lim not in LIM_SCHEMES:
            raise ValueError(f"unknown limiter {self.lim!r}")


@dataclasses.dataclass
class Case:
    name: str
    mesh: Mesh
    uv: np.ndarray          # (T, nl-1, 2)
    w: np.ndarray           # (N, nl)
    wi: np.ndarray
    we: np.ndarray
    tracers: np.ndarray     # (ntr, N, nl-1)
    schemes: List[Scheme]
    dt: float
    nsteps: int
    use_wsplit: bool = False  # dynamics%use_wsplit: w = we + wi, wi advected implicitly
    dt_node: Optional[np.ndarray] = None  # (N) time-step factor per node (None: 1); node n steps dt * dt_node[n]

    @property
    def ntracers(self) -> int:
        return int(self.tracers.shape[0])

    def node_level_updates_per_step(self) -> int:
        return self.ntracers * self.mesh.n_node_levels


def node_xy_m(mesh: Mesh) -> np.ndarray:
    return mesh.xy_m


def _periodic_dist2(mesh: Mesh, x0: float, y0: float) -> np.ndarray:
    xy = mesh.xy_m
    dx = xy[:, 0] - x0
    dy = xy[:, 1] - y0
    dx = dx - mesh.Lx * np.round(dx / mesh.Lx)
    dy = dy - mesh.Ly * np.round(dy / mesh.Ly)
    return dx * dx + dy * dy


def wet_mask(mesh: Mesh) -> np.ndarray:
    """(N, nl-1) True on wet node layers."""
    k = np.arange(1, mesh.nl)[None, :]
    return (k >= mesh.ulevels_nod2D[:, None]) & (k <= (mesh.nlevels_nod2D - 1)[:, None])


def gaussian_tracer(mesh: Mesh, amp: float, sigma: float, x0: float, y0: float,
                    base: float = 0.0) -> np.ndarray:
    """base + amp*exp(-r^2/(2 sigma^2)) in every wet layer, 0 below the bottom."""
    g = base + amp * np.exp(-_periodic_dist2(mesh, x0, y0) / (2.0 * sigma * sigma))
    vals = np.broadcast_to(g[:, None], (mesh.nod2D, mesh.nl - 1)).copy()
    return np.where(wet_mask(mesh), vals, 0.0)


def linear_salinity(mesh: Mesh) -> np.ndarray:
    s = 34.0 + 2.0 * mesh.xy_m[:, 1] / mesh.Ly
    vals = np.broadcast_to(s[:, None], (mesh.nod2D, mesh.nl - 1)).copy()
    return np.where(wet_mask(mesh), vals, 0.0)


def blob_tracer(mesh: Mesh, rng: np.random.Generator, n_blobs: int = 4) -> np.ndarray:
    g = np.zeros(mesh.nod2D)
    for _ in range(n_blobs):
        x0 = rng.uniform(0, mesh.Lx)
        y0 = rng.uniform(0, mesh.Ly)
        sig = rng.uniform(0.04, 0.15) * mesh.Lx
        g += rng.uniform(0.0, 1.0) * np.exp(-_periodic_dist2(mesh, x0, y0) / (2 * sig * sig))
    vals = np.broadcast_to(g[:, None], (mesh.nod2D, mesh.nl - 1)).copy()
    return np.where(wet_mask(mesh), vals, 0.0)


def streamfunction_velocity(mesh: Mesh, umax: float,
                            depth_scale: Optional[float] = None, taper_layers: float = 4.0,
                            psi: Optional[np.ndarray] = None) -> np.ndarray:
    """Element velocities (u, v) = (-d psi/dy, d psi/dx) of the P1 interpolant of a nodal
    streamfunction, by default psi = sin(2 pi x/Lx) sin(2 pi y/Ly) on a periodic mesh, scaled so
    that the largest element speed is ``umax``.

    Layer k uses psi_k = psi * clip((m(n) - k) / taper_layers, 0, 1), where m(n) is the number of
    layers in which every element around node n is wet: psi_k vanishes at every node touching a
    dry element of layer k (the layer's coastline, the no-normal-flow condition) and rises
    smoothly away from it.  The flux through each median-dual segment telescopes to psi
    differences, so the transport of every layer is exactly divergence free on the dual cells,
    also where the bathymetry cuts them: a uniform tracer stays uniform and concentrations stay
    within their initial range with w = 0.  On a mesh without bathymetry (every column full depth)
    psi_k = psi.  ``depth_scale`` multiplies layer k by exp(-|Z_k|/depth_scale).
    """
    if psi is None:
        xy = mesh.xy_m
        psi = np.sin(2 * math.pi * xy[:, 0] / mesh.Lx) * np.sin(2 * math.pi * xy[:, 1] / mesh.Ly)
    en = mesh.elem2D_nodes - 1
    g = mesh.gradient_sca

    def curl(p):
        dpdx = g[:, 0] * p[en[:, 0]] + g[:, 1] * p[en[:, 1]] + g[:, 2] * p[en[:, 2]]
        dpdy = g[:, 3] * p[en[:, 0]] + g[:, 4] * p[en[:, 1]] + g[:, 5] * p[en[:, 2]]
        return -dpdy, dpdx

    fac = np.ones(mesh.nl - 1)
    if depth_scale is not None:
        fac = np.exp(-np.abs(mesh.Z) / depth_scale)
    uv = np.zeros((mesh.elem2D, mesh.nl - 1, 2), dtype=np.float64)
    full = (mesh.nlevels_nod2D_min - 1).astype(np.float64)
    flat = bool(np.all(full == mesh.nl - 1))
    wet_e = mesh.nlevels - 1
    for k in range(mesh.nl - 1):
        pk = psi if flat else psi * np.clip((full - k) / taper_layers, 0.0, 1.0)
        uk, vk = curl(pk)
        on = wet_e >= k + 1
        uv[:, k, 0] = np.where(on, uk * fac[k], 0.0)
        uv[:, k, 1] = np.where(on, vk * fac[k], 0.0)
    uv *= umax / np.max(np.hypot(uv[:, :, 0], uv[:, :, 1]))
    return uv


def split_vertical_velocity(case: Case, cap: float) -> Case:
    """The case with its vertical velocity split as FESOM2's w_split does: the explicit part
    ``we = sign(w) * min(|w|, cap * max|w|)`` and the implicit rest ``wi = w - we``, with
    dynamics%use_wsplit on (0 < cap < 1 leaves the strongest interfaces partly implicit)."""
    w = np.asarray(case.w, dtype=np.float64)
    lim = cap * float(np.abs(w).max())
    we = np.sign(w) * np.minimum(np.abs(w), lim)
    wi = w - we
    return dataclasses.replace(case, we=we, wi=wi, use_wsplit=True)


def cfl_dt(mesh: Mesh, umax: float, cfl: float = 0.5) -> float:
    dx = min(mesh.Lx / mesh.nx, mesh.Ly / mesh.ny)
    return cfl * dx / umax


def analytic20(scheme: Optional[Scheme] = None, nsteps: int = 10) -> Case:
    """20x20 cells, 2000 km square, 10 levels x 10 m, T Gaussian + S linear, psi flow, w = 0."""
    Lx = Ly = 2.0e6
    mesh = periodic_mesh(20, 20, Lx, Ly, np.full(9, 10.0))
    umax = 0.1
    uv = streamfunction_velocity(mesh, umax)
    zeros = np.zeros((mesh.nod2D, mesh.nl))
    T = gaussian_tracer(mesh, 20.0, Lx / 10.0, 0.5 * Lx, 0.5 * Ly)
    S = linear_salinity(mesh)
    sc = scheme or Scheme()
    return Case("analytic20", mesh, uv, zeros, zeros.copy(), zeros.copy(), np.stack([T, S]),
                [sc, dataclasses.replace(sc)], cfl_dt(mesh, umax), nsteps)


def bathymetry_case(name: str, n: int, nl: int, seed: int, ntracers: int = 2,
                    permute: bool = True, nsteps: int = 100, umax: float = 0.1,
                    depth_scale: Optional[float] = 500.0, lev_lo: int = 10,
                    lev_hi: Optional[int] = None, L: float = 2.0e6,
                    scheme: Optional[Scheme] = None, w_cfl: float = 0.0, cavity: int = 0) -> Case:
    """n x n periodic mesh with nl levels stretched 5-250 m and smooth random bathymetry.

    ``w_cfl`` > 0 adds an interior vertical velocity whose largest vertical Courant number
    |w| dt / h_min equals w_cfl (surface and bottom interfaces stay at w = 0).  ``cavity`` > 0
    puts an ice-shelf cavity up to that many layers deep over the domain's centre (ulevels > 1);
    the velocity above an element's top is zero.
    """
    ss = np.random.SeedSequence(seed)
    s_bath, s_perm, s_tr, s_w = ss.spawn(4)
    layer_h = stretched_layers(nl - 1, 5.0, 250.0)
    hi = nl if lev_hi is None else lev_hi
    lev = smooth_bathymetry_levels(n, n, nl, min(lev_lo, hi), hi,
                                   int(s_bath.generate_state(1)[0]))
    ulev = cavity_ulevels(n, n, lev, cavity) if cavity else None
    mesh = periodic_mesh(n, n, L, L, layer_h, elem_nlevels=lev, elem_ulevels=ulev,
                         seed=int(s_perm.generate_state(1)[0]), permute=permute)
    uv = streamfunction_velocity(mesh, umax, depth_scale)
    if cavity:
        dry = np.arange(1, mesh.nl)[None, :] < mesh.ulevels[:, None]
        uv = np.where(dry[..., None] if uv.ndim == 3 and uv.shape[-1] == 2 else dry, 0.0, uv)
    w = np.zeros((mesh.nod2D, mesh.nl))
    dt = cfl_dt(mesh, umax)
    if w_cfl:
        w_amp = w_cfl * float(layer_h.min()) / dt
        # Interior-only vertical velocity (surface and bottom interfaces stay 0), smooth in x, y.
        rng = np.random.default_rng(s_w)
        xy = mesh.xy_m
        ph = rng.uniform(0, 2 * math.pi, 2)
        horiz = np.sin(2 * math.pi * xy[:, 0] / L + ph[0]) * np.cos(2 * math.pi * xy[:, 1] / L + ph[1])
        k = np.arange(1, mesh.nl + 1)[None, :]
        interior = (k > mesh.ulevels_nod2D[:, None]) & (k < mesh.nlevels_nod2D[:, None])
        prof = np.sin(math.pi * (k - 1) / (mesh.nl - 1))
        w = np.where(interior, w_amp * horiz[:, None] * prof, 0.0)
    rng = np.random.default_rng(s_tr)
    tr = [gaussian_tracer(mesh, 20.0, L / 10.0, 0.5 * L, 0.5 * L), linear_salinity(mesh)]
    while len(tr) < ntracers:
        tr.append(blob_tracer(mesh, rng))
    tracers = np.stack(tr[:ntracers])
    sc = scheme or Scheme()
    return Case(name, mesh, uv, w, w.copy(), w.copy(), tracers,
                [dataclasses.replace(sc) for _ in range(ntracers)], dt, nsteps)


def sphere_case(name: str, mesh: Mesh, seed: int = 0, ntracers: int = 2, nsteps: int = 10,
                umax: float = 0.1, scheme: Optional[Scheme] = None, w_cfl: float = 0.0,
                depth_scale: Optional[float] = 1000.0) -> Case:
    """A case on a mesh of the sphere (a FESOM2 file mesh, see mesh_io.sphere_mesh).

    Flow: the streamfunction psi = sin(2 lon) cos(lat)^2 (model coordinates), set to zero on the
    coastline (the nodes of boundary edges) so that no transport crosses it, layered as in
    :func:`streamfunction_velocity`.  T: a Gaussian blob of width 0.3 rad around a seeded ocean
    node on 10; S: 34 + 2 sin(lat); further tracers: seeded blobs.  dt: half the smallest edge
    length over ``umax``.  ``w_cfl`` adds an interior vertical velocity as in
    :func:`bathymetry_case`.
    """
    from .mesh_io import edge_lengths
    rng = np.random.default_rng(seed)
    lon, lat = mesh.coord_nod2D[:, 0], mesh.coord_nod2D[:, 1]
    coast = np.zeros(mesh.nod2D, dtype=bool)
    coast[(mesh.edges[mesh.edge_tri[:, 1] == 0] - 1).ravel()] = True
    psi = np.where(coast, 0.0, np.sin(2.0 * lon) * np.cos(lat) ** 2)
    uv = streamfunction_velocity(mesh, umax, depth_scale, psi=psi)
    dt = 0.5 * float(edge_lengths(mesh).min()) / umax
    w = np.zeros((mesh.nod2D, mesh.nl))
    if w_cfl:
        h_min = float(np.min(mesh.zbar[:-1] - mesh.zbar[1:]))
        ph = rng.uniform(0, 2 * math.pi, 2)
        horiz = np.sin(lon + ph[0]) * np.cos(2.0 * lat + ph[1])
        k = np.arange(1, mesh.nl + 1)[None, :]
        interior = (k > mesh.ulevels_nod2D[:, None]) & (k < mesh.nlevels_nod2D[:, None])
        prof = np.sin(math.pi * (k - 1) / (mesh.nl - 1))
        w = np.where(interior, w_cfl * h_min / dt * horiz[:, None] * prof, 0.0)
    wet = wet_mask(mesh)

    def blob(n0, sigma, amp, base):
        c = math.cos(lat[n0]) * np.cos(lat) * np.cos(lon - lon[n0]) + math.sin(lat[n0]) * np.sin(lat)
        d = np.arccos(np.clip(c, -1.0, 1.0))
        g = base + amp * np.exp(-d * d / (2.0 * sigma * sigma))
        return np.where(wet, np.broadcast_to(g[:, None], wet.shape), 0.0)

    tr = [blob(int(rng.integers(mesh.nod2D)), 0.3, 10.0, 10.0),
          np.where(wet, np.broadcast_to((34.0 + 2.0 * np.sin(lat))[:, None], wet.shape), 0.0)]
    while len(tr) < ntracers:
        tr.append(blob(int(rng.integers(mesh.nod2D)), rng.uniform(0.1, 0.5), rng.uniform(0.0, 1.0), 0.0))
    sc = scheme or Scheme()
    return Case(name, mesh, uv, w, w.copy(), w.copy(), np.stack(tr[:ntracers]),
                [dataclasses.replace(sc) for _ in range(ntracers)], dt, nsteps)


def mesh1000(nsteps: int = 100) -> Case:
    return bathymetry_case("mesh1000", 1000, 48, seed=1, ntracers=2, nsteps=nsteps)


def mesh500x64(nsteps: int = 200) -> Case:
    return bathymetry_case("mesh500x64", 500, 48, seed=3, ntracers=64, nsteps=nsteps)


def multimesh32(nsteps: int = 100, count: int = 32, seed: int = 4,
                size_range: Sequence[int] = (64, 512)) -> List[Case]:
    rng = np.random.default_rng(seed)
    cases = []
    lo, hi = size_range
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        nl = int(rng.integers(10, 81))
        umax = float(rng.uniform(0.05, 0.5))
        cases.append(bathymetry_case(f"multimesh32[{i}]", n, nl, seed=1000 + i, ntracers=4,
                                     nsteps=nsteps, umax=umax, lev_lo=min(10, nl), lev_hi=nl))
    return cases


def ensemble_cases(base: Case, members: int, seed: int = 0, wsplit_cap: Optional[float] = None) -> List[Case]:
    """``members`` cases on base's mesh object with base's time step and schemes, each with its own
    velocity and tracers (an ensemble; member 0 is base).  Member k > 0: the streamfunction
    sin(2 pi x/Lx + a) sin(2 pi y/Ly + b) with seeded phases, its speed scaled by a seeded factor
    in [0.5, 1] of base's largest speed (so base's dt stays within its CFL), the layering of
    :func:`streamfunction_velocity`; a vertical velocity where base has one, with seeded phases
    and amplitude up to base's largest |w|, on the same interior interfaces; base's tracers scaled
    by (1 + k / 10).  ``wsplit_cap`` splits every member's w as :func:`split_vertical_velocity`."""
    mesh = base.mesh
    rng = np.random.default_rng(np.random.SeedSequence(seed).spawn(1)[0])
    umax = float(np.max(np.hypot(base.uv[..., 0], base.uv[..., 1])))
    wmax = float(np.abs(base.w).max())
    xy = mesh.xy_m
    out = [base]
    for k in range(1, members):
        a, b = rng.uniform(0.0, 2.0 * math.pi, 2)
        psi = np.sin(2 * math.pi * xy[:, 0] / mesh.Lx + a) * np.sin(2 * math.pi * xy[:, 1] / mesh.Ly + b)
        scale = rng.uniform(0.5, 1.0)
        uv = streamfunction_velocity(mesh, umax * scale, 500.0, psi=psi)
        if mesh.ulevels.max() > 1:  # no transport above an element's top (a cavity roof)
            dry = np.arange(1, mesh.nl)[None, :] < mesh.ulevels[:, None]
            uv = np.where(dry[..., None], 0.0, uv)
        w = np.zeros_like(base.w)
        if wmax > 0.0:
            c, d = rng.uniform(0.0, 2.0 * math.pi, 2)
            horiz = np.sin(2 * math.pi * xy[:, 0] / mesh.Lx + c) * np.cos(2 * math.pi * xy[:, 1] / mesh.Ly + d)
            lev = np.arange(1, mesh.nl + 1)[None, :]
            interior = (lev > mesh.ulevels_nod2D[:, None]) & (lev < mesh.nlevels_nod2D[:, None])
            prof = np.sin(math.pi * (lev - 1) / (mesh.nl - 1))
            w = np.where(interior, wmax * rng.uniform(0.5, 1.0) * horiz[:, None] * prof, 0.0)
        c_k = Case(f"{base.name}_m{k}", mesh, uv, w, w.copy(), w.copy(), base.tracers * (1.0 + 0.1 * k),
                   [dataclasses.replace(sc) for sc in base.schemes], base.dt, base.nsteps)
        out.append(c_k)
    if wsplit_cap is not None:
        out = [split_vertical_velocity(c, wsplit_cap) for c in out]
    return out


def union_case(cases: Sequence[Case], name: str = "union", own_dt: bool = True) -> Case:
    """Independent cases with the same tracer count and schemes as one case on the disjoint union
    of their meshes.  own_dt: every member keeps its own time step (``dt_node`` holds the member's
    dt at each of its nodes and the union steps with dt = 1, so node n advances with exactly its
    member's dt); otherwise the union steps with the smallest member dt (stable for every member)."""
    cases = list(cases)
    ntr = cases[0].ntracers
    if any(c.ntracers != ntr for c in cases):
        raise ValueError("union_case: every member needs the same number of tracers")
    for c in cases[1:]:
        if [dataclasses.astuple(a) for a in c.schemes] != [dataclasses.astuple(a) for a in cases[0].schemes]:
            raise ValueError(f"union_case: member {c.name} has other schemes than {cases[0].name}; "
                             "the tracers of a union share their schemes")
        if c.use_wsplit != cases[0].use_wsplit:
            raise ValueError("union_case: every member needs the same use_wsplit")
    mesh = union_mesh([c.mesh for c in cases])
    L = mesh.nl - 1

    uv = np.concatenate([np.pad(c.uv, ((0, 0), (0, L - c.uv.shape[1]), (0, 0))) for c in cases])
    w = np.concatenate([np.pad(c.w, ((0, 0), (0, mesh.nl - c.w.shape[1]))) for c in cases])
    wi = np.concatenate([np.pad(c.wi, ((0, 0), (0, mesh.nl - c.wi.shape[1]))) for c in cases])
    we = np.concatenate([np.pad(c.we, ((0, 0), (0, mesh.nl - c.we.shape[1]))) for c in cases])
    tr = np.concatenate([np.pad(c.tracers, ((0, 0), (0, 0), (0, L - c.tracers.shape[2])))
                         for c in cases], axis=1)
    schemes = [dataclasses.replace(s) for s in cases[0].schemes]
    if own_dt:
        dt_node = np.concatenate([np.full(c.mesh.nod2D, float(c.dt)) for c in cases])
        return Case(name, mesh, uv, w, wi, we, tr, schemes, 1.0, cases[0].nsteps,
                    use_wsplit=cases[0].use_wsplit, dt_node=dt_node)
    return Case(name, mesh, uv, w, wi, we, tr, schemes, min(c.dt for c in cases), cases[0].nsteps,
                use_wsplit=cases[0].use_wsplit)


def multimesh32_union(nsteps: int = 100, **kw) -> Case:
    return union_case(multimesh32(nsteps, **kw), "multimesh32")
