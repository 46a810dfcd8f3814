"""Case bundles for oracle/_ref/ref_driver (test infrastructure only).

A bundle is a directory of raw little-endian arrays in Fortran order plus ``dims.txt``; the
reference driver reads it verbatim, so the reference kernels see exactly the arrays the GPU path
and the C restatement see.  Only tests/ and the golden-fixture script use this module.
"""

from __future__ import annotations

import os
import subprocess
from pathlib import Path
from typing import Dict, Optional, Sequence

import numpy as np

REPO = Path(__file__).resolve().parents[2]
REF_DIR = REPO / "oracle" / "_ref"
REF_DRIVER = REF_DIR / "ref_driver"
REF_ANALYTIC = REF_DIR / "fesom_tracer_analytic"

_INT_ARRAYS = ("elem2D_nodes", "edges", "edge_tri", "nod_in_elem2D", "nod_in_elem2D_num",
               "nlevels", "ulevels", "nlevels_nod2D", "ulevels_nod2D", "nlevels_nod2D_min",
               "ulevels_nod2D_max", "edge_up_dn_tri", "nboundary_lay")
_REAL_ARRAYS = ("elem_area", "area", "areasvol", "edge_dxdy", "edge_cross_dxdy", "elem_cos",
                "gradient_sca", "hnode", "hnode_new", "helem", "zbar_3d_n", "Z_3d_n")


def ref_available() -> bool:
    return REF_DRIVER.exists()


def write_bundle(path, case, save_steps: Sequence[int] = (), save_work: bool = False,
                 timing: bool = False, nsteps: Optional[int] = None, ab_mode: int = 0) -> Path:
    """ab_mode: 0 valuesAB := values every step, 1 frozen at the initial values, 2 FESOM2's AB2
    interpolation; dynamics%use_wsplit is taken from case.use_wsplit (default off)."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    m = case.mesh
    ns = case.nsteps if nsteps is None else nsteps
    lines = [f"{m.nl} {m.nod2D} {m.elem2D} {m.edge2D} {m.max_adj}",
             f"{case.ntracers} {ns} {case.dt!r}",
             f"{int(save_work)} {int(timing)} {int(bool(getattr(case, 'use_wsplit', False)))} {int(ab_mode)}",
             f"{len(save_steps)}"]
    if save_steps:
        lines.append(" ".join(str(int(s)) for s in save_steps))
    for sc in case.schemes:
        lines.append(f"'{sc.hor}' '{sc.ver}' '{sc.lim}' {float(sc.ph)!r} {float(sc.pv)!r}")
    (path / "dims.txt").write_text("\n".join(lines) + "\n")
    for name in _INT_ARRAYS:
        np.ascontiguousarray(getattr(m, name), dtype="<i4").tofile(path / f"{name}.bin")
    for name in _REAL_ARRAYS:
        np.ascontiguousarray(getattr(m, name), dtype="<f8").tofile(path / f"{name}.bin")
    np.ascontiguousarray(case.uv, dtype="<f8").tofile(path / "uv.bin")
    for name, arr in (("w", case.w), ("wi", case.wi), ("we", case.we)):
        np.ascontiguousarray(arr, dtype="<f8").tofile(path / f"{name}.bin")
    np.ascontiguousarray(case.tracers, dtype="<f8").tofile(path / "values.bin")
    return path


def run_ref(path, timeout: float = 600.0, threads: int = 1) -> subprocess.CompletedProcess:
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    return subprocess.run([str(REF_DRIVER), str(path)], capture_output=True, text=True,
                          timeout=timeout, env=env, check=True)


def read_values(path, case, step: int) -> np.ndarray:
    m = case.mesh
    a = np.fromfile(Path(path) / f"out_values_s{step:04d}.bin", dtype="<f8")
    return a.reshape(case.ntracers, m.nod2D, m.nl - 1)


def read_work(path, case) -> Dict[str, np.ndarray]:
    m = case.mesh
    L = m.nl - 1
    shapes = {"edge_up_dn_grad": (m.edge2D, L, 4), "fct_LO": (m.nod2D, L),
              "adv_flux_hor": (m.edge2D, L), "adv_flux_ver": (m.nod2D, m.nl),
              "dttf_h": (m.nod2D, L), "dttf_v": (m.nod2D, L)}
    out = {}
    for k, shp in shapes.items():
        f = Path(path) / f"out_w_{k}.bin"
        if f.exists():
            out[k] = np.fromfile(f, dtype="<f8").reshape(shp)
    return out


def read_timing(path) -> Dict[str, float]:
    out = {}
    for line in (Path(path) / "out_timing.txt").read_text().splitlines():
        k, v = line.split()
        out[k] = float(v)
    return out
