"""The dwarf's output formats for Python visualisation (lib/mesh_output.F90:19-162), so that the
reference's own scripts (visualization/plot_mesh.py, plot_scalars.py, plot_compare_precision.py)
read meshes and tracer fields written from this library.

Mesh directory (write_mesh_for_python, lib/mesh_output.F90:19-80):
  mesh_info.txt     ``nod2D=<n>``, ``elem2D=<n>``, ``nl=<n>``, one per line
  coord_nod2D.bin   float64 (2, nod2D) column-major, metres (coord_nod2D in radians x r_earth)
  elem2D_nodes.bin  int32 (3, elem2D) column-major, 0-based node numbers
  zbar.bin          float64 (nl,) level depths
  depth.bin         float64 (nod2D,) node depths
Scalar output (open_scalar_output / write_scalar_step / close_scalar_output, :82-162):
  scalar_info.txt   ``nod2D=``, ``nz=``, ``nsteps=``, ``num_tracers=``, ``wp=`` (the working real
                    kind, 8 for the dp build), ``tracer_<i>=<name>``
  <name>.bin        one float64 (nz, nod2D) column-major frame per write, appended (stream)
The reference names tracers 1 and 2 ``temperature`` and ``salinity`` and has no names beyond two;
here tracer i > 2 is ``tracer<i>``.  A frame of tracer values in this package's (nod2D, nz)
row-major layout has exactly the bytes of the reference's (nz, nod2D) column-major frame.
"""

from __future__ import annotations

import os
from pathlib import Path
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

R_EARTH = 6367500.0  # lib/oce_modules.F90:24


def tracer_names(num_tracers: int) -> List[str]:
    base = ["temperature", "salinity"]
    return [base[i] if i < 2 else f"tracer{i + 1}" for i in range(num_tracers)]


def _write_info(path: Path, items: Sequence[Tuple[str, object]]) -> None:
    # Fortran '(A,I0)' records: no padding, newline-terminated
    path.write_text("".join(f"{k}={v}\n" for k, v in items))


def write_mesh_arrays(output_dir, coord_m: np.ndarray, elem_nodes0: np.ndarray, zbar: np.ndarray,
                      depth: np.ndarray) -> Path:
    """The five mesh files from arrays already in the file's units: coord_m (nod2D, 2) metres,
    elem_nodes0 (elem2D, 3) 0-based, zbar (nl,), depth (nod2D,)."""
    d = Path(output_dir)
    d.mkdir(parents=True, exist_ok=True)
    coord_m = np.asarray(coord_m, dtype="<f8")
    elems = np.asarray(elem_nodes0, dtype="<i4")
    nod2D, elem2D, nl = coord_m.shape[0], elems.shape[0], int(np.asarray(zbar).shape[0])
    _write_info(d / "mesh_info.txt", [("nod2D", nod2D), ("elem2D", elem2D), ("nl", nl)])
    np.ascontiguousarray(coord_m).tofile(d / "coord_nod2D.bin")      # (nod2D, 2) C = (2, nod2D) F
    np.ascontiguousarray(elems).tofile(d / "elem2D_nodes.bin")       # (elem2D, 3) C = (3, elem2D) F
    np.asarray(zbar, dtype="<f8").tofile(d / "zbar.bin")
    np.asarray(depth, dtype="<f8").tofile(d / "depth.bin")
    return d


def node_depth(mesh) -> np.ndarray:
    """Depth of each node's bottom (zbar_3d_n at its deepest interface, <= 0): the dwarf's
    analytic mesh stores -max_depth, its depth for every node, in the same sign convention."""
    return mesh.zbar_3d_n[np.arange(mesh.nod2D), mesh.nlevels_nod2D - 1].astype(np.float64)


def write_mesh_for_python(mesh, output_dir, depth: Optional[np.ndarray] = None) -> Path:
    """write_mesh_for_python (lib/mesh_output.F90:19-80) for a Mesh of this package (whose
    connectivity is 1-based, as the reference's; the file is 0-based)."""
    coord_m = np.asarray(mesh.coord_nod2D, dtype=np.float64) * R_EARTH
    d = node_depth(mesh) if depth is None else depth
    return write_mesh_arrays(output_dir, coord_m, np.asarray(mesh.elem2D_nodes) - 1, mesh.zbar, d)


def _read_info(path: Path) -> Dict[str, str]:
    out = {}
    for line in path.read_text().splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            out[k.strip()] = v.strip()
    return out


def read_mesh_for_python(dirpath) -> Dict[str, np.ndarray]:
    """The mesh directory back as arrays: coord_m (nod2D, 2), elem_nodes0 (elem2D, 3), zbar, depth."""
    d = Path(dirpath)
    info = _read_info(d / "mesh_info.txt")
    nod2D, elem2D, nl = int(info["nod2D"]), int(info["elem2D"]), int(info["nl"])
    coord = np.fromfile(d / "coord_nod2D.bin", dtype="<f8").reshape(nod2D, 2)
    elems = np.fromfile(d / "elem2D_nodes.bin", dtype="<i4").reshape(elem2D, 3)
    zbar = np.fromfile(d / "zbar.bin", dtype="<f8")
    depth = np.fromfile(d / "depth.bin", dtype="<f8")
    if zbar.shape != (nl,) or depth.shape != (nod2D,):
        raise ValueError(f"{d}: file sizes do not match mesh_info.txt")
    return {"coord_m": coord, "elem_nodes0": elems, "zbar": zbar, "depth": depth}


class ScalarOutput:
    """open_scalar_output / write_scalar_step / close_scalar_output (lib/mesh_output.F90:82-162)."""

    def __init__(self, output_dir, nod2D: int, nz: int, nsteps: int, num_tracers: int, wp: int = 8,
                 names: Optional[Sequence[str]] = None):
        self.dir = Path(output_dir)
        self.dir.mkdir(parents=True, exist_ok=True)
        self.nod2D, self.nz, self.num_tracers = nod2D, nz, num_tracers
        self.names = list(names) if names is not None else tracer_names(num_tracers)
        if len(self.names) != num_tracers:
            raise ValueError("one name per tracer")
        items = [("nod2D", nod2D), ("nz", nz), ("nsteps", nsteps), ("num_tracers", num_tracers),
                 ("wp", wp)] + [(f"tracer_{i + 1}", n) for i, n in enumerate(self.names)]
        _write_info(self.dir / "scalar_info.txt", items)
        self._files = [open(self.dir / f"{n}.bin", "wb") for n in self.names]

    def write_step(self, fields: Iterable[np.ndarray]) -> None:
        """One frame per tracer; each field (nod2D, nz) float64 in reference node numbering."""
        fields = list(fields)
        if len(fields) != self.num_tracers:
            raise ValueError(f"{len(fields)} fields for {self.num_tracers} tracers")
        for f, a in zip(self._files, fields):
            a = np.asarray(a, dtype="<f8")
            if a.shape != (self.nod2D, self.nz):
                raise ValueError(f"field shape {a.shape} != {(self.nod2D, self.nz)}")
            f.write(np.ascontiguousarray(a).tobytes())

    def write_tracers(self, tracer) -> None:
        """The current values of every tracer of a FESOMTracerAdvection (copied to the host)."""
        self.write_step(tracer.get_all_values())

    def close(self) -> None:
        for f in self._files:
            f.close()
        self._files = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_scalar_output(dirpath) -> Tuple[Dict[str, str], Dict[str, np.ndarray]]:
    """scalar_info.txt and every tracer's frames as (nframes, nod2D, nz)."""
    d = Path(dirpath)
    info = _read_info(d / "scalar_info.txt")
    nod2D, nz, ntr = int(info["nod2D"]), int(info["nz"]), int(info["num_tracers"])
    out = {}
    for i in range(ntr):
        name = info[f"tracer_{i + 1}"]
        raw = np.fromfile(d / f"{name}.bin", dtype="<f8")
        if raw.size % (nod2D * nz):
            raise ValueError(f"{name}.bin holds a partial frame")
        out[name] = raw.reshape(-1, nod2D, nz)
    return info, out
