"""Writes tests/golden/pi_mesh.npz: the FESOM2 'pi' mesh files of the reference's own test data
(tests/data/pi/*.out: nod2d, elem2d, aux3d standard depths, nlvls, elvls, edgenum, edges,
edge_tri) as arrays, read by fesom2_tracer_dwarf_amd.mesh_io.read_mesh_files.  The files are
mesh data (node coordinates, connectivity, level counts); nothing else is taken.

    python tests/golden/make_pi_mesh.py [/root/reference/tests/data/pi]
"""
import sys
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO))

from fesom2_tracer_dwarf_amd.mesh_io import read_mesh_files  # noqa: E402

src = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/data/pi")
raw = read_mesh_files(src)
out = Path(__file__).resolve().parent / "pi_mesh.npz"
np.savez_compressed(out, **{k: (v.astype(np.int32) if v.dtype.kind == "i" else v) for k, v in raw.items()})
print(out, out.stat().st_size, "bytes")
