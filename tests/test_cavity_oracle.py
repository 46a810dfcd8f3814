"""Ice-shelf cavities (ulevels > 1): the CPU restatement (oracle/) against the reference Fortran.

The oracle restates every ulevels branch of the advection (segments A/B of the edge fluxes,
lib/oce_adv_tra_hor.F90:134-160; the vertical schemes from ulevels_nod2D, lib/oce_adv_tra_ver.F90;
the gradient prep above ulevels_nod2D_max, lib/oce_muscl_adv.F90:388-424; the limiter,
lib/oce_adv_tra_fct.F90:96-470; the implicit split).  These tests pin it on cavity meshes:
against fixtures written by the compiled reference (tests/golden/cavity.npz, make_golden.py) and,
where oracle/_ref is built, against the live reference on a fresh case.  The GPU library rejects
cavity meshes at set-up (tests/test_abi_exports.py).
"""
import hashlib
import sys
from pathlib import Path

import numpy as np
import pytest

import bundle
from fesom2_tracer_dwarf_amd import configs
from fesom2_tracer_dwarf_amd.configs import Scheme

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden  # noqa: E402

needs_ref = pytest.mark.skipif(not bundle.ref_available(),
                               reason="oracle/_ref not built (reference sources absent)")


def test_cavity_mesh_levels_consistent():
    """A node is wet from the shallowest element top to the deepest element bottom around it; no
    thickness above a top; every column keeps at least two wet layers."""
    case = make_golden.cavity_cases()["cavity_MFCT_QR4C_FCT"]
    m = case.mesh
    assert m.ulevels.max() > 1
    nie = m.nod_in_elem2D - 1
    ok = np.arange(nie.shape[1])[None, :] < m.nod_in_elem2D_num[:, None]
    ul = m.ulevels[np.where(ok, nie, 0)]
    assert np.array_equal(m.ulevels_nod2D, np.where(ok, ul, 1 << 30).min(axis=1))
    assert np.array_equal(m.ulevels_nod2D_max, np.where(ok, ul, 0).max(axis=1))
    assert np.all(m.nlevels - m.ulevels >= 2)
    lay = np.arange(1, m.nl)[None, :]
    assert np.all(m.hnode[lay < m.ulevels_nod2D[:, None]] == 0.0)
    assert np.all(m.helem[lay < m.ulevels[:, None]] == 0.0)
    assert np.all(case.uv[lay < m.ulevels[:, None]] == 0.0)


def test_oracle_equals_reference_fixture_cavity(oracle_mod):
    fx = np.load(GOLDEN / "cavity.npz")
    hashes = []
    for name, case in make_golden.cavity_cases().items():
        hashes.append(make_golden.case_hash(case))
        out, _ = oracle_mod.Oracle(case).run(save_steps=[3])
        assert np.array_equal(out[3], fx[name]), name
    assert str(fx["input_sha256"]) == hashlib.sha256("".join(hashes).encode()).hexdigest()


@needs_ref
@pytest.mark.parametrize("hor,ver,lim", [("MUSCL", "QR4C", "FCT"), ("MFCT", "PPM", "NONE"), ("UPW1", "UPW1", "FCT")])
def test_oracle_equals_live_reference_cavity(oracle_mod, tmp_path, hor, ver, lim):
    case = configs.bathymetry_case("cavlive", 13, 14, seed=23, nsteps=4, lev_lo=7, w_cfl=0.4, cavity=4,
                                   scheme=Scheme(hor, ver, lim, 0.9, 0.6))
    p = bundle.write_bundle(tmp_path / "cav", case, save_steps=[4])
    bundle.run_ref(p)
    out, _ = oracle_mod.Oracle(case).run(save_steps=[4])
    assert np.array_equal(out[4], bundle.read_values(p, case, 4))
