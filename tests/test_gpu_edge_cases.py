"""GPU parity at the edges of the input space: two- and three-layer columns (every vertical
scheme's surface, second and bottom rows at once), more tracers than one block holds (64 + 1)
and zero steps."""
import numpy as np
import pytest

from fesom2_tracer_dwarf_amd import configs
from fesom2_tracer_dwarf_amd.configs import Scheme

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def rel_linf(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def run(case, nsteps, kb=0):
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    t = FESOMTracerAdvection()
    t.set_block_width(kb)
    t.setup_case(case)
    t.advect_steps(nsteps, case.dt)
    out = t.get_all_values()
    t.cleanup()
    return out


@pytest.mark.parametrize("nl", [3, 4])
@pytest.mark.parametrize("hor,ver,lim", [("MFCT", "QR4C", "FCT"), ("MUSCL", "PPM", "FCT"), ("UPW1", "CDIFF", "NONE"),
                                         ("MFCT", "QR4C", "NONE"), ("MUSCL", "UPW1", "FCT")])
@pytest.mark.parametrize("kb", [1, 2])
def test_thin_columns_match_oracle(oracle_mod, nl, hor, ver, lim, kb):
    case = configs.bathymetry_case("thin", 8, nl, seed=5, nsteps=3, lev_lo=nl,
                                   scheme=Scheme(hor, ver, lim, 0.8, 0.7), w_cfl=0.3)
    print(f"case thin nl={nl} {hor}/{ver}/{lim} kb={kb}", flush=True)
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    got = run(case, 3, kb)
    assert rel_linf(got, ref[3]) <= RTOL


def test_65_tracers_two_blocks_match_oracle(oracle_mod):
    case = configs.bathymetry_case("t65", 6, 8, seed=8, ntracers=65, nsteps=2, lev_lo=4, w_cfl=0.3)
    print("case 65 tracers (blocks of 64 + 1)", flush=True)
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[2])
    got = run(case, 2)
    assert got.shape == ref[2].shape and rel_linf(got, ref[2]) <= RTOL


def test_zero_steps_keep_values():
    case = configs.bathymetry_case("z", 6, 8, seed=9, nsteps=1, lev_lo=4, w_cfl=0.3)
    print("case zero steps", flush=True)
    got = run(case, 0)
    assert np.array_equal(got, case.tracers)
