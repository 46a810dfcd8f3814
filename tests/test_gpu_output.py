"""Tracer frames written from the library in the dwarf's scalar-output format
(lib/mesh_output.F90:82-162) read back as the oracle's values after each step."""
import numpy as np
import pytest

from fesom2_tracer_dwarf_amd import configs, output

pytestmark = pytest.mark.gpu


def test_scalar_frames_from_library_match_oracle(oracle_mod, tmp_path):
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("out", 10, 12, seed=61, nsteps=3, lev_lo=4, w_cfl=0.3)
    print("case scalar output 3 steps", flush=True)
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[1, 2, 3])
    m = case.mesh
    t = FESOMTracerAdvection()
    t.setup_case(case)
    output.write_mesh_for_python(m, tmp_path)
    with output.ScalarOutput(tmp_path, m.nod2D, m.nl - 1, 3, case.ntracers) as so:
        so.write_tracers(t)
        for _ in range(3):
            t.advect_steps(1, case.dt)
            so.write_tracers(t)
    t.cleanup()
    info, frames = output.read_scalar_output(tmp_path)
    assert frames["temperature"].shape == (4, m.nod2D, m.nl - 1)
    assert np.array_equal(frames["temperature"][0], case.tracers[0])
    for k in (1, 2, 3):
        got = np.stack([frames["temperature"][k], frames["salinity"][k]])
        assert np.array_equal(got, ref[k]), k
