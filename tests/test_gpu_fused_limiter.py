"""The fused limiter kernel (k_limit: bounds, ratios, limiting and update in one kernel on node
tiles with two-ring halos, the values alternating between two buffers per step) against the two
kernels it replaces (k_bounds, k_final) bit for bit, and against the oracle (FP64 rel Linf <=
1e-12), through the C ABI: tracer blocks of width 1, 2 and 4, odd and even step counts (the
values end in the primary buffer), level chunks of the other kernels, mixed schemes in one block
(two passes per block and step), the vertical split, graph replays with eager work between them,
the do_oce_adv_tra drop-in (accumulated tendencies, kept fluxes), the reference's pi mesh (coasts,
node degrees up to 8) and the paths it must leave to the two kernels."""
import dataclasses

import numpy as np
import pytest

from fesom2_tracer_dwarf_amd import abi, configs
from fesom2_tracer_dwarf_amd.configs import Scheme

pytestmark = pytest.mark.gpu
RTOL = 1e-12

SCHEMES = [Scheme("MFCT", "QR4C", "FCT", 0.85, 0.75), Scheme("MUSCL", "PPM", "FCT", 0.9, 0.6),
           Scheme("UPW1", "CDIFF", "FCT")]


def say(msg):
    print(msg, flush=True)


def rel_linf(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def run(case, nsteps, kb, fused, lc=(16, 16, 16, 16), graph_chunks=None, wsplit=False):
    """nsteps (or graph replays of graph_chunks steps each) with the fused limiter on or off;
    returns (values, whether the fused limiter ran)."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    t = FESOMTracerAdvection()
    t.set_block_width(kb)
    t.set_launch_config(lc)
    t.set_fused_limiter(fused)
    t.setup_case(case)
    if wsplit:
        t.set_wsplit(True)
    active = t.fused_limiter_active()
    if graph_chunks:
        t.set_graph(True)
        done = 0
        while done < nsteps:
            t.advect_steps(graph_chunks, case.dt)
            t.stats(1)  # eager work between the replays
            done += graph_chunks
        t.set_graph(False)
    else:
        t.advect_steps(nsteps, case.dt)
    out = t.get_all_values()
    t.set_fused_limiter(False)
    t.set_block_width(0)
    t.cleanup()
    return out, active


@pytest.mark.parametrize("nsteps", [3, 4])
@pytest.mark.parametrize("kb", [1, 2, 4])
@pytest.mark.parametrize("si", range(len(SCHEMES)))
def test_fused_limiter_equals_two_kernels_and_oracle(oracle_mod, si, kb, nsteps):
    case = configs.bathymetry_case("bathy48f", 48, 45, seed=81, ntracers=3, nsteps=nsteps, lev_lo=3,
                                   scheme=SCHEMES[si], w_cfl=0.3)
    say(f"case bathy48f scheme {si} kb={kb} {nsteps} steps: fused limiter vs two kernels")
    fused, active = run(case, nsteps, kb, True)
    plain, off = run(case, nsteps, kb, False)
    say(f"  fused limiter active: {active}")
    assert not off
    if kb == 1:
        assert active  # the two rings of these 256-node tiles fit their staging slots
    assert np.array_equal(fused, plain)
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[nsteps])
    assert rel_linf(fused, ref[nsteps]) <= RTOL


def test_fused_limiter_mixed_schemes_and_level_chunks(oracle_mod):
    """Two schemes in one block (two passes of one block per step, both reading one buffer and
    writing the other) and short level chunks in the kernels before the limiter."""
    base = configs.bathymetry_case("bathy40m", 40, 44, seed=83, ntracers=4, nsteps=5, lev_lo=3, w_cfl=0.3)
    case = dataclasses.replace(base, schemes=[SCHEMES[0], SCHEMES[1], SCHEMES[0], SCHEMES[2]])
    say("case bathy40m mixed schemes kb=4 and kb=2, 5 steps, level chunks (8, 5, 7, 3)")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[5])
    for kb in (2, 4):
        fused, active = run(case, 5, kb, True, lc=(8, 5, 7, 3))
        plain, _ = run(case, 5, kb, False, lc=(8, 5, 7, 3))
        say(f"  kb={kb} active {active}: rel Linf {rel_linf(fused, ref[5]):.3e}")
        assert np.array_equal(fused, plain)
        assert rel_linf(fused, ref[5]) <= RTOL


def test_fused_limiter_graph_replays(oracle_mod):
    """Graph replays of 3 steps (an odd count: the captured sequence ends with the copy back to
    the primary buffer) with eager statistics between them equal 9 eager steps."""
    case = configs.bathymetry_case("bathy32g", 32, 30, seed=85, ntracers=2, nsteps=9, lev_lo=4, w_cfl=0.3)
    say("case bathy32g fused limiter, graph of 3 steps replayed 3 times")
    eager, active = run(case, 9, 1, True)
    graph, _ = run(case, 9, 1, True, graph_chunks=3)
    assert active
    assert np.array_equal(eager, graph)
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[9])
    assert rel_linf(graph, ref[9]) <= RTOL


def test_fused_limiter_with_vertical_split(oracle_mod):
    case = configs.split_vertical_velocity(
        configs.bathymetry_case("bathy30s", 30, 28, seed=87, ntracers=2, nsteps=3, lev_lo=4, w_cfl=1.2), 0.5)
    say("case bathy30s use_wsplit, fused limiter vs two kernels")
    fused, active = run(case, 3, 1, True, wsplit=True)
    plain, _ = run(case, 3, 1, False, wsplit=True)
    assert active
    assert np.array_equal(fused, plain)
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    assert rel_linf(fused, ref[3]) <= RTOL


@pytest.fixture(scope="module")
def pi_mesh():
    from pathlib import Path
    from fesom2_tracer_dwarf_amd import mesh_io
    with np.load(Path(__file__).resolve().parent / "golden" / "pi_mesh.npz") as d:
        return mesh_io.sphere_mesh({k: d[k] for k in d.files})


@pytest.mark.parametrize("kb", [1, 2])
def test_fused_limiter_pi_mesh(oracle_mod, pi_mesh, kb):
    """Coasts and nodes of degree 7 and 8 (entries past the six held in registers, for the tile
    nodes and the first-ring nodes)."""
    case = configs.sphere_case("pi", pi_mesh, seed=13, ntracers=2, nsteps=3, scheme=SCHEMES[1], w_cfl=0.3)
    say(f"case pi mesh kb={kb} fused limiter")
    fused, active = run(case, 3, kb, True)
    plain, _ = run(case, 3, kb, False)
    say(f"  active {active}")
    assert np.array_equal(fused, plain)
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    assert rel_linf(fused, ref[3]) <= RTOL


def test_fused_limiter_dropin(oracle_mod):
    """tracer_do_oce_adv_tra with the fused limiter: accumulated tendencies, fct_LO and the kept
    limited fluxes equal the oracle's, for both lanes of a block of two, valuesAB apart from the
    values."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy24d", 24, 22, seed=89, ntracers=2, nsteps=1, lev_lo=4, w_cfl=0.3)
    rng = np.random.default_rng(3)
    o = oracle_mod.Oracle(case)
    t = FESOMTracerAdvection()
    t.set_block_width(1)
    t.set_fused_limiter(True)
    t.setup_case(case)
    assert t.fused_limiter_active()
    say("case bathy24d drop-in with the fused limiter")
    for tid in (1, 2):
        vals = case.tracers[tid - 1]
        vab = vals + 0.01 * rng.standard_normal(vals.shape) * (vals != 0)
        ref = oracle_mod.DropIn(o).run(vals, vab, case.schemes[tid - 1], case.dt)
        t.set_values_ab(tid, vab)
        t.zero_tendencies()
        t.do_oce_adv_tra(case.dt, tid)
        h, v = t.get_tendencies()
        got = {"dttf_h": h, "dttf_v": v, "fct_LO": t.get_work("fct_LO"),
               "adv_flux_hor": t.get_work("adv_flux_hor"), "adv_flux_ver": t.get_work("adv_flux_ver")}
        for k, g in got.items():
            err = rel_linf(g, ref[k])
            say(f"  tracer {tid} {k}: rel Linf {err:.3e}")
            assert err <= RTOL, (tid, k)
    t.set_fused_limiter(False)
    t.set_block_width(0)
    t.cleanup()


def test_fused_limiter_left_to_two_kernels_where_it_does_not_apply():
    """Off by default; AB2 time levels and a pass without the limiter run the two kernels."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy16n", 16, 14, seed=91, ntracers=2, nsteps=1, lev_lo=3, w_cfl=0.3)
    t = FESOMTracerAdvection()
    t.set_block_width(1)
    t.setup_case(case)
    assert not t.fused_limiter_active()  # off by default
    t.set_fused_limiter(True)
    assert t.fused_limiter_active()
    t.set_fused_limiter(False)
    assert not t.fused_limiter_active()
    t.set_fused_limiter(True)
    t.set_ab_mode(2)
    assert not t.fused_limiter_active()
    t.set_ab_mode(0)
    abi.check(t.lib.tracer_set_scheme(2, b"MFCT", b"QR4C", b"NONE", 1.0, 1.0), "tracer_set_scheme")
    assert not t.fused_limiter_active()
    t.set_fused_limiter(False)
    t.set_block_width(0)
    t.cleanup()
