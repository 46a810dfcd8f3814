"""Ensemble members (tracer_set_ensemble): several velocity fields on one mesh, each tracer advected
by its member's velocity, the members of a tracer block sharing every mesh read.  Each member's
tracers are checked against the oracle run of that member alone (its own velocity and tracers,
FP64: rel Linf <= 1e-12, bitwise in practice).  Block layouts: tiled and untiled kernels, blocks
that hold one member and blocks whose lanes belong to different members, the implicit vertical
split, the drop-in on a member's tracer, and a velocity change of one member between calls.

The reference advects with one velocity per model instance (do_oce_adv_tra takes vel, w, w_i, w_e
per call, lib/oce_adv_tra_driver.F90:23-45); an ensemble member here is what one such instance
computes.
"""
import numpy as np
import pytest

from fesom2_tracer_dwarf_amd import configs
from fesom2_tracer_dwarf_amd.configs import Scheme

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def say(msg):
    print(msg, flush=True)


def rel_linf(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def members_case(nmem, scheme=None, wsplit_cap=None, nsteps=5):
    base = configs.bathymetry_case("ens14", 14, 16, seed=71, ntracers=2, nsteps=nsteps, lev_lo=4,
                                   w_cfl=0.3, scheme=scheme)
    return configs.ensemble_cases(base, nmem, seed=5, wsplit_cap=wsplit_cap)


def interleaved(cases):
    """Library order (member-major slot index per tracer id) that puts tracer i of every member
    next to each other, so a block's lanes belong to different members."""
    slots = [(k, i) for k, c in enumerate(cases) for i in range(c.ntracers)]
    pos = {s: j for j, s in enumerate(slots)}
    return [pos[(k, i)] for i in range(cases[0].ntracers) for k in range(len(cases))]


def check_members(t, cases, oracle_mod, nsteps, label):
    for k, c in enumerate(cases):
        ref, _ = oracle_mod.Oracle(c).run(nsteps=nsteps, save_steps=[nsteps])
        for i, tid in enumerate(t.member_tracers[k]):
            got = t.get_tracer_values(tid)
            err = rel_linf(got, ref[nsteps][i])
            say(f"  {label} member {k + 1} tracer {i + 1} (id {tid}): rel Linf {err:.3e} "
                f"bitwise {np.array_equal(got, ref[nsteps][i])}")
            assert err <= RTOL, (k, i)


@pytest.mark.parametrize("kb,tiles,mixed", [(1, True, False), (2, True, True), (2, True, False),
                                             (4, False, True), (8, True, True), (64, True, True)])
def test_ensemble_members_match_their_oracle_runs(oracle_mod, kb, tiles, mixed):
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    nsteps = 5
    cases = members_case(3, nsteps=nsteps)
    say(f"case ens14 3 members x 2 tracers MFCT/QR4C/FCT, width {kb}, tiles {tiles}, "
        f"mixed blocks {mixed}")
    t = FESOMTracerAdvection()
    t.set_block_width(kb)
    t.set_halo_tiles(tiles)
    t.set_launch_config((4, 4, 4, 4))
    t.setup_ensemble(cases, order=interleaved(cases) if mixed else None)
    assert t.ensemble() == 3
    t.advect_steps(nsteps, cases[0].dt)
    check_members(t, cases, oracle_mod, nsteps, f"width {kb}")
    t.cleanup()


@pytest.mark.parametrize("hor,ver,wsplit", [("MUSCL", "PPM", False), ("UPW1", "UPW1", False),
                                             ("MFCT", "QR4C", True), ("MFCT", "CDIFF", False)])
def test_ensemble_schemes_and_split(oracle_mod, hor, ver, wsplit):
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    nsteps = 4
    sc = Scheme(hor=hor, ver=ver, lim="FCT")
    cases = members_case(4, scheme=sc, wsplit_cap=0.5 if wsplit else None, nsteps=nsteps)
    say(f"case ens14 4 members {hor}/{ver}/FCT wsplit {wsplit}")
    for kb in (2, 64):
        t = FESOMTracerAdvection()
        t.set_block_width(kb)
        t.setup_ensemble(cases, order=interleaved(cases))
        t.advect_steps(nsteps, cases[0].dt)
        check_members(t, cases, oracle_mod, nsteps, f"width {kb}")
        t.cleanup()


def test_ensemble_without_limiter_and_graph(oracle_mod):
    """No limiter (the vertical-flux kernel), replayed from a captured graph."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    nsteps = 4
    cases = members_case(2, scheme=Scheme(hor="MFCT", ver="QR4C", lim="NONE"), nsteps=nsteps)
    say("case ens14 2 members MFCT/QR4C/NONE, graph")
    t = FESOMTracerAdvection()
    t.set_graph(True)
    t.setup_ensemble(cases, order=interleaved(cases))
    t.advect_steps(2, cases[0].dt)
    t.advect_steps(2, cases[0].dt)
    check_members(t, cases, oracle_mod, nsteps, "graph")
    t.set_graph(False)
    t.cleanup()


def test_ensemble_dropin_and_velocity_change(oracle_mod):
    """tracer_do_oce_adv_tra on a tracer of member 2 uses member 2's velocity; a new velocity of
    member 2 reaches the next step (and member 1 keeps its own)."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    cases = members_case(3, nsteps=1)
    say("case ens14 3 members drop-in on member 2, then a new velocity for member 2")
    t = FESOMTracerAdvection()
    t.setup_ensemble(cases, order=interleaved(cases))
    tid = t.member_tracers[1][0]
    c = cases[1]
    vals = c.tracers[0]
    ref = oracle_mod.DropIn(oracle_mod.Oracle(c)).run(vals, vals, c.schemes[0], c.dt)
    t.zero_tendencies()
    t.do_oce_adv_tra(c.dt, tid)
    h, v = t.get_tendencies()
    for k, g in {"dttf_h": h, "dttf_v": v}.items():
        err = rel_linf(g, ref[k])
        say(f"  drop-in {k}: rel Linf {err:.3e}")
        assert err <= RTOL, k
    # member 2 takes member 3's velocity: its tracers then follow the oracle of that velocity
    import dataclasses
    new2 = dataclasses.replace(cases[1], uv=cases[2].uv, w=cases[2].w, wi=cases[2].wi, we=cases[2].we)
    t.set_velocity_member(2, new2.uv, new2.w, new2.wi, new2.we)
    t.advect_steps(2, c.dt)
    check_members(t, [cases[0], new2, cases[2]], oracle_mod, 2, "after the change")
    t.cleanup()


def test_ensemble_conserves_and_differs():
    """Every member conserves its volume sum to FP64 round-off and the members differ."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    base = configs.bathymetry_case("ens24", 24, 20, seed=9, ntracers=2, nsteps=20, lev_lo=6)
    cases = configs.ensemble_cases(base, 8, seed=2)
    say("case ens24 8 members x 2 tracers conservation")
    t = FESOMTracerAdvection()
    t.set_block_width(16)
    t.setup_ensemble(cases, order=interleaved(cases))
    ids = [tid for m in t.member_tracers for tid in m]
    v0 = {tid: t.stats(tid)["volume_sum"] for tid in ids}
    t.advect_steps(20, base.dt)
    for tid in ids:
        v1 = t.stats(tid)["volume_sum"]
        assert abs(v1 - v0[tid]) <= 1e-13 * abs(v0[tid]), (tid, v0[tid], v1)
    a = t.get_tracer_values(t.member_tracers[1][0]) / 1.1
    b = t.get_tracer_values(t.member_tracers[2][0]) / 1.2
    assert rel_linf(a, b) > 1e-6
    t.cleanup()
