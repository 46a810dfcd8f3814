"""The do_oce_adv_tra drop-in through device pointers, the per-velocity edge transport, the
work-array read-back rules, stochastic rounding in captured graphs and per-member time steps of a
mesh union -- all through the C ABI against the oracle (FP64: rel Linf <= 1e-12) or against the
library's own eager path (bit for bit)."""
import dataclasses

import numpy as np
import pytest

from fesom2_tracer_dwarf_amd import abi, configs
from fesom2_tracer_dwarf_amd.configs import Scheme

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def say(msg):
    print(msg, flush=True)


def rel_linf(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def dropin_case():
    return configs.bathymetry_case("bathy16d", 16, 14, seed=5, nsteps=1, lev_lo=5, w_cfl=0.3)


def test_dropin_device_pointers_match_oracle(oracle_mod):
    """tracer_set_values_ab_device / tracer_get_tendencies_device (PyTorch tensors in reference
    order) give the oracle's tendencies for both lanes of a two-tracer block, with T's and S's
    calls sharing one edge transport."""
    import torch
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = dropin_case()
    rng = np.random.default_rng(1)
    o = oracle_mod.Oracle(case)
    t = FESOMTracerAdvection()
    t.setup_case(case)
    say("case bathy16d drop-in via device pointers, tracers 1 and 2")
    for tid in (1, 2):
        vals = case.tracers[tid - 1]
        vab = vals + 0.01 * rng.standard_normal(vals.shape) * (vals != 0)
        ref = oracle_mod.DropIn(o).run(vals, vab, case.schemes[tid - 1], case.dt)
        t.set_values_ab_tensor(tid, torch.from_numpy(vab).cuda())
        t.zero_tendencies()
        t.do_oce_adv_tra(case.dt, tid)
        h, v = t.get_tendencies_tensor()
        h2, v2 = t.get_tendencies()
        assert np.array_equal(h.cpu().numpy(), h2) and np.array_equal(v.cpu().numpy(), v2)
        for k, g in (("dttf_h", h2), ("dttf_v", v2)):
            err = rel_linf(g, ref[k])
            say(f"  tracer {tid} {k}: rel Linf {err:.3e}")
            assert err <= RTOL, (tid, k)
    t.cleanup()


def test_zeroed_tendencies_and_accumulation(oracle_mod):
    """tracer_zero_tendencies defers the zeroing to the next call (whose update writes its sums):
    a read-back right after it gives zeros, a call after it gives the oracle's tendencies, and a
    second call without zeroing adds its tendencies to the first."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = dropin_case()
    o = oracle_mod.Oracle(case)
    t = FESOMTracerAdvection()
    t.set_block_width(1)
    t.setup_case(case)
    say("case bathy16d zeroed tendencies, then one and two accumulating calls")
    vals = case.tracers[0]
    ref = oracle_mod.DropIn(o).run(vals, vals, case.schemes[0], case.dt)
    t.zero_tendencies()
    t.do_oce_adv_tra(case.dt, 1)
    h, v = t.get_tendencies()
    assert rel_linf(h, ref["dttf_h"]) <= RTOL and rel_linf(v, ref["dttf_v"]) <= RTOL
    t.zero_tendencies()
    h0, v0 = t.get_tendencies()
    assert not np.any(h0) and not np.any(v0)
    t.do_oce_adv_tra(case.dt, 1)
    t.do_oce_adv_tra(case.dt, 1)
    h2, v2 = t.get_tendencies()
    assert rel_linf(h2, 2 * ref["dttf_h"]) <= RTOL and rel_linf(v2, 2 * ref["dttf_v"]) <= RTOL
    t.zero_tendencies()
    t.do_oce_adv_tra(case.dt, 1)
    h1, v1 = t.get_tendencies()
    assert np.array_equal(h1, h) and np.array_equal(v1, v)
    t.set_block_width(0)
    t.cleanup()


def test_dropin_follows_a_new_velocity(oracle_mod):
    """The drop-in reuses the edge transport between calls with the same velocity and recomputes
    it after tracer_set_velocity_elem: a call after a velocity change gives the oracle's result for
    the new velocity."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = dropin_case()
    case2 = dataclasses.replace(case, uv=case.uv * -0.7, w=case.w * 0.5, wi=case.wi * 0.5, we=case.we * 0.5)
    say("case bathy16d drop-in, velocity changed between calls")
    t = FESOMTracerAdvection()
    t.setup_case(case)
    vals = case.tracers[0]
    for c in (case, case2, case):
        t.set_velocity(c.uv, c.w, c.wi, c.we)
        ref = oracle_mod.DropIn(oracle_mod.Oracle(c)).run(vals, vals, c.schemes[0], c.dt)
        for _ in range(2):  # a repeated call reuses the transport
            t.zero_tendencies()
            t.do_oce_adv_tra(c.dt, 1)
            h, v = t.get_tendencies()
            assert rel_linf(h, ref["dttf_h"]) <= RTOL and rel_linf(v, ref["dttf_v"]) <= RTOL
        assert rel_linf(t.get_work("fct_LO"), ref["fct_LO"]) <= RTOL
    t.cleanup()


def test_work_arrays_refused_after_advect():
    """tracer_get_work serves the last drop-in call only; after tracer_advect_steps it returns -1."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = dropin_case()
    t = FESOMTracerAdvection()
    t.setup_case(case)
    with pytest.raises(abi.TracerError):
        t.get_work("fct_LO")
    t.zero_tendencies()
    t.do_oce_adv_tra(case.dt, 1)
    t.get_work("fct_LO")
    t.advect_steps(1, case.dt)
    with pytest.raises(abi.TracerError):
        t.get_work("fct_LO")
    with pytest.raises(abi.TracerError):
        t.get_work("adv_flux_hor")
    t.cleanup()


def test_dropin_without_kept_work_gives_the_same_tendencies(oracle_mod):
    """tracer_set_keep_work(0): the drop-in skips the limited-flux stores; its tendencies and
    fct_LO equal the oracle's as before, and tracer_get_work refuses the flux arrays it did not
    keep until a call keeps them again."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = dropin_case()
    o = oracle_mod.Oracle(case)
    t = FESOMTracerAdvection()
    t.setup_case(case)
    say("case bathy16d drop-in with tracer_set_keep_work(0), tracers 1 and 2")
    for tid in (1, 2):
        vals = case.tracers[tid - 1]
        ref = oracle_mod.DropIn(o).run(vals, vals, case.schemes[tid - 1], case.dt)
        for keep in (False, True):
            t.set_keep_work(keep)
            t.zero_tendencies()
            t.do_oce_adv_tra(case.dt, tid)
            h, v = t.get_tendencies()
            assert rel_linf(h, ref["dttf_h"]) <= RTOL and rel_linf(v, ref["dttf_v"]) <= RTOL, (tid, keep)
            assert rel_linf(t.get_work("fct_LO"), ref["fct_LO"]) <= RTOL
            if keep:
                assert rel_linf(t.get_work("adv_flux_hor"), ref["adv_flux_hor"]) <= RTOL
                assert rel_linf(t.get_work("adv_flux_ver"), ref["adv_flux_ver"]) <= RTOL
            else:
                for name in ("adv_flux_hor", "adv_flux_ver"):
                    with pytest.raises(abi.TracerError):
                        t.get_work(name)
    t.cleanup()


@pytest.mark.parametrize("prec", [1, 2])
def test_stochastic_rounding_graph_replays_equal_eager_steps(prec):
    """Single-precision storage with stochastic rounding: 1-step calls replayed from a captured
    graph give bit for bit what eager 1-step calls give (the step counter that seeds the rounding
    lives in device memory and advances at every replay)."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy12s", 12, 12, seed=23, ntracers=3, nsteps=6, lev_lo=4, w_cfl=0.3)
    say(f"case bathy12s precision {prec} graph vs eager, 6 one-step calls")
    out = {}
    for graph in (False, True):
        t = FESOMTracerAdvection()
        t.set_precision(prec)
        t.setup_case(case)
        t.set_graph(graph)
        for s in range(6):
            t.advect_steps(1, case.dt)
            if s == 2:
                t.stats(1)  # eager work between replays
        out[graph] = t.get_all_values()
        t.set_graph(False)
        t.set_precision(0)
        t.cleanup()
    assert np.array_equal(out[False], out[True])


def test_union_members_keep_their_own_dt(oracle_mod):
    """Three meshes with different time steps in one union (a time-step factor per node, stepped
    with dt = 1) equal each member advanced alone with its own dt; FP64, widths 1 (tiled) and 4."""
    cs = [configs.bathymetry_case(f"m{i}", n, nl, seed=70 + i, ntracers=4, nsteps=3, lev_lo=3,
                                  lev_hi=nl, umax=u, w_cfl=0.3, scheme=Scheme("MFCT", "QR4C", "FCT", 0.9, 0.8))
          for i, (n, nl, u) in enumerate([(9, 11, 0.1), (14, 17, 0.3), (7, 8, 0.05)])]
    assert len({c.dt for c in cs}) == 3
    u = configs.union_case(cs)
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    for kb in (1, 4):
        say(f"case union of 3 meshes with own dt, kb={kb}")
        t = FESOMTracerAdvection()
        t.set_block_width(kb)
        t.setup_case(u)
        t.advect_steps(3, u.dt)
        got = t.get_all_values()
        t.set_block_width(0)
        t.cleanup()
        off = 0
        for c in cs:
            ref, _ = oracle_mod.Oracle(c).run(save_steps=[3])
            L = c.mesh.nl - 1
            part = got[:, off:off + c.mesh.nod2D]
            err = rel_linf(part[:, :, :L], ref[3])
            say(f"  member {c.name} dt={c.dt:.1f}: rel Linf {err:.3e}")
            assert err <= RTOL
            off += c.mesh.nod2D


@pytest.mark.parametrize("prec,bound,drift_bound", [(1, 1e-5, 1e-14), (2, 5e-3, 1e-10), (3, 2e-2, 1e-10)])
def test_compensated_storage_conserves_to_fp64_roundoff(oracle_mod, prec, bound, drift_bound):
    """Single-precision storage with the rounding remainders kept beside the values and the
    low-order increments (tracer_set_compensated): with FP32 storage the volume-weighted sum is
    conserved to FP64 round-off over 100 steps (stochastic rounding drifts by ~1e-9 here); with
    16-bit storage the FP32 remainder itself rounds at 2^-35 of a value, so the drift stays near
    1e-11 (stochastic rounding: ~1e-7 to 1e-5).  Values stay close to the FP64 oracle, and a
    checkpoint / restore and a block-width change keep the remainders."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy14c", 14, 16, seed=61, ntracers=3, nsteps=100, lev_lo=4, w_cfl=0.3)
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[100])
    t = FESOMTracerAdvection()
    t.set_precision(prec)
    t.set_compensated(True)
    t.setup_case(case)
    v0 = [t.stats(i + 1)["volume_sum"] for i in range(case.ntracers)]
    t.advect_steps(40, case.dt)
    t.checkpoint()
    t.advect_steps(10, case.dt)
    t.checkpoint(restore=True)
    t.set_block_width(1)
    t.advect_steps(60, case.dt)
    got = t.get_all_values()
    v1 = [t.stats(i + 1)["volume_sum"] for i in range(case.ntracers)]
    t.set_block_width(0)
    t.cleanup()
    drift = max(abs(b - a) / abs(a) for a, b in zip(v0, v1))
    err = rel_linf(got, ref[100])
    say(f"case bathy14c precision {prec} compensated: volume-sum drift {drift:.3e}, rel Linf vs FP64 {err:.3e}")
    assert drift <= drift_bound
    assert err <= bound
