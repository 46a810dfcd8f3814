"""tracer_do_oce_adv_tra_noalloc against tracer_do_oce_adv_tra, through the C ABI.

The variant that allocates nothing keeps every work array in the library's fixed device scratch
and takes the caller's device arrays (values, valuesAB in, tendencies accumulated) in reference
order.  Its tendencies equal the allocating drop-in's bit for bit, on two calls in a row, the
second on a smaller mesh (the scratch still holds the first call's contents), and a mesh that
needs more scratch than the fixed array is refused with a message.
"""
import ctypes

import numpy as np
import pytest

from fesom2_tracer_dwarf_amd import abi, configs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def noalloc_vs_dropin(case, tid, calls=2):
    """Tendencies of `calls` accumulating calls of each entry point (both start from zero)."""
    import torch
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    rng = np.random.default_rng(tid)
    vals = case.tracers[tid - 1]
    vab = vals + 0.01 * rng.standard_normal(vals.shape) * (vals != 0)
    t = FESOMTracerAdvection()
    t.setup_case(case)
    t.lib.tracer_init_for_advection()
    # the allocating drop-in: values and valuesAB set, tendencies zeroed once, then accumulated
    t.set_values(tid, vals)
    t.set_values_ab(tid, vab)
    t.zero_tendencies()
    for _ in range(calls):
        t.do_oce_adv_tra(case.dt, tid)
    h_ref, v_ref = t.get_tendencies()
    # the variant on caller device arrays
    dv, dab = torch.from_numpy(vals).cuda(), torch.from_numpy(vab).cuda()
    dh = torch.zeros_like(dv)
    dvv = torch.zeros_like(dv)
    torch.cuda.synchronize()
    for _ in range(calls):
        rc = t.lib.tracer_do_oce_adv_tra_noalloc(case.dt, tid, ctypes.c_void_p(dv.data_ptr()),
                                                 ctypes.c_void_p(dab.data_ptr()), ctypes.c_void_p(dh.data_ptr()),
                                                 ctypes.c_void_p(dvv.data_ptr()))
        abi.check(rc, "tracer_do_oce_adv_tra_noalloc")
    out = dh.cpu().numpy(), dvv.cpu().numpy(), h_ref, v_ref
    t.cleanup()
    return out


def test_noalloc_equals_dropin_twice_then_smaller():
    big = configs.bathymetry_case("bathy20n", 20, 16, seed=81, nsteps=1, lev_lo=4, w_cfl=0.3)
    small = configs.bathymetry_case("bathy9n", 9, 10, seed=82, nsteps=1, lev_lo=3, w_cfl=0.3)
    for case, tid in ((big, 1), (small, 2)):
        say(f"case {case.name} tracer {tid}: noalloc vs drop-in, two accumulating calls")
        gh, gv, wh, wv = noalloc_vs_dropin(case, tid)
        say(f"  max |dh| {np.max(np.abs(gh - wh)):.3e}, max |dv| {np.max(np.abs(gv - wv)):.3e}")
        assert np.any(wh) and np.any(wv)
        assert np.array_equal(gh, wh) and np.array_equal(gv, wv)


def test_noalloc_refuses_a_mesh_beyond_the_scratch():
    import torch
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy260n", 260, 48, seed=83, nsteps=1, lev_lo=10)
    say("case bathy260n (67600 nodes x 47 layers): beyond the scratch")
    t = FESOMTracerAdvection()
    t.setup_case(case)
    t.lib.tracer_init_for_advection()
    x = torch.zeros((case.mesh.nod2D, case.mesh.nl - 1), dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(x.data_ptr())
    assert t.lib.tracer_do_oce_adv_tra_noalloc(case.dt, 1, p, p, p, p) == -1
    assert b"scratch" in t.lib.tracer_last_error()
    t.cleanup()
