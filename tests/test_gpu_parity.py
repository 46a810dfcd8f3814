"""GPU parity: the HIP path (through the C ABI) against the CPU restatement of the reference.

FP64 bar: relative L-infinity <= 1e-12 per step (the kernels keep the reference's operation and
summation order, so the observed difference is normally exactly zero).
"""
import dataclasses
import itertools

import numpy as np
import pytest

from fesom2_tracer_dwarf_amd import configs
from fesom2_tracer_dwarf_amd.configs import Scheme

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def rel_linf(got, ref):
    scale = max(np.max(np.abs(ref)), 1e-300)
    return float(np.max(np.abs(got - ref)) / scale)


def gpu_run(case, nsteps, kb=0, lc=None, prec=0):
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    t = FESOMTracerAdvection()
    t.set_block_width(kb)
    t.set_launch_config(lc or (16, 16, 16, 16))
    t.set_precision(prec)
    t.setup_case(case)
    t.advect_steps(nsteps, case.dt)
    out = t.get_all_values()
    t.set_precision(0)
    t.set_block_width(0)
    t.cleanup()
    return out


def say(msg):
    print(msg, flush=True)


def test_analytic20_matches_oracle(oracle_mod):
    case = configs.analytic20()
    say("case analytic20 MFCT/QR4C/FCT 10 steps")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[10])
    got = gpu_run(case, 10)
    err = rel_linf(got, ref[10])
    say(f"analytic20 rel Linf {err:.3e} bitwise {np.array_equal(got, ref[10])}")
    assert err <= RTOL


SCHEMES = list(itertools.product(["UPW1", "MUSCL", "MFCT"], ["UPW1", "QR4C", "CDIFF", "PPM"],
                                 ["NONE", "FCT"]))


@pytest.mark.parametrize("hor,ver,lim", SCHEMES)
def test_schemes_bathymetry_matches_oracle(oracle_mod, hor, ver, lim):
    case = configs.bathymetry_case("bathy10", 10, 12, seed=11, nsteps=3, lev_lo=4,
                                   scheme=Scheme(hor, ver, lim, 0.85, 0.75), w_cfl=0.3)
    say(f"case bathy10 {hor}/{ver}/{lim} 3 steps")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    got = gpu_run(case, 3)
    err = rel_linf(got, ref[3])
    say(f"  rel Linf {err:.3e} bitwise {np.array_equal(got, ref[3])}")
    assert err <= RTOL


@pytest.mark.parametrize("tid", [1, 2])
def test_dropin_do_oce_adv_tra_matches_oracle(oracle_mod, tid):
    """tracer_do_oce_adv_tra: tendencies, low-order solution and limited fluxes, with a valuesAB
    different from values (tracer 2 is lane 1 of a two-tracer block)."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy16", 16, 14, seed=5, nsteps=1, lev_lo=5, w_cfl=0.3)
    rng = np.random.default_rng(0)
    vals = case.tracers[tid - 1]
    vab = vals + 0.01 * rng.standard_normal(vals.shape) * (vals != 0)
    sc = case.schemes[tid - 1]
    say(f"case bathy16 drop-in do_oce_adv_tra MFCT/QR4C/FCT tracer {tid}")
    o = oracle_mod.Oracle(case)
    ref = oracle_mod.DropIn(o).run(vals, vab, sc, case.dt)
    t = FESOMTracerAdvection()
    t.setup_case(case)
    assert t.block_width() == 2
    t.set_values_ab(tid, vab)
    t.zero_tendencies()
    t.do_oce_adv_tra(case.dt, tid)
    h, v = t.get_tendencies()
    got = {"dttf_h": h, "dttf_v": v, "fct_LO": t.get_work("fct_LO"),
           "adv_flux_hor": t.get_work("adv_flux_hor"), "adv_flux_ver": t.get_work("adv_flux_ver")}
    t.cleanup()
    for k, g in got.items():
        err = rel_linf(g, ref[k])
        say(f"  {k}: rel Linf {err:.3e} bitwise {np.array_equal(g, ref[k])}")
        assert err <= RTOL, k


def test_conservation_and_device_tensors():
    import torch
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy24", 24, 20, seed=2, nsteps=5, lev_lo=6)
    say("case bathy24 conservation + torch tensors")
    t = FESOMTracerAdvection()
    t.setup_case(case)
    v0 = [t.stats(i + 1)["volume_sum"] for i in range(case.ntracers)]
    t.advect_steps(5, case.dt)
    for i in range(case.ntracers):
        v1 = t.stats(i + 1)["volume_sum"]
        assert abs(v1 - v0[i]) <= 1e-13 * abs(v0[i]), (i, v0[i], v1)
    host = t.get_tracer_values(2)
    dev = t.get_values_tensor(2)
    assert np.array_equal(dev.cpu().numpy(), host)
    x = torch.from_numpy(case.tracers[0]).cuda()
    t.set_values_tensor(1, x)
    assert np.array_equal(t.get_tracer_values(1), case.tracers[0])
    st = t.get_statistics()
    assert st["min"] == case.tracers[0].min() and st["max"] == case.tracers[0].max()
    t.cleanup()


CHUNK_SCHEMES = [Scheme("MFCT", "QR4C", "FCT", 0.85, 0.75), Scheme("MUSCL", "PPM", "FCT", 0.9, 0.6),
                 Scheme("UPW1", "CDIFF", "NONE"), Scheme("MFCT", "QR4C", "NONE", 1.0, 1.0)]


@pytest.mark.parametrize("lc", [(1, 1, 1, 1), (2, 3, 5, 7), (4, 4, 4, 4), (64, 64, 64, 64)])
@pytest.mark.parametrize("kb", [1, 2, 4])
@pytest.mark.parametrize("si", range(len(CHUNK_SCHEMES)))
def test_level_chunks_and_block_widths(oracle_mod, lc, kb, si):
    """Results do not depend on how columns are split into level chunks or on how the three
    tracers are interleaved into blocks (kb=2: a full and a half block, kb=4: one unused lane)."""
    case = configs.bathymetry_case("bathy12", 12, 13, seed=17, ntracers=3, nsteps=2, lev_lo=3,
                                   scheme=CHUNK_SCHEMES[si], w_cfl=0.3)
    say(f"case bathy12 scheme {si} lc={lc} kb={kb}")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[2])
    got = gpu_run(case, 2, kb=kb, lc=lc)
    assert rel_linf(got, ref[2]) <= RTOL


def test_mixed_schemes_in_one_block(oracle_mod):
    """Four tracers with three different schemes in one block of width 4: one masked pass per
    scheme."""
    case = configs.bathymetry_case("bathy12m", 12, 13, seed=19, ntracers=4, nsteps=3, lev_lo=3,
                                   w_cfl=0.3)
    case.schemes = [Scheme("MFCT", "QR4C", "FCT", 0.85, 0.75), Scheme("UPW1", "PPM", "NONE"),
                    Scheme("MFCT", "QR4C", "FCT", 0.85, 0.75), Scheme("MUSCL", "CDIFF", "FCT", 0.9, 0.6)]
    say("case bathy12m mixed schemes kb=4")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    got = gpu_run(case, 3, kb=4)
    assert rel_linf(got, ref[3]) <= RTOL


def test_graph_replays_match_oracle(oracle_mod):
    """advect_steps captured into one HIP graph and replayed three times, with eager work
    (statistics, a value read-back) between the replays; every replay is checked."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy14g", 14, 12, seed=29, ntracers=3, nsteps=6, lev_lo=4,
                                   w_cfl=0.3)
    say("case bathy14g graph replays")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[2, 4, 6])
    t = FESOMTracerAdvection()
    t.set_graph(True)
    t.setup_case(case)
    for step in (2, 4, 6):
        t.advect_steps(2, case.dt)
        t.stats(1)
        got = t.get_all_values()
        say(f"  after {step} steps: rel Linf {rel_linf(got, ref[step]):.3e}")
        assert rel_linf(got, ref[step]) <= RTOL, step
    t.set_graph(False)
    t.cleanup()


def test_block_width_change_keeps_values(oracle_mod):
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy10w", 10, 11, seed=31, ntracers=3, nsteps=2, lev_lo=3,
                                   w_cfl=0.3)
    say("case bathy10w width 1 -> 4 -> 2")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[1, 2])
    t = FESOMTracerAdvection()
    t.set_block_width(1)
    t.setup_case(case)
    t.advect_steps(1, case.dt)
    t.set_block_width(4)
    assert rel_linf(t.get_all_values(), ref[1]) <= RTOL
    t.advect_steps(1, case.dt)
    t.set_block_width(2)
    got = t.get_all_values()
    t.set_block_width(0)
    t.cleanup()
    assert rel_linf(got, ref[2]) <= RTOL


@pytest.mark.parametrize("prec", [1, 2, 3])
def test_fp32_batch_of_64_equals_single_lanes(prec):
    """FP32 arithmetic (FP32, FP16 or BF16 storage): 64 tracers in one interleaved block give
    bit-identical results to the same tracers advanced one per block (each lane runs the same
    operations; the stochastic rounding is keyed by tracer number, cell and step)."""
    case = configs.bathymetry_case("bathy8x64", 8, 10, seed=37, ntracers=64, nsteps=2, lev_lo=3,
                                   w_cfl=0.3)
    say(f"case bathy8x64 precision {prec} kb=64 vs kb=1")
    a = gpu_run(case, 2, kb=64, prec=prec)
    b = gpu_run(case, 2, kb=1, prec=prec)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("hor,ver,lim", [("UPW1", "QR4C", "FCT"), ("MUSCL", "PPM", "FCT"), ("MFCT", "CDIFF", "NONE")])
def test_batch_of_64_equals_single_lanes_schemes(prec, hor, ver, lim):
    """The 64-wide kernels (one column per wavefront, next level's values loaded during the
    current one) against one tracer per block, bit for bit, for the other horizontal schemes and
    without the limiter, in FP64 and in FP32."""
    case = configs.bathymetry_case("bathy9x64", 9, 11, seed=39, ntracers=64, nsteps=2, lev_lo=3,
                                   scheme=Scheme(hor, ver, lim, 0.8, 0.7), w_cfl=0.3)
    say(f"case bathy9x64 {hor}/{ver}/{lim} precision {prec} kb=64 vs kb=1")
    a = gpu_run(case, 2, kb=64, prec=prec)
    b = gpu_run(case, 2, kb=1, prec=prec)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_uniform_field_preserved_every_precision(mode):
    """A uniform tracer on a full-depth mesh under a divergence-free flow stays uniform: to FP64
    round-off in FP64, and exactly in the single-precision modes (low-order fluxes in double,
    antidiffusive fluxes and gradients in difference form)."""
    case = configs.bathymetry_case("flat12", 12, 10, seed=41, ntracers=2, nsteps=20, lev_lo=10,
                                   lev_hi=10)
    wet = configs.wet_mask(case.mesh)
    case.tracers[:] = np.where(wet, 35.0, 0.0)[None]
    say(f"case flat12 uniform 35.0, precision mode {mode}")
    got = gpu_run(case, 20, prec=mode)
    err = float(np.max(np.abs(got[:, wet] - 35.0)))
    say(f"  max |T - 35| = {err:.3e}")
    assert err <= (1e-12 * 35.0 if mode == 0 else 0.0)


def test_fp32_200_steps_within_1e5_of_fp64(oracle_mod):
    """FP32 arithmetic and storage (FP64 accumulation, update, limiter ratios) over 200 steps
    against the FP64 reference restatement: relative L-infinity <= 1e-5."""
    case = configs.bathymetry_case("bathy20", 20, 16, seed=23, ntracers=2, nsteps=200, lev_lo=5,
                                   w_cfl=0.3)
    say("case bathy20 FP32 200 steps")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[200])
    got = gpu_run(case, 200, prec=1)
    err = rel_linf(got, ref[200])
    say(f"  rel Linf {err:.3e}")
    assert err <= 1e-5


PREC_BOUNDS = {1: 1e-6, 2: 5e-3, 3: 3e-2}  # rel L-inf after 5 steps vs FP64: FP32, FP16, BF16 storage


@pytest.mark.parametrize("hor,ver", [("MFCT", "QR4C"), ("UPW1", "PPM"), ("MUSCL", "CDIFF")])
def test_fp32_without_limiter_close_to_fp64(oracle_mod, hor, ver):
    """FP32 without the limiter (valuesAB is values, update in place: the upwind part added in
    double reads the values as they were before the update) against the FP64 restatement."""
    case = configs.bathymetry_case("bathy20n", 20, 16, seed=29, nsteps=5, lev_lo=5, w_cfl=0.3,
                                   scheme=Scheme(hor, ver, "NONE", 0.8, 0.7))
    say(f"case bathy20n {hor}/{ver}/NONE FP32 5 steps")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[5])
    got = gpu_run(case, 5, prec=1)
    err = rel_linf(got, ref[5])
    say(f"  rel Linf {err:.3e}")
    assert err <= PREC_BOUNDS[1]


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_mixed_precision_close_to_fp64(oracle_mod, mode):
    """FP32 arithmetic with FP32 / FP16 / BF16 storage against the FP64 reference restatement;
    the volume-weighted sum of the stored field is checked against its initial value."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy20", 20, 16, seed=23, nsteps=5, lev_lo=5, w_cfl=0.3)
    say(f"case bathy20 precision mode {mode}")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[5])
    t = FESOMTracerAdvection()
    t.set_launch_config((16, 16, 16, 16))
    t.set_precision(mode)
    t.setup_case(case)
    v0 = [t.stats(i + 1)["volume_sum"] for i in range(case.ntracers)]
    t.advect_steps(5, case.dt)
    got = t.get_all_values()
    v1 = [t.stats(i + 1)["volume_sum"] for i in range(case.ntracers)]
    t.lib.tracer_set_precision(0)
    t.cleanup()
    err = rel_linf(got, ref[5])
    drift = max(abs(a - b) / abs(a) for a, b in zip(v0, v1))
    say(f"  mode {mode}: rel Linf {err:.3e}, volume-sum drift {drift:.3e}")
    assert err <= PREC_BOUNDS[mode]
    assert drift <= {1: 1e-6, 2: 1e-3, 3: 1e-2}[mode]


def test_union_of_meshes_matches_oracle(oracle_mod):
    """Three independent meshes of different size and depth advanced as one disjoint union
    (the multi-mesh batch), FP64 against the oracle on the same union."""
    cs = [configs.bathymetry_case(f"m{i}", n, nl, seed=60 + i, ntracers=4, nsteps=3, lev_lo=3,
                                  lev_hi=nl, umax=u, w_cfl=0.3)
          for i, (n, nl, u) in enumerate([(9, 11, 0.1), (14, 17, 0.3), (7, 8, 0.05)])]
    u = configs.union_case(cs, own_dt=False)
    say("case union of 3 meshes, 4 tracers, FP64")
    ref, _ = oracle_mod.Oracle(u).run(save_steps=[3])
    got = gpu_run(u, 3, kb=4)
    assert rel_linf(got, ref[3]) <= RTOL


def test_autotune_and_checkpoint_keep_state(oracle_mod, tmp_path, monkeypatch):
    """The autotuner runs trial steps between a device checkpoint and its restore: afterwards the
    state is untouched and the tuned configuration gives the reference result."""
    from fesom2_tracer_dwarf_amd import autotune
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    monkeypatch.setenv("FESOMTRA_TUNE_CACHE", str(tmp_path / "tune.json"))
    case = configs.bathymetry_case("bathy13t", 13, 14, seed=43, ntracers=3, nsteps=3, lev_lo=4,
                                   w_cfl=0.3)
    say("case bathy13t autotune + checkpoint")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    t = FESOMTracerAdvection()
    t.setup_case(case)
    res = autotune.tune(t, case, 0, candidates=(2, 5, 13), steps=2)
    assert all(c in (2, 5, 13) for c in res["launch_config"]) and not res["cached"]
    assert tuple(t.launch_config()) == tuple(res["launch_config"])
    assert np.array_equal(t.get_all_values(), case.tracers)
    t.advect_steps(3, case.dt)
    got = t.get_all_values()
    again = autotune.tune(t, case, 0, candidates=(2, 5, 13), steps=2)
    assert again["cached"] and list(again["launch_config"]) == list(res["launch_config"])
    t.set_launch_config((16, 16, 16, 16))
    t.cleanup()
    assert rel_linf(got, ref[3]) <= RTOL


def test_layout_autotune_keeps_state(oracle_mod):
    """The layout autotuner (block width, tiles, limiter form, streams) runs trial steps between
    a device checkpoint and its restore: afterwards the values are the initial ones and the chosen
    layout gives the reference result."""
    from fesom2_tracer_dwarf_amd import autotune
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy40l", 40, 30, seed=47, ntracers=2, nsteps=3, lev_lo=4, w_cfl=0.3)
    say("case bathy40l layout autotune")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    t = FESOMTracerAdvection()
    t.setup_case(case)
    res = autotune.tune_layout(t, case, steps=2)
    say(f"  chose {res['block_width']}, tiles {res['tiles']}, fused limiter {res['fused_limiter']}, "
        f"two streams {res['concurrent']} of {sorted(res['step_ms'])}")
    assert any("/fused/" in k for k in res["step_ms"]) and any("one-stream" in k for k in res["step_ms"])
    assert t.block_width() == res["block_width"]
    assert np.array_equal(t.get_all_values(), case.tracers)
    t.advect_steps(3, case.dt)
    got = t.get_all_values()
    t.set_fused_limiter(False)
    t.set_concurrent(True)
    t.set_halo_tiles(True)
    t.set_block_width(0)
    t.cleanup()
    assert rel_linf(got, ref[3]) <= RTOL


FUSED_SCHEMES = [Scheme("MFCT", "QR4C", "FCT", 0.85, 0.75), Scheme("MUSCL", "PPM", "FCT", 0.9, 0.6),
                 Scheme("UPW1", "CDIFF", "FCT"), Scheme("MFCT", "UPW1", "FCT", 1.0, 1.0)]


@pytest.mark.parametrize("si", range(len(FUSED_SCHEMES)))
@pytest.mark.parametrize("nsteps", [1, 2, 3])
def test_fused_tiles_match_oracle(oracle_mod, si, nsteps):
    """The fused single-pass kernel (tiles with a three-ring halo exchanged in LDS) is
    bit-compatible with the reference on a mesh of several tiles, for odd and even step counts
    (the tracer alternates between two buffers)."""
    case = configs.bathymetry_case("bathy40f", 40, 14, seed=47, ntracers=2, nsteps=nsteps, lev_lo=3,
                                   scheme=FUSED_SCHEMES[si], w_cfl=0.3)
    say(f"case bathy40f fused scheme {si} steps {nsteps}")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[nsteps])
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    t = FESOMTracerAdvection()
    t.set_block_width(1)
    t.setup_case(case)
    t.set_fused(True)
    t.advect_steps(nsteps, case.dt)
    got = t.get_all_values()
    t.set_fused(False)
    t.set_block_width(0)
    t.cleanup()
    err = rel_linf(got, ref[nsteps])
    say(f"  rel Linf {err:.3e} bitwise {np.array_equal(got, ref[nsteps])}")
    assert err <= RTOL


def test_fused_graph_and_mixed_limiters(oracle_mod):
    """Fused FCT tracers next to a tracer without limiter (multi-kernel path), replayed from a HIP
    graph with eager work between replays."""
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy36g", 36, 12, seed=53, ntracers=3, nsteps=6, lev_lo=3,
                                   w_cfl=0.3)
    case.schemes[1] = Scheme("MUSCL", "QR4C", "NONE", 0.9, 0.8)
    say("case bathy36g fused + graph + non-limited tracer")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3, 6])
    t = FESOMTracerAdvection()
    t.set_block_width(1)
    t.set_graph(True)
    t.setup_case(case)
    t.set_fused(True)
    for step in (3, 6):
        t.advect_steps(3, case.dt)
        t.stats(2)
        got = t.get_all_values()
        say(f"  after {step} steps: rel Linf {rel_linf(got, ref[step]):.3e}")
        assert rel_linf(got, ref[step]) <= RTOL, step
    t.set_fused(False)
    t.set_graph(False)
    t.set_block_width(0)
    t.cleanup()


def test_autotune_does_not_change_single_precision_results():
    """The checkpoint restores the stochastic-rounding step counter with the tracers, so a tuned
    run gives bit-identical FP32 results to an untuned one."""
    from fesom2_tracer_dwarf_amd import autotune
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    case = configs.bathymetry_case("bathy12r", 12, 11, seed=59, ntracers=4, nsteps=4, lev_lo=3,
                                   w_cfl=0.3)
    say("case bathy12r FP32 with and without autotune")
    outs = []
    for tune in (False, True):
        t = FESOMTracerAdvection()
        t.set_precision(1)
        t.setup_case(case)
        if tune:
            autotune.tune(t, case, 1, candidates=(3, 7), steps=2, use_cache=False)
        t.advect_steps(4, case.dt)
        outs.append(t.get_all_values())
        t.set_precision(0)
        t.cleanup()
    assert np.array_equal(outs[0], outs[1])


TILE_SCHEMES = [Scheme("MFCT", "QR4C", "FCT", 0.85, 0.75), Scheme("MUSCL", "PPM", "FCT", 0.9, 0.6),
                Scheme("UPW1", "CDIFF", "FCT"), Scheme("MFCT", "UPW1", "FCT", 1.0, 1.0),
                Scheme("MUSCL", "QR4C", "NONE", 0.8, 0.7), Scheme("MFCT", "PPM", "NONE", 1.0, 1.0)]


def _run_tiles(case, nsteps, kb, tiles, prec=0, lc=(16, 5, 7, 3)):
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    t = FESOMTracerAdvection()
    t.set_block_width(kb)
    t.set_halo_tiles(tiles)
    t.set_launch_config(lc)
    t.set_precision(prec)
    t.setup_case(case)
    ntiles = t.halo_tiles()
    t.advect_steps(nsteps, case.dt)
    out = t.get_all_values()
    t.set_precision(0)
    t.set_block_width(0)
    t.set_halo_tiles(True)
    t.cleanup()
    return out, ntiles


@pytest.mark.parametrize("kb", [1, 2, 4])
@pytest.mark.parametrize("si", range(len(TILE_SCHEMES)))
def test_halo_tiles_match_untiled_and_oracle(oracle_mod, kb, si):
    """The tiled kernels (edge tiles, and with FCT node tiles with halos; level chunks) give the
    same bits as the untiled kernels and as the reference restatement."""
    case = configs.bathymetry_case("bathy48", 48, 45, seed=71, ntracers=3, nsteps=3, lev_lo=3,
                                   scheme=TILE_SCHEMES[si], w_cfl=0.3)
    say(f"case bathy48 scheme {si} kb={kb} tiled vs untiled")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    tiled, nt = _run_tiles(case, 3, kb, True)
    plain, n0 = _run_tiles(case, 3, kb, False)
    assert nt == -(-case.mesh.nod2D // (256 // kb)) and n0 == 0
    assert np.array_equal(tiled, plain)
    assert rel_linf(tiled, ref[3]) <= RTOL


@pytest.mark.parametrize("lim", ["FCT", "NONE"])
@pytest.mark.parametrize("prec", [1, 3])
def test_halo_tiles_single_precision_bitwise(prec, lim):
    case = configs.bathymetry_case("bathy40p", 40, 37, seed=73, ntracers=2, nsteps=4, lev_lo=3,
                                   w_cfl=0.3, scheme=Scheme("MFCT", "QR4C", lim, 0.85, 0.75))
    say(f"case bathy40p precision {prec} limiter {lim} tiled vs untiled")
    tiled, nt = _run_tiles(case, 4, 2, True, prec=prec)
    plain, _ = _run_tiles(case, 4, 2, False, prec=prec)
    assert nt > 0
    assert np.array_equal(tiled, plain)


def test_no_halo_tiles_for_wide_blocks():
    case = configs.bathymetry_case("bathy20w", 20, 18, seed=75, ntracers=8, nsteps=1, lev_lo=3,
                                   w_cfl=0.3)
    _, nt = _run_tiles(case, 1, 8, True)
    assert nt == 0


# --- FESOM2 'pi' mesh of the reference's test data: sphere, coasts (455 boundary edges, edges
# without up/down-wind triangles), node degrees up to 8, 47 layers of 5-250 m.
PI_SCHEMES = [Scheme("MFCT", "QR4C", "FCT", 0.85, 0.75), Scheme("MUSCL", "PPM", "FCT", 0.9, 0.6),
              Scheme("UPW1", "CDIFF", "NONE"), Scheme("MFCT", "QR4C", "NONE", 1.0, 1.0)]


@pytest.fixture(scope="module")
def pi_mesh():
    from pathlib import Path
    from fesom2_tracer_dwarf_amd import mesh_io
    with np.load(Path(__file__).resolve().parent / "golden" / "pi_mesh.npz") as d:
        return mesh_io.sphere_mesh({k: d[k] for k in d.files})


@pytest.mark.parametrize("kb,tiles", [(0, True), (1, True), (0, False)])
@pytest.mark.parametrize("si", range(len(PI_SCHEMES)))
def test_pi_mesh_matches_oracle(oracle_mod, pi_mesh, si, kb, tiles):
    case = configs.sphere_case("pi", pi_mesh, seed=7, ntracers=2, nsteps=3, scheme=PI_SCHEMES[si], w_cfl=0.3)
    say(f"case pi mesh scheme {si} kb={kb} tiles={tiles}")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    got, _ = _run_tiles(case, 3, kb, tiles, lc=(16, 16, 16, 16))
    err = rel_linf(got, ref[3])
    say(f"pi mesh rel Linf {err:.3e}")
    assert err <= RTOL


def test_pi_mesh_wide_block_matches_oracle(oracle_mod, pi_mesh):
    """Four tracers in a 64-wide block (one column per wavefront): nodes of degree 7 and 8 go
    beyond the six node->edge entries held in registers."""
    case = configs.sphere_case("pi4", pi_mesh, seed=9, ntracers=4, nsteps=3, w_cfl=0.3)
    say("case pi mesh 4 tracers kb=64")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    got = gpu_run(case, 3, kb=64)
    assert rel_linf(got, ref[3]) <= RTOL


def test_pi_mesh_single_precision(oracle_mod, pi_mesh):
    case = configs.sphere_case("pi32", pi_mesh, seed=11, ntracers=2, nsteps=5)
    say("case pi mesh FP32 5 steps")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[5])
    got = gpu_run(case, 5, prec=1)
    err = rel_linf(got, ref[5])
    say(f"pi mesh FP32 rel Linf {err:.3e}")
    assert err <= 1e-5


def test_mesh_read_by_library_matches_oracle(oracle_mod, pi_mesh, tmp_path):
    """tracer_load_mesh_from_files on the pi mesh directory: the library derives the geometry
    itself (C++), the oracle runs on the Python-derived mesh.  The derivations agree in every
    discrete choice (orientation, element lists, up/down-wind triangles) and to the last bits in
    the floating-point geometry (the trigonometric functions of the two sides may differ by an
    ulp), so the results agree to 1e-10."""
    from pathlib import Path
    from fesom2_tracer_dwarf_amd import mesh_io
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    with np.load(Path(__file__).resolve().parent / "golden" / "pi_mesh.npz") as d:
        raw = {k: d[k] for k in d.files}
    mesh_io.write_mesh_files(raw, tmp_path / "pi")
    case = configs.sphere_case("pi_files", pi_mesh, seed=15, ntracers=2, nsteps=3, w_cfl=0.3)
    say("case pi mesh read by the library")
    ref, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    t = FESOMTracerAdvection()
    t.setup_case(case, mesh_files=str(tmp_path / "pi"))
    t.advect_steps(3, case.dt)
    got = t.get_all_values()
    t.cleanup()
    err = rel_linf(got, ref[3])
    say(f"pi mesh via files rel Linf {err:.3e}")
    assert err <= 1e-10
