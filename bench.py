"""Benchmark: node-level tracer updates/s of the FESOM2 tracer advection step on one MI355X.

Headline (default): the plain FP64 step on a 1000x1000 doubly periodic mesh (1M nodes, 2M
triangles, 3M edges), 48 levels stretched 5-250 m, smooth random bathymetry with 10-48 levels,
randomly permuted numbering, T and S advected with MFCT / QR4C / FCT at CFL 0.5; one step advances
both tracers.  Beside it (same JSON line, key "batched") the mixed-precision batched form: 64
tracers on a 500x500x48 mesh in one interleaved block, FP32 flux arithmetic and storage with FP64
accumulation, update, limiter ratios and sums.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5] [--workload mesh1000]
    python bench.py --workload multimesh32          # 32 uneven meshes, one HIP graph per run
    python bench.py --precision-study               # 1000-step FP32/FP16/BF16 errors vs FP64, all modes

Prints one JSON line.  ``value`` = tracers x wet node-levels x steps / wall time of the timed
steps (inputs resident in HBM).  ``roofline`` relates the algorithmic bytes of one step (every
input read once, the tracers written once, at storage precision) to the mean HIP-event duration
of the whole step on the library's stream (the edge-transport kernel k_qflux and every pass's
four kernels); ``traffic`` is the rocprof-measured HBM traffic of one step
(profiles/traffic_<workload>.json, used only when it was measured on the current kernel sources:
the file carries their hash).  ``cpu_baseline`` times the reference Fortran (built from its
sources into oracle/_ref, OpenMP) on one tracer-step of the same mesh, or, when that build is
absent, the C restatement on one thread, saying so.

    python bench.py --rocprof-summary DIR --workload mesh1000   # rocprof CSVs -> profiles/
"""

from __future__ import annotations

import argparse
import csv
import dataclasses
import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
PREC_NAMES = {0: "f64", 1: "f32", 2: "f32/f16-storage", 3: "f32/bf16-storage"}
PREC_BYTES = {0: (8, 8), 1: (4, 8), 2: (2, 8), 3: (2, 8)}  # (tracer storage, mesh) bytes; mesh is FP64 always


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(mesh, k_tracers: int = 1, s_tr: int = 8, s_mesh: int = 8, members: int = 1) -> int:
    """Bytes one pass must move at least: every distinct input once, every output once.

    K tracers in and out, per node-layer mesh arrays (hnode, hnode_new, areasvol, Z_3d_n),
    per node-interface arrays (area, w, zbar_3d_n), per element-layer uv(2) and helem, and the
    per-edge / per-element / per-node connectivity, geometry and level integers.  An ensemble of
    `members` velocities adds each further member's uv and w.
    """
    N, T, E = mesh.nod2D, mesh.elem2D, mesh.edge2D
    NL = mesh.n_node_levels
    TL = mesh.n_elem_levels
    adj = float(np.mean(mesh.nod_in_elem2D_num))
    return int(k_tracers * 2 * s_tr * NL
               + s_mesh * (4 * NL + 3 * (NL + N) + 3 * TL)
               + (members - 1) * s_mesh * (2 * TL + NL + N)
               + E * (16 + 6 * s_mesh + 8)
               + T * (20 + 8 * s_mesh)
               + N * (20 + 4 * adj))


def port_baseline(case, sample_tracer_steps: int = 1):
    """The CPU restatement (oracle/, C, one thread) on one tracer-step of the full mesh."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle_py
    o = oracle_py.Oracle(case)
    vals = np.ascontiguousarray(case.tracers[0]).copy()
    t0 = time.perf_counter()
    for _ in range(sample_tracer_steps):
        o.step(vals, case.schemes[0], case.dt)
    dt = time.perf_counter() - t0
    ups = case.mesh.n_node_levels * sample_tracer_steps / dt
    return {"value": ups, "unit": "node-level updates/s", "cores": 1, "kind": "port",
            "sample": f"{sample_tracer_steps} tracer-step(s) of tracer T on the full "
                      f"{case.name} mesh, CPU restatement of the reference (oracle/), "
                      f"1 thread, {dt:.2f} s"}


def reference_baseline(case, exe: Path, threads: int):
    """The reference Fortran (lib/*.F90 built from its sources with flang -O2 -fopenmp, its own
    OpenMP regions and lock-based scatter) on one tracer-step of the full mesh: gradients +
    fill_up_dn_grad, do_oce_adv_tra and the caller's update, timed with MPI_Wtime."""
    import dataclasses
    import resource
    import subprocess
    import tempfile
    sys.path.insert(0, str(REPO / "oracle" / "ref"))
    import bundle
    one = dataclasses.replace(case, tracers=case.tracers[:1], schemes=case.schemes[:1], nsteps=1)

    def unlimited_stack():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    with tempfile.TemporaryDirectory(prefix="fesomtra_ref_") as td:
        p = bundle.write_bundle(Path(td) / "b", one, timing=True)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_STACKSIZE="512M")
        subprocess.run([str(exe), str(p)], capture_output=True, text=True, env=env, timeout=900,
                       check=True, preexec_fn=unlimited_stack)
        t = bundle.read_timing(p)
    total = t["seconds_gradients"] + t["seconds_advection"] + t["seconds_update"]
    nl = t["node_levels_per_tracer"] * t["tracer_steps"]
    return {"value": nl / total, "unit": "node-level updates/s", "cores": threads, "kind": "reference",
            "sample": f"1 tracer-step of tracer T on the full {case.name} mesh ({int(nl)} node-levels), "
                      f"reference Fortran lib/*.F90 (flang -O2 -fdefault-real-8 -fopenmp, "
                      f"{threads} OpenMP threads): gradients+fill_up_dn_grad "
                      f"{t['seconds_gradients']:.2f} s, do_oce_adv_tra {t['seconds_advection']:.2f} s, "
                      f"update {t['seconds_update']:.2f} s; advection-only rate "
                      f"{nl / t['seconds_advection']:.3e}/s"}


def host_cpus():
    """(threads to use, description): the CPUs this process may run on, capped by OMP_NUM_THREADS
    where the environment sets it (the CPU share a one-GPU job is given on a shared node)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail, int(env)) if env.isdigit() and int(env) > 0 else avail
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    desc = (f"{threads} threads on a host with {os.cpu_count()} logical CPUs ({avail} available to "
            f"this process{', OMP_NUM_THREADS=' + env if env else ''}){'; ' + model if model else ''}")
    return threads, desc


def cpu_baseline(case):
    exe = REPO / "oracle" / "_ref" / "ref_driver_omp"
    threads, desc = host_cpus()
    if exe.exists():
        out = reference_baseline(case, exe, threads)
        out["host"] = desc
        return out
    why = (f"the reference build {exe.relative_to(REPO)} is absent (oracle/ref/build_ref.sh builds it "
           "from the reference sources); timing the C restatement instead")
    log("CPU BASELINE: " + why)
    out = port_baseline(case)
    out["host"] = desc
    out["note"] = why
    return out


def kernel_source_hash() -> str:
    """Hash of the kernel sources and build recipe: a stored traffic figure applies to one build."""
    h = hashlib.sha256()
    csrc = REPO / "fesom2_tracer_dwarf_amd" / "csrc"
    for f in sorted(list(csrc.glob("*.hip")) + list(csrc.glob("*.hpp")) + list(csrc.glob("*.cpp")) + list(csrc.glob("*.inc"))
                    + [REPO / "fesom2_tracer_dwarf_amd" / "_build.py"]):
        h.update(f.name.encode())
        h.update(f.read_bytes())
    return h.hexdigest()[:16]


def traffic_from_profiles(name: str, passes: int = None):
    """HBM bytes of one step from profiles/traffic_<name>.json, if measured on these sources and,
    when `passes` is given, with as many passes per step (tracers / block width) as this run."""
    f = REPO / "profiles" / f"traffic_{name}.json"
    if not f.exists():
        return None, f"no {f.relative_to(REPO)}"
    d = json.loads(f.read_text())
    if d.get("source_hash") != kernel_source_hash():
        return None, (f"{f.relative_to(REPO)} was measured on other kernel sources "
                      f"({d.get('source_hash')} != {kernel_source_hash()})")
    if passes is not None and d.get("passes_per_step") not in (None, passes):
        return None, (f"{f.relative_to(REPO)} was measured with {d['passes_per_step']} passes per step, "
                      f"this run makes {passes}")
    return d.get("hbm_bytes_per_step"), f"{f.relative_to(REPO)} (source hash {d['source_hash']})"


def make_case(name: str, nsteps=None):
    from fesom2_tracer_dwarf_amd import configs
    if name == "mesh1000":
        return configs.mesh1000(nsteps or 100)
    if name == "mesh500x64":
        return configs.mesh500x64(nsteps or 200)
    if name == "multimesh32":
        return configs.multimesh32_union(nsteps or 100)
    if name == "analytic20":
        return configs.analytic20(nsteps=nsteps or 10)
    raise ValueError(name)


def run(case, prec: int, kb: int, steps: int, warmup: int, graph: bool = False, lc=None,
        stage_steps: int = 10, compare=None, keep: bool = False, stochastic: bool = True,
        tune: bool = True, fused: bool = False, compensated: bool = False, members=None,
        fused_limiter: bool = False, wsplit: bool = False, concurrent: bool = True):
    """Advance `case` on the GPU; returns throughput, device time, per-stage times, drift.

    keep: also return the final values (list per tracer); compare: final values of a reference
    run of the same steps, against which relative L-inf / L2 over the wet layers are reported.
    members: the cases of an ensemble on case's mesh (case holds all their tracers, member by
    member), set up with tracer_set_ensemble."""
    import torch
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    m = case.mesh
    t = FESOMTracerAdvection()
    t.set_precision(prec)
    t.set_block_width(kb)
    t.set_graph(graph)
    t.set_rounding(stochastic)
    t.set_compensated(compensated)
    t.set_fused_limiter(fused_limiter)
    t.set_concurrent(concurrent)
    if lc:
        t.set_launch_config(lc)
    if members:
        t.setup_ensemble(members)
    else:
        t.setup_case(case)
    if wsplit:
        t.set_wsplit(True)
    if fused:
        t.set_fused(True)
    tuned, layout = None, None
    if tune and not lc and not fused:
        from fesom2_tracer_dwarf_amd import autotune
        if kb == 0:
            layout = autotune.tune_layout(t, case)
            log(f"autotuned layout: block width {layout['block_width']}, tiles {layout['tiles']}, "
                f"fused limiter {layout['fused_limiter']}, two streams {layout['concurrent']} "
                f"(ms/step {layout['step_ms']})")
        tuned = autotune.tune(t, case, prec, use_cache=False)
        log(f"autotuned levels per thread {tuned['launch_config']} (stage ms {[round(x, 3) for x in tuned['stage_ms']]})")
    width = t.block_width()
    limiter = t.fused_limiter_active()
    v0 = [t.stats(i + 1)["volume_sum"] for i in range(case.ntracers)]
    if warmup:
        t.advect_steps(warmup, case.dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = t.advect_steps(steps, case.dt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    drift = max(abs(t.stats(i + 1)["volume_sum"] - v0[i]) / abs(v0[i]) for i in range(case.ntracers))
    kept, err = None, None
    if keep or compare is not None:
        from fesom2_tracer_dwarf_amd import configs
        wet = configs.wet_mask(m)
        kept = [] if keep else None
        linf, l2n, l2d, rmax = 0.0, 0.0, 0.0, 0.0
        for i in range(case.ntracers):
            v = t.get_tracer_values(i + 1)
            if keep:
                kept.append(v)
            if compare is not None:
                d = (v - compare[i])[wet]
                r = compare[i][wet]
                linf = max(linf, float(np.max(np.abs(d))) / float(np.max(np.abs(r))))
                l2n += float(np.sum(d * d))
                l2d += float(np.sum(r * r))
        if compare is not None:
            err = {"rel_linf": linf, "rel_l2": float(np.sqrt(l2n / l2d))}
    # per-stage HIP-event timing: a separate short run outside the timed region (after the
    # values above were read)
    t.lib.tracer_set_kernel_timing(1)
    t.advect_steps(stage_steps, case.dt)
    stage_ms, passes = t.kernel_times()
    qflux_ms, step_ms, nsteps_timed = t.step_times()
    t.lib.tracer_set_kernel_timing(0)
    cfg = t.launch_config()
    tiles = t.halo_tiles()
    t.set_graph(False)
    t.set_fused(False)
    t.set_wsplit(False)
    t.set_precision(0)
    t.set_block_width(0)
    t.cleanup()
    updates = case.ntracers * m.n_node_levels * steps
    return {"value": updates / wall, "wall_s": wall, "dev_ms": dev_ms, "stage_ms": stage_ms,
            "passes_timed": passes, "block_width": width, "drift": drift, "launch_config": cfg,
            "values": kept, "error": err, "autotune": tuned, "layout": layout, "qflux_ms": qflux_ms,
            "step_ms": step_ms, "steps_timed": nsteps_timed, "fused_limiter": limiter, "steps": steps,
            "tiles": tiles}


def roofline(case, prec: int, r: dict, name: str, members: int = 1):
    """The step as the unit: algorithmic bytes of one step of every tracer over the mean time of a
    step of the timed run, from the HIP events the library records around it on its own stream
    (tracer_last_elapsed_ms; k_qflux and every pass, the low-order kernels on the second stream
    joined into it).  stage_ms are per-kernel HIP-event times of a separate run with the stages
    timed one after another."""
    s_tr, s_m = PREC_BYTES[prec]
    k = case.ntracers
    B = algorithmic_bytes(case.mesh, k, s_tr, s_m, members)
    step_ms = r["dev_ms"] / r["steps"]
    achieved = B / (step_ms * 1e-3) / 1e9
    st = r["stage_ms"]
    passes = max(1, round(r["passes_timed"] / max(r["steps_timed"], 1)))
    traffic, source = traffic_from_profiles(name, passes)
    fused = not (st[0] or st[1] or st[2])
    lim = r.get("fused_limiter", False)
    if fused:
        label, stages = "fused pass (k_tile)", {"k_tile": round(st[3], 4)}
    elif lim:
        label = "pass (k_edge + k_lo + k_limit)"
        stages = {"k_edge": round(st[0], 4), "k_lo": round(st[1], 4), "k_limit": round(st[2] + st[3], 4)}
    else:
        label = "pass (k_edge + k_lo + k_bounds + k_final)"
        stages = {"k_edge": round(st[0], 4), "k_lo": round(st[1], 4), "k_bounds": round(st[2], 4),
                  "k_final": round(st[3], 4)}
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": source,
            "kernel": ("one step: k_qflux + " + (f"{passes} x " if passes > 1 else "") + label +
                       f" over {k} tracer(s) in blocks of {r['block_width']}, {PREC_NAMES[prec]}"),
            "algorithmic_bytes_per_step": B, "tracers_per_step": k, "step_ms": round(step_ms, 4),
            "step_ms_stages_serial": round(r["step_ms"], 4),
            "stage_ms": {"k_qflux": round(r["qflux_ms"], 4), **stages},
            "passes_per_step": passes}


# Block width of the 32-mesh union: at width 1 every tile of the union fits the tiled kernels'
# staging space (256 nodes, their halo and edges); at width 4 the 64-node tiles cut along the
# members' edges into slivers whose halos overflow it, and the union runs untiled (24.4 against
# 20.8 ms per step of 4 tracers).
MULTIMESH_WIDTH = 1

PROFILE_ARGS = ("--steps 5 --warmup 1 --no-cpu-baseline --no-batched --no-dropin --no-multimesh --no-precision "
                "--no-vertical --no-autotune --concurrent 0 "
                "--lc 48,48,48,48")


def profile_args(workload: str) -> str:
    """The bench options of the profiled runs: one stream (the kernels' own durations, as the
    timed runs' autotuned layout); for the 64-tracer and the multi-mesh workloads the compensated
    storage of their bench lines; the multi-mesh steps launched directly rather than replayed as
    a graph (the same kernels, each dispatch counted on its own)."""
    if workload == "mesh500x64":
        return PROFILE_ARGS + " --compensated"
    if workload == "multimesh32":
        return PROFILE_ARGS + " --compensated --no-graph"
    return PROFILE_ARGS


def _rocprof_rows(d: Path, suffix: str):
    files = sorted(d.rglob(f"*{suffix}"))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    rows = []
    for f in files:
        with open(f, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _short(name: str) -> str:
    """void fesomtra::k_edge<true, 2, ...>(...) -> fesomtra::k_edge"""
    n = name.split("<")[0].split("(")[0].strip()
    return n[5:] if n.startswith("void ") else n


def _counter_means(d: Path, counter: str):
    """Mean value (KiB) per dispatch of `counter` for each kernel name."""
    acc = {}
    for r in _rocprof_rows(d, "counter_collection.csv"):
        if r.get("Counter_Name") != counter:
            continue
        k = _short(r["Kernel_Name"])
        acc.setdefault(k, []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


SQ_PASSES = ("SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_VMEM "
             "SQ_ACTIVE_INST_LDS SQ_INSTS_VALU",
             "SQ_WAVES SQ_INSTS_VMEM_RD SQ_INSTS_LDS SQ_INSTS_VALU_FMA_F64 SQ_INSTS_VALU_ADD_F64 SQ_INSTS_VALU_MUL_F64 "
             "SQ_INSTS_VALU_TRANS_F64 SQ_INSTS_SALU")


def sq_summary(args):
    """profiles/sq_<workload>.txt from the two rocprofv3 SQ counter passes DIR/s1 and DIR/s2 of
    `bench.py --workload W PROFILE_ARGS` (counters summed over the GPU, averaged over launches)."""
    d = Path(args.sq_summary)
    w = args.workload
    acc = {}
    for sub in ("s1", "s2"):
        for r in _rocprof_rows(d / sub, "counter_collection.csv"):
            k = _short(r["Kernel_Name"])
            if not k.startswith("fesomtra::"):
                continue
            acc.setdefault((k, r["Kernel_Name"]), {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    pa = profile_args(w)
    lines = [f"# {w}: SQ counters of bench.py ({pa})", "",
             "Commands (two counter passes of the same program):"]
    for i, c in enumerate(SQ_PASSES, 1):
        lines.append(f"    rocprofv3 --pmc {c} -f csv -d <dir>/s{i} -o s{i} -- python -u bench.py --workload {w} {pa}")
    lines += [f"    python bench.py --sq-summary <dir> --workload {w}",
              f"Kernel source hash {kernel_source_hash()}", "",
              "Per kernel instantiation: fraction of wavefront cycles spent waiting (any dependency / for an",
              "instruction to issue), cycles with a VALU, vector-memory or LDS instruction active per wavefront",
              "cycle, and instructions per wavefront (VALU, of which FP64 add/mul/fma/transcendental; vector",
              "loads; LDS; scalar).", "",
              f"{'kernel':<60}{'wait':>6}{'waitI':>6}{'valu':>6}{'vmem':>6}{'lds':>6}{'VALU/w':>8}{'F64%':>6}"
              f"{'VMRD/w':>8}{'LDS/w':>7}{'SALU/w':>8}"]
    f64n = ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_TRANS_F64")
    for (k, full), c in sorted(acc.items()):
        m = {n: sum(v) / len(v) for n, v in c.items()}
        wc, nw = m.get("SQ_WAVE_CYCLES", 0.0), m.get("SQ_WAVES", 0.0)
        if wc <= 0 or nw <= 0:
            continue
        valu = m.get("SQ_INSTS_VALU", 0.0)
        name = full.split("(")[0].replace("void ", "")[:59]
        lines.append(f"{name:<60}{m.get('SQ_WAIT_ANY', 0) / wc:>6.2f}{m.get('SQ_WAIT_INST_ANY', 0) / wc:>6.2f}"
                     f"{m.get('SQ_ACTIVE_INST_VALU', 0) / wc:>6.2f}{m.get('SQ_ACTIVE_INST_VMEM', 0) / wc:>6.2f}"
                     f"{m.get('SQ_ACTIVE_INST_LDS', 0) / wc:>6.2f}{valu / nw:>8.0f}"
                     f"{100 * sum(m.get(x, 0) for x in f64n) / max(valu, 1):>6.1f}{m.get('SQ_INSTS_VMEM_RD', 0) / nw:>8.0f}"
                     f"{m.get('SQ_INSTS_LDS', 0) / nw:>7.0f}{m.get('SQ_INSTS_SALU', 0) / nw:>8.0f}")
    out = REPO / "profiles" / f"sq_{w}.txt"
    out.write_text("\n".join(lines) + "\n")
    print("\n".join(lines))


def rocprof_summary(args):
    """profiles/kernels_<workload>.txt and profiles/traffic_<workload>.json from rocprofv3 output
    directories under DIR: k (kernel trace, --stats), f (--pmc FETCH_SIZE), w (--pmc WRITE_SIZE) of
    `bench.py --workload W PROFILE_ARGS`, and cf / cw (the same counters on --copy-calibration,
    which copies 1 GiB, to scale the counters to bytes)."""
    import subprocess
    d = Path(args.rocprof_summary)
    w = args.workload
    stats = {}
    for r in _rocprof_rows(d / "k", "kernel_stats.csv"):
        k = _short(r["Name"])
        calls = int(r["Calls"])
        tot = float(r["TotalDurationNs"])
        c0 = stats.get(k, (0, 0.0, float("inf"), 0.0))
        stats[k] = (c0[0] + calls, c0[1] + tot, min(c0[2], float(r["MinNs"])), max(c0[3], float(r["MaxNs"])))
    fetch, write = _counter_means(d / "f", "FETCH_SIZE"), _counter_means(d / "w", "WRITE_SIZE")
    copy_bytes = float(1 << 30)
    cf = _counter_means(d / "cf", "FETCH_SIZE")
    cw = _counter_means(d / "cw", "WRITE_SIZE")
    ck = max(cf, key=cf.get)  # the copy kernel dominates the calibration run
    fscale = copy_bytes / (cf[ck] * 1024.0)
    wscale = copy_bytes / (cw[max(cw, key=cw.get)] * 1024.0)
    # every kernel the step launches (the transport, the gradient pre-pass, the pass kernels, the
    # rounding counter); relayouts and statistics belong to set-up and read-back
    step = ["fesomtra::k_qflux", "fesomtra::k_mgrad", "fesomtra::k_edge", "fesomtra::k_lo", "fesomtra::k_vimpl",
            "fesomtra::k_cavity_aux", "fesomtra::k_bounds", "fesomtra::k_final", "fesomtra::k_limit", "fesomtra::k_vert",
            "fesomtra::k_copy", "fesomtra::k_seed", "fesomtra::k_ab2"]
    pa = profile_args(w)
    lines = [f"# {w}: rocprofv3 kernel trace and HBM counters of bench.py ({pa})", "",
             "Commands (each a run of its own; the copy calibration moves 1 GiB):",
             f"    rocprofv3 --kernel-trace --stats -f csv -d <dir>/k -o k -- python -u bench.py --workload {w} {pa}",
             f"    rocprofv3 --pmc FETCH_SIZE -f csv -d <dir>/f -o f -- python -u bench.py --workload {w} {pa}",
             f"    rocprofv3 --pmc WRITE_SIZE -f csv -d <dir>/w -o w -- python -u bench.py --workload {w} {pa}",
             "    rocprofv3 --pmc FETCH_SIZE -f csv -d <dir>/cf -o cf -- python -u bench.py --copy-calibration",
             "    rocprofv3 --pmc WRITE_SIZE -f csv -d <dir>/cw -o cw -- python -u bench.py --copy-calibration",
             f"    python bench.py --rocprof-summary <dir> --workload {w}",
             f"Counter scale from the copy: FETCH_SIZE x{fscale:.3f}, WRITE_SIZE x{wscale:.3f} (bytes per KiB counted / 1024)",
             f"Kernel source hash {kernel_source_hash()}", "",
             f"{'kernel':<28}{'calls':>7}{'avg_us':>10}{'min_us':>10}{'max_us':>10}{'HBM_read_MB':>13}{'HBM_write_MB':>14}"]
    rows = sorted(stats.items(), key=lambda kv: -kv[1][1])
    for k, (calls, tot, mn, mx) in rows:
        rd = fetch.get(k, 0.0) * 1024 * fscale / 1e6
        wr = write.get(k, 0.0) * 1024 * wscale / 1e6
        lines.append(f"{k:<28}{calls:>7}{tot / calls / 1e3:>10.1f}{mn / 1e3:>10.1f}{mx / 1e3:>10.1f}{rd:>13.1f}{wr:>14.1f}")
    present = [k for k in step if k in stats]
    # per step: k_qflux once, the pass kernels as often per step as they ran per k_qflux call
    nq = stats["fesomtra::k_qflux"][0]
    us, rd, wr = 0.0, 0.0, 0.0
    for k in present:
        calls, tot = stats[k][0], stats[k][1]
        per = calls / nq
        us += tot / nq / 1e3
        rd += fetch.get(k, 0.0) * 1024 * fscale * per
        wr += write.get(k, 0.0) * 1024 * wscale * per
    lines += ["", f"one step ({' + '.join(k.split('::')[1] for k in present)}): {us:.1f} us, "
              f"HBM read {rd / 1e6:.1f} MB + write {wr / 1e6:.1f} MB = {(rd + wr) / 1e6:.1f} MB"]
    try:
        commit = subprocess.run(["git", "-C", str(REPO), "rev-parse", "--short", "HEAD"], capture_output=True,
                                text=True).stdout.strip()
    except OSError:
        commit = ""
    out = REPO / "profiles"
    (out / f"kernels_{w}.txt").write_text("\n".join(lines) + "\n")
    (out / f"traffic_{w}.json").write_text(json.dumps({
        "workload": w, "hbm_bytes_per_step": int(rd + wr), "hbm_read_bytes_per_step": int(rd),
        "hbm_write_bytes_per_step": int(wr), "step_kernels": present, "step_us": us,
        "passes_per_step": round(stats["fesomtra::k_edge"][0] / nq) if "fesomtra::k_edge" in stats else None,
        "fetch_scale": fscale, "write_scale": wscale, "source_hash": kernel_source_hash(),
        "commit_before_measurement": commit, "source": f"profiles/kernels_{w}.txt"}, indent=1) + "\n")
    print("\n".join(lines))


def dropin_summary(d: Path, steps: int) -> None:
    """profiles/kernels_dropin_mesh1000.txt from the kernel trace DIR/k of `bench.py --dropin-only
    --steps S`: the kernels' durations, and over the timed steps (from the end of the warm-up step's
    last update kernel to the end of the last one) the elapsed time and the time the GPU was busy
    (union of the kernel intervals: the low-order kernel runs on a second stream beside the edge
    kernels, so their durations overlap and do not add up)."""
    stats = {}
    for r in _rocprof_rows(d / "k", "kernel_stats.csv"):
        k = _short(r["Name"])
        calls, tot = int(r["Calls"]), float(r["TotalDurationNs"])
        c0 = stats.get(k, (0, 0.0, float("inf"), 0.0))
        stats[k] = (c0[0] + calls, c0[1] + tot, min(c0[2], float(r["MinNs"])), max(c0[3], float(r["MaxNs"])))
    trace = sorted(_rocprof_rows(d / "k", "kernel_trace.csv"), key=lambda r: int(r["Start_Timestamp"]))
    fin = [r for r in trace if _short(r["Kernel_Name"]) == "fesomtra::k_final"]
    ntr = (len(fin) // (steps + 1)) or 1  # tracers per step (one warm-up step before the timed ones)
    t0, t1 = int(fin[ntr - 1]["End_Timestamp"]), int(fin[-1]["End_Timestamp"])
    win = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), _short(r["Kernel_Name"])) for r in trace
           if int(r["Start_Timestamp"]) >= t0 and int(r["End_Timestamp"]) <= t1]
    busy, cs, ce = 0, None, None
    for st, en, _ in sorted(win):
        if ce is None or st > ce:
            busy += 0 if ce is None else ce - cs
            cs, ce = st, en
        else:
            ce = max(ce, en)
    busy += 0 if ce is None else ce - cs
    nsteps = (len(fin) - ntr) // ntr
    per_kernel = {}
    for st, en, k in win:
        per_kernel.setdefault(k, []).append(en - st)
    cmd = f"python -u bench.py --dropin-only --steps {steps}"
    lines = [f"# do_oce_adv_tra drop-in sequence for T and S, 1000x1000x48 FP64, block width 1 ({cmd})", "",
             "Commands:",
             f"    rocprofv3 --kernel-trace --stats -f csv -d <dir>/k -o k -- {cmd}",
             f"    python bench.py --dropin-summary <dir> --steps {steps}",
             f"Kernel source hash {kernel_source_hash()}", "",
             "Whole run (set-up and teardown copies included):",
             f"{'kernel':<34}{'calls':>7}{'avg_us':>10}{'min_us':>10}{'max_us':>10}"]
    for k, (calls, tot, mn, mx) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"{k:<34}{calls:>7}{tot / calls / 1e3:>10.1f}{mn / 1e3:>10.1f}{mx / 1e3:>10.1f}")
    lines += ["", f"Timed steps ({nsteps} steps of {ntr} tracers, one do_oce_adv_tra call per tracer):",
              f"{'kernel':<34}{'calls':>7}{'avg_us':>10}"]
    for k, v in sorted(per_kernel.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"{k:<34}{len(v):>7}{sum(v) / len(v) / 1e3:>10.1f}")
    lines += ["", f"elapsed {(t1 - t0) / 1e6 / max(nsteps, 1):.3f} ms per step, GPU busy "
              f"{busy / 1e6 / max(nsteps, 1):.3f} ms per step (union of kernel intervals); the velocity is "
              "unchanged, so k_qflux ran once, before the timed steps"]
    (REPO / "profiles" / "kernels_dropin_mesh1000.txt").write_text("\n".join(lines) + "\n")
    print("\n".join(lines))


def dropin_timing(case, kb: int, steps: int = 10, keep_work: bool = False):
    """The reference's per-tracer do_oce_adv_tra call sequence on device arrays in reference order
    (the Fortran call sequence of the integration notes with the device-pointer entry points), per
    step and tracer: tracer_set_values_device (the caller's current values), tracer_zero_tendencies,
    tracer_do_oce_adv_tra and tracer_get_tendencies_device; the caller's own update of its arrays
    is not included.  keep_work: the call also stores the limited fluxes for tracer_get_work.
    Returns ms per step (all tracers) over `steps` steps after one warm-up step."""
    import ctypes
    import torch
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    t = FESOMTracerAdvection()
    t.set_block_width(kb)
    t.setup_case(case)
    t.set_keep_work(keep_work)
    vals = [torch.from_numpy(np.ascontiguousarray(case.tracers[i])).cuda() for i in range(case.ntracers)]
    dh = torch.empty_like(vals[0])
    dv = torch.empty_like(vals[0])
    L = t.lib
    nl, N = case.mesh.nl, case.mesh.nod2D
    P = ctypes.c_void_p

    def step():
        for i in range(case.ntracers):
            L.tracer_set_values_device(i + 1, P(vals[i].data_ptr()), nl, N)
            L.tracer_zero_tendencies()
            L.tracer_do_oce_adv_tra(case.dt, i + 1)
            L.tracer_get_tendencies_device(P(dh.data_ptr()), P(dv.data_ptr()), nl - 1, N)

    torch.cuda.synchronize()
    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    w = t.block_width()
    t.set_keep_work(True)
    t.set_block_width(0)
    t.cleanup()
    return ms, w


def describe(case):
    m = case.mesh
    sc = case.schemes[0]
    size = f"{m.nx}x{m.ny}" if m.nx else f"{m.nod2D} nodes"
    return (f"{case.name}: {size} periodic, nl={m.nl}, {case.ntracers} tracers, "
            f"{sc.hor}/{sc.ver}/{sc.lim}")


STUDY_ALL = ((1, True, False), (1, False, False), (1, False, True), (2, True, False), (2, False, False),
             (2, False, True), (3, True, False), (3, False, False), (3, False, True))
# the bench's own C2 line: each storage with the remainders kept (compensated) and with stochastic
# rounding
STUDY_BENCH = ((1, False, True), (2, False, True), (3, False, True), (1, True, False), (2, True, False),
               (3, True, False))


def precision_study(case, steps: int, modes=STUDY_ALL):
    """FP32 arithmetic with FP32 / FP16 / BF16 storage for `steps` steps against the FP64 GPU path
    (itself bit-identical to the reference restatement) on `case`: per mode (precision,
    stochastic rounding, compensated storage) and tracer the rel L2 / Linf over the wet layers and
    the drift of the volume-weighted sum."""
    from fesom2_tracer_dwarf_amd import configs
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection
    wet = configs.wet_mask(case.mesh)
    log(f"precision study on {describe(case)}, {steps} steps")
    results = {}
    ref = None
    for prec, sr, comp in ((0, True, False),) + tuple(modes):
        t = FESOMTracerAdvection()
        t.set_precision(prec)
        t.set_rounding(sr)
        t.set_compensated(comp)
        t.setup_case(case)
        v0 = [t.stats(i + 1)["volume_sum"] for i in range(case.ntracers)]
        ms = t.advect_steps(steps, case.dt)
        vals = t.get_all_values()
        v1 = [t.stats(i + 1)["volume_sum"] for i in range(case.ntracers)]
        t.set_precision(0)
        t.set_compensated(False)
        t.cleanup()
        if prec == 0:
            ref = vals
        per = []
        for i in range(case.ntracers):
            d = (vals[i] - ref[i])[wet]
            rr = ref[i][wet]
            per.append({"tracer": i + 1,
                        "rel_l2": float(np.sqrt(np.sum(d * d) / np.sum(rr * rr))),
                        "rel_linf": float(np.max(np.abs(d)) / np.max(np.abs(rr))),
                        "volume_sum_rel_drift": abs(v1[i] - v0[i]) / abs(v0[i])})
        key = PREC_NAMES[prec] + ("" if prec == 0 else (", compensated" if comp else
                                                          (", stochastic rounding" if sr else ", nearest")))
        results[key] = {"device_ms": ms, "tracers": per}
        log(f"{key}: {ms:.1f} ms, {per}")
    return {"study": "precision", "workload": describe(case), "steps": steps,
            "reference": "FP64 GPU path", "results": results}


def size_sweep(sizes, steps: int, warmup: int) -> None:
    """FP64 throughput against mesh size: n x n periodic meshes with 48 levels from the generator
    and seed of the 1000x1000x48 workload (random node numbering, smooth bathymetry), 2 tracers
    MFCT/QR4C/FCT, 48 levels per thread, the timed steps replayed as one HIP graph.  One JSON
    line per size on stdout."""
    from fesom2_tracer_dwarf_amd import configs
    for n in sizes:
        t0 = time.perf_counter()
        case = configs.bathymetry_case(f"mesh{n}", n, 48, seed=1, ntracers=2, nsteps=steps)
        m = case.mesh
        log(f"size {n}: N_L={m.n_node_levels} (generated in {time.perf_counter() - t0:.1f} s)")
        r = run(case, 0, 0, steps, warmup, graph=True, lc=[48, 48, 48, 48], tune=False)
        row = {"n": n, "node_levels": m.n_node_levels, "ms_per_step": r["wall_s"] * 1e3 / steps,
               "updates_per_s": r["value"], "stage_ms": [round(x, 3) for x in r["stage_ms"]],
               "qflux_ms": r["qflux_ms"], "block_width": r["block_width"]}
        log(f"size {n}: {row['ms_per_step']:.3f} ms/step, {r['value']:.3e} updates/s, stages {row['stage_ms']}")
        print(json.dumps(row), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: workload's)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mesh1000",
                    choices=["mesh1000", "mesh500x64", "multimesh32", "analytic20"])
    ap.add_argument("--precision", type=int, default=None, choices=[0, 1, 2, 3],
                    help="0 FP64, 1 FP32, 2 FP16 storage, 3 BF16 storage (default: workload's)")
    ap.add_argument("--block-width", type=int, default=0,
                    help="tracers interleaved per block and advanced per kernel pass (0: all)")
    ap.add_argument("--graph", action="store_true", help="replay the timed steps as one HIP graph")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every step's kernels directly, also for the workloads replayed as a graph "
                         "by default (profiled counter runs)")
    ap.add_argument("--fused", type=int, default=0, choices=[0, 1],
                    help="fused single-pass kernel for FP64 FCT tracers, one tracer per block")
    ap.add_argument("--fused-limiter", type=int, default=0, choices=[0, 1],
                    help="limiter bounds, ratios and update in one kernel where it applies (FP64, width <= 4)")
    ap.add_argument("--concurrent", type=int, default=1, choices=[0, 1],
                    help="low-order kernels on a second stream beside the edge-flux kernels")
    ap.add_argument("--lc", default="", help="levels per thread: edge,lo,bounds,final")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch configuration instead of autotuning it")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the mixed-precision batched companion run")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the timing of the per-tracer do_oce_adv_tra call sequence")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the 1000-step FP32/FP16/BF16 storage companion (C2) on the main mesh")
    ap.add_argument("--no-vertical", action="store_true",
                    help="skip the companion run of the main mesh with a vertical velocity")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the ensemble companion run (16 velocity members on the batched mesh)")
    ap.add_argument("--no-multimesh", action="store_true",
                    help="skip the 32-mesh union companion run")
    ap.add_argument("--batched-steps", type=int, default=200)
    ap.add_argument("--compensated", action="store_true",
                    help="single precision: keep each stored value's rounding remainder (tracer_set_compensated)")
    ap.add_argument("--rounding", default="stochastic", choices=["stochastic", "nearest"],
                    help="rounding of the tracer update in the single-precision modes")
    ap.add_argument("--no-batched-accuracy", action="store_true",
                    help="skip the FP64 run of the batched workload used for its error")
    ap.add_argument("--precision-study", action="store_true")
    ap.add_argument("--study-steps", type=int, default=1000)
    ap.add_argument("--size-sweep", default="",
                    help="comma-separated n: FP64 throughput on n x n x 48 meshes of the 1000x1000x48 "
                         "workload's generator (2 tracers, 48 levels per thread, HIP graph of 20 steps)")
    ap.add_argument("--dropin-only", action="store_true",
                    help="time only the do_oce_adv_tra drop-in sequence for T and S on the 1000x1000x48 "
                         "workload (block width 1, device arrays)")
    ap.add_argument("--dropin-summary", default="",
                    help="summarise the rocprofv3 kernel trace DIR/k of --dropin-only into "
                         "profiles/kernels_dropin_mesh1000.txt (no GPU)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (rocprof counter scale)")
    ap.add_argument("--rocprof-summary", default="",
                    help="summarise rocprofv3 CSV directories DIR/{k,f,w,cf,cw} into profiles/ (no GPU)")
    ap.add_argument("--sq-summary", default="",
                    help="summarise the two SQ counter passes DIR/{s1,s2} into profiles/sq_<workload>.txt (no GPU)")
    args = ap.parse_args()
    if args.rocprof_summary:
        rocprof_summary(args)
        return
    if args.sq_summary:
        sq_summary(args)
        return
    if args.dropin_summary:
        dropin_summary(Path(args.dropin_summary), args.steps or 10)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo design is documented, not built)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8,
                          "launches": 3}))
        return

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.precision_study:
        from fesom2_tracer_dwarf_amd import configs
        print(json.dumps(precision_study(configs.mesh1000(args.study_steps), args.study_steps)), flush=True)
        return
    if args.dropin_only:
        case = make_case("mesh1000")
        steps = args.steps or 10
        ms, w = dropin_timing(case, 1, steps=steps)
        log(f"drop-in sequence, block width {w}: {ms:.3f} ms per step")
        print(json.dumps({"workload": describe(case), "dropin_ms_per_step": ms, "block_width": w,
                          "steps": steps, "tracers": case.ntracers}), flush=True)
        return
    if args.size_sweep:
        size_sweep([int(x) for x in args.size_sweep.split(",")], args.steps or 20, args.warmup)
        return

    defaults = {"mesh1000": (100, 0, 0, False), "mesh500x64": (200, 1, 64, False),
                "multimesh32": (100, 1, MULTIMESH_WIDTH, True), "analytic20": (10, 0, 0, False)}
    d_steps, d_prec, d_kb, d_graph = defaults[args.workload]
    steps = args.steps or d_steps
    prec = d_prec if args.precision is None else args.precision
    fused = bool(args.fused) and prec == 0
    kb = 1 if fused else (args.block_width or d_kb)
    graph = (args.graph or d_graph) and not args.no_graph
    lc = [int(x) for x in args.lc.split(",")] if args.lc else None

    t0 = time.perf_counter()
    case = make_case(args.workload)
    m = case.mesh
    log(f"workload {describe(case)}: N={m.nod2D} T={m.elem2D} E={m.edge2D} N_L={m.n_node_levels} "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    r = run(case, prec, kb, steps, args.warmup, graph, lc, tune=not args.no_autotune, fused=fused,
            stochastic=args.rounding == "stochastic", fused_limiter=bool(args.fused_limiter),
            concurrent=bool(args.concurrent), compensated=args.compensated and prec != 0)
    log(f"timed {steps} steps: wall {r['wall_s'] * 1e3:.2f} ms, device {r['dev_ms']:.2f} ms, "
        f"{r['value']:.3e} updates/s, stages {r['stage_ms']}, block width {r['block_width']}, "
        f"halo tiles {r['tiles']}")
    roof = roofline(case, prec, r, args.workload)

    batched = None
    if args.workload == "mesh1000" and not args.no_batched:
        bcase = make_case("mesh500x64")
        log(f"batched companion {describe(bcase)}")
        ref64 = None
        if not args.no_batched_accuracy:
            r64 = run(bcase, 0, 64, args.batched_steps, 0, keep=True, tune=not args.no_autotune)
            ref64 = r64["values"]
            log(f"batched FP64 reference: {r64['value']:.3e} updates/s")
        # compensated FP32 storage (remainders kept: the sum conserved to FP64 round-off), and the
        # plain FP32 storage with the chosen rounding beside it
        rb = run(bcase, 1, 64, args.batched_steps, 0, compare=ref64, tune=not args.no_autotune,
                 compensated=True)
        rs = run(bcase, 1, 64, args.batched_steps, 0, compare=ref64, tune=not args.no_autotune,
                 stochastic=args.rounding == "stochastic")
        ref64 = None
        batched = {"workload": describe(bcase), "precision": "FP32 arithmetic and storage with the "
                   "rounding remainders kept (compensated), FP64 accumulation / update / limiter "
                   "ratios / sums", "block_width": rb["block_width"],
                   "steps": args.batched_steps, "value": rb["value"], "unit": "node-level updates/s",
                   "ms_per_step": rb["wall_s"] * 1e3 / args.batched_steps,
                   "vs_fp64_value": rb["value"] / r["value"],
                   "volume_sum_rel_drift": rb["drift"],
                   "error_vs_fp64": rb["error"],
                   "fp64_same_workload_value": None if args.no_batched_accuracy else r64["value"],
                   "vs_fp64_same_workload": None if args.no_batched_accuracy else rb["value"] / r64["value"],
                   "roofline": roofline(bcase, 1, rb, "mesh500x64"),
                   "plain_storage": {"rounding": args.rounding, "value": rs["value"],
                                     "ms_per_step": rs["wall_s"] * 1e3 / args.batched_steps,
                                     "volume_sum_rel_drift": rs["drift"], "error_vs_fp64": rs["error"]}}
        log(f"batched: {rb['value']:.3e} updates/s ({rb['value'] / r['value']:.2f}x FP64), drift {rb['drift']:.2e}; "
            f"plain storage {rs['value']:.3e}, drift {rs['drift']:.2e}")

    ensemble = None
    if args.workload == "mesh1000" and not args.no_batched and not args.no_ensemble:
        # the batched workload's mesh with 16 members of 4 tracers, each member its own velocity
        from fesom2_tracer_dwarf_amd import configs
        base = dataclasses.replace(bcase, tracers=bcase.tracers[:4], schemes=bcase.schemes[:4])
        members = configs.ensemble_cases(base, 16, seed=7)
        ecase = dataclasses.replace(base, name="mesh500_ensemble16x4",
                                    tracers=np.concatenate([c.tracers for c in members]),
                                    schemes=[sc for c in members for sc in c.schemes])
        re_ = run(ecase, 1, 64, args.batched_steps, 0, tune=not args.no_autotune, compensated=True,
                  members=members)
        ensemble = {"workload": "the batched workload's mesh and schemes, 16 ensemble members x 4 "
                                "tracers in one block of 64 lanes, each member with its own velocity "
                                "(tracer_set_ensemble)",
                    "members": 16, "tracers_per_member": 4, "precision": batched["precision"],
                    "steps": args.batched_steps, "value": re_["value"], "unit": "node-level updates/s",
                    "ms_per_step": re_["wall_s"] * 1e3 / args.batched_steps,
                    "vs_shared_velocity": re_["value"] / batched["value"],
                    "volume_sum_rel_drift": re_["drift"], "step_ms_kernels": re_["step_ms"],
                    "qflux_ms": re_["qflux_ms"],
                    "roofline": roofline(ecase, 1, re_, "mesh500_ensemble16x4", members=16)}
        log(f"ensemble: {re_['value']:.3e} updates/s ({ensemble['vs_shared_velocity']:.2f}x the shared "
            f"velocity), drift {re_['drift']:.2e}")

    multimesh = None
    if args.workload == "mesh1000" and not args.no_multimesh:
        t0 = time.perf_counter()
        mcase = make_case("multimesh32")
        log(f"multi-mesh companion {describe(mcase)} (generated in {time.perf_counter() - t0:.1f} s)")
        rm = run(mcase, 1, MULTIMESH_WIDTH, 100, 5, graph=True, tune=not args.no_autotune, compensated=True)
        dts = sorted(set(np.round(mcase.dt_node, 3).tolist()))
        multimesh = {"workload": describe(mcase), "meshes": 32, "nodes": mcase.mesh.nod2D,
                     "node_levels_per_tracer": mcase.mesh.n_node_levels, "tracers": mcase.ntracers,
                     "member_dt_s": [dts[0], dts[-1]], "precision": "FP32, compensated storage",
                     "launch": "the 32 meshes as one disjoint union, every member with its own dt; "
                               "100 steps replayed as one HIP graph",
                     "steps": 100, "value": rm["value"], "unit": "node-level updates/s",
                     "ms_per_step": rm["wall_s"] * 1e3 / 100, "volume_sum_rel_drift": rm["drift"],
                     "step_ms_kernels": rm["step_ms"], "block_width": rm["block_width"], "halo_tiles": rm["tiles"],
                     "roofline": roofline(mcase, 1, rm, "multimesh32")}
        log(f"multi-mesh: {rm['value']:.3e} updates/s")

    precision = None
    if args.workload == "mesh1000" and not args.no_precision:
        precision = precision_study(case, args.study_steps, STUDY_BENCH)

    vertical = None
    if args.workload == "mesh1000" and not args.no_vertical:
        # the main workload's mesh with an interior vertical velocity (largest vertical Courant
        # number 0.3), so the vertical schemes and the vertical limiter do real work
        from fesom2_tracer_dwarf_amd import configs
        from fesom2_tracer_dwarf_amd.configs import Scheme
        t0 = time.perf_counter()
        vcase = configs.bathymetry_case("mesh1000w", 1000, 48, seed=1, ntracers=2, nsteps=20, w_cfl=0.3)
        log(f"vertical companion: the main mesh with w (generated in {time.perf_counter() - t0:.1f} s)")
        vertical = {"workload": "the main workload's mesh and tracers with an interior vertical velocity "
                                "(largest vertical Courant number 0.3), FP64, block of 2, the main run's "
                                "launch configuration", "steps": 20}
        lc_main = list(r["launch_config"])
        for label, sc, split in (("MFCT/QR4C/FCT", Scheme(), False),
                                 ("MUSCL/PPM/FCT", Scheme("MUSCL", "PPM", "FCT", 0.9, 0.6), False),
                                 ("MFCT/QR4C/FCT use_wsplit", Scheme(), True)):
            c = dataclasses.replace(vcase, schemes=[sc, sc])
            if split:
                c = configs.split_vertical_velocity(c, 0.5)
            rv = run(c, 0, 0, 20, 2, lc=lc_main, tune=False, wsplit=split)
            vertical[label] = {"ms_per_step": rv["wall_s"] * 1e3 / 20, "value": rv["value"],
                               "volume_sum_rel_drift": rv["drift"]}
            log(f"vertical {label}: {rv['wall_s'] * 1e3 / 20:.3f} ms per step, drift {rv['drift']:.2e}")

    dropin = None
    if args.workload == "mesh1000" and not args.no_dropin:
        dropin = {"sequence": "per tracer: tracer_set_values_device, tracer_zero_tendencies, "
                              "tracer_do_oce_adv_tra, tracer_get_tendencies_device (device arrays in "
                              "reference order); tendencies only (tracer_set_keep_work(0)) unless "
                              "the key says keep_work", "tracers": case.ntracers}
        for kbw in (1, 0):
            ms, w = dropin_timing(case, kbw)
            dropin[f"ms_per_step_block_width_{w}"] = ms
            log(f"drop-in sequence, block width {w}: {ms:.3f} ms per step")
        best = min(v for k, v in dropin.items() if k.startswith("ms_per_step"))
        ms, w = dropin_timing(case, 1, keep_work=True)
        dropin[f"keep_work_ms_per_step_block_width_{w}"] = ms
        log(f"drop-in sequence keeping the limited fluxes, block width {w}: {ms:.3f} ms per step")
        dropin["vs_advect_steps"] = best / (r["wall_s"] * 1e3 / steps)

    cpu = None
    if not args.no_cpu_baseline and args.workload in ("mesh1000", "analytic20"):
        cpu = cpu_baseline(case)

    out = {
        "metric": "tracer node-level updates/s, FP64 vs mixed-prec batched; % HBM roofline",
        "value": r["value"],
        "unit": "node-level updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall_s"] * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": PREC_NAMES[prec],
        "data": "synthetic (seeded mesh, bathymetry, tracers and psi velocity)",
        "config": {"workload": describe(case), "nodes": m.nod2D, "elements": m.elem2D,
                   "edges": m.edge2D, "node_levels_per_tracer": m.n_node_levels,
                   "tracers": case.ntracers, "block_width": r["block_width"],
                   "launch_config": list(r["launch_config"]), "hip_graph": graph, "fused": fused,
                   "fused_limiter": r["fused_limiter"], "concurrent_low_order": bool(args.concurrent),
                   "dt_s": case.dt, "parallelism": "1 GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "volume_sum_rel_drift": r["drift"],
        "layout_autotune": r["layout"],
        "batched": batched,
        "multimesh": multimesh,
        "ensemble": ensemble,
        "precision": precision,
        "vertical": vertical,
        "dropin": dropin,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
