"""Benchmark: node-level tracer updates/s of the FP64 tracer advection step on one MI355X.

Workload (default): 1000x1000 doubly periodic mesh (1M nodes, 2M triangles, 3M edges), 48 levels
stretched 5-250 m, smooth random bathymetry with 10-48 levels, randomly permuted numbering,
T and S advected with MFCT / QR4C / FCT in FP64 at CFL 0.5.  One step advances both tracers.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5]

Prints one JSON line.  ``value`` = tracers x wet node-levels x steps / wall time of the timed
steps (inputs resident in HBM).  ``roofline`` relates the algorithmic bytes of one tracer-step
(every input read once, the tracer written once, at storage precision) to the mean HIP-event
duration of the step's kernels; ``traffic`` is the rocprof-measured HBM traffic of the same step
(profiles/traffic_mesh1000.json).  ``cpu_baseline`` times the CPU restatement of the reference
on one tracer-step of the same mesh, single-threaded.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def algorithmic_bytes(mesh, k_tracers: int = 1, s_tr: int = 8, s_mesh: int = 8) -> int:
    """Bytes one step must move at least: every distinct input once, every output once.

    K tracers in and out, per node-layer mesh arrays (hnode, hnode_new, areasvol, Z_3d_n),
    per node-interface arrays (area, w, zbar_3d_n), per element-layer uv(2) and helem, and the
    per-edge / per-element / per-node connectivity, geometry and level integers.
    """
    N, T, E = mesh.nod2D, mesh.elem2D, mesh.edge2D
    NL = mesh.n_node_levels
    TL = mesh.n_elem_levels
    adj = float(np.mean(mesh.nod_in_elem2D_num))
    return int(k_tracers * 2 * s_tr * NL
               + s_mesh * (4 * NL + 3 * (NL + N) + 3 * TL)
               + E * (16 + 6 * s_mesh + 8)
               + T * (20 + 8 * s_mesh)
               + N * (20 + 4 * adj))


def cpu_baseline(case, sample_tracer_steps: int = 1):
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


def traffic_from_profiles(name: str):
    f = REPO / "profiles" / f"traffic_{name}.json"
    if not f.exists():
        return None, None
    d = json.loads(f.read_text())
    return d.get("hbm_bytes_per_launch"), d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mesh1000", choices=["mesh1000", "analytic20"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tracers-per-pass", type=int, default=1,
                    help="tracers advanced by one kernel pass (1 or 2)")
    ap.add_argument("--lc", default="", help="levels per thread: edge,lo,bounds,final")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (rocprof counter scale)")
    args = ap.parse_args()
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
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo design is documented, not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    from fesom2_tracer_dwarf_amd import configs
    from fesom2_tracer_dwarf_amd.tracer import FESOMTracerAdvection

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    case = configs.mesh1000() if args.workload == "mesh1000" else configs.analytic20()
    m = case.mesh
    print(f"[bench] workload {case.name}: N={m.nod2D} T={m.elem2D} E={m.edge2D} nl={m.nl} "
          f"N_L={m.n_node_levels} tracers={case.ntracers}", file=sys.stderr, flush=True)

    t = FESOMTracerAdvection()
    t.lib.tracer_set_tracers_per_pass(args.tracers_per_pass)
    if args.lc:
        lc = [int(x) for x in args.lc.split(",")]
        t.lib.tracer_set_launch_config(*lc)
    t.setup_case(case)
    v0 = [t.stats(i + 1)["volume_sum"] for i in range(case.ntracers)]
    t.advect_steps(args.warmup, case.dt)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = t.advect_steps(args.steps, case.dt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    updates = case.ntracers * m.n_node_levels * args.steps
    value = updates / wall
    drift = [abs(t.stats(i + 1)["volume_sum"] - v0[i]) / abs(v0[i]) for i in range(case.ntracers)]
    print(f"[bench] timed {args.steps} steps: wall {wall * 1e3:.2f} ms, device {dev_ms:.2f} ms",
          file=sys.stderr, flush=True)

    # Per-stage HIP-event timing (separate short run, outside the timed region).
    import ctypes
    t.lib.tracer_set_kernel_timing(1)
    t.advect_steps(min(args.steps, 10), case.dt)
    stage = (ctypes.c_double * 4)()
    cnt = ctypes.c_long()
    t.lib.tracer_get_kernel_times(stage, ctypes.byref(cnt))
    t.lib.tracer_set_kernel_timing(0)
    stage_ms = [float(x) for x in stage]
    step_ms = sum(stage_ms)
    t.cleanup()

    # one launch sequence advances `kt` tracers: its algorithmic bytes are those of K = kt
    kt = min(args.tracers_per_pass, case.ntracers)
    B = algorithmic_bytes(m, kt)
    achieved = B / (step_ms * 1e-3) / 1e9
    traffic, prof = traffic_from_profiles(case.name)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": f"one FP64 step of {kt} tracer(s): k_edge + k_lo + k_bounds + k_final",
            "algorithmic_bytes_per_launch": B,
            "tracers_per_launch": kt,
            "stage_ms_per_launch": {"k_edge": round(stage_ms[0], 4),
                                         "k_lo": round(stage_ms[1], 4),
                                         "k_bounds": round(stage_ms[2], 4),
                                         "k_final": round(stage_ms[3], 4)}}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(case)

    out = {
        "metric": "tracer node-level updates/s, FP64 vs mixed-prec batched; % HBM roofline",
        "value": value,
        "unit": "node-level updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded mesh, bathymetry, tracers and psi velocity)",
        "config": {"workload": f"{case.name}: {m.nx}x{m.ny} periodic, nl={m.nl}, "
                               f"{case.ntracers} tracers, MFCT/QR4C/FCT",
                   "nodes": m.nod2D, "elements": m.elem2D, "edges": m.edge2D,
                   "node_levels_per_tracer": m.n_node_levels, "tracers": case.ntracers,
                   "dt_s": case.dt, "parallelism": "1 GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "volume_sum_rel_drift": max(drift),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
