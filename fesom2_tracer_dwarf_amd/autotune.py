"""Autotuner of the launch configuration: the tracer layout (block width, LDS tiles) and the levels
per thread of each of the four kernels.

``tune_layout`` times whole steps (the device time of a few steps, HIP events on the library's
stream) for each candidate tracer block width -- how many tracers one kernel pass advances,
interleaved in HBM -- and, for the narrow blocks (width <= 4), with and without the LDS tiles of
the edge and node kernels and, on tiles, with the limiter as two kernels or as the fused kernel
where it applies, each with the low-order kernels on a second stream or on the main one; it keeps
the fastest.  The tracers are carried over exactly when the layout changes and the trial steps are
undone with a device checkpoint at every candidate.

Each column is split into chunks of ``lc`` layers, one thread per chunk and tracer lane.  Short
chunks expose more parallelism; long chunks avoid recomputing the vertical stencils at chunk
boundaries and keep more work per thread.  The best split depends on the mesh depth, the tracer
block width and the precision, so it is measured: every candidate is applied to all four kernels
at once, the per-kernel HIP-event times are recorded, and each kernel keeps its own fastest
candidate (the kernels are timed separately, so one sweep tunes all four).  The tracer state is
checkpointed on the device before the trials and restored afterwards.

Results are cached per workload signature in a JSON file (``FESOMTRA_TUNE_CACHE``, default
``~/.cache/fesomtra/autotune.json``).
"""

from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Dict, Iterable, Optional, Sequence

CANDIDATES = (8, 16, 24, 48)


def _cache_path() -> Path:
    return Path(os.environ.get("FESOMTRA_TUNE_CACHE",
                               Path.home() / ".cache" / "fesomtra" / "autotune.json"))


def signature(case, precision: int, block_width: int) -> str:
    m = case.mesh
    sc = case.schemes[0]
    return (f"N{m.nod2D}-E{m.edge2D}-nl{m.nl}-NL{m.n_node_levels}-tr{case.ntracers}-kb{block_width}"
            f"-p{precision}-{sc.hor}-{sc.ver}-{sc.lim}")


def width_candidates(ntracers: int) -> list:
    """Block widths worth timing: powers of two from max(1, ntr/8) up to min(64, next power of
    two of ntr)."""
    top = 1
    while top < min(ntracers, 64):
        top *= 2
    lo = max(1, top // 8)
    out, w = [], lo
    while w <= top:
        out.append(w)
        w *= 2
    return out


def tune_layout(tracer, case, widths: Optional[Sequence[int]] = None, steps: int = 3,
                verbose: bool = False) -> Dict:
    """Chooses and applies the fastest tracer layout for the case set up in ``tracer``: block
    width, tiles, limiter form and stream use.  Returns {"block_width", "tiles", "fused_limiter",
    "concurrent", "step_ms": {"<width>/<tiles>/<limiter>/<streams>": ms}}."""
    lib = tracer.lib
    tracer.set_graph(False)
    widths = list(widths or width_candidates(case.ntracers))
    table = {}
    for w in widths:
        for tiles in ((True, False) if w <= 4 else (True,)):
            for fused in ((False, True) if (w <= 4 and tiles) else (False,)):
                tracer.set_block_width(w)
                tracer.set_halo_tiles(tiles)
                tracer.set_fused_limiter(fused)
                if fused and not tracer.fused_limiter_active():
                    continue  # the two rings of some tile do not fit, or a pass has no limiter
                for conc in (True, False):
                    tracer.set_concurrent(conc)
                    if lib.tracer_checkpoint(0) != 0:
                        raise RuntimeError("tracer_checkpoint failed")
                    tracer.advect_steps(1, case.dt)  # warm-up (and tile set-up)
                    ms = tracer.advect_steps(steps, case.dt) / steps
                    if lib.tracer_checkpoint(1) != 0:
                        raise RuntimeError("tracer_checkpoint restore failed")
                    table[(w, tiles, fused, conc)] = ms
                    if verbose:
                        print(f"[autotune] width {w} tiles {tiles} fused limiter {fused} two streams {conc}: "
                              f"{ms:.4f} ms/step", flush=True)
    (w, tiles, fused, conc) = min(table, key=table.get)
    tracer.set_block_width(w)
    tracer.set_halo_tiles(tiles)
    tracer.set_fused_limiter(fused)
    tracer.set_concurrent(conc)
    return {"block_width": w, "tiles": tiles, "fused_limiter": fused, "concurrent": conc,
            "step_ms": {f"{k[0]}/{'tiles' if k[1] else 'untiled'}/{'fused' if k[2] else 'two-kernel'}"
                        f"/{'two-streams' if k[3] else 'one-stream'}": v for k, v in table.items()}}


def tune(tracer, case, precision: int, candidates: Sequence[int] = CANDIDATES, steps: int = 3,
         use_cache: bool = True, verbose: bool = False) -> Dict:
    """Chooses and applies the fastest levels-per-thread of each kernel for the case already set
    up in ``tracer`` (a FESOMTracerAdvection).  Returns {"launch_config", "stage_ms", "cached"}."""
    key = signature(case, precision, tracer.block_width())
    path = _cache_path()
    if use_cache and path.exists():
        try:
            hit = json.loads(path.read_text()).get(key)
        except (OSError, ValueError):
            hit = None
        if hit:
            tracer.set_launch_config(hit["launch_config"])
            return dict(hit, cached=True)
    lib = tracer.lib
    tracer.set_graph(False)
    if lib.tracer_checkpoint(0) != 0:
        raise RuntimeError("tracer_checkpoint failed")
    table = {}
    for lc in candidates:
        tracer.set_launch_config((lc, lc, lc, lc))
        tracer.advect_steps(1, case.dt)  # warm-up
        lib.tracer_set_kernel_timing(1)
        tracer.advect_steps(steps, case.dt)
        table[lc], _ = tracer.kernel_times()
        lib.tracer_set_kernel_timing(0)
        if verbose:
            print(f"[autotune] lc={lc}: {[round(x, 4) for x in table[lc]]} ms", flush=True)
    if lib.tracer_checkpoint(1) != 0:
        raise RuntimeError("tracer_checkpoint restore failed")
    best = [min(candidates, key=lambda c: table[c][k]) for k in range(4)]
    stage = [table[best[k]][k] for k in range(4)]
    tracer.set_launch_config(best)
    out = {"launch_config": best, "stage_ms": stage,
           "sweep_ms": {str(c): table[c] for c in candidates}}
    if use_cache:
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            data = json.loads(path.read_text()) if path.exists() else {}
            data[key] = out
            path.write_text(json.dumps(data, indent=1))
        except (OSError, ValueError):
            pass
    return dict(out, cached=False)
