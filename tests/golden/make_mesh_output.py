"""Writes tests/golden/mesh_output_ref/: the files the compiled reference driver writes with
--save-mesh --save-scalars (lib/mesh_output.F90 through src/fesom_analytic.F90) for its periodic
6 x 5 analytic mesh with 4 levels and its default 10 steps.  Test data for the output formats
(tests/test_output.py); needs oracle/_ref built by oracle/ref/build_ref.sh.

    python tests/golden/make_mesh_output.py
"""
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
ANALYTIC = REPO / "oracle" / "_ref" / "fesom_tracer_analytic"
ARGS = ["6", "5", "4", "--periodic", "--save-mesh", "--save-scalars"]


def main():
    if not ANALYTIC.exists():
        raise SystemExit("oracle/_ref/fesom_tracer_analytic not built: run oracle/ref/build_ref.sh")
    dst = HERE / "mesh_output_ref"
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "out"
        r = subprocess.run([str(ANALYTIC), *ARGS, "--output-dir", str(out)], cwd=td,
                           capture_output=True, text=True, timeout=120, check=True)
        if dst.exists():
            shutil.rmtree(dst)
        shutil.copytree(out, dst)
        # the driver's console lines for steps 1-10 (T min, max, plain sum), for the test
        lines = [l.strip() for l in r.stdout.splitlines() if "T min, max, sum" in l]
        (dst / "console_steps.txt").write_text("\n".join(lines) + "\n")
    print("wrote", dst, sorted(p.name for p in dst.iterdir()), file=sys.stderr)


if __name__ == "__main__":
    main()
