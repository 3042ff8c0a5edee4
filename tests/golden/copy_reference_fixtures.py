"""Copy the reference's frozen test DATA files (no scripts, no sources) into tests/golden/femm/.

Run once in a container that has the reference tree; the copies travel with the
repository so tests never read the reference at run time.  Every file is an input
or an expected output held by the reference's own tests (stock-FEMM solutions,
B-H tables and MATLAB-generated B-H samples, generic mesh artifacts).
"""
import hashlib
import shutil
import sys
from pathlib import Path

SOURCE = Path("/root/reference/gpu_solver/tests/fixtures")
TARGET = Path(__file__).resolve().parent / "femm"
FILES = [
    "35PN230_v1/35PN230.tab",
    "35PN230_v1/reference_samples.txt",
    "35PN230_v1/smoothing_reference.txt",
    "linear_square_v1/linear_square.ans",
    "linear_square_v1/linear_square.circuit.txt",
    "nonlinear_pm_coil_v1/nonlinear_pm_coil_planar.ans",
    "nonlinear_pm_coil_v1/nonlinear_pm_coil_planar.summary.txt",
    "nonlinear_pm_coil_v1/nonlinear_pm_coil_planar.postprocess.txt",
    "periodic_strip_v1/periodic_strip.gpu.json",
    "periodic_strip_v1/periodic_strip.ans",
    "periodic_strip_v1/periodic_strip.summary.txt",
    "antiperiodic_strip_v1/antiperiodic_strip.gpu.json",
    "antiperiodic_strip_v1/antiperiodic_strip.ans",
    "antiperiodic_strip_v1/antiperiodic_strip.summary.txt",
    "axisymmetric_coil_v1/axisymmetric_coil.gpu.json",
    "axisymmetric_coil_v1/axisymmetric_coil.ans",
    "axisymmetric_coil_v1/axisymmetric_coil.summary.txt",
]


def main():
    if not SOURCE.is_dir():
        sys.exit("reference fixtures not present")
    lines = []
    for rel in FILES:
        dst = TARGET / rel
        dst.parent.mkdir(parents=True, exist_ok=True)
        shutil.copyfile(SOURCE / rel, dst)
        lines.append(f"{hashlib.sha256(dst.read_bytes()).hexdigest()}  {rel}")
    (TARGET / "SHA256SUMS").write_text("\n".join(lines) + "\n", encoding="utf-8")


if __name__ == "__main__":
    main()
