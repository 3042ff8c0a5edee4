"""Wire-contract adapters from Python: one request file in, one response file out.

    python -m femm_gpu_amd.sample REQUEST.json RESPONSE.json          (--solve)
    python -m femm_gpu_amd.sample --motor-single-sample REQUEST RESPONSE
    python -m femm_gpu_amd.sample --motor-batch REQUEST RESPONSE

The adapters are the library's C++ command line (``femm_cli_main``, include/femm_wire.h;
femm-gpu_amd/csrc/host): strict request and artifact parsing, the artifact checks in the
reference's order, one batched solve per artifact through the C ABI, the weighted-stress and
air-gap postprocess, the reference's response keys, order and 17-digit numbers, and atomic
publication.  Exit status 0 when every response says PASS, 1 otherwise, 2 for a wrong
command line.  ``bin/femm_gpu_amd`` is the same command line as an executable.
"""

from __future__ import annotations

import sys

from .wire import cli

PLANAR_DC = "gpu_femm_planar_dc_sample_v1"
GENERIC = "gpu_femm_magnetostatic_sample_v1"
MOTOR = "gpu_femm_motor_sample_v1"
MOTOR_BATCH = "gpu_femm_motor_batch_v1"


def run_sample(request_path, response_path) -> int:
    """PlanarDcSingleSampleAdapter (planar-DC and magnetostatic sample requests)."""
    return cli(["--solve", request_path, response_path])


def run_motor_sample(request_path, response_path) -> int:
    """MotorSingleSampleAdapter (gpu_femm_motor_sample_v1)."""
    return cli(["--motor-single-sample", request_path, response_path])


def run_motor_batch(request_path, response_path) -> int:
    """MotorBatchAdapter (gpu_femm_motor_batch_v1)."""
    return cli(["--motor-batch", request_path, response_path])


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else list(argv)
    if len(argv) == 2 and not argv[0].startswith("--"):
        argv = ["--solve"] + argv
    sys.stdout.flush()
    sys.stderr.flush()
    return cli(argv)


if __name__ == "__main__":
    sys.exit(main())
