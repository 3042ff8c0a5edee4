"""femm_newton_batch_device_noalloc (no device allocation, workspace in the library's fixed
device array) against femm_newton_batch_device, through the C ABI: two calls in a row, the
second with a smaller batch, each equal to the existing entry point's output bit for bit (same
kernels, same launch shapes, zeroed workspace); the reference FP64 form and the chosen mixed
block-Jacobi safeguarded-acceptance form; a batch beyond the array and a plan that is not yet
resident are rejected with their status codes and launch nothing."""

import sys

import numpy as np
import pytest

from femm_gpu_amd import meshes
from femm_gpu_amd.model import accepted_options
from femm_gpu_amd.solver import RESULT_DTYPE, Plan, mixed_config, reference_config

pytestmark = pytest.mark.gpu

RESULT_BYTES = RESULT_DTYPE.itemsize


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _buffers(torch, plan, currents):
    B = currents.shape[0]
    return dict(cur=torch.from_numpy(np.ascontiguousarray(currents)).to("cuda"),
                a=torch.full((B, plan.N), np.nan, dtype=torch.float64, device="cuda"),
                flux=torch.full((B, plan.C), np.nan, dtype=torch.float64, device="cuda"),
                res=torch.zeros(B * RESULT_BYTES, dtype=torch.uint8, device="cuda"))


def _results(buf):
    return np.frombuffer(buf["res"].cpu().numpy().tobytes(), dtype=RESULT_DTYPE)


@pytest.mark.parametrize("form", ["fp64", "mixed_block4"])
def test_noalloc_equals_device_entry_point(form):
    import torch
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    say(f"case no-alloc solve ({form}): motor 160 cells, 3 items then 1 item")
    cfg = reference_config() if form == "fp64" else \
        mixed_config(block_size=4, inexact_forcing=0.01)
    model, _ = meshes.motor_sector(angular_cells=160)
    plan = Plan(model)
    opts = accepted_options()
    first = meshes.motor_currents(3, seed=61)
    second = meshes.motor_currents(1, seed=67)
    for k, cur in enumerate((first, second)):
        ref = _buffers(torch, plan, cur)
        plan.solve_batch_device(ref["cur"].data_ptr(), cur.shape[0], ref["a"].data_ptr(),
                                ref["flux"].data_ptr(), ref["res"].data_ptr(),
                                stream.cuda_stream, opts, cfg)
        out = _buffers(torch, plan, cur)
        st = plan.solve_batch_device_noalloc(out["cur"].data_ptr(), cur.shape[0],
                                             out["a"].data_ptr(), out["flux"].data_ptr(),
                                             out["res"].data_ptr(), stream.cuda_stream, opts, cfg)
        torch.cuda.synchronize()
        assert st == 0
        r, q = _results(ref), _results(out)
        say(f"  call {k + 1}: {cur.shape[0]} items, status {q['status'].tolist()}, newton "
            f"{q['newton_iterations'].tolist()} (existing {r['newton_iterations'].tolist()})")
        assert np.all(q["status"] == 0)
        assert np.array_equal(q, r)
        assert torch.equal(out["a"], ref["a"]) and torch.equal(out["flux"], ref["flux"])
    plan.close()


def test_noalloc_rejects_oversize_batches_and_cold_plans():
    import torch
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    say("case no-alloc limits: cold plan, oversize batch")
    model, _ = meshes.motor_sector(angular_cells=160)
    plan = Plan(model)
    cur = meshes.motor_currents(1, seed=3)
    buf = _buffers(torch, plan, cur)
    # A fresh plan is not resident on the device: uploading it would allocate.
    st = plan.solve_batch_device_noalloc(buf["cur"].data_ptr(), 1, buf["a"].data_ptr(),
                                         buf["flux"].data_ptr(), buf["res"].data_ptr(),
                                         stream.cuda_stream, accepted_options(),
                                         reference_config())
    assert st == 2
    plan.solve_batch_device(buf["cur"].data_ptr(), 1, buf["a"].data_ptr(),
                            buf["flux"].data_ptr(), buf["res"].data_ptr(), stream.cuda_stream,
                            accepted_options(), reference_config())
    torch.cuda.synchronize()
    # About 9 MB of mixed block workspace per item of this mesh: 64 items exceed 128 MiB.
    big = meshes.motor_currents(64, seed=5)
    bb = _buffers(torch, plan, big)
    st = plan.solve_batch_device_noalloc(bb["cur"].data_ptr(), 64, bb["a"].data_ptr(),
                                         bb["flux"].data_ptr(), bb["res"].data_ptr(),
                                         stream.cuda_stream, accepted_options(),
                                         mixed_config(block_size=4))
    torch.cuda.synchronize()
    assert st == 7
    assert torch.isnan(bb["a"]).all()          # nothing was launched
    plan.close()
