"""Group solves: meshes of uneven size in one cooperative launch (femm_newton_group).

Per item the group launch runs the same solve as Plan.solve_batch with the same number of
workgroups per item, so its outputs must be bitwise identical to the per-plan solves
(A, flux and every result field); against the FP64 oracle the usual bars hold (A within
1e-9 relative L2, flux within 1e-7 relative, step < 1e-6 and equation residual <= 1e-8).
"""

import sys

import numpy as np
import pytest

import oracle_lib as ol
from femm_gpu_amd import meshes
from femm_gpu_amd.model import accepted_options
from femm_gpu_amd.solver import RESULT_DTYPE, Group, Plan, mixed_config, reference_config

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _rel_l2(x, y):
    return float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))


def _ring(cells, seed, bh):
    return meshes.square_coil_ring(cells=cells, seed=seed, ring_bh=bh)


class _Entry:
    """Device buffers of one group entry and the per-plan reference solve."""

    def __init__(self, cells, items, partitions, seed, bh=True):
        self.model, base = _ring(cells, seed, bh)
        self.plan = Plan(self.model)
        scale = np.linspace(1.0, 0.4, items)[:, None]
        self.currents = np.ascontiguousarray(base[None, :] * scale)
        self.items, self.partitions = items, partitions
        self.cur = torch.from_numpy(self.currents).cuda()
        self.a = torch.zeros((items, self.plan.N), dtype=torch.float64, device="cuda")
        self.flux = torch.zeros((items, self.plan.C), dtype=torch.float64, device="cuda")
        self.res = torch.zeros(items * RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")

    def spec(self):
        return dict(plan=self.plan, items=self.items, partitions=self.partitions,
                    currents=self.cur.data_ptr(), a=self.a.data_ptr(), flux=self.flux.data_ptr(),
                    results=self.res.data_ptr())

    def outputs(self):
        return (self.a.cpu().numpy(), self.flux.cpu().numpy(),
                np.frombuffer(self.res.cpu().numpy().tobytes(), dtype=RESULT_DTYPE))


def _group_vs_plans(entries, cfg, opts):
    group = Group([e.spec() for e in entries])
    stream = torch.cuda.current_stream()
    group.solve(stream.cuda_stream, opts, cfg())
    torch.cuda.synchronize()
    for e in entries:
        a, flux, res = e.outputs()
        single = e.plan.solve_batch(e.currents, opts, cfg(partitions=e.partitions))
        say(f"  {e.model.name}: items {e.items} x {e.partitions} wg, newton "
            f"{res['newton_iterations'].tolist()} status {res['status'].tolist()}")
        assert np.array_equal(a, single["a"]), e.model.name
        assert np.array_equal(flux, single["flux"]), e.model.name
        assert res.tobytes() == single["results"].tobytes(), e.model.name
    return group


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_group_matches_per_plan_solves(precision):
    say(f"case group vs per-plan solves ({precision}): rings 8/12/16/24 cells")
    cfg = reference_config if precision == "fp64" else mixed_config
    entries = [_Entry(8, 2, 1, 1), _Entry(12, 1, 2, 2), _Entry(16, 3, 1, 3), _Entry(24, 1, 4, 4)]
    opts = accepted_options()
    group = _group_vs_plans(entries, cfg, opts)
    # A second launch reuses the uploaded table and gives the same bits.
    before = [e.outputs() for e in entries]
    group.solve(torch.cuda.current_stream().cuda_stream, opts, cfg())
    torch.cuda.synchronize()
    for e, (a, flux, res) in zip(entries, before):
        a2, flux2, res2 = e.outputs()
        assert np.array_equal(a, a2) and np.array_equal(flux, flux2)
        assert res.tobytes() == res2.tobytes()
    group.close()
    for e in entries:
        e.plan.close()


def test_group_matches_oracle():
    say("case group vs oracle: rings 10 (x2) and 20 (x1, 2 wg), FP64 reference form")
    entries = [_Entry(10, 2, 1, 7), _Entry(20, 1, 2, 8)]
    opts = accepted_options()
    group = Group([e.spec() for e in entries])
    group.solve(torch.cuda.current_stream().cuda_stream, opts, reference_config())
    torch.cuda.synchronize()
    for e in entries:
        a, flux, res = e.outputs()
        for b in range(e.items):
            ref = ol.newton_solve(e.model, e.currents[b], opts)
            say(f"  {e.model.name} item {b}: newton {res['newton_iterations'][b]} (oracle "
                f"{ref['newton']}) dA {_rel_l2(a[b], ref['a']):.2e}")
            assert res["status"][b] == 0 and ref["status"] == 0
            assert res["relative_step"][b] < 1e-6 and res["equation_residual"][b] <= 1e-8
            assert res["newton_iterations"][b] == ref["newton"]
            assert _rel_l2(a[b], ref["a"]) <= 1e-9
            assert np.abs(flux[b] - ref["flux"]).max() <= 1e-7 * np.abs(ref["flux"]).max()
    group.close()
    for e in entries:
        e.plan.close()


def test_group_autotune_fits_and_solves():
    say("case group autotune: linear rings (mu_r 1000) of 8..128 cells, one item each")
    cells = [8, 11, 16, 23, 32, 45, 64, 90, 128]
    entries = [_Entry(c, 1, 1, 20 + i, bh=False) for i, c in enumerate(cells)]
    group = Group([e.spec() for e in entries])
    parts, blocks = group.autotune(reference_config())
    say(f"  partitions {parts}, blocks {blocks}")
    assert blocks == sum(p * e.items for p, e in zip(parts, entries))
    assert all(p >= 1 and (p & (p - 1)) == 0 for p in parts)
    assert parts[-1] >= parts[0]
    for e, p in zip(entries, parts):
        e.partitions = p
    _group_vs_plans(entries, reference_config, accepted_options())
    for e in entries:
        _, _, res = e.outputs()
        assert res["status"][0] == 0
        assert res["relative_step"][0] < 1e-6 and res["equation_residual"][0] <= 1e-8
    group.close()
    for e in entries:
        e.plan.close()


@pytest.mark.parametrize("forcing", [0.0, 0.01])
def test_group_safeguarded_matches_per_plan_solves(forcing):
    """femm_newton_group_safeguarded: every entry's items bit for bit the per-plan
    femm_newton_batch_safeguarded solve at the same partition count (one and several
    workgroups per item), exact and with inexact forcing; every item passes both gates."""
    say(f"case group safeguarded vs per-plan (forcing {forcing}): rings 8/12/16/24 cells")
    entries = [_Entry(8, 2, 1, 1), _Entry(12, 1, 2, 2), _Entry(16, 3, 1, 3), _Entry(24, 1, 4, 4)]
    opts = accepted_options()
    group = Group([e.spec() for e in entries])
    group.solve(torch.cuda.current_stream().cuda_stream, opts,
                reference_config(inexact_forcing=forcing), safeguarded=True)
    torch.cuda.synchronize()
    for e in entries:
        a, flux, res = e.outputs()
        single = e.plan.solve_batch(e.currents, opts,
                                    reference_config(partitions=e.partitions,
                                                     inexact_forcing=forcing),
                                    safeguarded=True)
        say(f"  {e.model.name}: items {e.items} x {e.partitions} wg, newton "
            f"{res['newton_iterations'].tolist()} pcg {res['linear_iterations'].tolist()}")
        assert np.all(res["status"] == 0)
        assert np.all(res["relative_step"] < 1e-6) and np.all(res["equation_residual"] <= 1e-8)
        assert np.array_equal(a, single["a"]), e.model.name
        assert np.array_equal(flux, single["flux"]), e.model.name
        assert res.tobytes() == single["results"].tobytes(), e.model.name
    group.close()
    for e in entries:
        e.plan.close()


@pytest.mark.parametrize("form", ["driver_fp64", "safeguarded_mixed_block4"])
def test_group_graph_replays_equal_eager_solves(form):
    """femm_group_capture / femm_group_replay: every replay re-reads the currents and rewrites
    the outputs exactly as the eager group solve does (bitwise), with eager work between the
    replays; a replay after a plan's workspace grew is refused until the group is captured
    again."""
    say(f"case group graph ({form}): rings 8/12/20 cells, replays vs eager solves")
    safeguarded = form.startswith("safeguarded")
    cfg = (lambda **k: mixed_config(block_size=4, inexact_forcing=0.01, **k)) if safeguarded \
        else reference_config
    opts = accepted_options()
    entries = [_Entry(8, 2, 1, 31), _Entry(12, 1, 2, 32), _Entry(20, 2, 2, 33)]
    group = Group([e.spec() for e in entries])
    stream = torch.cuda.current_stream()

    def run(scale, replay):
        for e in entries:
            e.cur.copy_(torch.from_numpy(e.currents * scale).cuda())
            e.a.fill_(float("nan"))
            e.flux.fill_(float("nan"))
            e.res.fill_(255)
        if replay:
            group.replay(stream.cuda_stream)
        else:
            group.solve(stream.cuda_stream, opts, cfg(), safeguarded=safeguarded)
        torch.cuda.synchronize()
        return [e.outputs() for e in entries]

    eager = {s: run(s, False) for s in (1.0, 0.7)}
    for e, (a, flux, res) in zip(entries, eager[1.0]):
        say(f"  {e.model.name}: items {e.items} x {e.partitions} wg, status "
            f"{res['status'].tolist()} newton {res['newton_iterations'].tolist()}")
        assert (res["status"] == 0).all()
    group.capture(opts, cfg(), safeguarded=safeguarded)
    for k, scale in enumerate((1.0, 0.7, 1.0)):
        got = run(scale, True)
        for e, (a, flux, res), (ra, rflux, rres) in zip(entries, got, eager[scale]):
            assert np.array_equal(a, ra) and np.array_equal(flux, rflux), (k, e.model.name)
            assert res.tobytes() == rres.tobytes(), (k, e.model.name)
        # eager work between replays: a host solve on one of the group's plans
        entries[0].plan.solve_batch(entries[0].currents, opts, cfg(partitions=1))
        say(f"  replay {k} (scale {scale}): bitwise equal to the eager solve")
    # A larger batch reallocates the plan's workspace: the graph no longer matches it.
    big = np.repeat(entries[2].currents, 3, axis=0)
    entries[2].plan.solve_batch(big, opts, cfg(partitions=1))
    with pytest.raises(RuntimeError):
        group.replay(stream.cuda_stream)
    group.capture(opts, cfg(), safeguarded=safeguarded)
    got = run(0.7, True)
    for (a, flux, res), (ra, rflux, rres) in zip(got, eager[0.7]):
        assert np.array_equal(a, ra) and np.array_equal(flux, rflux)
        assert res.tobytes() == rres.tobytes()
    say("  refused after the workspace grew; replay after a new capture equal")
    group.close()
    for e in entries:
        e.plan.close()
