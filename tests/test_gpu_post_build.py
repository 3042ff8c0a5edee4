"""Weighted-stress postprocess plan built on the GPU (femm_plan_set_postprocess).

The plan (element classes, edge and fixed-value rules, mask Laplace solve on the solve plan's
pattern, mask gradients, node incidence, probe location) is built by kernels; the oracle
restates the reference's host build (weighted_stress.cu:24-183, field_postprocess.cpp:221-339).
Every case compares the status with the oracle's; accepted cases compare the node mask
bitwise and the loads and radial B of one state (relative 1e-12, they differ from the
oracle's only by summation order).  The rule cases pin the oracle too: their expected status
is written out.
"""

import math
import sys

import numpy as np
import pytest

import femm_fixtures as ff
import oracle_lib as ol
from femm_gpu_amd import meshes
from femm_gpu_amd.model import MU0, Material, Model
from femm_gpu_amd.solver import FemmError, Plan

pytestmark = pytest.mark.gpu

OK, INVALID_ARGUMENT, BOUNDARY_INVALID, INVALID_MATERIAL = 0, 2, 5, 15
IRON, BODY, COIL, AIR = 0, 1, 2, 3


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def grid(n, label_of, h=1e-3, dirichlet=True, constraints=None):
    """n x n cells of side h centred on the origin, two CCW triangles per cell, labels from
    label_of(i, j) of the cell; Dirichlet A = 0 on the outer boundary."""
    idx = lambda i, j: j * (n + 1) + i
    xy = [((i - n / 2) * h, (j - n / 2) * h) for j in range(n + 1) for i in range(n + 1)]
    tri, lab = [], []
    for j in range(n):
        for i in range(n):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)
            tri += [[a, b, c], [a, c, d]]
            lab += [label_of(i, j)] * 2
    mats = [Material(1.0 / (MU0 * 1000.0)), Material(1.0 / (MU0 * 1000.0)),
            Material(1.0 / MU0), Material(1.0 / MU0)]
    border = [idx(i, j) for j in range(n + 1) for i in range(n + 1)
              if i in (0, n) or j in (0, n)] if dirichlet else []
    return Model(np.array(xy), np.array(tri), np.array(lab), mats, [], np.array(border, np.int32),
                 np.zeros(len(border)), constraints, 1, 1.0, name="grid")


def body_in_air(i, j, n=12):
    if 4 <= i < 8 and 4 <= j < 8:
        return BODY
    if i in (0, n - 1) or j in (0, n - 1):
        return IRON
    return AIR


def _state(model):
    x, y = model.node_xy[:, 0], model.node_xy[:, 1]
    return (0.3 + 2.0 * x - 1.5 * y + 40.0 * x * y + 25.0 * x * x)[None, :].copy()


def _compare(model, post, oracle_labels, expect=None, state=None):
    """GPU build vs oracle on one model; returns the status."""
    plan = Plan(model)
    try:
        plan.set_postprocess(**post)
        st = OK
    except FemmError as e:
        st = e.code
    a = _state(model) if state is None else state
    bx, by, _ = plan.finalize_batch(a)
    sl, al, sg, ag = oracle_labels
    ref_st, ref_loads, ref_br, ref_mask = ol.weighted_stress(
        model, a[0], bx[0], by[0], sl, al, sg, ag, post.get("radius_m", 0.0),
        post.get("angles_rad", ()))
    say(f"  status gpu {st} oracle {ref_st} (expected {expect})")
    assert st == ref_st
    if expect is not None:
        assert st == expect
    if st == OK:
        mask = plan.postprocess_mask()
        assert np.array_equal(mask, ref_mask)
        loads, radial = plan.postprocess_batch(a)
        scale = max(np.abs(ref_loads).max(), 1e-300)
        assert np.all(np.abs(loads[0] - ref_loads) <= 1e-12 * scale + 1e-300)
        if len(ref_br):
            assert np.all(np.abs(radial[0] - ref_br) <= 1e-12 * np.abs(ref_br).max() + 1e-300)
        say(f"  mask nodes at 1: {int(mask.sum())}, loads {loads[0]}")
    plan.close()
    return st


def _labels(sel=BODY, air=AIR):
    return dict(selected_label=sel, air_label=air), (sel, air, -1, -1)


@pytest.mark.parametrize("case", [
    "body_in_air", "edge_touch", "dirichlet_touch", "vertex_other_first",
    "vertex_body_first", "no_body", "all_fixed", "probe_outside", "probe_in_iron",
    "radius_nan", "periodic_pairs", "periodic_pair_to_boundary"])
def test_rules_match_oracle(case):
    say(f"case post build rule: {case}")
    n = 12
    post, labels = _labels()
    angles = [0.1, 1.3, 2.9, 4.4]
    post.update(radius_m=4.5e-3, angles_rad=angles, max_mask_iterations=2000)
    expect = OK
    if case == "body_in_air":
        model = grid(n, lambda i, j: body_in_air(i, j, n))
    elif case == "edge_touch":
        model = grid(n, lambda i, j: IRON if (i == 8 and j == 5) else body_in_air(i, j, n))
        expect = INVALID_MATERIAL
    elif case == "dirichlet_touch":
        model = grid(n, lambda i, j: BODY if (i < 4 and 4 <= j < 8) else AIR)
        expect = BOUNDARY_INVALID
    elif case == "vertex_other_first":
        # iron cell (3, 3) meets body cell (4, 4) at one node only; iron comes first
        model = grid(n, lambda i, j: IRON if (i, j) == (3, 3) else
                     (BODY if (i, j) == (4, 4) else AIR))
        expect = BOUNDARY_INVALID
    elif case == "vertex_body_first":
        model = grid(n, lambda i, j: BODY if (i, j) == (3, 3) else
                     (IRON if (i, j) == (4, 4) else AIR))
        post.update(radius_m=4.0e-3)
    elif case == "no_body":
        model = grid(n, lambda i, j: AIR if body_in_air(i, j, n) == BODY else body_in_air(i, j, n))
        expect = INVALID_MATERIAL
    elif case == "all_fixed":
        # every node has a fixed value: body below, one air row whose nodes all lie on the
        # body or the iron, iron above; Dirichlet on the top side only (iron nodes)
        model = grid(n, lambda i, j: BODY if j < 5 else (AIR if j == 5 else IRON),
                     dirichlet=False)
        model.dirichlet_nodes = np.arange(n * (n + 1), (n + 1) * (n + 1), dtype=np.int32)
        model.dirichlet_values = np.zeros(n + 1)
        post.update(angles_rad=[])
    elif case == "probe_outside":
        model = grid(n, lambda i, j: body_in_air(i, j, n))
        post.update(radius_m=9e-3)
        expect = INVALID_ARGUMENT
    elif case == "probe_in_iron":
        model = grid(n, lambda i, j: body_in_air(i, j, n))
        post.update(radius_m=5.8e-3, angles_rad=[0.0])
        expect = INVALID_ARGUMENT
    elif case == "radius_nan":
        model = grid(n, lambda i, j: body_in_air(i, j, n))
        post.update(radius_m=math.nan)
        expect = INVALID_ARGUMENT
    elif case == "periodic_pairs":
        # left and right sides tied anti-periodically (the mask ties them with sign +1)
        cons = [[j * (n + 1), j * (n + 1) + n, -1] for j in range(1, n)]
        model = grid(n, lambda i, j: COIL if i == 0 or i == n - 1 else
                     (BODY if 4 <= i < 8 and 4 <= j < 8 else AIR), dirichlet=False,
                     constraints=np.array(cons, np.int32))
        bottom_top = [i for i in range(n + 1)] + [n * (n + 1) + i for i in range(n + 1)]
        model.dirichlet_nodes = np.array(bottom_top, np.int32)
        model.dirichlet_values = np.zeros(len(bottom_top))
        post, labels = _labels(air=COIL)
        post.update(radius_m=3.0e-3, angles_rad=[0.3, 2.0], max_mask_iterations=2000)
        model.triangle_material[model.triangle_material == AIR] = COIL
    elif case == "periodic_pair_to_boundary":
        # a body node tied to a Dirichlet node: its class is held at 0 and 1
        cons = [[6 * (n + 1) + 4, 0, 1]]
        model = grid(n, lambda i, j: body_in_air(i, j, n), constraints=np.array(cons, np.int32))
        expect = BOUNDARY_INVALID
    _compare(model, post, labels, expect)


def test_pm_coil_and_ring_fixtures_match_oracle():
    say("case post build fixtures: pm_coil (frozen FEMM loads) and the uniform-field ring")
    model, _, _, _ = ff.nonlinear_pm_coil()
    post = ff.nonlinear_pm_coil_postprocess()
    _compare(model, dict(selected_label=post["selected_label"], air_label=3,
                         radius_m=post["radius_m"], angles_rad=post["angles"]),
             (post["selected_label"], 3, -1, -1), OK)
    ring = ff.postprocess_ring()
    _compare(ring, dict(selected_label=1, air_label=3, radius_m=0.75,
                        angles_rad=[0.0, 0.5 * math.pi]), (1, 3, -1, -1), OK)


def test_motor_sector_plan_matches_oracle():
    """Anti-periodic 12-slot sector (rotor group vs air group): the mask ties the sector
    edges with sign +1; mask bitwise, loads and 32 probes of one state vs the oracle."""
    say("case post build motor sector (160 angular cells), groups")
    model, _ = meshes.motor_sector(angular_cells=160)
    post = meshes.motor_postprocess()
    post.update(max_mask_iterations=20000)
    st = _compare(model, post, (-1, -1, post["selected_group"], post["air_group"]), OK)
    assert st == OK


def test_rebuild_replaces_plan():
    """A second set_postprocess replaces the first plan (new selection, fewer probes), and a
    failed build leaves no plan behind."""
    say("case post build: rebuild and failed rebuild")
    n = 12
    model = grid(n, lambda i, j: body_in_air(i, j, n))
    plan = Plan(model)
    a = _state(model)
    plan.set_postprocess(selected_label=BODY, air_label=AIR, radius_m=4.5e-3,
                         angles_rad=[0.1, 1.3, 2.9], max_mask_iterations=2000)
    first = plan.postprocess_batch(a)
    plan.set_postprocess(selected_label=BODY, air_label=AIR, radius_m=4.5e-3,
                         angles_rad=[1.3], max_mask_iterations=2000)
    second = plan.postprocess_batch(a)
    assert np.array_equal(first[0], second[0]) and second[1].shape == (1, 1)
    assert np.array_equal(first[1][:, 1], second[1][:, 0])
    with pytest.raises(FemmError):
        plan.set_postprocess(selected_label=BODY, air_label=AIR, radius_m=9e-3,
                             angles_rad=[0.0], max_mask_iterations=2000)
    with pytest.raises(FemmError):
        plan.postprocess_batch(a)
    plan.close()
