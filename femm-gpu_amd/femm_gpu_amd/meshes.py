"""Seeded mesh generators for the benchmark and parity workloads.

All generators return a :class:`~femm_gpu_amd.model.Model` in SI units with
counter-clockwise P1 triangles.  They are preprocessing only (the reference
meshes with Triangle outside the solve); node counts and seeds are recorded in
``Model.name``.

* :func:`square_coil_ring`  -- planar linear air box with a square coil and an
  iron ring (structured grid, jittered interior).
* :func:`motor_sector`      -- planar 12-slot / 4-pole surface-PM motor, one pole
  pitch (90 deg) with anti-periodic sector boundaries, B-H iron, 3 phase coils.
* :func:`c_core`            -- planar C-core with an air gap, an NdFeB block and two
  coils, Delaunay-meshed from seeded points.
* :func:`solenoid_rz`       -- axisymmetric solenoid with a plunger.
"""

from __future__ import annotations

import math

import numpy as np

from .model import AXISYMMETRIC, MU0, PLANAR, Material, Model


def tanh_bh_curve(ms=1.9, a=300.0, h_max=2e5, points=40):
    """40 monotone knots of B = Ms tanh(H/a) + mu0 H on H in [0, h_max] (H = 0 included).

    H knots are H = 0 followed by a geometric grid from a/30 to h_max, so the knee
    near H ~ a is resolved.
    """
    h = np.concatenate([[0.0], np.geomspace(a / 30.0, h_max, points - 1)])
    b = ms * np.tanh(h / a) + MU0 * h
    return b, h


def _area(xy, tri):
    p = xy[tri]
    return 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                        - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))


def _ccw(xy, tri):
    p = xy[tri]
    det = ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
           - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    tri = tri.copy()
    flip = det < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    return tri


def _structured_square(cells, side, seed, jitter=0.2):
    n = cells + 1
    h = side / cells
    ij = np.stack(np.meshgrid(np.arange(n), np.arange(n), indexing="xy"), -1).reshape(-1, 2)
    xy = (ij * h - side / 2.0).astype(np.float64)
    interior = (ij[:, 0] > 0) & (ij[:, 0] < cells) & (ij[:, 1] > 0) & (ij[:, 1] < cells)
    rng = np.random.default_rng(seed)
    xy[interior] += rng.uniform(-jitter, jitter, size=(int(interior.sum()), 2)) * h
    tri = []
    for y in range(cells):
        for x in range(cells):
            ll = y * n + x
            lr, ul = ll + 1, ll + n
            ur = ul + 1
            tri.append((ll, lr, ur))      # fixtures.cpp:33-42 split
            tri.append((ll, ur, ul))
    boundary = np.nonzero(~interior)[0].astype(np.int32)
    return xy, np.asarray(tri, np.int32), boundary


def square_coil_ring(cells=64, seed=12044, side=0.2, coil_side=0.04, ring_in=0.03, ring_out=0.04,
                     coil_j=1e6, ring_mu=1000.0):
    """Air box (side 0.2 m) on a cells x cells grid, 2 triangles per cell, interior nodes
    jittered by U(-0.2, 0.2) h.  Centred square coil (J = 1e6 A/m^2 as turns = area and
    1e6 A) and an iron ring (mu_r = 1000) by element centroid.  A = 0 on the box.

    Returns (model, currents)."""
    xy, tri, boundary = _structured_square(cells, side, seed)
    c = xy[tri].mean(axis=1)
    r = np.hypot(c[:, 0], c[:, 1])
    coil = (np.abs(c[:, 0]) <= coil_side / 2) & (np.abs(c[:, 1]) <= coil_side / 2)
    ring = (r >= ring_in) & (r <= ring_out) & ~coil
    label = np.where(coil, 1, np.where(ring, 2, 0)).astype(np.int32)
    area = _area(xy, tri)
    coil_area = float(area[coil].sum())
    mats = [Material(1.0 / MU0, group=0),
            Material(1.0 / MU0, source_j_per_a=coil_area / coil_area, circuit=0, group=1),
            Material(1.0 / (MU0 * ring_mu), group=2)]
    model = Model(xy, tri, label, mats, [], boundary, np.zeros(len(boundary)), None, 1, 1.0,
                  PLANAR, name=f"square_coil_ring_{cells}x{cells}_seed{seed}")
    return model, np.array([coil_j])


# Motor geometry (mm) of the 12-slot / 4-pole surface-PM sector.
_MOTOR_BANDS = [  # (r_start, r_end, intervals)
    (10.0, 30.0, 20),    # rotor iron
    (30.0, 34.0, 12),    # magnets / inter-pole air
    (34.0, 35.0, 6),     # air gap
    (35.0, 37.0, 8),     # tooth tips / slot openings
    (37.0, 58.0, 50),    # slots and teeth
    (58.0, 75.0, 25),    # stator back iron
]


def motor_sector(angular_cells=480, seed=20251203, ms=1.9, hc=(9e5, 9e5), fill=(1.0, 1.0, 1.0),
                 jitter=0.15, radial_scale=1):
    """One pole pitch (90 deg) of a 12-slot / 4-pole surface-PM machine, anti-periodic.

    Structured polar grid: radial bands of _MOTOR_BANDS (x radial_scale) by
    `angular_cells` sectors (must be a multiple of 160 so every slot / magnet edge lies
    on a grid line); interior nodes jittered by U(-jitter, jitter) of the local spacing.
    The default (480) gives 122 x 481 = 58,682 nodes and 116,160 triangles.

    Materials: rotor and stator iron with B = Ms tanh(H/300) + mu0 H (40 knots), two
    parallel-magnetised NdFeB blocks (mu_r 1.05, Hc from `hc`, angles 27 and 63 deg),
    air, and three slot coils carrying phases A, -C, B (turns = +-area times the slot
    fill factor, so the phase current is the current density in A/m^2).  A = 0 on the
    shaft (r = 10 mm) and outer (r = 75 mm) arcs; nodes on theta = 0 and theta = 90 deg
    are anti-periodic pairs.  Returns (model, nominal currents)."""
    if angular_cells % 160:
        raise ValueError("angular_cells must be a multiple of 160")
    radii = [np.array([_MOTOR_BANDS[0][0]])]
    for r0, r1, k in _MOTOR_BANDS:
        radii.append(np.linspace(r0, r1, k * radial_scale + 1)[1:])
    r_nodes = np.concatenate(radii) * 1e-3
    nr, nt = len(r_nodes), angular_cells + 1
    theta = np.linspace(0.0, 0.5 * math.pi, nt)
    R, TH = np.meshgrid(r_nodes, theta, indexing="ij")       # node (i, j) = i * nt + j
    rng = np.random.default_rng(seed)
    dr = np.gradient(r_nodes)[:, None] * np.ones((1, nt))
    dth = (theta[1] - theta[0]) * np.ones((nr, nt))
    interior = np.zeros((nr, nt), bool)
    interior[1:-1, 1:-1] = True
    # Keep every material-interface grid line straight so labels follow the geometry.
    band_edges_mm = {b[0] for b in _MOTOR_BANDS} | {b[1] for b in _MOTOR_BANDS} | {38.0, 57.0}
    r_mm = r_nodes * 1e3
    on_radial_edge = np.array([any(abs(r - e) < 1e-9 for e in band_edges_mm) for r in r_mm])
    interior[on_radial_edge, :] = False
    cell_deg = 90.0 / angular_cells
    edge_deg = [9, 13.5, 16.5, 21, 39, 43.5, 46.5, 51, 69, 73.5, 76.5, 81, 45]
    for d in edge_deg:
        interior[:, int(round(d / cell_deg))] = False
    Rj = R + np.where(interior, rng.uniform(-jitter, jitter, R.shape) * dr, 0.0)
    THj = TH + np.where(interior, rng.uniform(-jitter, jitter, R.shape) * dth, 0.0)
    xy = np.stack([Rj * np.cos(THj), Rj * np.sin(THj)], -1).reshape(-1, 2)

    # cells (i, j): radial interval i between r_nodes[i], r_nodes[i+1]; angular j.
    rc = 0.5 * (r_mm[:-1] + r_mm[1:])[:, None]
    tc = ((np.arange(angular_cells) + 0.5) * cell_deg)[None, :]
    rc, tc = np.broadcast_arrays(rc, tc)
    AIR, ROTOR, STATOR, MAG1, MAG2, COIL1, COIL2, COIL3 = range(8)
    label = np.full(rc.shape, AIR, np.int32)
    label[rc < 30] = ROTOR
    label[(rc > 30) & (rc < 34) & (tc > 9) & (tc < 45)] = MAG1
    label[(rc > 30) & (rc < 34) & (tc > 45) & (tc < 81)] = MAG2
    stator = rc > 35
    slot_centres = (15.0, 45.0, 75.0)
    in_slot = np.zeros(rc.shape, bool)
    in_opening = np.zeros(rc.shape, bool)
    for s in slot_centres:
        in_slot |= (rc > 37) & (rc < 58) & (np.abs(tc - s) < 6)
        in_opening |= (rc > 35) & (rc < 37) & (np.abs(tc - s) < 1.5)
    label[stator & ~in_slot & ~in_opening] = STATOR
    for k, s in enumerate(slot_centres):
        coil = (rc > 38) & (rc < 57) & (np.abs(tc - s) < 6)
        label[coil] = COIL1 + k
    tri = []
    tlabel = []
    for i in range(nr - 1):
        for j in range(angular_cells):
            a, b = i * nt + j, i * nt + j + 1
            c, d = (i + 1) * nt + j, (i + 1) * nt + j + 1
            # alternate the diagonal so the grid has no preferred direction
            if (i + j) % 2 == 0:
                tri += [(a, c, d), (a, d, b)]
            else:
                tri += [(a, c, b), (b, c, d)]
            tlabel += [label[i, j]] * 2
    tri = _ccw(xy, np.asarray(tri, np.int32))
    tlabel = np.asarray(tlabel, np.int32)
    area = _area(xy, tri)
    b, h = tanh_bh_curve(ms=ms)
    mats = [Material(1.0 / MU0, group=0),
            Material(1.0, bh_curve=0, group=1),
            Material(1.0, bh_curve=0, group=2),
            Material(1.0 / (MU0 * 1.05), hc=hc[0], magnetization_deg=27.0, group=3),
            Material(1.0 / (MU0 * 1.05), hc=hc[1], magnetization_deg=63.0, group=4)]
    signs = (1.0, -1.0, 1.0)   # phases A, -C, B in the three slots of one pole pitch
    for k in range(3):
        coil_area = float(area[tlabel == COIL1 + k].sum())
        turns = signs[k] * fill[k] * coil_area
        mats.append(Material(1.0 / MU0, source_j_per_a=turns / coil_area, circuit=k,
                             group=5 + k))
    outer = np.concatenate([np.arange(nt), (nr - 1) * nt + np.arange(nt)]).astype(np.int32)
    pairs = [(i * nt, i * nt + nt - 1, -1) for i in range(1, nr - 1)]
    model = Model(xy, tri, tlabel, mats, [(b, h)], outer, np.zeros(len(outer)),
                  np.asarray(pairs, np.int32), 3, 0.05, PLANAR,
                  name=f"motor_sector_12s4p_{nr}x{nt}_seed{seed}")
    return model, np.array([5e6, -3e6, -2e6])


def motor_currents(count, seed, low=-8e6, high=8e6):
    """Per-sample phase current densities ~ U(low, high) A/m^2 (seeded)."""
    return np.random.default_rng(seed).uniform(low, high, size=(count, 3))


def c_core(edge_iron=0.5e-3, edge_air=2e-3, seed=2, scale=1.0):
    """C-shaped iron core with a 2 mm air gap, an NdFeB block (mu_r 1.05, Hc 9e5 A/m along
    x) in the back limb and two coils (+-5e5 A/m^2) around it, inside an air box.

    Seeded points (iron spacing `edge_iron`, air spacing growing from the iron spacing to
    `edge_air` with distance from the core) are Delaunay-triangulated; labels follow the
    element centroid.  The default gives ~60k nodes.  Returns (model, currents)."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    s = scale * 1e-3
    outer = 70 * s          # core outer half-size
    inner = 40 * s          # window half-size
    gap = 1 * s             # half of the 2 mm gap (right limb, centred on y = 0)
    box = 160 * s           # air box half-size
    mag = (-70 * s, -40 * s, -10 * s, 10 * s)   # x0, x1, y0, y1 of the NdFeB block

    def in_core(p):
        x, y = p[..., 0], p[..., 1]
        ring = (np.maximum(np.abs(x), np.abs(y)) <= outer) & ~(
            (np.abs(x) < inner) & (np.abs(y) < inner))
        gap_cut = (x > inner) & (np.abs(y) < gap)
        return ring & ~gap_cut

    def dist_core(p):
        x, y = np.abs(p[..., 0]), np.abs(p[..., 1])
        d_out = np.hypot(np.maximum(x - outer, 0), np.maximum(y - outer, 0))
        return d_out

    pts = []
    # iron + magnet: jittered lattice at edge_iron
    h = edge_iron
    g = np.arange(-outer, outer + h / 2, h)
    X, Y = np.meshgrid(g, g)
    P = np.stack([X.ravel(), Y.ravel()], -1)
    P = P[in_core(P)]
    P = P + rng.uniform(-0.15, 0.15, P.shape) * h
    pts.append(P)
    # window and near-field air: graded rejection sampling on a fine lattice
    hmin, hmax = edge_iron, edge_air
    g2 = np.arange(-box, box + hmin / 2, hmin)
    X, Y = np.meshgrid(g2, g2)
    Q = np.stack([X.ravel(), Y.ravel()], -1)
    Q = Q[~in_core(Q)]
    local = np.minimum(hmax, hmin + 0.08 * dist_core(Q))
    keep = rng.uniform(0, 1, len(Q)) < (hmin / local) ** 2
    Q = Q[keep] + rng.uniform(-0.2, 0.2, (int(keep.sum()), 2)) * hmin
    Q = Q[(np.abs(Q[:, 0]) < box - hmax / 2) & (np.abs(Q[:, 1]) < box - hmax / 2)]
    pts.append(Q)
    # box boundary
    nb = int(round(2 * box / hmax))
    edge = np.linspace(-box, box, nb + 1)
    B = np.concatenate([np.stack([edge, np.full_like(edge, -box)], -1),
                        np.stack([edge, np.full_like(edge, box)], -1),
                        np.stack([np.full_like(edge[1:-1], -box), edge[1:-1]], -1),
                        np.stack([np.full_like(edge[1:-1], box), edge[1:-1]], -1)])
    pts.append(B)
    xy = np.concatenate(pts)
    xy = np.unique(np.round(xy / (1e-6 * s)) * (1e-6 * s), axis=0)
    tri = Delaunay(xy).simplices.astype(np.int32)
    tri = _ccw(xy, tri)
    tri = tri[_area(xy, tri) > 1e-6 * edge_iron ** 2]
    c = xy[tri].mean(axis=1)
    AIR, IRON, MAG, COIL_P, COIL_N = range(5)
    label = np.full(len(tri), AIR, np.int32)
    label[in_core(c)] = IRON
    in_mag = (c[:, 0] >= mag[0]) & (c[:, 0] <= mag[1]) & (c[:, 1] >= mag[2]) & (c[:, 1] <= mag[3])
    label[in_mag] = MAG
    # coils wound around the top limb: one side in the window, one outside the core
    cp = (c[:, 0] > -20 * s) & (c[:, 0] < 20 * s) & (c[:, 1] > 28 * s) & (c[:, 1] < 38 * s)
    cn = (c[:, 0] > -20 * s) & (c[:, 0] < 20 * s) & (c[:, 1] > 72 * s) & (c[:, 1] < 82 * s)
    label[cp] = COIL_P
    label[cn] = COIL_N
    area = _area(xy, tri)
    bcurve, hcurve = tanh_bh_curve()
    mats = [Material(1.0 / MU0, group=0), Material(1.0, bh_curve=0, group=1),
            Material(1.0 / (MU0 * 1.05), hc=9e5, magnetization_deg=0.0, group=2)]
    for k, lab in enumerate((COIL_P, COIL_N)):
        a = float(area[label == lab].sum())
        mats.append(Material(1.0 / MU0, source_j_per_a=a / a, circuit=k, group=3 + k))
    used = np.unique(tri)
    remap = -np.ones(len(xy), np.int64)
    remap[used] = np.arange(len(used))
    xy = xy[used]
    tri = remap[tri].astype(np.int32)
    on_box = (np.abs(np.abs(xy[:, 0]) - box) < 1e-12) | (np.abs(np.abs(xy[:, 1]) - box) < 1e-12)
    dn = np.nonzero(on_box)[0].astype(np.int32)
    model = Model(xy, tri, label, mats, [(bcurve, hcurve)], dn, np.zeros(len(dn)), None, 2, 0.05,
                  PLANAR, name=f"c_core_{len(xy)}_seed{seed}")
    return model, np.array([5e5, -5e5])


def solenoid_rz(radial_cells=64, axial_cells=128, seed=4, plunger_mu=None):
    """Axisymmetric solenoid with an iron plunger (r-z, x = r).  Structured grid on
    r in [0, 40] mm, z in [-40, 40] mm; plunger r < 5 mm, |z| < 20 mm (B-H iron);
    coil 8 mm < r < 14 mm, |z| < 15 mm, J = 2e6 A/m^2.  A = 0 on the axis and box."""
    nr, nz = radial_cells + 1, axial_cells + 1
    r = np.linspace(0.0, 40e-3, nr)
    z = np.linspace(-40e-3, 40e-3, nz)
    R, Z = np.meshgrid(r, z, indexing="xy")
    xy = np.stack([R.ravel(), Z.ravel()], -1)
    rng = np.random.default_rng(seed)
    ii = np.arange(nr * nz)
    col, row = ii % nr, ii // nr
    interior = (col > 0) & (col < nr - 1) & (row > 0) & (row < nz - 1)
    xy[interior] += rng.uniform(-0.1, 0.1, (int(interior.sum()), 2)) * np.array(
        [r[1] - r[0], z[1] - z[0]])
    tri = []
    for y in range(nz - 1):
        for x in range(nr - 1):
            ll = y * nr + x
            tri += [(ll, ll + 1, ll + nr + 1), (ll, ll + nr + 1, ll + nr)]
    tri = _ccw(xy, np.asarray(tri, np.int32))
    c = xy[tri].mean(axis=1)
    label = np.zeros(len(tri), np.int32)
    label[(c[:, 0] < 5e-3) & (np.abs(c[:, 1]) < 20e-3)] = 1
    coil = (c[:, 0] > 8e-3) & (c[:, 0] < 14e-3) & (np.abs(c[:, 1]) < 15e-3)
    label[coil] = 2
    b, h = tanh_bh_curve()
    mats = [Material(1.0 / MU0), Material(1.0, bh_curve=0) if plunger_mu is None else
            Material(1.0 / (MU0 * plunger_mu)), Material(1.0 / MU0, source_j_per_a=1.0, circuit=0)]
    boundary = np.nonzero(~interior)[0].astype(np.int32)
    return (Model(xy, tri, label, mats, [(b, h)], boundary, np.zeros(len(boundary)), None, 1, 1.0,
                  AXISYMMETRIC, name=f"solenoid_rz_{nr}x{nz}_seed{seed}"), np.array([2e6]))
