"""Seeded mesh generators for the benchmark and parity workloads.

All generators return a :class:`~femm_gpu_amd.model.Model` in SI units with
counter-clockwise P1 triangles.  They are preprocessing only (the reference
meshes with Triangle outside the solve); node counts and seeds are recorded in
``Model.name``.

* :func:`square_coil_ring`  -- planar linear air box with a square coil and an
  iron ring (structured grid, jittered interior).
* :func:`motor_sector`      -- planar 12-slot / 4-pole surface-PM motor, one pole
  pitch (90 deg) with anti-periodic sector boundaries, B-H iron, 3 phase coils
  (:class:`MotorFamily` meshes once and varies materials / rotor position per sample;
  :func:`study_family` / :func:`study_samples` define the 11-input learned study).
* :func:`c_core`            -- planar C-core with an air gap, an NdFeB block and two
  coils, Delaunay-meshed from seeded points.
* :func:`solenoid_rz`       -- axisymmetric solenoid with a plunger.
"""

from __future__ import annotations

import math

import numpy as np

from .model import AXISYMMETRIC, MU0, PLANAR, AirGap, Material, Model


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
                     coil_j=1e6, ring_mu=1000.0, ring_bh=False):
    """Air box (side 0.2 m) on a cells x cells grid, 2 triangles per cell, interior nodes
    jittered by U(-0.2, 0.2) h.  Centred square coil (J = 1e6 A/m^2 as turns = area and
    1e6 A) and an iron ring (mu_r = 1000) by element centroid.  A = 0 on the box.
    ring_bh: the ring follows the tanh B-H curve (Ms = 1.9 T, a = 300 A/m) instead of
    mu_r, which the coil's ~1,600 A-turns drive into saturation (a nonlinear solve).

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
            Material(1.0, bh_curve=0, group=2) if ring_bh else
            Material(1.0 / (MU0 * ring_mu), group=2)]
    curves = [tanh_bh_curve()] if ring_bh else []
    model = Model(xy, tri, label, mats, curves, boundary, np.zeros(len(boundary)), None, 1, 1.0,
                  PLANAR, name=f"square_coil_ring_{cells}x{cells}_seed{seed}"
                  + ("_bh" if ring_bh else ""))
    return model, np.array([coil_j])


# Motor geometry (mm) of the 12-slot / 4-pole surface-PM sector.
SHAFT_SLEEVE_MM = 11.0   # free space between the shaft (10 mm, A = 0) and the rotor iron
AIRGAP_PROBE_MM = 34.5   # air-gap B probe circle (gap 34..35 mm)
ROTOR_GROUP, AIR_GROUP = 1, 0
_MOTOR_BANDS = [  # (r_start, r_end, intervals)
    (10.0, 30.0, 20),    # rotor iron
    (30.0, 34.0, 12),    # magnets / inter-pole air
    (34.0, 35.0, 6),     # air gap
    (35.0, 37.0, 8),     # tooth tips / slot openings
    (37.0, 58.0, 50),    # slots and teeth
    (58.0, 75.0, 25),    # stator back iron
]


class MotorFamily:
    """One pole pitch (90 deg) of a 12-slot / 4-pole surface-PM machine, anti-periodic,
    meshed once; :meth:`model` assigns materials for one sample.

    Structured polar grid: radial bands of _MOTOR_BANDS (x radial_scale) by
    `angular_cells` sectors (a multiple of 160 so every slot / magnet edge lies on a grid
    line); interior nodes jittered by U(-jitter, jitter) of the local spacing, except on
    material-interface lines and, with `structured_magnet_band`, everywhere inside the
    magnet band (so the magnets can be moved by whole cells on the same mesh).

    Slots of half-width 6 deg (openings 1.5 deg) at `slot_centres_deg` carry the coils of
    `slot_phases` (circuit, sign) pairs.  The default (15, 45, 75) gives three full slots
    with phases A, -C, B; (0, 30, 60, 90) cuts the sector through slot centres, giving two
    half slots (A, and its anti-periodic image -A) and two full slots (-C, B), i.e. four
    coil regions.  Rotor: a non-magnetic shaft sleeve (free space) between 10 and 11 mm,
    iron from 11 to 30 mm and two parallel-magnetised NdFeB blocks (mu_r 1.05) between 30
    and 34 mm spanning rotor angles 9..45 and 45..81 deg with magnetisation angles 27 and
    63 deg (both turning with the rotor).  A = 0 on the shaft (r = 10 mm) and outer
    (r = 75 mm) arcs; nodes on theta = 0 and 90 deg are anti-periodic pairs.

    FEMM block groups: 0 free space (air gap, slot openings, inter-pole gaps, shaft sleeve),
    1 the rotor body (iron and both magnets: the body the weighted-stress torque acts on;
    it touches no Dirichlet node), 2 stator iron, 5 + k coil region k."""

    def __init__(self, angular_cells=480, seed=20251203, jitter=0.15, radial_scale=1,
                 slot_centres_deg=(15.0, 45.0, 75.0), slot_phases=((0, 1.0), (1, -1.0), (2, 1.0)),
                 structured_magnet_band=False):
        if angular_cells % 160:
            raise ValueError("angular_cells must be a multiple of 160")
        if len(slot_phases) != len(slot_centres_deg):
            raise ValueError("one (circuit, sign) per slot")
        self.angular_cells, self.seed = angular_cells, seed
        self.slot_centres = tuple(float(c) for c in slot_centres_deg)
        self.slot_phases = tuple((int(c), float(g)) for c, g in slot_phases)
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
        band_edges_mm = ({b[0] for b in _MOTOR_BANDS} | {b[1] for b in _MOTOR_BANDS}
                         | {SHAFT_SLEEVE_MM, 38.0, 57.0})
        r_mm = r_nodes * 1e3
        on_radial_edge = np.array([any(abs(r - e) < 1e-9 for e in band_edges_mm) for r in r_mm])
        interior[on_radial_edge, :] = False
        if structured_magnet_band:
            interior[(r_mm > 30.0 - 1e-9) & (r_mm < 34.0 + 1e-9), :] = False
        cell_deg = 90.0 / angular_cells
        self.cell_deg = cell_deg
        edge_deg = {9.0, 45.0, 81.0}
        for c in self.slot_centres:
            edge_deg |= {c - 6.0, c + 6.0, c - 1.5, c + 1.5}
        for d in sorted(edge_deg):
            if 0.0 < d < 90.0:
                interior[:, int(round(d / cell_deg))] = False
        Rj = R + np.where(interior, rng.uniform(-jitter, jitter, R.shape) * dr, 0.0)
        THj = TH + np.where(interior, rng.uniform(-jitter, jitter, R.shape) * dth, 0.0)
        self.xy = np.stack([Rj * np.cos(THj), Rj * np.sin(THj)], -1).reshape(-1, 2)

        # cells (i, j): radial interval i between r_nodes[i], r_nodes[i+1]; angular j.
        rc = 0.5 * (r_mm[:-1] + r_mm[1:])[:, None]
        tc = ((np.arange(angular_cells) + 0.5) * cell_deg)[None, :]
        self.rc, self.tc = np.broadcast_arrays(rc, tc)
        tri = []
        cell = []
        for i in range(nr - 1):
            for j in range(angular_cells):
                a, b = i * nt + j, i * nt + j + 1
                c, d = (i + 1) * nt + j, (i + 1) * nt + j + 1
                # alternate the diagonal so the grid has no preferred direction
                if (i + j) % 2 == 0:
                    tri += [(a, c, d), (a, d, b)]
                else:
                    tri += [(a, c, b), (b, c, d)]
                cell += [i * angular_cells + j] * 2
        self.tri = _ccw(self.xy, np.asarray(tri, np.int32))
        self.cell_of_tri = np.asarray(cell, np.int64)
        self.area = _area(self.xy, self.tri)
        self.outer = np.concatenate([np.arange(nt), (nr - 1) * nt + np.arange(nt)]).astype(np.int32)
        self.pairs = np.asarray([(i * nt, i * nt + nt - 1, -1) for i in range(1, nr - 1)], np.int32)
        self.nr, self.nt = nr, nt

    AIR, ROTOR, STATOR, MAG1, MAG2, COIL0 = range(6)

    @property
    def coil_regions(self):
        return len(self.slot_centres)

    def labels(self, rotor_steps=0):
        """Per-triangle material labels with the rotor turned by `rotor_steps` cells."""
        rc, tc = self.rc, self.tc
        phi = tc - rotor_steps * self.cell_deg          # rotor-frame angle
        label = np.full(rc.shape, self.AIR, np.int32)
        label[(rc > SHAFT_SLEEVE_MM) & (rc < 30)] = self.ROTOR
        band = (rc > 30) & (rc < 34)
        label[band & (phi > 9) & (phi < 45)] = self.MAG1
        label[band & (phi > 45) & (phi < 81)] = self.MAG2
        stator = rc > 35
        in_slot = np.zeros(rc.shape, bool)
        in_opening = np.zeros(rc.shape, bool)
        for s in self.slot_centres:
            in_slot |= (rc > 37) & (rc < 58) & (np.abs(tc - s) < 6)
            in_opening |= (rc > 35) & (rc < 37) & (np.abs(tc - s) < 1.5)
        label[stator & ~in_slot & ~in_opening] = self.STATOR
        for k, s in enumerate(self.slot_centres):
            coil = (rc > 38) & (rc < 57) & (np.abs(tc - s) < 6)
            label[coil] = self.COIL0 + k
        return label.reshape(-1)[self.cell_of_tri].astype(np.int32)

    def model(self, hc=(9e5, 9e5), ms=1.9, fill=None, rotor_steps=0, name=None):
        """The sample's model: magnet coercivities `hc`, B-H saturation `ms` (both irons),
        coil-region fill factors `fill` (turns = sign * fill * region area, so a phase
        current in A/m^2 times the fill is the region's current density) and the rotor
        turned by `rotor_steps` angular cells (magnetisation turning with it)."""
        fill = tuple(fill) if fill is not None else (1.0,) * self.coil_regions
        if len(fill) != self.coil_regions:
            raise ValueError(f"{self.coil_regions} fill factors expected")
        tlabel = self.labels(rotor_steps)
        b, h = tanh_bh_curve(ms=ms)
        turn = rotor_steps * self.cell_deg
        mats = [Material(1.0 / MU0, group=0),
                Material(1.0, bh_curve=0, group=1),
                Material(1.0, bh_curve=0, group=2),
                Material(1.0 / (MU0 * 1.05), hc=hc[0], magnetization_deg=27.0 + turn, group=1),
                Material(1.0 / (MU0 * 1.05), hc=hc[1], magnetization_deg=63.0 + turn, group=1)]
        for k, (circuit, sign) in enumerate(self.slot_phases):
            coil_area = float(self.area[tlabel == self.COIL0 + k].sum())
            turns = sign * fill[k] * coil_area
            mats.append(Material(1.0 / MU0, source_j_per_a=turns / coil_area, circuit=circuit,
                                 group=5 + k))
        circuits = 1 + max(c for c, _ in self.slot_phases)
        return Model(self.xy, self.tri, tlabel, mats, [(b, h)], self.outer,
                     np.zeros(len(self.outer)), self.pairs, circuits, 0.05, PLANAR,
                     name=name or f"motor_sector_12s4p_{self.nr}x{self.nt}_seed{self.seed}")


def motor_sector(angular_cells=480, seed=20251203, ms=1.9, hc=(9e5, 9e5), fill=(1.0, 1.0, 1.0),
                 jitter=0.15, radial_scale=1):
    """The benchmark motor: :class:`MotorFamily` with three full slots (phases A, -C, B),
    the rotor at its reference position.  The default (480 cells) gives 122 x 481 = 58,682
    nodes and 116,160 triangles.  Returns (model, nominal currents)."""
    family = MotorFamily(angular_cells=angular_cells, seed=seed, jitter=jitter,
                         radial_scale=radial_scale)
    return family.model(hc=hc, ms=ms, fill=fill), np.array([5e6, -3e6, -2e6])


STUDY_SLOTS = (0.0, 30.0, 60.0, 90.0)
STUDY_PHASES = ((0, 1.0), (1, -1.0), (2, 1.0), (0, -1.0))
STUDY_ROTOR_STEPS = 15          # rotor index k in [-15, 15] turns the rotor by 3 k cells


def study_family(angular_cells=480, seed=20251203):
    """Motor family of the learned warm-start study: sector cut through slot centres (four
    coil regions), structured magnet band (rotor positions on one mesh)."""
    return MotorFamily(angular_cells=angular_cells, seed=seed, slot_centres_deg=STUDY_SLOTS,
                       slot_phases=STUDY_PHASES, structured_magnet_band=True)


def study_samples(count, seed):
    """Seeded per-sample inputs (11 columns): 3 phase current densities U(-8e6, 8e6) A/m^2,
    2 magnet coercivities U(7e5, 1e6) A/m, B-H saturation scale U(0.9, 1.1), 4 coil-region
    fill factors U(0.4, 0.6), rotor index integer U[-15, 15] (rotor turned by 3 index cells)."""
    rng = np.random.default_rng(seed)
    x = np.empty((count, 11))
    x[:, 0:3] = rng.uniform(-8e6, 8e6, (count, 3))
    x[:, 3:5] = rng.uniform(7e5, 1e6, (count, 2))
    x[:, 5] = rng.uniform(0.9, 1.1, count)
    x[:, 6:10] = rng.uniform(0.4, 0.6, (count, 4))
    x[:, 10] = rng.integers(-STUDY_ROTOR_STEPS, STUDY_ROTOR_STEPS + 1, count)
    return x


def study_model(family, sample):
    """Model of one study sample (row of :func:`study_samples`)."""
    return family.model(hc=(sample[3], sample[4]), ms=1.9 * sample[5], fill=tuple(sample[6:10]),
                        rotor_steps=3 * int(round(sample[10])))


def motor_currents(count, seed, low=-8e6, high=8e6):
    """Per-sample phase current densities ~ U(low, high) A/m^2 (seeded)."""
    return np.random.default_rng(seed).uniform(low, high, size=(count, 3))


def c_core(edge_iron=0.5e-3, edge_air=2e-3, seed=2, scale=0.75):
    """C-shaped iron core with a 2 mm air gap, an NdFeB block (mu_r 1.05, Hc 9e5 A/m along
    x) in the back limb and two coils (+-5e5 A/m^2) around it, inside an air box.

    Seeded points (iron spacing `edge_iron`, air spacing growing from the iron spacing to
    `edge_air` with distance from the core) are Delaunay-triangulated; labels follow the
    element centroid.  Geometry (mm, times `scale`): core outer half-size 70, window
    half-size 40, air box half-size 160, magnet 30 x 20 in the left limb; the default
    scale 0.75 gives 62,237 nodes with 0.5 mm iron / 2 mm air edges, and edge_iron = 0.1 mm
    gives 1,163,107 nodes.  Returns (model, currents)."""
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


def motor_postprocess(probes=32):
    """Weighted-stress torque on the rotor body (group 1) against free space (group 0) and
    radial air-gap B at `probes` angles (k + 1/2) * 90 / probes deg on the 34.5 mm circle:
    keyword arguments of ``Plan.set_postprocess`` for every :class:`MotorFamily` model."""
    angles = (np.arange(probes) + 0.5) * (0.5 * math.pi / probes)
    return dict(selected_group=ROTOR_GROUP, air_group=AIR_GROUP,
                radius_m=AIRGAP_PROBE_MM * 1e-3, angles_rad=angles)


class AgeMachine:
    """Full 360 deg 12-slot / 4-pole surface-PM machine on polar grids of `cells` equal
    angular cells, its air gap (34..35 mm) either a native air-gap element joining the
    separately meshed rotor and stator rings (gap="age", FEMM's sliding band) or
    `gap_intervals` layers of triangles (gap="triangles", a conforming mesh of the same
    machine).  No jitter: every ring is uniform, as FEMM requires of the rings an air-gap
    element joins.

    Rotor (rotor frame): A = 0 on the shaft circle (10 mm), free space to 11 mm, B-H iron to
    30 mm, four parallel-magnetised NdFeB poles (mu_r 1.05) between 30 and 34 mm spanning
    90 k + 9 .. 90 k + 81 deg, magnetised along 90 k + 45 deg, alternately outward and
    inward.  Stator: slot openings 35..37 mm (1.5 deg half-width), 12 slots 37..58 mm of
    6 deg half-width centred at 15 + 30 m deg with coils 38..57 mm of phases A, -C, B, -A,
    C, -B (repeating; 3 circuits, J per ampere = +-1), B-H iron elsewhere, A = 0 on the
    outer circle (75 mm).  Labels follow element centroids.

    Groups: 0 free space, 1 rotor body (iron and magnets), 2 stator iron, 5 + m slot m."""

    ROTOR_RINGS_MM = ((10.0, 11.0, 1), (11.0, 30.0, 5), (30.0, 34.0, 2))
    STATOR_RINGS_MM = ((35.0, 37.0, 1), (37.0, 38.0, 1), (38.0, 57.0, 4), (57.0, 58.0, 1),
                       (58.0, 75.0, 3))
    PHASES = ((0, 1.0), (2, -1.0), (1, 1.0), (0, -1.0), (2, 1.0), (1, -1.0))

    def __init__(self, cells=96, gap="age", gap_intervals=2, radial_scale=1, sector=False):
        if gap not in ("age", "triangles"):
            raise ValueError("gap is 'age' or 'triangles'")
        self.cells, self.gap, self.sector = int(cells), gap, bool(sector)
        # sector: one pole pitch (90 deg), anti-periodic edges; else the full circle
        self.span = 90.0 if self.sector else 360.0

        def radii(bands):
            r = [bands[0][0]]
            for r0, r1, k in bands:
                r += list(np.linspace(r0, r1, k * radial_scale + 1)[1:])
            return np.asarray(r) * 1e-3

        rotor_r, stator_r = radii(self.ROTOR_RINGS_MM), radii(self.STATOR_RINGS_MM)
        if gap == "triangles":
            gap_r = np.linspace(34.0, 35.0, gap_intervals + 1)[1:-1] * 1e-3
            grids = [np.concatenate([rotor_r, gap_r, stator_r])]
        else:
            grids = [rotor_r, stator_r]
        n = self.cells
        m = n + 1 if self.sector else n            # nodes per ring
        theta = np.arange(m) * (math.radians(self.span) / n)
        xy, tri, rc, tc, offsets, pairs = [], [], [], [], [], []
        base = 0
        for r in grids:
            offsets.append(base)
            R, TH = np.meshgrid(r, theta, indexing="ij")
            xy.append(np.stack([R * np.cos(TH), R * np.sin(TH)], -1).reshape(-1, 2))
            for i in range(len(r) - 1):
                for j in range(n):
                    a, b = base + i * m + j, base + i * m + (j + 1) % m
                    c, d = base + (i + 1) * m + j, base + (i + 1) * m + (j + 1) % m
                    if (i + j) % 2 == 0:
                        tri += [(a, c, d), (a, d, b)]
                    else:
                        tri += [(a, c, b), (b, c, d)]
                    rc += [0.5 * (r[i] + r[i + 1]) * 1e3] * 2
                    tc += [(j + 0.5) * self.span / n] * 2
            if self.sector:   # the edge at 90 deg is the anti-periodic image of the one at 0
                for i in range(len(r)):
                    shaft_ring = r is grids[0] and i == 0
                    outer_ring = r is grids[-1] and i == len(r) - 1
                    if not (shaft_ring or outer_ring):
                        pairs.append((base + i * m, base + i * m + n, -1))
            base += len(r) * m
        self.xy = np.concatenate(xy)
        self.tri = _ccw(self.xy, np.asarray(tri, np.int32))
        self.rc, self.tc = np.asarray(rc), np.asarray(tc)
        self.area = _area(self.xy, self.tri)
        self.pairs = np.asarray(pairs, np.int32).reshape(-1, 3)
        last = grids[-1]
        shaft = offsets[0] + np.arange(m)
        outer = offsets[-1] + (len(last) - 1) * m + np.arange(m)
        self.dirichlet = np.concatenate([shaft, outer]).astype(np.int32)
        if gap == "age":
            self.inner_ring = offsets[0] + (len(rotor_r) - 1) * m + np.arange(n)
            self.outer_ring = offsets[1] + np.arange(n)

    AIR, ROTOR, STATOR, MAG0, COIL0 = 0, 1, 2, 3, 7

    def labels(self, rotor_steps=0):
        """Per-triangle labels with the rotor's magnets turned by `rotor_steps` cells."""
        rc, tc = self.rc, self.tc
        phi = np.mod(tc - rotor_steps * self.span / self.cells, 360.0)
        label = np.full(rc.shape, self.AIR, np.int32)
        label[(rc > 11.0) & (rc < 30.0)] = self.ROTOR
        band = (rc > 30.0) & (rc < 34.0)
        for k in range(4):
            label[band & (phi > 90 * k + 9) & (phi < 90 * k + 81)] = self.MAG0 + k
        slot = np.rint((tc - 15.0) / 30.0).astype(int) % 12
        dist = np.abs(np.mod(tc - (15.0 + 30.0 * slot) + 180.0, 360.0) - 180.0)
        stator = rc > 35.0
        in_slot = (rc > 37.0) & (rc < 58.0) & (dist < 6.0)
        in_opening = (rc > 35.0) & (rc < 37.0) & (dist < 1.5)
        label[stator & ~in_slot & ~in_opening] = self.STATOR
        coil = (rc > 38.0) & (rc < 57.0) & (dist < 6.0)
        label[coil] = self.COIL0 + slot[coil]
        return label

    def air_gap(self):
        """The sliding-band air-gap element at rotor angle 0: record k holds the inner ring's
        nodes k - 1 and k and the outer ring's (FEMM's preprocessor, femm/writepoly.cpp:
        1649-1667), shifts 0 (node 0 of both rings at angle 0).  In the sector the ring
        positions -1 and `cells` are the anti-periodic images of nodes cells - 1 and 0
        (weight -1)."""
        n = self.cells
        k = np.arange(n + 1)
        prev, cur = (k - 1) % n, k % n
        nodes = np.stack([self.inner_ring[prev], self.inner_ring[cur], self.outer_ring[prev],
                          self.outer_ring[cur]], 1)
        weights = np.ones((n + 1, 4))
        if self.sector:
            weights[0, [0, 2]] = -1.0
            weights[n, [1, 3]] = -1.0
        return AirGap(inner_radius_m=34e-3, outer_radius_m=35e-3, arc_length_deg=self.span,
                      nodes=nodes, weights=weights, antiperiodic=self.sector)

    def model(self, hc=9e5, ms=1.9, rotor_steps=0, name=None):
        tlabel = self.labels(rotor_steps)
        b, h = tanh_bh_curve(ms=ms)
        mats = [Material(1.0 / MU0, group=0), Material(1.0, bh_curve=0, group=1),
                Material(1.0, bh_curve=0, group=2)]
        turn = rotor_steps * self.span / self.cells
        for k in range(4):
            mats.append(Material(1.0 / (MU0 * 1.05), hc=hc,
                                 magnetization_deg=90.0 * k + 45.0 + 180.0 * (k % 2) + turn,
                                 group=1))
        for m in range(12):
            circuit, sign = self.PHASES[m % 6]
            mats.append(Material(1.0 / MU0, source_j_per_a=sign, circuit=circuit, group=5 + m))
        gaps = [self.air_gap()] if self.gap == "age" else []
        kind = "sector" if self.sector else "full"
        return Model(self.xy, self.tri, tlabel, mats, [(b, h)], self.dirichlet,
                     np.zeros(len(self.dirichlet)), self.pairs, 3, 0.05, PLANAR,
                     name=name or f"age_machine_{kind}_{self.gap}_{self.cells}", air_gaps=gaps)


def age_machine(cells=96, gap="age", sector=False, **kw):
    """:class:`AgeMachine` model at rotor angle 0."""
    return AgeMachine(cells, gap, sector=sector).model(**kw)
