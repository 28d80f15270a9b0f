"""The study driver over the lattices: a restatement of crates/cfd/src/tunnel.rs
(`plan`, `estimate`, `finest_fit`, `run_with`, `sig`) and the reference
cases of crates/cfd/src/validation.rs.

As in the reference, a still part with 48 or more cells across its largest
size runs on the refined D3Q27 solver (`MultiLattice`, nested boxes round
the part, tunnel.rs:645-687, `SolverOptions.refine`); spinning parts and
coarser still parts step one uniform D3Q19 lattice.

Units: the surface is in metres; `unit_m` converts the spin centre (model
units, e.g. mm) to metres; results are SI (the conventions of units.py).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np

from . import (FaceBc, Geometry, Lattice, LatticeSetup, Probes, SliceDef, StepIn, Storage, voxelize_spin,
               voxelize_static)
from . import units as U
from .mlattice import LevelBox, MultiLattice, MultiSetup, voxelize_static27
from .shapes import TriMesh, disk, sphere, volume
from .workloads import rotation

# tunnel.rs:145-148: measured MLUPS of the reference (for its time estimates
# and its per-call batch size, which also sets when convergence is checked)
MLUPS_FP32 = 2800.0
MLUPS_FP16 = 4500.0
MLUPS_REFINED = 1400.0     # the refined D3Q27 solver, all levels
THIN_MAX_BOOST = 8
SIDE_MIN_FRONT = 1.5
REFINE_THIN = 2.0
# tunnel.rs:462-466: nested refinement round the part beyond its bounding
# box, in frontal sizes: (upstream, downstream, sideways) for levels 1, 2, 3
REFINE_EXT = ((1.2, 3.0, 1.0), (0.8, 1.6, 0.75), (0.65, 1.0, 0.65))


def still_levels(cells_across: int) -> int:
    """tunnel.rs:467-475: levels for a still part; the coarsest keeps at
    least 24 cells across the part's largest size, at most 4 levels."""
    nl = 1
    while nl < 4 and cells_across / float(1 << nl) >= 24.0:
        nl += 1
    return nl


# ---------------------------------------------------------------- configuration (simlab_core config.rs)
@dataclass
class Air:
    density: float = 1.204               # sea level, 20 C (config.rs:146-151)
    kinematic_viscosity: float = 1.516e-5


@dataclass
class Spin:
    rpm: float
    axis: tuple = (0.0, 0.0, 1.0)
    center_mm: tuple = (0.0, 0.0, 0.0)


@dataclass
class WindConfig:
    wind_dir: tuple = (1.0, 0.0, 0.0)
    speed_m_s: float = 10.0
    air: Air = field(default_factory=Air)
    spin: Optional[Spin] = None
    quality: object = "normal"           # "draft" | "normal" | "fine" | int (custom cells across)
    duration: Optional[float] = None


@dataclass
class Margins:
    upstream: float
    downstream: float
    side: float


@dataclass
class SolverOptions:
    """tunnel.rs:161-236 (defaults as there)."""
    cells_across: Optional[int] = None
    storage: Optional[Storage] = None
    memory_budget: Optional[int] = None
    max_lattice_speed: float = 0.075
    max_tip_speed: float = 0.10
    max_tau: float = 1.0
    wind_margins: Margins = field(default_factory=lambda: Margins(2.5, 5.0, 3.0))
    still_margins: Margins = field(default_factory=lambda: Margins(2.5, 2.5, 1.0))
    smagorinsky: float = 0.10
    tolerance: float = 0.005
    spin_ramp_revs: float = 0.5
    vertex_offset: float = 1.0
    side_bc: str = "open"
    sponge_sigma: float = 0.1
    sponge_widths: tuple = (0.4, 1.0, 0.4)
    absorb_velocity: bool = False
    total_pressure_inlet: bool = True
    wall_model: bool = False
    refine: bool = True


def cells_across(q) -> int:
    """tunnel.rs:63-70."""
    if isinstance(q, int):
        return min(max(q, 8), 4096)
    return {"draft": 64, "normal": 128, "fine": 192}[q]


def default_duration(q, still_spin: bool) -> float:
    """tunnel.rs:73-80: flow passes (wind) or revolutions (still air)."""
    d, n, f = (8.0, 12.0, 16.0) if still_spin else (20.0, 40.0, 60.0)
    if q == "draft":
        return d
    if q == "fine":
        return f
    return n


def thin_cells_target(q) -> int:
    """tunnel.rs:414-421."""
    return 0 if isinstance(q, int) else {"draft": 16, "normal": 24, "fine": 32}[q]


def cells_for_thin(n_q: int, n_thin: int, l_ref: float, thin: float) -> int:
    """tunnel.rs:425-431."""
    if thin <= 0.0 or l_ref <= 0.0:
        return n_q
    want = int(math.ceil(n_thin * l_ref / thin))
    return min(max(want, n_q), n_q * THIN_MAX_BOOST)


def thin_cost_cap(q) -> float:
    """tunnel.rs:441-447."""
    return {"draft": 4.0e11, "normal": 4.0e12}.get(q if not isinstance(q, int) else "", 2.0e13)


def sig(x: float, n: int) -> str:
    """tunnel.rs:1421-1430: x with n significant figures, plain decimals
    (halves round away from zero)."""
    if x == 0.0 or not math.isfinite(x):
        if math.isnan(x):
            return "NaN"
        if math.isinf(x):
            return "inf" if x > 0 else "-inf"
        return "-0" if math.copysign(1.0, x) < 0 else "0"
    mag = math.floor(math.log10(abs(x)))
    decimals = max(n - 1 - mag, 0)
    factor = 10.0 ** (n - 1 - mag)
    v = x * factor
    rounded = math.copysign(math.floor(abs(v) + 0.5), v) / factor
    return f"{rounded:.{decimals}f}"


# ---------------------------------------------------------------- geometry helpers
def _norm(v):
    return float(np.linalg.norm(v))


def _normalize(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def frame(main) -> np.ndarray:
    """tunnel.rs:495-501: rows ex (main), ey, ez of the lattice frame."""
    ex = _normalize(main)
    up = np.array([1.0, 0.0, 0.0]) if abs(ex[2]) > 0.95 else np.array([0.0, 0.0, 1.0])
    ez = _normalize(up - ex * (up @ ex))
    ey = np.cross(ez, ex)
    return np.array([ex, ey, ez])


_FRONTAL_CACHE: dict = {}


def frontal_area(m: TriMesh, r: np.ndarray, res: int = 1024) -> float:
    """tunnel.rs:504-539: projected area on the plane across r[0], rasterized
    at res pixels across the larger extent (cached per mesh and frame, since
    planning calls it for every grid it tries)."""
    key = (id(m), len(m.positions), len(m.triangles), np.asarray(r, np.float64).tobytes(), res)
    hit = _FRONTAL_CACHE.get(key)
    if hit is not None and hit[0] is m:
        return hit[1]
    a = _frontal_area(m, r, res)
    if len(_FRONTAL_CACHE) > 64:
        _FRONTAL_CACHE.clear()
    _FRONTAL_CACHE[key] = (m, a)
    return a


def _frontal_area(m: TriMesh, r: np.ndarray, res: int) -> float:
    P = np.asarray(m.positions, np.float64)
    pts = np.stack([P @ r[1], P @ r[2]], axis=1)
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    size = max(hi[0] - lo[0], hi[1] - lo[1])
    if size <= 0.0:
        return 0.0
    h = size / res
    nu, nv = int(math.ceil((hi[0] - lo[0]) / h)) + 1, int(math.ceil((hi[1] - lo[1]) / h)) + 1
    img = np.zeros((nv, nu), bool)
    T = np.asarray(m.triangles, np.int64)
    for t in T:
        a, b, c = pts[t[0]], pts[t[1]], pts[t[2]]
        i0 = max(int(math.floor((min(a[0], b[0], c[0]) - lo[0]) / h)), 0)
        i1 = min(int(math.ceil((max(a[0], b[0], c[0]) - lo[0]) / h)), nu - 1)
        j0 = max(int(math.floor((min(a[1], b[1], c[1]) - lo[1]) / h)), 0)
        j1 = min(int(math.ceil((max(a[1], b[1], c[1]) - lo[1]) / h)), nv - 1)
        if i1 < i0 or j1 < j0:
            continue
        xs = lo[0] + (np.arange(i0, i1 + 1) + 0.5) * h
        ys = lo[1] + (np.arange(j0, j1 + 1) + 0.5) * h
        X, Y = np.meshgrid(xs, ys)

        def e(p, q):
            return (q[0] - p[0]) * (Y - p[1]) - (q[1] - p[1]) * (X - p[0])

        e0, e1, e2 = e(a, b), e(b, c), e(c, a)
        inside = ((e0 >= 0) & (e1 >= 0) & (e2 >= 0)) | ((e0 <= 0) & (e1 <= 0) & (e2 <= 0))
        img[j0:j1 + 1, i0:i1 + 1] |= inside
    return float(img.sum()) * h * h


def _triangle_areas(m: TriMesh) -> np.ndarray:
    P = np.asarray(m.positions, np.float64)
    T = np.asarray(m.triangles, np.int64)
    return 0.5 * np.linalg.norm(np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]]), axis=1)


# ---------------------------------------------------------------- planning (tunnel.rs:349-412, 541-812)
@dataclass
class Plan:
    r: np.ndarray
    pref: np.ndarray
    dx: float
    dt: float
    dims: tuple
    center: tuple
    u_lat: float
    omega_lat: np.ndarray
    nu_lat: float
    spin: bool
    still: bool
    cells_across: int
    d_front: float
    frontal_area: float
    steps_unit: float
    duration: float
    storage: Storage
    memory: int
    blockage: float
    sponge: tuple
    faces: tuple
    rotor_radius: float
    axis_lat: np.ndarray
    axis_model: np.ndarray
    box_lo: tuple
    box_n: tuple
    est_records: int
    updates_per_step: float
    d_thin: float
    thin: Optional[tuple] = None    # (target, wanted) when a slender part's grid was sized for its thickness
    # refined grid (still parts): coarsest dims, nested boxes, the finest
    # level's origin in global finest cells, cells per level (tunnel.rs:455-460)
    multi: Optional["MultiPlan"] = None

    def cost(self) -> float:
        return self.updates_per_step * math.ceil(self.duration * self.steps_unit)

    def cells_thin(self) -> float:
        return self.d_thin / self.dx


@dataclass
class MultiPlan:
    dims0: tuple
    boxes: list
    org_f: tuple
    cells: list


def _bytes_per_cell(s: Storage) -> int:
    return 19 * s.pop_bytes() + 4


# Most cells one lattice holds whatever the memory: 32-bit cell indices with
# signed neighbour offsets (simlab_cfd_max_cells), for every storage. The
# reference's limit (tunnel.rs:140-142) is its GPU buffer size, half of 4 GiB
# per population direction: 537 M cells in FP32, which on one 288 GB GPU is
# not the binding limit, so FP32 is kept up to this one.
MAX_CELLS_LAYOUT = (1 << 31) - 1


def max_cells_layout(s: Storage) -> int:
    """tunnel.rs:140-142, with this lattice's limit."""
    return MAX_CELLS_LAYOUT


def plan(cfg: WindConfig, m: TriMesh, unit_m: float, opt: SolverOptions, budget: Optional[int] = None) -> Plan:
    if not m.triangles:
        raise ValueError("the part has no surface")
    speed = cfg.speed_m_s
    if not (speed >= 0.0) or not math.isfinite(speed):
        raise ValueError("wind speed must be zero or positive")
    spin = cfg.spin if (cfg.spin is not None and cfg.spin.rpm != 0.0 and _norm(cfg.spin.axis) > 0.0) else None
    if speed == 0.0 and spin is None:
        raise ValueError("still air needs a spinning part; set a wind speed or a spin")
    if speed > 0.0 and _norm(cfg.wind_dir) == 0.0:
        raise ValueError("the wind direction is zero")
    still = speed == 0.0
    r = frame(spin.axis) if still else frame(cfg.wind_dir)
    P = np.asarray(m.positions, np.float64)
    blo, bhi = P.min(axis=0), P.max(axis=0)
    pref = np.asarray(spin.center_mm, np.float64) * unit_m if spin else 0.5 * (blo + bhi)
    V = P - pref
    Q = V @ r.T
    lo, hi = Q.min(axis=0), Q.max(axis=0)
    axis_model = _normalize(spin.axis) if spin else np.array([1.0, 0.0, 0.0])
    a = V @ axis_model
    rotor_radius = float(np.sqrt(np.maximum((V * V).sum(axis=1) - a * a, 0.0)).max())
    alo, ahi = float(a.min()), float(a.max())
    axis_lat = r @ axis_model
    if spin:
        for k in range(3):
            e = rotor_radius * math.sqrt(max(1.0 - axis_lat[k] ** 2, 0.0))
            p0, p1 = alo * axis_lat[k], ahi * axis_lat[k]
            lo[k] = min(p0, p1) - e
            hi[k] = max(p0, p1) + e
    ext = hi - lo
    l_ref = float(ext.max())
    if l_ref <= 0.0:
        raise ValueError("the part has no size")
    d_thin = max(min(ext[1], ext[2]), l_ref * 1e-3) if spin is None else l_ref
    thin = None
    if opt.cells_across is not None:
        n_across = opt.cells_across
    else:
        n_q = cells_across(cfg.quality)
        n_t = thin_cells_target(cfg.quality)
        want = cells_for_thin(n_q, n_t, l_ref, d_thin)
        if want > n_q:
            # the finest grid up to `want` within the budget and the time cap,
            # tried coarser by ~8 % a step (tunnel.rs:605-621)
            bud = budget if budget is not None else opt.memory_budget
            base = plan(cfg, m, unit_m, replace(opt, cells_across=n_q), bud)
            t_cap = max(thin_cost_cap(cfg.quality), base.cost())
            n = want
            while True:
                p = plan(cfg, m, unit_m, replace(opt, cells_across=n), bud)
                fits = (bud is None or p.memory <= bud) and p.cost() <= t_cap
                if fits or n <= n_q:
                    p = p if fits else base
                    p.thin = (n_t, want)
                    return p
                n = max(int(n * 0.92), n_q)
        n_across = n_q
    dx = l_ref / n_across
    d_front = max(2.0 * rotor_radius if still else max(ext[1], ext[2]), l_ref * 1e-3)
    d_box = min(d_front, REFINE_THIN * d_thin) if spin is None else d_front
    mg = opt.still_margins if still else opt.wind_margins
    side = mg.side * d_front if still else max(mg.side * d_box, SIDE_MIN_FRONT * d_front)
    dom_lo = np.array([lo[0] - mg.upstream * d_front, lo[1] - side, lo[2] - side])
    dom_hi = np.array([hi[0] + mg.downstream * d_front, hi[1] + side, hi[2] + side])
    dims = tuple(int(math.ceil((dom_hi[k] - dom_lo[k]) / dx)) + 1 for k in range(3))
    center = tuple(float(-dom_lo[k] / dx) for k in range(3))
    # still parts: nested refinement, the finest level holding the part (tunnel.rs:643-687)
    multi = None
    if spin is None and opt.refine:
        nl = still_levels(n_across)
        if nl > 1:
            s0 = float(1 << (nl - 1))
            dx0 = dx * s0
            dims0 = tuple(int(math.ceil((dom_hi[k] - dom_lo[k]) / dx0)) + 1 for k in range(3))
            dims = tuple(int((dims0[k] - 1) * s0) + 1 for k in range(3))
            lo_c = [(center[k] + lo[k] / dx) / s0 for k in range(3)]
            hi_c = [(center[k] + hi[k] / dx) / s0 for k in range(3)]
            pd = list(dims0)
            boxes = []
            org = [0.0, 0.0, 0.0]
            cells = [int(np.prod(np.asarray(dims0, np.int64)))]
            for l in range(1, nl):
                sc = float(1 << (nl - l))          # finest cells per cell of level l - 1
                df = d_box / dx / sc
                eu, ed, es = REFINE_EXT[l - 1]
                e_lo = (eu * df, es * df, es * df)
                e_hi = (ed * df, es * df, es * df)
                blo, bn = [0] * 3, [0] * 3
                for k in range(3):
                    a0 = max(int(math.floor(lo_c[k] - e_lo[k])), 3)
                    b0 = min(int(math.ceil(hi_c[k] + e_hi[k])), pd[k] - 4)
                    blo[k] = a0
                    bn[k] = max(b0 - a0 + 1, 8)
                for k in range(3):
                    org[k] += blo[k] * sc
                    lo_c[k] = 2.0 * (lo_c[k] - blo[k])
                    hi_c[k] = 2.0 * (hi_c[k] - blo[k])
                    pd[k] = 2 * (bn[k] - 1) + 1
                cells.append(int(np.prod(np.asarray(pd, np.int64))))
                boxes.append(LevelBox(lo=tuple(blo), n=tuple(bn)))
            multi = MultiPlan(dims0=dims0, boxes=boxes, org_f=tuple(org), cells=cells)
    # time step from the Mach limit, then the viscosity limit (tunnel.rs:690-700)
    omega = spin.rpm * 2.0 * math.pi / 60.0 if spin else 0.0
    tip = abs(omega) * rotor_radius
    umax = speed + tip
    ulat_max = opt.max_tip_speed if spin else opt.max_lattice_speed
    dt = ulat_max * dx / umax
    nu = cfg.air.kinematic_viscosity
    nu_max = (opt.max_tau - 0.5) / 3.0
    if nu * dt / (dx * dx) > nu_max:
        dt = nu_max * dx * dx / nu
    nu_lat = nu * dt / (dx * dx)
    u_lat = speed * dt / dx
    omega_lat = axis_lat * (omega * dt)
    steps_unit = 2.0 * math.pi / (abs(omega) * dt) if still else l_ref / speed / dt
    duration = max(cfg.duration if cfg.duration is not None else default_duration(cfg.quality, still), 1.0)
    n = int(np.prod(np.asarray(dims, np.int64)))
    if spin:
        bl, bn = [0] * 3, [0] * 3
        for k in range(3):
            a0 = int(math.floor(center[k] + lo[k] / dx - 3.0))
            b0 = int(math.ceil(center[k] + hi[k] / dx + 3.0))
            bl[k] = max(a0, 1)
            bn[k] = max(min(b0, dims[k] - 2) - bl[k] + 1, 1)
        box_lo, box_n = tuple(bl), tuple(bn)
    else:
        box_lo, box_n = (0, 0, 0), (0, 0, 0)
    area = float(_triangle_areas(m).sum())
    est_records = int(np.prod(box_n)) if spin else int(2.0 * area / (dx * dx)) + 1000
    faces_cells = 2 * (dims[0] * dims[1] + dims[1] * dims[2] + dims[0] * dims[2])
    bud = budget if budget is not None else opt.memory_budget

    def mem_for(s):
        if multi is not None:
            # A and B population sets, flags and moments per level (tunnel.rs:726-733)
            pops = sum(c * (2 * 27 * s.pop_bytes() + 4 + 8) for c in multi.cells)
            return pops + est_records * 24 * 4 * 2 + 64 * 1024 * 1024
        return n * _bytes_per_cell(s) + est_records * 40 * 4 + faces_cells * 44 + 64 * 1024 * 1024

    # the refined solver keeps 32-bit cell indices per level
    n_lay = max(multi.cells) if multi is not None else n
    lay_max = (1 << 32) - 2 if multi is not None else max_cells_layout(Storage.FP32)
    if opt.storage is not None:
        storage = opt.storage
    elif (bud is not None and mem_for(Storage.FP32) > bud) or n_lay > lay_max:
        storage = Storage.FP16
    else:
        storage = Storage.FP32
    memory = mem_for(storage) if n_lay <= lay_max else (1 << 62)
    frontal = math.pi * rotor_radius ** 2 if still else frontal_area(m, r)
    blockage = frontal / ((dom_hi[1] - dom_lo[1]) * (dom_hi[2] - dom_lo[2]))

    # sponge widths in cells of the coarsest level (tunnel.rs:757-761)
    dxs = dx * ((dims[0] - 1) / (multi.dims0[0] - 1)) if multi is not None else dx

    def cl(x, least):
        return max(int(round(x / dxs)), least)

    if still:
        s = cl(opt.sponge_widths[2] * d_front, 8)
        sponge = (s,) * 6
        faces = (FaceBc.OPEN,) * 6
    else:
        sbc = FaceBc.VELOCITY if opt.side_bc == "velocity" else FaceBc.OPEN
        s_side = cl(opt.sponge_widths[2] * d_front, 6)
        sponge = (cl(opt.sponge_widths[0] * d_front, 6), cl(opt.sponge_widths[1] * d_front, 8), s_side, s_side, s_side, s_side)
        faces = (FaceBc.VELOCITY, FaceBc.OPEN, sbc, sbc, sbc, sbc)
    # cell updates per finest step, all levels (tunnel.rs:800-806)
    if multi is not None:
        nl = len(multi.cells)
        ups = sum(c * float(1 << l) for l, c in enumerate(multi.cells)) / float(1 << (nl - 1))
    else:
        ups = float(n)
    return Plan(r=r, pref=pref, dx=dx, dt=dt, dims=dims, center=center, u_lat=u_lat, omega_lat=omega_lat, nu_lat=nu_lat,
                spin=spin is not None, still=still, cells_across=n_across, d_front=d_front, frontal_area=frontal,
                steps_unit=steps_unit, duration=duration, storage=storage, memory=memory, blockage=blockage,
                sponge=sponge, faces=faces, rotor_radius=rotor_radius, axis_lat=axis_lat, axis_model=axis_model,
                box_lo=box_lo, box_n=box_n, est_records=est_records, updates_per_step=ups, d_thin=d_thin,
                thin=thin, multi=multi)


def _thin_warning(p: Plan) -> Optional[str]:
    """tunnel.rs:1396-1418."""
    if p.spin:
        return None
    have = p.cells_thin()
    thin_mm = p.d_thin * 1e3
    if p.thin is not None and have + 0.5 < p.thin[0]:
        return (f"This part is thin across the wind ({sig(thin_mm, 2)} mm), and the grid has only {have:.0f} cells across "
                f"it instead of the {p.thin[0]} this detail setting asks for (it would need {p.thin[1]} cells across the "
                f"whole part; memory and run time allow {p.cells_across}), so its drag is less certain.")
    if p.thin is None and have < 8.0:
        return (f"This part is thin across the wind ({sig(thin_mm, 2)} mm) and the grid has only {have:.0f} cells across "
                f"it, so its drag is rough; a finer grid would be more accurate.")
    return None


@dataclass
class Estimate:
    fits: bool
    cells: int
    grid: tuple
    cells_across: int
    memory_bytes: int
    budget_bytes: int
    storage: Storage
    steps: int
    seconds: float
    note: str
    thin_mm: float
    cells_thin: float
    thin_warning: Optional[str]


def estimate(cfg: WindConfig, m: TriMesh, unit_m: float, budget_bytes: int, opt: Optional[SolverOptions] = None) -> Estimate:
    """tunnel.rs:815-866."""
    opt = opt or SolverOptions()
    p = plan(cfg, m, unit_m, opt, budget_bytes)
    cells = sum(p.multi.cells) if p.multi is not None else int(np.prod(np.asarray(p.dims, np.int64)))
    steps = int(math.ceil(p.duration * p.steps_unit))
    if p.multi is not None:
        mlups = MLUPS_REFINED
    else:
        mlups = MLUPS_FP16 if p.storage == Storage.FP16 else MLUPS_FP32
    seconds = steps * p.updates_per_step / (mlups * 1e6) * (1.15 if p.spin else 1.0) + 2.0 + p.est_records * 2e-6
    levels = f" near it, {len(p.multi.cells)} grid levels" if p.multi is not None else ""
    note = (f"{p.cells_across} cells across the part ({p.dx * 1e3:.2f} mm cells{levels}), {p.dims[0]} x {p.dims[1]} x {p.dims[2]} "
            f"grid, {'FP16' if p.storage == Storage.FP16 else 'FP32'} storage, {p.duration:.0f} "
            f"{'revolutions' if p.still else 'flow passes'}")
    return Estimate(fits=p.memory <= budget_bytes, cells=cells, grid=p.dims, cells_across=p.cells_across,
                    memory_bytes=p.memory, budget_bytes=budget_bytes, storage=p.storage, steps=steps, seconds=seconds,
                    note=note, thin_mm=0.0 if p.spin else p.d_thin * 1e3, cells_thin=0.0 if p.spin else p.cells_thin(),
                    thin_warning=_thin_warning(p))


def finest_fit(cfg: WindConfig, m: TriMesh, unit_m: float, budget_bytes: int, opt: Optional[SolverOptions] = None) -> int:
    """tunnel.rs:880-894: the most cells across that fit the budget (bisection)."""
    opt = opt or SolverOptions()

    def fits(c):
        return plan(cfg, m, unit_m, replace(opt, cells_across=c), budget_bytes).memory <= budget_bytes

    lo, hi = 8, 4097
    if not fits(lo):
        raise ValueError("even the coarsest grid does not fit the memory budget")
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if fits(mid):
            lo = mid
        else:
            hi = mid
    return min(lo, 4096)


# ---------------------------------------------------------------- the run (tunnel.rs:902-1392)
@dataclass
class SliceField:
    name: str
    label: str
    unit: str
    min: float
    max: float
    data: np.ndarray


@dataclass
class FlowSlice:
    """A plane of window-averaged flow through the part (tunnel.rs:1259-1312):
    origin and the two edge vectors in model units, nu x nv pixels (u fastest)."""
    name: str
    label: str
    origin_mm: np.ndarray
    u_mm: np.ndarray
    v_mm: np.ndarray
    nu: int
    nv: int
    fields: List[SliceField]

    def field(self, name: str) -> SliceField:
        return next(f for f in self.fields if f.name == name)


@dataclass
class WindSummary:
    """core config.rs WindSummary as tunnel.rs:1355-1390 fills it."""
    headline: str
    force_n: np.ndarray
    moment_nm: np.ndarray
    drag_n: float
    lift_n: float
    side_n: float
    cd: Optional[float]
    ref_area_mm2: float
    thrust_n: Optional[float]
    torque_nm: Optional[float]
    power_w: Optional[float]
    reynolds: float
    cells: int
    steps: int
    mlups: float
    seconds: float
    converged: bool
    warnings: List[str]
    grid: tuple
    cells_across: int
    dx_mm: float
    dt_s: float
    mach_lattice: float
    tau: float
    blockage: float
    storage: Storage
    memory_bytes: int
    averaged_over: float
    simulated: float
    time_unit: str
    q_ref_pa: float
    thin_mm: float
    cells_thin: float
    kind: str = "wind"
    setup_s: float = 0.0        # voxelization, probes and lattice construction


@dataclass
class WindResult:
    summary: WindSummary
    surface_pressure_pa: np.ndarray
    surface_cp: np.ndarray
    slices: List[FlowSlice]
    history: list


def _vertex_normals(m: TriMesh) -> np.ndarray:
    P = np.asarray(m.positions, np.float64)
    T = np.asarray(m.triangles, np.int64)
    nrm = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    vn = np.zeros_like(P)
    for k in range(3):
        np.add.at(vn, T[:, k], nrm)
    ln = np.linalg.norm(vn, axis=1, keepdims=True)
    return vn / np.where(ln > 0, ln, 1.0)


def _check_thin(warnings: List[str], solid_cells: float, m: TriMesh, dx: float) -> None:
    """tunnel.rs:1432-1445."""
    vol = abs(volume(m))
    if vol > 0.0:
        ratio = solid_cells * dx ** 3 / vol
        if ratio < 0.6:
            warnings.append(f"parts of the model are thinner than a grid cell ({ratio * 100.0:.0f} % of its volume became "
                            "solid cells); thin walls still block the air, but use a finer grid for accurate forces")
    else:
        warnings.append("the surface is not closed; inside and outside were guessed by voting along three directions")


def _window_check(w: np.ndarray, spin: bool, axis_l: np.ndarray):
    """tunnel.rs:1143-1188: half-window disagreement, the block-mean standard
    error about a straight-line trend, and whether the halves differ by more
    than their noise, worst over the monitored signals (drag; or thrust and
    torque of a spinning part)."""
    half = len(w) // 2
    sigs = [w[:, :3] @ axis_l, w[:, 3:] @ axis_l] if spin else [w[:, 0]]
    worst = se_rel = 0.0
    trend = False
    nbk = 8
    kk = np.arange(nbk) - (nbk - 1) / 2.0
    for sg in sigs:
        m1, m2 = sg[:half].mean(), sg[half:].mean()
        sc = 0.5 * (abs(m1) + abs(m2))
        bl = len(sg) // nbk
        bm = sg[:bl * nbk].reshape(nbk, bl).mean(axis=1)
        mm = bm.mean()
        slope = (kk * (bm - mm)).sum() / (kk * kk).sum()
        sd = math.sqrt(((bm - mm - slope * kk) ** 2).sum() / (nbk - 2))
        se = sd / math.sqrt(nbk)
        if sc > 0.0:
            worst = max(worst, abs(m2 - m1) / 2.0 / sc)
            se_rel = max(se_rel, se / sc)
            trend = trend or abs(m2 - m1) > 6.0 * se
    return worst, se_rel, trend, half


class Cancelled(Exception):
    """A run stopped by its caller (core progress.rs:44-50, "stopped")."""

    def __init__(self):
        super().__init__("stopped")


class Progress:
    """What a run reports to and asks of its caller (core progress.rs:6-13):
    `report(fraction, message)` after every batch, `metric(name, step,
    value)` for the live plots (thrust_n / torque_nm of a spinning part,
    drag_n / lift_n in a wind), and `cancelled()`, asked after the
    voxelization and after every batch; a true answer ends the run with
    `Cancelled`. The base class reports nothing and never cancels."""

    def report(self, fraction: float, message: str) -> None:
        pass

    def metric(self, name: str, step: int, value: float) -> None:
        pass

    def cancelled(self) -> bool:
        return False


class _CallbackProgress(Progress):
    """A plain `progress(fraction, message)` callable as a Progress."""

    def __init__(self, fn):
        self.fn = fn

    def report(self, fraction, message):
        self.fn(fraction, message)


def _as_progress(progress) -> Progress:
    if progress is None:
        return Progress()
    if isinstance(progress, Progress) or hasattr(progress, "cancelled"):
        return progress
    return _CallbackProgress(progress)


def run_with(cfg: WindConfig, surface: TriMesh, unit_m: float, opt: Optional[SolverOptions] = None,
             progress=None) -> WindResult:
    """Runs the wind tunnel (tunnel.rs:902-1392). `surface` in metres;
    `unit_m` metres per model unit, used for the spin centre and the slice
    positions. `progress` is a Progress (or a plain `progress(fraction,
    message)` callable): reports and live metrics after every batch; a run
    it cancels raises `Cancelled` (tunnel.rs:947-949, 1139-1141)."""
    t_start = time.perf_counter()
    progress = _as_progress(progress)
    opt = opt or SolverOptions()
    p = plan(cfg, surface, unit_m, opt, None)
    warnings: List[str] = []
    P = np.asarray(surface.positions, np.float64)
    q = (P - p.pref) @ p.r.T
    to_body = q / p.dx
    if p.spin:
        # occupancy 8x finer than the lattice, capped at 64 M samples (tunnel.rs:914-925)
        vol = float(np.prod(to_body.max(axis=0) - to_body.min(axis=0) + 2.0))
        h = max((vol / 64e6) ** (1.0 / 3.0), 0.125)
        vox = voxelize_spin(to_body, surface.triangles, tuple(p.axis_lat), h)
        solid = float(sum(bin(int(b)).count("1") for b in vox.occ_bits)) * vox.occ_h ** 3
        _check_thin(warnings, solid, surface, p.dx)
        geom = Geometry.spin(vox, p.box_lo, p.box_n)
    elif p.multi is not None:
        # the finest level's cells, in its own coordinates (tunnel.rs:929-938)
        b = p.multi.boxes[-1]
        fd = tuple(2 * (int(b.n[k]) - 1) + 1 for k in range(3))
        vox = voxelize_static27(np.asarray(p.center) + to_body - np.asarray(p.multi.org_f), surface.triangles, fd)
        _check_thin(warnings, float(len(vox.solid)), surface, p.dx)
        geom = Geometry.static(vox)
    else:
        vox = voxelize_static(np.asarray(p.center) + to_body, surface.triangles, p.dims)
        _check_thin(warnings, float(len(vox.solid)), surface, p.dx)
        geom = Geometry.static(vox)
    if progress.cancelled():
        raise Cancelled()
    flip = -1.0 if volume(surface) < 0.0 else 1.0
    normals = (_vertex_normals(surface) * flip) @ p.r.T
    nx, ny, nz = p.dims
    if p.multi is not None:
        # slices through the part centre in global finest cells; each pixel
        # from the finest level covering it, with a pitch that keeps a slice
        # under ~4 M pixels (tunnel.rs:966-989)
        pitch = 1
        while max((nx // pitch) * (nz // pitch), (ny // pitch) * (nz // pitch)) > 4_000_000:
            pitch *= 2

        def snap(x):
            return max(int(round(x / pitch)) * pitch, pitch)

        def span(n):
            return max((n - 1 - pitch) // pitch, 1)

        xc = snap(min(max(p.center[0], 1.0), nx - 2.0))
        yc = snap(min(max(p.center[1], 1.0), ny - 2.0))
        slices = [SliceDef((pitch, yc, pitch), (pitch, 0, 0), (0, 0, pitch), span(nx), span(nz)),
                  SliceDef((xc, pitch, pitch), (0, pitch, 0), (0, 0, pitch), span(ny), span(nz))]
    else:
        xc = max(int(round(min(max(p.center[0], 1.0), nx - 2.0))), 1)
        yc = max(int(round(min(max(p.center[1], 1.0), ny - 2.0))), 1)
        slices = [SliceDef((1, yc, 1), (1, 0, 0), (0, 0, 1), max(nx - 2, 1), max(nz - 2, 1)),
                  SliceDef((xc, 1, 1), (0, 1, 0), (0, 0, 1), max(ny - 2, 1), max(nz - 2, 1))]
    probes = Probes([(tuple(b), tuple(nv)) for b, nv in zip(to_body, normals)], slices, opt.vertex_offset)
    u_inf = (p.u_lat, 0.0, 0.0)
    if p.multi is not None:
        nl = len(p.multi.cells)
        msetup = MultiSetup(dims0=p.multi.dims0, boxes=list(p.multi.boxes), faces=p.faces,
                            nu0=p.nu_lat / float(1 << (nl - 1)), smagorinsky=0.0, u_inf=u_inf, u_init=u_inf,
                            sponge=p.sponge, sponge_tau=1.0, sponge_sigma=opt.sponge_sigma,
                            absorb_velocity=opt.absorb_velocity, total_pressure_inlet=opt.total_pressure_inlet,
                            storage=p.storage,
                            center=tuple(p.center[k] - p.multi.org_f[k] for k in range(3)))
        lat = MultiLattice(msetup, geom, probes)
    else:
        setup = LatticeSetup(dims=p.dims, faces=p.faces, nu=p.nu_lat, smagorinsky=opt.smagorinsky, u_inf=u_inf,
                             u_init=u_inf, sponge=p.sponge, sponge_tau=1.0, sponge_sigma=opt.sponge_sigma,
                             absorb_velocity=opt.absorb_velocity and not p.still,
                             total_pressure_inlet=opt.total_pressure_inlet, storage=p.storage, center=p.center,
                             wall_model=opt.wall_model)
        lat = Lattice(setup, geom, probes)
    ss = lat.sub_steps() if p.multi is not None else 1
    t_setup = time.perf_counter() - t_start
    try:
        if lat.storage != p.storage:
            warnings.append("this GPU has no half-precision storage; used FP32")
        if lat.storage == Storage.FP16:
            warnings.append("FP16 storage (to fit the memory budget): slightly noisier forces")

        total_steps = -(-int(math.ceil(p.duration * p.steps_unit)) // ss) * ss
        omega_mag = _norm(p.omega_lat)
        ramp_steps = int(math.ceil(opt.spin_ramp_revs * 2.0 * math.pi / omega_mag)) if p.spin else 0
        unit_steps = p.steps_unit
        # wind: skip the first 35 % of the run, then average at least 10
        # passes; still air: 3 revolutions, then at least 2 (tunnel.rs:1047-1058)
        tr_units, min_win_units = (3.0, 2.0) if p.still else (0.35 * p.duration, 10.0)
        w0 = int(min(tr_units, 0.35 * p.duration) * unit_steps)
        w0 = max(w0, ramp_steps + int(0.5 * unit_steps))
        min_win = int(min(min_win_units, 0.3 * p.duration) * unit_steps)
        w0 = min(w0, total_steps // 2)
        sample_every = max(int(math.floor(unit_steps / 100.0)), 1)
        batch = -(-min(max(int(0.25 * MLUPS_FP32 * 1e6 / p.updates_per_step), 16), 4096) // ss) * ss

        hist = np.zeros((total_steps, 6))
        theta = 0.0
        axis_l = _normalize(p.axis_lat)
        rot_prev = rotation(axis_l, 0.0)
        step, converged, sampling = 0, False, False
        fscale = U.force_scale(cfg.air.density, p.dx, p.dt)
        mscale = fscale * p.dx
        history = []
        omega_dot = float(p.omega_lat @ axis_l)
        t_loop = time.perf_counter()
        while step < total_steps:
            nb = min(batch, total_steps - step)
            ins = []
            for k in range(nb):
                s = step + k
                if p.spin:
                    if s < ramp_steps:
                        x = (s + 0.5) / ramp_steps
                        f = x * x * (3.0 - 2.0 * x)
                    else:
                        f = 1.0
                    # signed: a negative rpm turns the part the other way about +axis
                    theta += f * omega_dot
                    rot = rotation(axis_l, theta)
                    st = StepIn(rot, rot_prev, tuple(p.omega_lat * f), False)
                    rot_prev = rot
                else:
                    st = StepIn.still(False)
                if s + 1 > w0 and (s + 1 - w0) % sample_every == 0:
                    st.sample = True
                ins.append(st)
            if not sampling and step + nb > w0:
                lat.reset_samples()
                sampling = True
            h = lat.run(ins)
            hist[step:step + nb] = h
            step += nb
            fm = h.mean(axis=0)
            f_model = U.to_model(p.r, fm[:3]) * fscale
            m_model = U.to_model(p.r, fm[3:]) * mscale
            history.append((step / unit_steps, f_model, m_model))
            if p.spin:
                progress.metric("thrust_n", step, float(np.dot(f_model, p.axis_model)))
                progress.metric("torque_nm", step, float(np.dot(m_model, p.axis_model)))
            if not p.still:
                progress.metric("drag_n", step, float(np.dot(f_model, p.r[0])))
                progress.metric("lift_n", step, float(np.dot(f_model, p.r[2])))
            el = time.perf_counter() - t_loop
            progress.report(0.02 + 0.96 * step / total_steps,
                            f"Solving: {step} of {total_steps} steps ({step / unit_steps:.1f} "
                            f"{'revolutions' if p.still else 'flow passes'}, {step * p.updates_per_step / el / 1e6:.0f} MLUPs)")
            if progress.cancelled():
                raise Cancelled()
            if not np.all(np.isfinite(hist[:step])):
                raise RuntimeError("the flow solution blew up (unstable); try a finer grid or a lower speed")
            if step > w0 + min_win:
                worst, se_rel, trend, half = _window_check(hist[w0:step], p.spin, axis_l)
                if worst < opt.tolerance and se_rel < opt.tolerance:
                    converged = True
                    break
                # still drifting: restart the window at its middle, never in
                # the last 30 % of the run (tunnel.rs:1190-1200)
                if worst > 3.0 * opt.tolerance and trend and step - w0 >= min_win and step < 0.7 * total_steps:
                    w0 += half
                    lat.reset_samples()
        loop_s = time.perf_counter() - t_loop
        mlups = step * p.updates_per_step / loop_s / 1e6
        hist = hist[:step]
        avg = hist[min(w0, max(len(hist) - 1, 0)):].mean(axis=0)
        rho, speed = cfg.air.density, cfg.speed_m_s
        spin = cfg.spin if p.spin else None
        s = U.summarize(avg, p.r, rho, p.dx, p.dt, speed=speed, frontal_area=p.frontal_area,
                        spin_axis=p.axis_model if spin else None, rpm=spin.rpm if spin else 0.0,
                        omega_lat=p.omega_lat, rotor_radius=p.rotor_radius)
        q_ref = s["q_ref"]
        if p.spin:
            # blade Reynolds number at 75 % radius, the mean chord from the
            # surface area (both faces of two blades: A = 4 R c)
            chord = float(_triangle_areas(surface).sum()) / (4.0 * p.rotor_radius)
            reynolds = (0.75 * omega_mag / p.dt * p.rotor_radius + speed) * chord / cfg.air.kinematic_viscosity
        else:
            reynolds = speed * p.d_front / cfg.air.kinematic_viscosity

        vel = p.dx / p.dt
        pconv = rho * vel * vel / 3.0
        pres = ((lat.vertex_means().astype(np.float64) - 1.0) * pconv).astype(np.float32)
        cp = (pres.astype(np.float64) / q_ref).astype(np.float32)
        sm = lat.slice_means().astype(np.float64)
        out_slices: List[FlowSlice] = []
        off = 0
        for si, sd in enumerate(slices):
            npx = sd.nu * sd.nv
            px = sm[off:off + npx]
            off += npx
            su, sv = np.asarray(sd.u, np.float64), np.asarray(sd.v, np.float64)
            ua, va = su / np.linalg.norm(su), sv / np.linalg.norm(sv)
            corner = np.asarray(sd.origin, np.float64) - 0.5 * (su + sv)
            origin = p.pref + U.to_model(p.r, (corner - np.asarray(p.center)) * p.dx)
            u_mm = U.to_model(p.r, su) * (sd.nu * p.dx) / unit_m
            v_mm = U.to_model(p.r, sv) * (sd.nv * p.dx) / unit_m
            uvw = px[:, 1:]

            def mk(name, label, unit, data):
                data = data.astype(np.float32)
                fin = data[np.isfinite(data)]
                return SliceField(name, label, unit, float(fin.min()) if fin.size else math.inf,
                                  float(fin.max()) if fin.size else -math.inf, data)

            if si == 0:
                name, label = "along", "Along the spin axis" if p.still else "Along the wind"
                ul, vl = ("axial" if p.still else "streamwise"), "up"
            else:
                name, label = "across", ("Across the axis, through the rotor" if p.still
                                         else "Across the wind, through the part")
                ul, vl = "sideways", "up"
            out_slices.append(FlowSlice(name, label, origin / unit_m, u_mm, v_mm, sd.nu, sd.nv, [
                mk("speed", "Air speed", "m/s", np.linalg.norm(uvw, axis=1) * vel),
                mk("pressure", "Pressure (relative to far field)", "Pa", (px[:, 0] - 1.0) * pconv),
                mk("u", f"In-plane velocity, {ul}", "m/s", (uvw @ ua) * vel),
                mk("v", f"In-plane velocity, {vl}", "m/s", (uvw @ va) * vel)]))

        if not converged:
            warnings.append(f"The forces had not settled to within {opt.tolerance * 100.0:.1f}% by the end of the run, "
                            "so the numbers are averages over its last part.")
        if p.blockage > 0.06:
            warnings.append(f"The part fills {p.blockage * 100.0:.1f}% of the tunnel's cross-section, so drag may read "
                            "a little high.")
        tw = _thin_warning(p)
        if tw:
            warnings.append(tw)
        tau0 = 0.5 + 3.0 * p.nu_lat
        wall_re = reynolds if p.spin else speed * p.dx / cfg.air.kinematic_viscosity
        if tau0 < 0.502 and wall_re > 300.0:
            warnings.append("The thin layer of air hugging the surface is finer than one grid cell here (cell Reynolds "
                            f"number {wall_re:.0f}), so surface friction and where the flow breaks away are approximate.")
        thrust, torque, power = s["thrust_n"], s["torque_nm"], s["power_w"]
        drag = s["drag_n"]
        if thrust is not None:
            headline = (f"Thrust {sig(abs(thrust), 3)} N ({sig(abs(thrust) / 9.80665e-3, 3)} g), torque "
                        f"{sig(abs(torque) * 1e3, 3)} mN·m, {sig(abs(power), 3)} W at {sig(abs(spin.rpm), 4)} rpm")
        elif s["cd"] is not None:
            headline = f"Drag {sig(drag, 3)} N at {sig(speed, 3)} m/s (Cd {sig(s['cd'], 3)})"
        else:
            headline = f"Drag {sig(drag, 3)} N at {sig(speed, 3)} m/s"
        summary = WindSummary(
            headline=headline, force_n=s["force_n"], moment_nm=s["moment_nm"], drag_n=drag, lift_n=s["lift_n"],
            side_n=s["side_n"], cd=s["cd"], ref_area_mm2=p.frontal_area * 1e6, thrust_n=thrust, torque_nm=torque,
            power_w=power, reynolds=reynolds,
            cells=sum(p.multi.cells) if p.multi is not None else int(np.prod(np.asarray(p.dims, np.int64))), steps=step,
            mlups=mlups, seconds=time.perf_counter() - t_start, converged=converged, warnings=warnings, grid=p.dims,
            cells_across=p.cells_across, dx_mm=p.dx * 1e3, dt_s=p.dt,
            mach_lattice=(p.u_lat + omega_mag * p.rotor_radius / p.dx) * math.sqrt(3.0), tau=tau0,
            blockage=p.blockage, storage=lat.storage, memory_bytes=lat.memory_bytes,
            averaged_over=max(step - w0, 0) / unit_steps, simulated=step / unit_steps,
            time_unit="revolutions" if p.still else "passes", q_ref_pa=q_ref,
            thin_mm=0.0 if p.spin else p.d_thin * 1e3, cells_thin=0.0 if p.spin else p.cells_thin(),
            setup_s=t_setup)
        return WindResult(summary, pres, cp, out_slices, history)
    finally:
        lat.close()


# ---------------------------------------------------------------- JSON forms (serde of core config.rs and tunnel.rs)
_STORAGE_NAME = {Storage.FP32: "Fp32", Storage.FP16: "Fp16", Storage.BF16: "Bf16", Storage.FP64: "Fp64"}


def _vec3(v, what: str) -> tuple:
    if not isinstance(v, (list, tuple)) or len(v) != 3:
        raise ValueError(f"{what} must be three numbers")
    return tuple(float(x) for x in v)


def config_from_json(obj) -> WindConfig:
    """A WindConfig from its JSON form (config.rs:137-188): wind_dir and
    speed_m_s required; air, spin, quality ({"type": "draft" | "normal" |
    "fine" | "custom", "cells_across": n}) and duration optional."""
    if isinstance(obj, (str, bytes)):
        import json
        obj = json.loads(obj)
    for k in ("wind_dir", "speed_m_s"):
        if k not in obj:
            raise ValueError(f"missing field `{k}`")
    cfg = WindConfig(wind_dir=_vec3(obj["wind_dir"], "wind_dir"), speed_m_s=float(obj["speed_m_s"]))
    air = obj.get("air")
    if air is not None:
        cfg.air = Air(float(air["density"]), float(air["kinematic_viscosity"]))
    spin = obj.get("spin")
    if spin is not None:
        for k in ("rpm", "axis", "center_mm"):
            if k not in spin:
                raise ValueError(f"missing field `{k}` in spin")
        cfg.spin = Spin(float(spin["rpm"]), _vec3(spin["axis"], "spin.axis"), _vec3(spin["center_mm"], "spin.center_mm"))
    q = obj.get("quality")
    if q is not None:
        t = q.get("type")
        if t == "custom":
            cfg.quality = int(q["cells_across"])
        elif t in ("draft", "normal", "fine"):
            cfg.quality = t
        else:
            raise ValueError(f"unknown quality `{t}`")
    d = obj.get("duration")
    cfg.duration = None if d is None else float(d)
    return cfg


def config_to_json(cfg: WindConfig) -> dict:
    q = {"type": "custom", "cells_across": int(cfg.quality)} if isinstance(cfg.quality, int) else {"type": cfg.quality}
    return {"wind_dir": list(cfg.wind_dir), "speed_m_s": cfg.speed_m_s,
            "air": {"density": cfg.air.density, "kinematic_viscosity": cfg.air.kinematic_viscosity},
            "spin": None if cfg.spin is None else {"rpm": cfg.spin.rpm, "axis": list(cfg.spin.axis),
                                                   "center_mm": list(cfg.spin.center_mm)},
            "quality": q, "duration": cfg.duration}


def _num(x):
    if x is None:
        return None
    x = float(x)
    return x if math.isfinite(x) else None


_SUMMARY_FIELDS = ("kind", "headline", "force_n", "moment_nm", "drag_n", "lift_n", "side_n", "cd", "ref_area_mm2",
                   "thrust_n", "torque_nm", "power_w", "reynolds", "cells", "steps", "mlups", "seconds", "converged",
                   "warnings", "grid", "cells_across", "dx_mm", "dt_s", "mach_lattice", "tau", "blockage", "storage",
                   "memory_bytes", "averaged_over", "simulated", "time_unit", "q_ref_pa", "thin_mm", "cells_thin")


def summary_to_json(s: WindSummary) -> dict:
    """The WindSummary fields in their serde order (tunnel.rs:262-304)."""
    out = {}
    for k in _SUMMARY_FIELDS:
        v = getattr(s, k)
        if k == "storage":
            v = _STORAGE_NAME[Storage(v)]
        elif k in ("force_n", "moment_nm"):
            v = [_num(x) for x in v]
        elif k == "grid":
            v = [int(x) for x in v]
        elif isinstance(v, (bool, str, list)) or v is None:
            pass
        elif k in ("cells", "steps", "cells_across", "memory_bytes"):
            v = int(v)
        else:
            v = _num(v)
        out[k] = v
    return out


def estimate_to_json(e: Estimate) -> dict:
    """tunnel.rs:325-345 serde form."""
    return {"fits": bool(e.fits), "cells": int(e.cells), "grid": [int(x) for x in e.grid],
            "cells_across": int(e.cells_across), "memory_bytes": int(e.memory_bytes),
            "budget_bytes": int(e.budget_bytes), "storage": _STORAGE_NAME[Storage(e.storage)], "steps": int(e.steps),
            "seconds": float(e.seconds), "note": e.note, "thin_mm": float(e.thin_mm), "cells_thin": float(e.cells_thin),
            "thin_warning": e.thin_warning}


def slice_to_json(sl: FlowSlice) -> dict:
    """FlowSlice serde form, field data included (NaN as null)."""
    return {"name": sl.name, "label": sl.label, "origin_mm": [float(x) for x in sl.origin_mm],
            "u_mm": [float(x) for x in sl.u_mm], "v_mm": [float(x) for x in sl.v_mm], "nu": int(sl.nu),
            "nv": int(sl.nv),
            "fields": [{"name": f.name, "label": f.label, "unit": f.unit, "min": _num(f.min), "max": _num(f.max),
                        "data": [_num(x) for x in np.asarray(f.data, np.float64)]} for f in sl.fields]}


# ---------------------------------------------------------------- result files (engine wind.rs:12-15, 440-497)
def _range(v: np.ndarray):
    """wind.rs:452-461: finite min and max, (0, 0) when none."""
    f = v[np.isfinite(v)]
    return (float(f.min()), float(f.max())) if f.size else (0.0, 0.0)


def safe_name(name: str) -> bool:
    """engine api.rs:392-394."""
    return bool(name) and all(c.isascii() and (c.isalnum() or c in "_.-") for c in name) and ".." not in name


def write_display_mesh(directory: str, prefix: str, m: TriMesh, patch=None) -> None:
    """core mesh.rs:176-198: positions f32 x3, area-weighted vertex normals
    f32 x3, indices u32 x3, patch u32 per triangle (0 when none), little-endian."""
    import os
    os.makedirs(directory, exist_ok=True)
    P = np.asarray(m.positions, np.float64)
    vn = _vertex_normals(m) if len(m.triangles) else np.zeros_like(P)
    np.asarray(P, "<f4").tofile(os.path.join(directory, f"{prefix}.positions.f32"))
    np.asarray(vn, "<f4").tofile(os.path.join(directory, f"{prefix}.normals.f32"))
    np.asarray(m.triangles, "<u4").reshape(-1).tofile(os.path.join(directory, f"{prefix}.indices.u32"))
    pt = np.zeros(len(m.triangles), "<u4") if patch is None else np.asarray(patch, "<u4")
    pt.tofile(os.path.join(directory, f"{prefix}.patch.u32"))


def write_result(directory: str, r: WindResult, surface_mm: Optional[TriMesh] = None):
    """The result files of a run (wind.rs:12-15, 465-497): the model surface
    as mesh.* (when given, in model mm), field.pressure.f32 and field.cp.f32
    per surface vertex, and each slice as slice.<name>.json plus
    slice.<name>.<field>.f32 (v-major rows of nu values, NaN inside the part).
    Returns the run's (fields, slices) lists."""
    import json
    import os
    os.makedirs(directory, exist_ok=True)
    if surface_mm is not None:
        write_display_mesh(directory, "mesh", surface_mm)
    fields = []
    pres = np.asarray(r.surface_pressure_pa, "<f4")
    if np.isfinite(pres).any():
        lo, hi = _range(pres)
        pres.tofile(os.path.join(directory, "field.pressure.f32"))
        m = max(abs(lo), abs(hi))
        fields.append({"name": "pressure", "label": "Air pressure", "unit": "Pa", "min": -m, "max": m,
                       "data_min": lo, "data_max": hi, "scale": "linear", "diverging": True})
    cp = np.asarray(r.surface_cp, "<f4")
    if np.isfinite(cp).any():
        lo, hi = _range(cp)
        cp.tofile(os.path.join(directory, "field.cp.f32"))
        fields.append({"name": "cp", "label": "Pressure coefficient", "unit": "", "min": lo, "max": hi,
                       "scale": "linear", "diverging": True})
    slices = []
    for sl in r.slices:
        if not sl.fields:
            continue
        infos = []
        for f in sl.fields:
            if not safe_name(f.name):
                continue
            lo, hi = _range(np.asarray(f.data))
            np.asarray(f.data, "<f4").tofile(os.path.join(directory, f"slice.{sl.name}.{f.name}.f32"))
            infos.append({"name": f.name, "label": f.label, "unit": f.unit, "min": lo, "max": hi})
        meta = {"name": sl.name, "label": sl.label, "origin_mm": [float(x) for x in sl.origin_mm],
                "u_mm": [float(x) for x in sl.u_mm], "v_mm": [float(x) for x in sl.v_mm], "nu": int(sl.nu),
                "nv": int(sl.nv), "fields": infos}
        with open(os.path.join(directory, f"slice.{sl.name}.json"), "w") as fh:
            json.dump(meta, fh, separators=(",", ":"))
        slices.append({"name": sl.name, "label": sl.label})
    return fields, slices


def read_slice(directory: str, name: str):
    """Reads slice.<name>.json and its fields back: (meta, {field: (nv, nu) array})."""
    import json
    import os
    with open(os.path.join(directory, f"slice.{name}.json")) as fh:
        meta = json.load(fh)
    data = {f["name"]: np.fromfile(os.path.join(directory, f"slice.{name}.{f['name']}.f32"), "<f4").reshape(
        meta["nv"], meta["nu"]) for f in meta["fields"]}
    return meta, data


# ---------------------------------------------------------------- reference cases (validation.rs)
SPHERE_D = 0.05
SPHERE_U = 10.0


def sphere_cd_cgw(re: float) -> float:
    """validation.rs:10-23: Clift, Grace & Weber (1978)."""
    w = math.log10(re)
    if re <= 20.0:
        return 24.0 / re * (1.0 + 0.1315 * re ** (0.82 - 0.05 * w))
    if re <= 260.0:
        return 24.0 / re * (1.0 + 0.1935 * re ** 0.6305)
    if re <= 1500.0:
        return 10.0 ** (1.6435 - 1.1242 * w + 0.1558 * w * w)
    if re <= 1.2e4:
        return 10.0 ** (-2.4571 + 2.5558 * w - 0.9295 * w * w + 0.1049 * w ** 3)
    return 10.0 ** (-1.9181 + 0.6370 * w - 0.0636 * w * w)


def sphere_cd_sn(re: float) -> float:
    """validation.rs:26-28: Schiller & Naumann (1933)."""
    return 24.0 / re * (1.0 + 0.15 * re ** 0.687)


def sphere_case(re: float, cells_per_d: int):
    """validation.rs:34-45: a 50 mm sphere in a 10 m/s wind along +x."""
    m = sphere([0.0, 0.0, 0.0], SPHERE_D / 2.0, 5)
    cfg = WindConfig(wind_dir=(1.0, 0.0, 0.0), speed_m_s=SPHERE_U,
                     air=Air(density=1.2, kinematic_viscosity=SPHERE_U * SPHERE_D / re), quality=int(cells_per_d))
    return cfg, m


def sphere_drag(re: float, cells_per_d: int, duration: float, opt: Optional[SolverOptions] = None, progress=None):
    """validation.rs:47-55: Cd with the exact sphere area (the rasterized one
    is within 0.1 %)."""
    cfg, m = sphere_case(re, cells_per_d)
    cfg.duration = duration
    r = run_with(cfg, m, 1e-3, opt, progress)
    a = math.pi * SPHERE_D * SPHERE_D / 4.0
    return r.summary.drag_n / (0.5 * cfg.air.density * SPHERE_U * SPHERE_U * a), r


def disk_drag(re: float, cells_per_d: int, duration: float, opt: Optional[SolverOptions] = None, progress=None):
    """validation.rs:59-66: a disk of the sphere's diameter, 5 % thick, normal to the wind."""
    cfg, _ = sphere_case(re, cells_per_d)
    cfg.duration = duration
    r = run_with(cfg, disk([0.0, 0.0, 0.0], SPHERE_D, 0.05 * SPHERE_D), 1e-3, opt, progress)
    a = math.pi * SPHERE_D * SPHERE_D / 4.0
    return r.summary.drag_n / (0.5 * cfg.air.density * SPHERE_U * SPHERE_U * a), r


def rotating_sphere(re: float, cells_per_d: int, revs: float, opt: Optional[SolverOptions] = None, progress=None):
    """validation.rs:70-88: torque on a 25 mm sphere spinning at 600 rpm in
    still air at rotational Reynolds number omega R^2 / nu, and Stokes'
    8 pi mu R^3 omega."""
    rad = 0.025
    m = sphere([0.0, 0.0, 0.0], rad, 5)
    rpm = 600.0
    omega = rpm * 2.0 * math.pi / 60.0
    nu = omega * rad * rad / re
    air = Air(density=1.2, kinematic_viscosity=nu)
    cfg = WindConfig(wind_dir=(0.0, 0.0, 1.0), speed_m_s=0.0, air=air, spin=Spin(rpm, (0.0, 0.0, 1.0)),
                     quality=int(cells_per_d), duration=revs)
    r = run_with(cfg, m, 1e-3, opt, progress)
    stokes = 8.0 * math.pi * air.density * nu * rad ** 3 * omega
    return (r.summary.torque_nm or 0.0), stokes, r


def prop_hover(spec, rpm: float, cells: int, revs: float, opt: Optional[SolverOptions] = None, progress=None) -> WindResult:
    """validation.rs:91-102: the test propeller spinning in still air about +z."""
    cfg = WindConfig(wind_dir=(0.0, 0.0, 1.0), speed_m_s=0.0, air=Air(), spin=Spin(rpm, (0.0, 0.0, 1.0)),
                     quality=int(cells), duration=revs)
    return run_with(cfg, spec.mesh(), 1e-3, opt, progress)
