"""Named workloads: the reference's own test lattices (crates/cfd/tests/
channel.rs:22-73, momentum.rs:9-37, examples/bench.rs) and the sphere-in-a-
channel cases the throughput and accuracy figures are quoted on. Each
builder returns (LatticeSetup, Geometry, meta) ready for `Lattice`."""
from __future__ import annotations

import math

from . import FaceBc, Geometry, LatticeSetup, Storage, voxelize_static
from .shapes import cuboid, cylinder, merge, sphere


def static_geometry(mesh, dims) -> Geometry:
    return Geometry.static(voxelize_static(mesh.positions, mesh.triangles, dims))


def poiseuille(h: float, storage: Storage = Storage.FP32, umax: float = 0.02):
    """Plane Poiseuille flow between plates off the grid (channel.rs:39-73):
    3 x ny x 3, all faces periodic, plates at y = 1.37 and 1.37 + h."""
    ya = 1.37
    yb = ya + h
    ny = int(math.ceil(yb)) + 2
    nx = nz = 3
    lo = cuboid([-5.0, -10.0, -5.0], [nx + 5.0, ya, nz + 5.0])
    hi = cuboid([-5.0, yb, -5.0], [nx + 5.0, ny + 10.0, nz + 5.0])
    nu = 0.1
    g = umax * 8.0 * nu / (h * h)
    setup = LatticeSetup(dims=(nx, ny, nz), faces=(FaceBc.PERIODIC,) * 6, nu=nu, body_force=(g, 0.0, 0.0),
                         storage=storage, debug_dump=True)
    geom = static_geometry(merge(lo, hi), (nx, ny, nz))
    steps = int(3.0 * h * h / nu)
    return setup, geom, dict(ya=ya, yb=yb, ny=ny, nx=nx, nz=nz, nu=nu, g=g, steps=steps)


def poiseuille_error(meta, mac, hist_last):
    """L2 relative error of u_x on the x=1, z=1 column and the wall-force balance (channel.rs:56-72)."""
    import numpy as np
    ya, yb, ny, nx, nu, g = meta["ya"], meta["yb"], meta["ny"], meta["nx"], meta["nu"], meta["g"]
    e2 = r2 = 0.0
    for y in range(ny):
        yf = float(y)
        if yf <= ya or yf >= yb:
            continue
        exact = g / (2.0 * nu) * (yf - ya) * (yb - yf)
        c = (1 * ny + y) * nx + 1
        u = float(mac[c][1])
        e2 += (u - exact) ** 2
        r2 += exact * exact
    rho = mac[:, 0]
    fluid = float(np.sum(rho[~np.isnan(rho)].astype(np.float64)))
    bal = (hist_last[0] - g * fluid) / (g * fluid)
    return math.sqrt(e2 / r2), bal


def ball_box(n: int, r: float, nu: float, u0: float, g: float, faces, storage: Storage = Storage.FP32):
    """momentum.rs:9-37: an n^3 box with a sphere at (c+0.3, c-0.6, c+0.17)."""
    c = n / 2.0
    m = sphere([c + 0.3, c - 0.6, c + 0.17], r, 5)
    setup = LatticeSetup(dims=(n, n, n), faces=tuple(faces), nu=nu, u_inf=(u0, 0.0, 0.0), u_init=(u0, 0.0, 0.0),
                         body_force=(g, 0.0, 0.0), storage=storage, center=(c, c, c), debug_dump=True)
    return setup, static_geometry(m, (n, n, n))


def channel_walls(dims):
    """Solid slabs whose surfaces sit half a cell inside the y and z edges
    (y, z = 0.5 and n - 1.5): halfway bounce-back channel walls."""
    nx, ny, nz = dims
    return [
        cuboid([-5.0, -10.0, -5.0], [nx + 5.0, 0.5, nz + 5.0]),
        cuboid([-5.0, ny - 1.5, -5.0], [nx + 5.0, ny + 10.0, nz + 5.0]),
        cuboid([-5.0, -5.0, -10.0], [nx + 5.0, ny + 5.0, 0.5]),
        cuboid([-5.0, -5.0, nz - 1.5], [nx + 5.0, ny + 5.0, nz + 10.0]),
    ]


def sphere_channel(dims=(128, 64, 64), radius=8.0, x0=32.0, re=100.0, u=0.05, cs=0.0,
                   storage: Storage = Storage.FP32, debug_dump=False, subdiv=5):
    """Flow past a sphere in a walled channel: -x velocity inlet (u), +x open
    outlet, y/z walls as halfway bounce-back slabs (periodic behind them),
    viscosity from Re on the sphere diameter. 128x64x64 r=8 Re=100 is the
    reference-sized case; 256x128x128 r=16 Re=1000 Cs=0.1 the large one."""
    nx, ny, nz = dims
    cy, cz = (ny - 1) / 2.0, (nz - 1) / 2.0
    nu = u * 2.0 * radius / re
    mesh = merge(sphere([x0, cy, cz], radius, subdiv), *channel_walls(dims))
    setup = LatticeSetup(dims=tuple(dims),
                         faces=(FaceBc.VELOCITY, FaceBc.OPEN, FaceBc.PERIODIC, FaceBc.PERIODIC, FaceBc.PERIODIC, FaceBc.PERIODIC),
                         nu=nu, smagorinsky=cs, u_inf=(u, 0.0, 0.0), u_init=(u, 0.0, 0.0), storage=storage,
                         center=(x0, cy, cz), debug_dump=debug_dump, total_pressure_inlet=False)
    geom = static_geometry(mesh, dims)
    cd_scale = 1.0 / (0.5 * u * u * math.pi * radius * radius)
    return setup, geom, dict(nu=nu, tau0=0.5 + 3.0 * nu, cd_scale=cd_scale, radius=radius, u=u, re=re)


def sphere_channel_small(storage: Storage = Storage.FP32, debug_dump=False):
    """128x64x64, sphere r=8 at x=32, u=0.05, Re=100, Cs=0 (2000 steps in full)."""
    return sphere_channel((128, 64, 64), 8.0, 32.0, 100.0, 0.05, 0.0, storage, debug_dump)


def sphere_channel_large(storage: Storage = Storage.FP32, debug_dump=False):
    """256x128x128, sphere r=16 at x=64, u=0.05, Re=1000, Smagorinsky Cs=0.1 (5000 steps in full)."""
    return sphere_channel((256, 128, 128), 16.0, 64.0, 1000.0, 0.05, 0.1, storage, debug_dump)


def porous_box(dims=(128, 64, 64), box=24, x0=40.0, u=0.04, re=200.0, seed=20261203,
               storage: Storage = Storage.FP32, debug_dump=False):
    """Drag-sensitivity case: a seeded random solid fraction in [0, 1] inside a
    box x box x box cube centred at (x0, ny/2, nz/2), zero elsewhere, under
    partial bounce-back; velocity inlet u at -x, open outlet at +x, periodic
    in y and z; nu from Re on the box size. Returns (setup, phi, meta); phi
    is cells float32, x fastest."""
    import numpy as np
    nx, ny, nz = dims
    nu = u * box / re
    setup = LatticeSetup(dims=dims, faces=(FaceBc.VELOCITY, FaceBc.OPEN, FaceBc.PERIODIC, FaceBc.PERIODIC,
                                           FaceBc.PERIODIC, FaceBc.PERIODIC),
                         nu=nu, u_inf=(u, 0.0, 0.0), u_init=(u, 0.0, 0.0), storage=storage, debug_dump=debug_dump)
    rng = np.random.default_rng(seed)
    phi = np.zeros((nz, ny, nx), np.float32)
    lo = [int(round(x0 - box / 2)), ny // 2 - box // 2, nz // 2 - box // 2]
    phi[lo[2]:lo[2] + box, lo[1]:lo[1] + box, lo[0]:lo[0] + box] = rng.random((box, box, box), dtype=np.float32)
    return setup, phi.reshape(-1), dict(nu=nu, tau0=0.5 + 3.0 * nu, box_lo=tuple(lo), box=box, u=u, re=re, seed=seed)


def drone_arm_mesh(scale: float = 1.0, center=(0.0, 0.0, 0.0), seed: int = 20261203):
    """Synthetic drone-frame arm: a box hub of 40 x 40 x 12 cells (x, y, z)
    joined to 4 cylinders (radius 6-12 cells, 90 cells long) whose axes lie
    within 30 degrees of the y-axis, two toward +y and two toward -y; seeded,
    times `scale`. Overlaps count as one solid (non-zero winding)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    cx, cy, cz = center
    s = scale
    parts = [cuboid([cx - 20 * s, cy - 20 * s, cz - 6 * s], [cx + 20 * s, cy + 20 * s, cz + 6 * s])]
    for k in range(4):
        sign = 1.0 if k < 2 else -1.0
        tilt = math.radians(rng.uniform(0.0, 30.0))
        az = rng.uniform(0.0, 2.0 * math.pi)
        d = np.array([math.sin(tilt) * math.cos(az), sign * math.cos(tilt), math.sin(tilt) * math.sin(az)])
        r = rng.uniform(6.0, 12.0) * s
        off = np.array([(-1) ** k * 10.0 * s, 0.0, 0.0])
        p0 = np.array([cx, cy, cz]) + off
        p1 = p0 + 90.0 * s * d
        parts.append(cylinder(list(p0), list(p1), r, 32))
    return merge(*parts)


def drone_arm(dims=(512, 256, 256), u=0.06, cs=0.16, re=2000.0, storage: Storage = Storage.FP16, scale=None,
              seed: int = 20261203, debug_dump=False):
    """The drone-arm tunnel: the arm (scaled with the domain's y size unless
    `scale` is given) at (0.3 nx, ny/2, nz/2), velocity inlet u at -x, open
    outlet at +x, open sides; Smagorinsky Cs; nu from Re on the hub width
    (nu = u * 40 * scale / Re)."""
    nx, ny, nz = dims
    s = (ny / 256.0) if scale is None else scale
    mesh = drone_arm_mesh(s, (0.3 * nx, ny / 2.0, nz / 2.0), seed)
    nu = u * 40.0 * s / re
    setup = LatticeSetup(dims=dims, faces=(FaceBc.VELOCITY, FaceBc.OPEN, FaceBc.OPEN, FaceBc.OPEN, FaceBc.OPEN, FaceBc.OPEN),
                         nu=nu, smagorinsky=cs, u_inf=(u, 0.0, 0.0), u_init=(u, 0.0, 0.0), storage=storage,
                         center=(0.3 * nx, ny / 2.0, nz / 2.0), debug_dump=debug_dump)
    return setup, static_geometry(mesh, dims), dict(nu=nu, tau0=0.5 + 3.0 * nu, scale=s, u=u, re=re, seed=seed)


def rotation(a, t: float):
    """tunnel.rs:484-493: rotation matrix about unit axis `a` by angle t
    (rows of the matrix; its columns are the rotated basis)."""
    c, s = math.cos(t), math.sin(t)
    k = 1.0 - c
    return ((c + a[0] * a[0] * k, a[0] * a[1] * k - a[2] * s, a[0] * a[2] * k + a[1] * s),
            (a[1] * a[0] * k + a[2] * s, c + a[1] * a[1] * k, a[1] * a[2] * k - a[0] * s),
            (a[2] * a[0] * k - a[1] * s, a[2] * a[1] * k + a[0] * s, c + a[2] * a[2] * k))


class SpinDriver:
    """Per-step inputs of a spinning part (tunnel.rs:1067-1100): the angular
    speed ramps in with a smoothstep over `ramp_revs` revolutions, the angle
    accumulates f * |omega| per step, each step carries its rotation, the
    previous step's rotation and omega scaled by f."""

    def __init__(self, axis, omega: float, ramp_revs: float = 0.5):
        n = math.sqrt(sum(a * a for a in axis))
        self.axis = tuple(a / n for a in axis)
        self.omega = float(omega)
        self.ramp = int(math.ceil(ramp_revs * 2.0 * math.pi / abs(omega))) if ramp_revs > 0 else 0
        self.step = 0
        self.theta = 0.0
        self.rot_prev = rotation(self.axis, 0.0)

    def next(self, n: int):
        from . import StepIn
        out = []
        for _ in range(int(n)):
            s = self.step
            if s < self.ramp:
                x = (s + 0.5) / self.ramp
                f = x * x * (3.0 - 2.0 * x)
            else:
                f = 1.0
            self.theta += f * self.omega
            rot = rotation(self.axis, self.theta)
            out.append(StepIn(rot, self.rot_prev, tuple(f * self.omega * a for a in self.axis), False))
            self.rot_prev = rot
            self.step += 1
        return out


def spin_box(center, vox, axis_k: int, dims, pad: int = 3):
    """The swept box of a part spinning about lattice axis `axis_k` through
    `center` (tunnel.rs:709-721): radius + pad across, the axial extent +
    pad along, kept one cell inside the domain."""
    lo, hi = [0.0] * 3, [0.0] * 3
    for k in range(3):
        if k == axis_k:
            lo[k], hi[k] = vox.axial[0], vox.axial[1]
        else:
            lo[k], hi[k] = -vox.radius, vox.radius
    bl, bn = [0] * 3, [0] * 3
    for k in range(3):
        a = int(math.floor(center[k] + lo[k] - pad))
        b = int(math.ceil(center[k] + hi[k] + pad))
        bl[k] = max(a, 1)
        bn[k] = max(min(b, dims[k] - 2) - bl[k] + 1, 1)
    return tuple(bl), tuple(bn)


def spinning_prop(dims=(384, 384, 384), radius: float = 120.0, tip: float = 0.08, u_axial: float = 0.01,
                  cs: float = 0.16, nu: float = 0.001, occ_h: float = 0.125, storage: Storage = Storage.FP16,
                  seed: int = 20261203, center=None, debug_dump=False):
    """The spinning-prop tunnel: a seeded two-blade prop (chord 30 -> 12,
    twist 35 -> 10 degrees, thickness 8 % of the chord, all scaled with
    radius / 120) spinning about +z at omega = tip / radius rad/step, axial
    inflow u_axial through a velocity face at -z, open faces elsewhere;
    body-frame occupancy at spacing occ_h (voxel.rs:509-553)."""
    from . import voxelize_spin
    from .shapes import tapered_prop
    s = radius / 120.0
    mesh = tapered_prop(radius, 30.0 * s, 12.0 * s, 35.0, 10.0, 0.08, seed=seed)
    vox = voxelize_spin(mesh.positions, mesh.triangles, (0.0, 0.0, 1.0), occ_h)
    c = tuple(d / 2.0 for d in dims) if center is None else tuple(center)
    box_lo, box_n = spin_box(c, vox, 2, dims)
    setup = LatticeSetup(dims=dims, faces=(FaceBc.OPEN, FaceBc.OPEN, FaceBc.OPEN, FaceBc.OPEN, FaceBc.VELOCITY, FaceBc.OPEN),
                         nu=nu, smagorinsky=cs, u_inf=(0.0, 0.0, u_axial), u_init=(0.0, 0.0, u_axial), storage=storage,
                         center=c, debug_dump=debug_dump)
    omega = tip / radius
    meta = dict(nu=nu, tau0=0.5 + 3.0 * nu, omega=omega, radius=radius, tip=tip, u_axial=u_axial, seed=seed,
                ntri=int(len(vox.tris)), box_lo=box_lo, box_n=box_n, occ_h=occ_h, axis=(0.0, 0.0, 1.0))
    return setup, Geometry.spin(vox, box_lo, box_n), meta


def orbit_sphere(r: float, ro: float, u: float, re: float, fixed: bool = False, spin_path: bool = False,
                 storage: Storage = Storage.FP32):
    """crates/cfd/examples/orbit_sphere.rs: a sphere of radius r orbiting
    about +z at radius ro with speed u through still fluid (open faces, a
    10-cell absorbing layer all round), or, with `fixed`, the same sphere
    held in a uniform stream u (velocity inlet at -x) in a domain of the
    same size, voxelized statically or, with `spin_path`, through the
    spinning-part path at rest. Returns (setup, geometry, meta) with meta
    holding the Stokes drag 3 pi nu d u and, for the orbit, omega = u / ro."""
    from . import voxelize_spin
    nu = u * 2.0 * r / re
    nxy, nz = int(2.0 * (ro + r) + 40.0), int(2.0 * r + 40.0)
    c = (nxy / 2.0, nxy / 2.0, nz / 2.0)
    kw = dict(nu=nu, smagorinsky=0.0, sponge=(10,) * 6, sponge_tau=1.0, sponge_sigma=0.1, storage=storage)
    meta = dict(nu=nu, stokes=3.0 * math.pi * nu * 2.0 * r * u, omega=u / ro)
    if fixed:
        dims = (nxy, int(2.0 * r + 40.0) * 2, nz)
        cc = (dims[0] / 2.0, dims[1] / 2.0, nz / 2.0)
        if spin_path:
            m = sphere([0.0, 0.0, 0.0], r, 5)
            vox = voxelize_spin(m.positions, m.triangles, (0.0, 0.0, 1.0), 0.125)
            b = int(r) + 4
            geom = Geometry.spin(vox, tuple(int(v) - b for v in cc), (2 * b + 1,) * 3)
        else:
            geom = static_geometry(sphere(list(cc), r, 5), dims)
        setup = LatticeSetup(dims=dims, faces=(FaceBc.VELOCITY,) + (FaceBc.OPEN,) * 5, u_inf=(u, 0.0, 0.0),
                             u_init=(u, 0.0, 0.0), total_pressure_inlet=True, center=cc, **kw)
        return setup, geom, meta
    m = sphere([ro, 0.0, 0.0], r, 5)
    vox = voxelize_spin(m.positions, m.triangles, (0.0, 0.0, 1.0), 0.125)
    b = int(math.ceil(ro + r + 3.0))
    geom = Geometry.spin(vox, (int(c[0]) - b, int(c[1]) - b, int(c[2]) - int(r) - 3), (2 * b + 1, 2 * b + 1, int(2 * r) + 7))
    setup = LatticeSetup(dims=(nxy, nxy, nz), faces=(FaceBc.OPEN,) * 6, center=c, **kw)
    return setup, geom, meta


def naca0012(c: float, alpha_deg: float, re: float, smagorinsky: float = 0.1, storage: Storage = Storage.FP32):
    """crates/cfd/examples/airfoil.rs: a NACA 0012 section of chord c cells,
    extruded through a 4-cell periodic span, at angle of attack alpha (about
    the quarter chord, nose up for alpha > 0) in a 12c x 8c tunnel: velocity
    inlet 0.075 at -x, open faces in x and y with absorbing layers, the
    quarter chord at (4c, 4c). Lift is F_y, drag F_x; meta['q'] is the
    dynamic pressure times chord times span, per step."""
    import numpy as np
    from .shapes import TriMesh, fix_orientation
    u = 0.075
    nu = u * c / re
    nx, ny, nz = int(12.0 * c), int(8.0 * c), 4
    x0, y0 = 4.0 * c, 4.0 * c
    a = math.radians(alpha_deg)
    nsec = 200
    sec = []
    for i in range(nsec):
        th = 2.0 * math.pi * i / nsec
        xc = 0.5 * (1.0 + math.cos(th))
        t = 0.6 * (0.2969 * math.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2 + 0.2843 * xc ** 3 - 0.1036 * xc ** 4)
        yc = t if th < math.pi else -t
        px, py = (xc - 0.25) * c, yc * c
        sec.append([x0 + px * math.cos(a) + py * math.sin(a), y0 - px * math.sin(a) + py * math.cos(a)])
    verts = [[p[0], p[1], z] for z in (-3.0, nz + 3.0) for p in sec]
    n = nsec
    tris = []
    for i in range(n):
        j = (i + 1) % n
        tris += [[i, j, n + j], [i, n + j, n + i]]
    for i in range(1, n - 1):
        tris += [[0, i + 1, i], [n, n + i, n + i + 1]]
    m = TriMesh(verts, tris)
    fix_orientation(m)
    dims = (nx, ny, nz)
    sp = int(0.5 * c)
    setup = LatticeSetup(dims=dims, faces=(FaceBc.VELOCITY, FaceBc.OPEN, FaceBc.OPEN, FaceBc.OPEN, FaceBc.PERIODIC,
                                           FaceBc.PERIODIC),
                         nu=nu, smagorinsky=smagorinsky, u_inf=(u, 0.0, 0.0), u_init=(u, 0.0, 0.0),
                         sponge=(sp, 2 * sp, sp, sp, 0, 0), sponge_tau=1.0, sponge_sigma=0.1,
                         total_pressure_inlet=True, storage=storage, center=(x0, y0, 0.0))
    return setup, static_geometry(m, dims), dict(u=u, nu=nu, q=0.5 * u * u * c * nz, steps_per_pass=c / u)
