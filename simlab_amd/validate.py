"""Validation runs from the command line (crates/cfd/examples/validate.rs,
same cases and options):

  python -m simlab_amd.validate sphere <Re> <cells per D> <flow passes>
  python -m simlab_amd.validate rot <Re> <cells per D> <revolutions>
  python -m simlab_amd.validate prop <rpm> <cells across> <revolutions>

Options after the numbers: fp16, fp32, p0, absu, velsides, side=, up=, down=,
smag=, u=, tol=, tau=, radial=, axial=, voff=, sigma=, spin=, spout=,
spside=. QUIET=1 hides the progress lines, HIST=1 prints the sphere's
coefficient history.
"""
from __future__ import annotations

import math
import os
import sys

from . import Storage
from . import tunnel as T
from .shapes import PropSpec


def options(args) -> "T.SolverOptions":
    """validate.rs:27-71: words and key=value pairs anywhere in the arguments."""
    o = T.SolverOptions()
    wm, sm = o.wind_margins, o.still_margins
    widths = list(o.sponge_widths)
    for s in args:
        if s == "fp16":
            o.storage = Storage.FP16
        elif s == "fp32":
            o.storage = Storage.FP32
        elif s == "p0":
            o.total_pressure_inlet = True
        elif s == "absu":
            o.absorb_velocity = True
        elif s == "velsides":
            o.side_bc = "velocity"
        if "=" in s:
            k, v = s.split("=", 1)
            v = float(v)
            if k == "side":
                wm = T.Margins(wm.upstream, wm.downstream, v)
            elif k == "up":
                wm = T.Margins(v, wm.downstream, wm.side)
            elif k == "down":
                wm = T.Margins(wm.upstream, v, wm.side)
            elif k == "smag":
                o.smagorinsky = v
            elif k == "u":
                o.max_lattice_speed = v
                o.max_tip_speed = v
            elif k == "tol":
                o.tolerance = v
            elif k == "tau":
                o.max_tau = v
            elif k == "radial":
                sm = T.Margins(sm.upstream, sm.downstream, v)
            elif k == "axial":
                sm = T.Margins(v, v, sm.side)
            elif k == "voff":
                o.vertex_offset = v
            elif k == "sigma":
                o.sponge_sigma = v
            elif k == "spin":
                widths[0] = v
            elif k == "spout":
                widths[1] = v
            elif k == "spside":
                widths[2] = v
    o.wind_margins, o.still_margins, o.sponge_widths = wm, sm, tuple(widths)
    return o


def _progress(f: float, msg: str) -> None:
    if not os.environ.get("QUIET"):
        sys.stderr.write(f"[{f * 100.0:5.1f}%] {msg}\n")
        sys.stderr.flush()


def main(argv=None) -> int:
    a = sys.argv[1:] if argv is None else list(argv)
    if not a:
        raise SystemExit(__doc__)
    o = options(a)
    case = a[0]
    if case == "sphere":
        re, d, dur = float(a[1]), int(float(a[2])), float(a[3])
        cd, r = T.sphere_drag(re, d, dur, o, _progress)
        s = r.summary
        refc = T.sphere_cd_cgw(re)
        print(f"SPHERE Re={re:g} D={d} cells: Cd={cd:.4f} ref(CGW)={refc:.4f} SN={T.sphere_cd_sn(re):.4f} "
              f"err={(cd / refc - 1.0) * 100.0:+.2f}% lift={s.lift_n / s.drag_n:.2e} conv={str(s.converged).lower()} "
              f"passes={s.simulated:.1f} avg={s.averaged_over:.1f} grid={list(s.grid)} blockage={s.blockage * 100.0:.2f}% "
              f"tau={s.tau:.4f} MLUPs={s.mlups:.0f} {s.storage.name} t={s.seconds:.0f}s")
        for w in s.warnings:
            print(f"  warning: {w}")
        if os.environ.get("HIST"):
            q = 0.5 * 1.2 * T.SPHERE_U ** 2 * math.pi * T.SPHERE_D ** 2 / 4.0
            for t, fv, _ in r.history[::max(len(r.history) // 40, 1)]:
                print(f"  t={t:.2f} Cd={fv[0] / q:.4f} Cl={fv[2] / q:.4f} Cs={fv[1] / q:.4f}")
    elif case == "rot":
        re, d, revs = float(a[1]), int(float(a[2])), float(a[3])
        t, stokes, r = T.rotating_sphere(re, d, revs, o, _progress)
        s = r.summary
        print(f"ROTSPHERE Re={re:g} D={d}: torque={t:.4e} stokes={stokes:.4e} ratio={-t / stokes:.4f} "
              f"err={(-t / stokes - 1.0) * 100.0:+.2f}% thrust={(s.thrust_n or 0.0):.2e} conv={str(s.converged).lower()} "
              f"grid={list(s.grid)} tau={s.tau:.3f} MLUPs={s.mlups:.0f}")
    elif case == "prop":
        rpm, cells, revs = float(a[1]), int(float(a[2])), float(a[3])
        r = T.prop_hover(PropSpec.five_inch(), rpm, cells, revs, o, _progress)
        s = r.summary
        print(f"PROP rpm={rpm:g} cells={cells}: thrust={s.thrust_n:.4f} N torque={s.torque_nm:.4e} N·m "
              f"power={s.power_w:.2f} W conv={str(s.converged).lower()} revs={s.simulated:.1f} avg={s.averaged_over:.1f} "
              f"grid={list(s.grid)} dx={s.dx_mm:.3f}mm tau={s.tau:.5f} Ma={s.mach_lattice:.3f} MLUPs={s.mlups:.0f} "
              f"mem={s.memory_bytes / 1e9:.2f}GB t={s.seconds:.0f}s")
        for w in s.warnings:
            print(f"  warning: {w}")
        for t, fv, mv in r.history[::max(len(r.history) // 20, 1)]:
            print(f"  t={t:.2f} Fz={fv[2]:.4f} Mz={mv[2]:.4e}")
    else:
        raise SystemExit("unknown case")
    return 0


if __name__ == "__main__":
    sys.exit(main())
