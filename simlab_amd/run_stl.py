"""Runs the wind tunnel on an STL from the command line and prints the
summary as JSON (crates/cfd/examples/run_stl.rs, same arguments):

  python -m simlab_amd.run_stl part.stl config.json [options]

`config.json` is a WindConfig (a file path or the JSON itself), e.g.
  {"wind_dir":[1,0,0],"speed_m_s":15,"quality":{"type":"normal"}}
Options:
  --unit-mm <x>     millimetres per STL unit (default 1)
  --fp16 | --fp32   population storage (default: FP32 unless --budget-gb forces FP16)
  --budget-gb <x>   memory budget for the estimate and storage choice
  --estimate        only print the estimate
  --out <dir>       write surface pressure, slices and history there
  --history         print the force history
  --axial <d> --radial <d>   still-air margins in rotor diameters (default 2.5, 1)
  --tip <u>         lattice speed of the blade tip (default 0.1)
  --smag <c>        Smagorinsky constant (default 0.1)
  --wind-margins <up,down,side>  wind domain margins in frontal sizes (default 2.5,5,3)
  --no-refine       accepted; every grid here is one uniform D3Q19 grid
  --velocity-inlet  fix the inlet velocity instead of its total pressure
  --cells <n>       cells across the part's largest size (overrides the quality's)
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

from . import Storage
from . import tunnel as T
from .stl import read_stl


def _progress(f: float, msg: str) -> None:
    sys.stderr.write(f"[{f * 100.0:5.1f}%] {msg}\n")
    sys.stderr.flush()


def parse(argv):
    """The options of run_stl.rs:44-100: (stl path, config, unit_mm,
    SolverOptions, budget bytes or None, estimate only, out dir, history)."""
    if len(argv) < 2:
        raise SystemExit("usage: run_stl <part.stl> <config.json | json> [--unit-mm x] [--fp16] [--budget-gb x] "
                         "[--estimate] [--out dir]")
    stl, ctext = argv[0], argv[1]
    if os.path.exists(ctext):
        with open(ctext) as fh:
            ctext = fh.read()
    cfg = T.config_from_json(json.loads(ctext))
    opt = T.SolverOptions()
    unit_mm, budget, only_est, out, history = 1.0, None, False, None, False
    i = 2
    while i < len(argv):
        a = argv[i]

        def val():
            if i + 1 >= len(argv):
                raise SystemExit(f"option {a} needs a value")
            return argv[i + 1]

        if a == "--unit-mm":
            unit_mm = float(val()); i += 1
        elif a == "--fp16":
            opt.storage = Storage.FP16
        elif a == "--fp32":
            opt.storage = Storage.FP32
        elif a == "--budget-gb":
            budget = int(float(val()) * 1e9); i += 1
        elif a == "--estimate":
            only_est = True
        elif a == "--axial":
            v = float(val()); i += 1
            opt.still_margins = T.Margins(v, v, opt.still_margins.side)
        elif a == "--radial":
            opt.still_margins = T.Margins(opt.still_margins.upstream, opt.still_margins.downstream, float(val())); i += 1
        elif a == "--tip":
            opt.max_tip_speed = float(val()); i += 1
        elif a == "--wind-margins":
            v = [float(x) for x in val().split(",")]; i += 1
            if len(v) != 3:
                raise SystemExit("--wind-margins takes up,down,side")
            opt.wind_margins = T.Margins(v[0], v[1], v[2])
        elif a == "--velocity-inlet":
            opt.total_pressure_inlet = False
        elif a == "--no-refine":
            pass
        elif a == "--cells":
            opt.cells_across = int(val()); i += 1
        elif a == "--smag":
            opt.smagorinsky = float(val()); i += 1
        elif a == "--history":
            history = True
        elif a == "--out":
            out = val(); i += 1
        else:
            raise SystemExit(f"unknown option {a}")
        i += 1
    opt.memory_budget = budget
    return stl, cfg, unit_mm, opt, budget, only_est, out, history


def write_outputs(directory: str, r: "T.WindResult") -> None:
    """run_stl.rs:119-140: pressure_pa.f32, cp.f32, slice.<name>.json (with
    its field data) and slice.<name>.<field>.f32, history.csv."""
    os.makedirs(directory, exist_ok=True)
    np.asarray(r.surface_pressure_pa, "<f4").tofile(os.path.join(directory, "pressure_pa.f32"))
    np.asarray(r.surface_cp, "<f4").tofile(os.path.join(directory, "cp.f32"))
    for sl in r.slices:
        with open(os.path.join(directory, f"slice.{sl.name}.json"), "w") as fh:
            json.dump(T.slice_to_json(sl), fh, indent=2)
        for f in sl.fields:
            np.asarray(f.data, "<f4").tofile(os.path.join(directory, f"slice.{sl.name}.{f.name}.f32"))
    with open(os.path.join(directory, "history.csv"), "w") as fh:
        fh.write("t,fx,fy,fz,mx,my,mz\n")
        for t, fv, mv in r.history:
            fh.write(",".join(repr(float(x)) for x in (t, *fv, *mv)) + "\n")


def main(argv=None) -> int:
    stl, cfg, unit_mm, opt, budget, only_est, out, history = parse(sys.argv[1:] if argv is None else argv)
    with open(stl, "rb") as fh:
        model = read_stl(fh.read())
    unit_m = unit_mm * 1e-3
    surface = model.scaled(unit_m)
    est = T.estimate(cfg, surface, unit_m, budget if budget is not None else (1 << 62), opt)
    sys.stderr.write(f"estimate: {json.dumps(T.estimate_to_json(est))}\n")
    if only_est:
        print(json.dumps(T.estimate_to_json(est), indent=2))
        return 0
    r = T.run_with(cfg, surface, unit_m, opt, _progress)
    print(json.dumps(T.summary_to_json(r.summary), indent=2))
    if history:
        for t, fv, mv in r.history:
            print(f"t={t:.3f} F=[{fv[0]:.5f}, {fv[1]:.5f}, {fv[2]:.5f}] M=[{mv[0]:.6f}, {mv[1]:.6f}, {mv[2]:.6f}]")
    if out:
        write_outputs(out, r)
    return 0


if __name__ == "__main__":
    sys.exit(main())
