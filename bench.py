#!/usr/bin/env python3
"""Throughput and accuracy of the MI355X D3Q19 lattice on the 256x128x128
sphere-in-a-channel workload (sphere r=16 at x=64, Re 1000 on its diameter,
Smagorinsky Cs=0.1, -x velocity inlet u=0.05, +x open outlet, halfway
bounce-back channel walls), the configuration the headline metric is quoted
on.

One step = one full lattice update (face kernel + stream-collide with the
wall records and the FP64 force reduction). MLUPS counts all nx*ny*nz cells
per step, as the reference does (crates/cfd/examples/bench.rs:25).

Reported (one JSON line):
  value            FP16-storage MLUPS (wall clock over the timed steps)
  roofline         the stream-collide kernel (k_step, one launch per step):
                   algorithmic bytes per launch (2*19*s + 4 bytes per fluid
                   cell it updates, s = 2 for FP16) / its average duration
                   from HIP events its own dispatch records
                   (hipExtLaunchKernel, no packets between kernels);
                   storages.*.roofline_frac is the whole-step figure
                   (MLUPS over all cells x bytes per cell / 8 TB/s)
  fp32 / bf16      the same for FP32 and BF16 storage
  drag             time-averaged drag coefficient of a 5000-step run per
                   storage and its error against the FP64 run
  cpu_baseline     the CPU oracle (C++ restatement of the reference kernels)
                   on a bounded sample of the same workload
  n384             the 384^3 cases (sphere channel, spinning prop), FP16 and
                   FP32, with their whole-step roofline fractions
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS and % HBM roofline (FP16-store vs FP32) at stated drag error vs FP64"
HBM_PEAK_GBS = 8000.0
WORKLOAD = "sphere-channel 256x128x128, sphere r=16 at x=64, Re=1000, Cs=0.1, u=0.05, halfway-bounce-back walls"


def bytes_per_cell(storage):
    return 2 * 19 * storage.pop_bytes() + 4


def settle(make, settle_ms):
    """Setup, untimed: a second lattice of the same workload, built by
    `make`, steps for about `settle_ms` of wall time just before the timed
    lattice's warm-up, so that a short warm-up starts with the GPU at its
    working clocks (the same kernels and tiles; the timed lattice itself
    starts its warm-up as built). The GPU idles meanwhile only for the
    Python between the two."""
    if settle_ms <= 0:
        return
    lat = make()
    t0 = time.perf_counter()
    while (time.perf_counter() - t0) * 1e3 < settle_ms:
        lat.run_still(20)
    del lat


def timed_run(sa, W, storage, steps, warmup, settle_ms=0):
    setup, geom, meta = W.sphere_channel_large(storage)
    lat = sa.Lattice(setup, geom)
    settle(lambda: sa.Lattice(setup, geom), settle_ms)
    lat.run_still(warmup)
    t0 = time.perf_counter()
    lat.run_still(steps)                      # returns after the stream is synchronized
    wall = time.perf_counter() - t0
    # second pass: the stream-collide kernel's own start and end, recorded by
    # its dispatch (hipExtLaunchKernel), and events around each chunk
    lat.set_timing(True)
    lat.run_still(steps)
    step_ms, main_ms, launches = lat.timing()
    n = setup.cells()
    res = {
        "mlups": n * steps / wall / 1e6,
        "ms_per_step": wall / steps * 1e3,
        "device_ms_per_step": step_ms / steps,
        "kernel_us": main_ms / launches * 1e3,
        "nrec": lat.nrec,
    }
    res["bulk_cells"] = lat.bulk_cells
    # cells one launch updates: every fluid cell (bulk and wall-record cells)
    res["kernel_cells"] = lat.bulk_cells + lat.nrec
    res["tile"] = list(lat.tile())
    b = bytes_per_cell(storage) * res["kernel_cells"]   # algorithmic bytes of one launch
    res["kernel_bytes"] = b
    res["kernel_gbs"] = b / (main_ms / launches * 1e-3) / 1e9
    res["kernel_roofline_frac"] = res["kernel_gbs"] / HBM_PEAK_GBS
    res["roofline_frac"] = res["mlups"] * 1e6 * bytes_per_cell(storage) / (HBM_PEAK_GBS * 1e9)
    del lat
    return res, n


def large_runs(sa, W, steps, warmup, cases=("sphere", "prop"), storages=("fp16", "fp32"), settle_ms=0):
    """The 384^3 cases: the sphere channel scaled to 384^3 (sphere r=24 at
    x=96, Re 1000, Cs 0.1) and the spinning two-blade prop (radius 120,
    tip speed 0.08, axial inflow 0.01, Cs 0.16), FP16 and FP32 storage;
    MLUPS over all cells from the wall clock, the bulk kernel's algorithmic
    GB/s from its HIP events."""
    out = {}
    sts = [(n, {"fp16": sa.Storage.FP16, "fp32": sa.Storage.FP32}[n]) for n in storages]
    for name, st in (sts if "sphere" in cases else []):
        print(f"# 384^3 sphere channel {name}", file=sys.stderr, flush=True)
        setup, geom, meta = W.sphere_channel((384, 384, 384), 24.0, 96.0, 1000.0, 0.05, 0.1, st)
        lat = sa.Lattice(setup, geom)
        settle(lambda: sa.Lattice(setup, geom), settle_ms)
        lat.run_still(warmup)
        t0 = time.perf_counter()
        lat.run_still(steps)
        wall = time.perf_counter() - t0
        lat.set_timing(True)
        lat.run_still(steps)
        step_ms, main_ms, launches = lat.timing()
        gbs = bytes_per_cell(st) * (lat.bulk_cells + lat.nrec) / (main_ms / launches * 1e-3) / 1e9
        mlups = setup.cells() * steps / wall / 1e6
        out[f"sphere384_{name}"] = {
            "mlups": round(mlups, 1), "ms_per_step": round(wall / steps * 1e3, 4),
            "roofline_frac": round(mlups * 1e6 * bytes_per_cell(st) / (HBM_PEAK_GBS * 1e9), 4),
            "kernel_gbs": round(gbs, 1), "kernel_roofline_frac": round(gbs / HBM_PEAK_GBS, 4),
            "tile": list(lat.tile())}
        del lat
    for name, st in (sts if "prop" in cases else []):
        print(f"# 384^3 spinning prop {name}", file=sys.stderr, flush=True)
        setup, geom, meta = W.spinning_prop(storage=st)
        lat = sa.Lattice(setup, geom)
        settle(lambda: sa.Lattice(setup, geom), settle_ms)
        drv = W.SpinDriver(meta["axis"], meta["omega"])
        lat.run(drv.next(warmup))
        batch = drv.next(steps)
        t0 = time.perf_counter()
        h = lat.run(batch)
        wall = time.perf_counter() - t0
        mlups = setup.cells() * steps / wall / 1e6
        out[f"prop384_{name}"] = {
            "mlups": round(mlups, 1), "ms_per_step": round(wall / steps * 1e3, 4),
            "roofline_frac": round(mlups * 1e6 * bytes_per_cell(st) / (HBM_PEAK_GBS * 1e9), 4),
            "box_cells": int(lat.nrec), "triangles": meta["ntri"],
            "thrust_z_last": float(h[-1, 2]), "torque_z_last": float(h[-1, 5])}
        del lat
    if "sphere384_fp16" in out and "sphere384_fp32" in out:
        out["sphere384_fp16_vs_fp32"] = round(out["sphere384_fp16"]["mlups"] / out["sphere384_fp32"]["mlups"], 3)
    return out


def study_prop(sa, W, steps):
    """The spinning prop at 384^3 (tip speed 0.08, axial inflow 0.01, Cs
    0.16, smoothstep spin-up over half a revolution): thrust (F_z) and
    torque (M_z) averaged over the second half of the run, FP16 and FP32
    storage against FP64 storage (the oracle's arithmetic on the GPU)."""
    out = {}
    for name, st in (("fp64", sa.Storage.FP64), ("fp32", sa.Storage.FP32), ("fp16", sa.Storage.FP16)):
        setup, geom, meta = W.spinning_prop(storage=st)
        lat = sa.Lattice(setup, geom)
        drv = W.SpinDriver(meta["axis"], meta["omega"])
        hs, t0 = [], time.perf_counter()
        for done in range(0, steps, 2000):
            hs.append(lat.run(drv.next(min(2000, steps - done))))
            print(f"# prop {name}: {done + len(hs[-1])} / {steps} steps, {time.perf_counter() - t0:.0f} s",
                  file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        import numpy as np
        h = np.concatenate(hs)
        w = h[steps // 2:]
        out[name] = {"thrust": float(w[:, 2].mean()), "torque": float(w[:, 5].mean()),
                     "side_force": float(np.hypot(w[:, 0].mean(), w[:, 1].mean())),
                     "mlups": round(setup.cells() * steps / wall / 1e6, 1), "revolutions": drv.theta / (2 * 3.141592653589793)}
        del lat
    for name in ("fp32", "fp16"):
        out[name]["thrust_err_vs_fp64"] = out[name]["thrust"] / out["fp64"]["thrust"] - 1.0
        out[name]["torque_err_vs_fp64"] = out[name]["torque"] / out["fp64"]["torque"] - 1.0
    out["window"] = f"mean over steps {steps // 2}..{steps}"
    return out


def study_arm(sa, W, steps, batch_steps):
    """The drone-arm tunnel 512x256x256 (u 0.06, Cs 0.16): drag per storage
    against FP64; then the arm at 128^3, 192^3, 256^3 and 384^3 in FP16,
    autotuned, stepped one after the other and as one batch."""
    import numpy as np
    out, drag = {}, {}
    runs = (("fp64", sa.Storage.FP64, 2000), ("fp32", sa.Storage.FP32, 2000), ("fp16", sa.Storage.FP16, 2000),
            ("fp16_calls_2500", sa.Storage.FP16, 2500))
    for name, st, call in runs:
        setup, geom, meta = W.drone_arm(storage=st)
        lat = sa.Lattice(setup, geom)
        t0 = time.perf_counter()
        hs = []
        for done in range(0, steps, call):
            hs.append(lat.run_still(min(call, steps - done)))
            print(f"# arm {name}: {done + len(hs[-1])} / {steps} steps", file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        drag[name] = np.concatenate(hs)[:, 0]
        out[name] = {"drag": float(drag[name][steps // 2:].mean()), "mlups": round(setup.cells() * steps / wall / 1e6, 1),
                     "nrec": lat.nrec, "halfway_cells": lat.halfway_cells, "steps_per_call": call}
        del lat
    d64 = drag["fp64"][steps // 2:]
    for name in ("fp32", "fp16", "fp16_calls_2500"):
        out[name]["drag_err_vs_fp64"] = out[name]["drag"] / out["fp64"]["drag"] - 1.0
        # the step-by-step drag against FP64's: median and 99th percentile of
        # the relative difference over the window
        rel = np.abs(drag[name][steps // 2:] / d64 - 1.0)
        out[name]["drag_step_err_p50"] = float(np.median(rel))
        out[name]["drag_step_err_p99"] = float(np.percentile(rel, 99))
    # how unsteady the drag is: FP64's own mean over 80 % of the window
    # against the whole window's, and its standard deviation
    out["fp64"]["window_spread"] = float(drag["fp64"][steps // 2:steps // 2 + 4 * steps // 10].mean() / d64.mean() - 1.0)
    out["fp64"]["std_over_mean"] = float(d64.std() / d64.mean())
    for label, sizes, k in (("batch", (128, 192, 256, 384), batch_steps),
                            ("batch_small", (64, 80, 96, 112, 128, 64, 96, 128), 4 * batch_steps)):
        lats, cells, tiles = [], 0, {}
        for n in sizes:
            setup, geom, meta = W.drone_arm(dims=(n, n, n), storage=sa.Storage.FP16, scale=n / 256.0)
            lat = sa.Lattice(setup, geom)
            t, us = lat.autotune(steps=20, use_cache=True)
            tiles[str(n)] = {"tile": list(t), "step_us": round(us, 1)}
            lats.append(lat)
            cells += setup.cells()
        for lat in lats:
            lat.run_still(10)
        t0 = time.perf_counter()
        for lat in lats:
            lat.run_still(k)
        seq = time.perf_counter() - t0
        b = sa.Batch(lats)
        b.run_still(10)
        t0 = time.perf_counter()
        b.run_still(k)
        bat = time.perf_counter() - t0
        out[label] = {"sizes": list(sizes), "tiles": tiles, "steps": k,
                      "sequential_mlups": round(cells * k / seq / 1e6, 1),
                      "batched_mlups": round(cells * k / bat / 1e6, 1)}
        del b, lats
    out["window"] = f"mean over steps {steps // 2}..{steps}"
    return out


def study_batch(sa, W, steps, sizes=(128, 192, 256, 384)):
    """The drone arm at several sizes in FP16, autotuned: MLUPS over all the
    members' cells stepped one lattice after the other, as one batch with one
    kernel per step (the default), and as a batch with a special and a bulk
    kernel on two streams (members over 4 Mi cells stepping alone)."""
    lats, cells, tiles = [], 0, {}
    for n in sizes:
        setup, geom, meta = W.drone_arm(dims=(n, n, n), storage=sa.Storage.FP16, scale=n / 256.0)
        lat = sa.Lattice(setup, geom)
        t, us = lat.autotune(steps=20, use_cache=True)
        tiles[str(n)] = {"tile": list(t), "step_us": round(us, 1)}
        lats.append(lat)
        cells += setup.cells()
    out = {"sizes": list(sizes), "tiles": tiles, "steps": steps}
    for lat in lats:
        lat.run_still(20)
    reps = []
    for _ in range(3):
        t0 = time.perf_counter()
        for lat in lats:
            lat.run_still(steps)
        reps.append(cells * steps / (time.perf_counter() - t0) / 1e6)
    out["sequential_mlups"] = round(max(reps), 1)
    for mode in ("fused", "split"):
        os.environ["SIMLAB_CFD_BATCH"] = mode
        b = sa.Batch(lats)
        b.run_still(20)
        reps = []
        for _ in range(3):
            t0 = time.perf_counter()
            b.run_still(steps)
            reps.append(cells * steps / (time.perf_counter() - t0) / 1e6)
        out[f"batched_{mode}_mlups"] = round(max(reps), 1)
        del b
    os.environ.pop("SIMLAB_CFD_BATCH", None)
    out["best_of"] = 3
    return out


def study_refined(sa, steps):
    """The refined D3Q27 solver on the study driver's plan of a sphere at
    Re 100 with 48 and 96 cells across (two and three levels, sponges,
    total-pressure inlet): cell updates per second over all levels (each
    level's cells times its steps), FP32 and FP16."""
    from simlab_amd import tunnel as T
    from simlab_amd import mlattice as ml
    import numpy as np
    out = {}
    for across in (48, 96):
        cfg, m = T.sphere_case(100.0, across)
        p = T.plan(cfg, m, 1e-3, T.SolverOptions())
        mp = p.multi
        P = np.asarray(m.positions, np.float64)
        to_lat = np.asarray(p.center) + ((P - p.pref) @ p.r.T) / p.dx - np.asarray(mp.org_f)
        b = mp.boxes[-1]
        fd = tuple(2 * (int(b.n[k]) - 1) + 1 for k in range(3))
        vox = ml.voxelize_static27(to_lat, m.triangles, fd)
        nl = len(mp.cells)
        for name, st in (("fp32", sa.Storage.FP32), ("fp16", sa.Storage.FP16)):
            s = ml.MultiSetup(dims0=mp.dims0, boxes=list(mp.boxes), faces=p.faces, nu0=p.nu_lat / float(1 << (nl - 1)),
                              u_inf=(p.u_lat, 0, 0), u_init=(p.u_lat, 0, 0), sponge=p.sponge, sponge_tau=1.0,
                              sponge_sigma=0.1, total_pressure_inlet=True, storage=st,
                              center=tuple(p.center[k] - mp.org_f[k] for k in range(3)))
            lat = ml.MultiLattice(s, sa.Geometry.static(vox))
            ss = lat.sub_steps()
            lat.run_still(20 * ss)
            n = max(1, steps // ss) * ss
            t0 = time.perf_counter()
            lat.run_still(n)
            wall = time.perf_counter() - t0
            upd = sum(c * float(1 << l) for l, c in enumerate(mp.cells)) / float(ss) * n
            out[f"sphere{across}_{name}"] = {"levels": nl, "level_cells": mp.cells, "nrec": lat.nrec,
                                             "fine_steps": n, "mlups": round(upd / wall / 1e6, 1),
                                             "ms_per_coarse_step": round(wall / (n / ss) * 1e3, 4)}
            print(f"# refined sphere {across} {name}: {out[f'sphere{across}_{name}']['mlups']} MLUPS",
                  file=sys.stderr, flush=True)
            lat.close()
    return out


def drag_run(sa, W, storage, steps):
    setup, geom, meta = W.sphere_channel_large(storage)
    lat = sa.Lattice(setup, geom)
    h = lat.run_still(steps)
    m1 = lat.mass()
    window = h[steps // 2:]
    cd = float(window[:, 0].mean() * meta["cd_scale"])
    cl = float(window[:, 2].mean() * meta["cd_scale"])
    del lat
    return {"cd": cd, "cl": cl, "cd_last": float(h[-1, 0] * meta["cd_scale"]), "mass_end": m1}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sa, W, target_s=5.0, repeats=3):
    """The CPU oracle on the bench workload: median MLUPS of `repeats` timed
    samples of about target_s seconds each, on the host threads this process
    may use (OMP_NUM_THREADS where it is set, as on a machine that gives
    each GPU its share of the host's cores; else every core)."""
    import statistics
    from oracle import pyoracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    setup, geom, meta = W.sphere_channel_large(sa.Storage.FP32)
    lat = O.OracleLattice(setup, geom, store=O.F32, threads=threads)
    lat.run_still(1)
    t0 = time.perf_counter()
    lat.run_still(2)
    per = (time.perf_counter() - t0) / 2
    k = max(2, min(200, int(target_s / max(per, 1e-6))))
    n = setup.cells()
    vals = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        lat.run_still(k)
        vals.append(n * k / (time.perf_counter() - t0) / 1e6)
    return {"value": round(statistics.median(vals), 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "host_cpus": os.cpu_count(), "samples": [round(v, 2) for v in vals],
            "sample": f"C++ oracle (restatement of the reference kernels), FP32 storage, {WORKLOAD}: median of "
                      f"{repeats} runs of {k} steps after 3 warm-up steps, {threads} threads"}


TRAFFIC_PROFILE = os.path.join("profiles", "stream_collide_sphere256_fp16_traffic.json")


def traffic_from_profile(storage_name):
    """HBM bytes per stream-collide launch from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload, for the storage it was
    taken with (None if absent or for another storage)."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE)) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("storage", "fp16") != storage_name:
        return None
    return d.get("bytes_per_launch")


def pattern_probe(reps):
    """Memory ceilings of the streaming access patterns at 384^3 FP16 (no
    arithmetic): AA even and odd steps in place, a two-lattice pull step and
    an aligned out-of-place step (simlab_cfd_probe_stream_patterns), beside
    a plain device copy (copy_calibration's kernel)."""
    import ctypes as C
    import simlab_amd as sa
    lib = sa.load_library()
    lib.simlab_cfd_probe_stream_patterns.restype = C.c_int
    lib.simlab_cfd_probe_stream_patterns.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
    g = (C.c_double * 4)()
    rc = lib.simlab_cfd_probe_stream_patterns(384, 384, 384, max(int(reps), 1), g)
    if rc != 0:
        raise RuntimeError(lib.simlab_cfd_last_error().decode())
    out = {"grid": [384, 384, 384], "bytes_per_cell": 76, "unit": "GB/s",
           "aa_even_in_place": round(g[0], 1), "aa_odd_in_place": round(g[1], 1),
           "aa_mean": round(2.0 / (1.0 / g[0] + 1.0 / g[1]), 1),
           "two_lattice_pull": round(g[2], 1), "aligned_out_of_place": round(g[3], 1)}
    return out


def copy_calibration(reps):
    """A kernel of known traffic: out = x * 1 over 2^28 floats reads and writes
    exactly 1 GiB each per launch."""
    import torch
    n = 1 << 28
    x = torch.ones(n, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    reps = max(1, min(reps, 20))
    t0 = time.perf_counter()
    for _ in range(reps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_calibration": {"read_bytes": 4 * n, "write_bytes": 4 * n, "launches": reps + 3,
                                           "ms_per_launch": dt * 1e3, "gbs": 8 * n / dt / 1e9}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    # 500 untimed steps (about 40 ms) bring the GPU to its sustained clock;
    # after 20 the timed steps read about 5 % lower at the same kernel time
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--drag-steps", type=int, default=5000)
    ap.add_argument("--settle-ms", type=float, default=100.0,
                    help="untimed stepping of a second lattice of the same workload for about this many ms just "
                         "before the warm-up (brings the GPU to its working clocks; 0: none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-drag", action="store_true")
    ap.add_argument("--no-384", action="store_true", help="skip the 384^3 sphere-channel and spinning-prop timings")
    ap.add_argument("--study", choices=["prop", "arm", "batch", "refined"], default=None,
                    help="accuracy studies instead of the benchmark line: the 384^3 spinning prop's thrust and "
                         "torque per storage (prop), or the 512x256x256 drone arm's drag per storage plus the "
                         "batched arm at four sizes (arm), the batch timings alone (batch), or the refined D3Q27 "
                         "solver's MLUPS on the study plans of a sphere at 48 and 96 cells across (refined); "
                         "one JSON object")
    ap.add_argument("--study-steps", type=int, default=0, help="steps of a study run (prop 20000, arm 10000)")
    ap.add_argument("--only-384", action="store_true",
                    help="only the 384^3 timings (one JSON object of them; for profiling those kernels alone)")
    ap.add_argument("--only", default="", help="comma list of storages to time (fp16,fp32,bf16); with --only-384 "
                                                  "fp16,fp32")
    ap.add_argument("--case", default="sphere,prop", help="with --only-384: comma list of the 384^3 cases to time "
                                                         "(sphere, prop)")
    ap.add_argument("--pattern-probe", action="store_true",
                    help="only measure the memory ceilings of the streaming access patterns at 384^3 FP16 (AA even "
                         "and odd steps in place, two-lattice pull, aligned out of place), with no arithmetic")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (the reference kernel that scales the "
                         "FETCH_SIZE/WRITE_SIZE counters of a profile into bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the lattice is single-device (the multi-GPU slab split is a design "
              f"sketch, not built); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.steps)
        return
    if args.pattern_probe:
        print(json.dumps({"pattern_probe": pattern_probe(min(args.steps, 50))}), flush=True)
        return

    import simlab_amd as sa
    from simlab_amd import workloads as W

    info = sa.gpu_info()
    print(f"# {info}", file=sys.stderr, flush=True)
    if args.study == "prop":
        print(json.dumps({"study_prop384": study_prop(sa, W, args.study_steps or 20000)}), flush=True)
        return
    if args.study == "refined":
        print(json.dumps({"study_refined": study_refined(sa, args.study_steps or 400)}), flush=True)
        return
    if args.study == "batch":
        print(json.dumps({"study_batch": study_batch(sa, W, args.study_steps or 300)}), flush=True)
        return
    if args.study == "arm":
        print(json.dumps({"study_arm": study_arm(sa, W, args.study_steps or 10000, 500)}), flush=True)
        return
    if args.only_384:
        cases = tuple(c for c in args.case.split(",") if c)
        sts = tuple(c for c in (args.only.split(",") if args.only else ["fp16", "fp32"]) if c)
        print(json.dumps({"n384": large_runs(sa, W, args.steps, args.warmup, cases, sts, args.settle_ms)}), flush=True)
        return
    storages = [s for s in (args.only.split(",") if args.only else ["fp16", "fp32", "bf16"]) if s]
    results = {}
    n = None
    for name in storages:
        st = {"fp16": sa.Storage.FP16, "fp32": sa.Storage.FP32, "bf16": sa.Storage.BF16}[name]
        print(f"# timing {name}: {args.warmup} warm-up + {args.steps} steps", file=sys.stderr, flush=True)
        results[name], n = timed_run(sa, W, st, args.steps, args.warmup, args.settle_ms)
        print(f"#   {name}: {results[name]['mlups']:.0f} MLUPS, kernel {results[name]['kernel_us']:.1f} us, "
              f"{results[name]['kernel_gbs']:.0f} GB/s", file=sys.stderr, flush=True)

    drag = None
    if not args.no_drag:
        drag = {}
        for name, st in (("fp64", sa.Storage.FP64), ("fp32", sa.Storage.FP32), ("fp16", sa.Storage.FP16), ("bf16", sa.Storage.BF16)):
            print(f"# drag run {name}: {args.drag_steps} steps", file=sys.stderr, flush=True)
            drag[name] = drag_run(sa, W, st, args.drag_steps)
        for name in ("fp32", "fp16", "bf16"):
            drag[name]["cd_err_vs_fp64"] = drag[name]["cd"] / drag["fp64"]["cd"] - 1.0
        drag["window"] = f"mean over steps {args.drag_steps // 2}..{args.drag_steps}"

    large = None
    if not args.no_384:
        large = large_runs(sa, W, max(20, min(args.steps, 200)), 10, settle_ms=args.settle_ms)

    cpu = None
    if not args.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(sa, W)

    head = results.get("fp16") or next(iter(results.values()))
    head_name = "fp16" if "fp16" in results else next(iter(results))
    head_st = {"fp16": sa.Storage.FP16, "fp32": sa.Storage.FP32, "bf16": sa.Storage.BF16}[head_name]
    out = {
        "metric": METRIC,
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_ms": args.settle_ms,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "storage": head_name,
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "dims": [256, 128, 128], "cells": n, "storage": head_name,
                   "bytes_per_cell": bytes_per_cell(head_st), "parallelism": "single-gpu"},
        "roofline": {
            "bound": "hbm",
            "achieved": round(head["kernel_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(head["kernel_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile(head_name),
            "kernel": "k_step (one launch per step: stream-collide of every fluid cell, wall records, halfway "
                      "bounce-back, FP64 force partials)",
            "kernel_us": round(head["kernel_us"], 2),
            "bytes_per_launch": head["kernel_bytes"],
            "bytes": f"{bytes_per_cell(head_st)} B (19 loads + 19 stores of {head_st.pop_bytes()} B + a 4 B flag) x "
                     f"{head['kernel_cells']} fluid cells the launch updates",
            "traffic_source": f"{TRAFFIC_PROFILE} (rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE per launch)",
        },
        "cpu_baseline": cpu,
        "storages": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()} for k, v in results.items()},
    }
    if "fp16" in results and "fp32" in results:
        out["fp16_vs_fp32_speedup"] = round(results["fp16"]["mlups"] / results["fp32"]["mlups"], 3)
    if drag is not None:
        out["drag"] = drag
    if large is not None:
        out["n384"] = large
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
