#!/usr/bin/env python
"""Benchmark of the staggered-grid FD time step on one MI355X.

Default workload: plain elastic 2D modelling, 2001 x 1001 cells at dx = 5 m
(2081 x 1081 padded with a 40-cell PML), 8th-order stencil, fp32, explosive
20 Hz Ricker, 2001 receivers recording vx and vz, dt = 0.5 ms, nt = 6000;
Vp = 1500->4000 m/s linear gradient times 20 seeded Gaussian anomalies,
Vs = Vp/1.73, Gardner density.  All inputs are synthetic and resident in HBM
before the timed region.

One "step" is one FD time step of that workload (velocity half step, stress
half step, source injection, two receiver gathers); --steps K times K
consecutive steps (K <= nt, default the full nt = 6000) after --warmup W
untimed steps (state and records are zeroed between warm-up and timing).
The time loop runs on the device as replayed hipGraph chunks.

Prints one JSON line: value = padded cells x K / wall seconds in
Gcell-updates/s, the roofline of the dominant kernel (elastic stress half
step: 44 B/cell + 32*lpml*(nx+nz) B of PML memory traffic per launch, timed
with HIP events on the launch stream) and the reference CPU path timed on
this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s (% HBM peak); FWI gradient shots/s, batched adjoint vs plain"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--graph-chunk", type=int, default=128,
                    help="largest hipGraph of the time loop (graphs of 1, 2, 4, ... steps up to it are "
                         "captured before the timed region; K steps replay its binary decomposition)")
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=40, help="time steps of the CPU sample")
    ap.add_argument("--cpu-start", type=int, default=3000,
                    help="first step of the CPU sample (the GPU run's state at that step is its start)")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="most threads of the shot-parallel CPU figure (the box's CPU share)")
    ap.add_argument("--cpu-config-steps", type=int, default=10,
                    help="time steps of the single-thread CPU samples of configs 3 and 5")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fwi", action="store_true", help="skip the FWI shots/s leg")
    ap.add_argument("--legs", default="a3d,e2dfwi,e3d",
                    help="extra legs, comma separated (empty: none): a3d (acoustic 3D modelling + FWI "
                         "gradient, 440^3 padded, nt 2500), e2dfwi (128 ragged elastic FWI windows, "
                         "nt 4000) and e3d (elastic 3D kernels and one FwiElastic3D gradient); about 25 s, "
                         "60 s and 10 s")
    ap.add_argument("--e2dfwi-shots", type=int, default=128)
    ap.add_argument("--e2dfwi-eager", action="store_true",
                    help="also time the elastic batch with eager launches (about 50 s more)")
    ap.add_argument("--fwi-plain-shots", type=int, default=8,
                    help="shots timed on the plain per-shot path (shots/s is per shot)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the roofline kernel timing (20 modelling steps for a non-zero state, then "
                         "--kernel-reps back-to-back launches of each half-step kernel): the launches a "
                         "rocprofv3 --kernel-trace --stats run of this command summarizes")
    ap.add_argument("--pmc-a3d", action="store_true",
                    help="no timing: the acoustic 3D workload of the a3d leg (440^3 padded, order 4), 20 "
                         "modelling steps, 10 launches of each half-step kernel and 5 calibration copies "
                         "(torch.neg of 256 MiB), for rocprofv3 --pmc passes")
    ap.add_argument("--pmc-dispatch", action="store_true",
                    help="no timing: 100 modelling steps, 10 launches of each half-step kernel and 5 "
                         "calibration copies (torch.neg of 256 MiB), for rocprofv3 --pmc passes")
    return ap.parse_args()


def pmc_dispatch():
    import torch
    from rockseis_amd import _lib as R
    from rockseis_amd import autotune
    mod, cfg, _ = build_workload()
    mod.prepare()
    autotune.tune_waves(mod.waves, cfg["order"], nshot=1)  # the bench's tile shape
    mod.launch(0, 100)
    torch.cuda.synchronize()
    g, s = mod._args[0], mod._args[1]
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for which in (0, 1):
        R.check(R.lib().rsg_repeat_step(R.KIND_E2D, which, C.byref(g), C.byref(s), 10, st))
    torch.cuda.synchronize()
    x = torch.ones(64 * 1024 * 1024, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print("pmc dispatch done", flush=True)


def _a3d_model():
    """The acoustic 3D modelling run of the a3d leg, prepared (not run)."""
    import numpy as np
    from rockseis_amd import core, synth
    n, lpml, dx, nt, dt, order = 400, 20, 10.0, 2500, 1e-3, 4
    vp = synth.layered_anomaly_vp3d(n, n, n, dx, dx, dx, nanom=10, seed=44)
    rho = synth.gardner(vp)
    m = core.ModelAcoustic3D(n, n, n, lpml, dx, dx, dx)
    m.Vp[:], m.R[:] = vp, rho
    src = core.Data3D(1, nt, dt, 0.0, "P")
    src.scoords[:] = np.float32([[n // 2 * dx, n // 2 * dx, 10.0]])
    src.data[:] = synth.ricker(nt, dt, 12.0)
    src.makeMap(m, core.SMAP)
    gx, gy = np.meshgrid(np.arange(200) * 20.0, np.arange(200) * 20.0)
    gc = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, 10.0)], 1).astype(np.float32)
    rec = core.Data3D(gc.shape[0], nt, dt, 0.0, "P")
    rec.gcoords[:] = gc
    rec.makeMap(m, core.GMAP)
    m.staggerModels()
    mod = core.ModellingAcoustic3D(m, src, order, 1)
    mod.setRecP(rec)
    return mod, m, src, rec, gc, (n, lpml, dx, nt, dt, order)


def pmc_a3d():
    import torch
    from rockseis_amd import _lib as R
    mod = _a3d_model()[0]
    mod.prepare()
    mod.launch(0, 20)
    torch.cuda.synchronize()
    g, s = mod._args[0], mod._args[1]
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for which in (0, 1):
        R.check(R.lib().rsg_repeat_step(R.KIND_A3D, which, C.byref(g), C.byref(s), 10, st))
    torch.cuda.synchronize()
    x = torch.ones(64 * 1024 * 1024, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print("pmc a3d dispatch done", flush=True)


def roofline_only(reps):
    import torch
    from rockseis_amd import autotune
    mod, cfg, _ = build_workload()
    mod.prepare()
    tuned = autotune.tune_waves(mod.waves, cfg["order"], nshot=1)
    mod.launch(0, 20)  # a non-zero state; runs this short keep the separate I/O kernels
    torch.cuda.synchronize()
    t_vel, t_str = kernel_time_ms(mod, 0, reps), kernel_time_ms(mod, 1, reps)
    return {"rows_per_thread": tuned["rz2d"], "pml_tile_cost": tuned.get("pml_cost"), "kernel_reps": reps,
            "e2d_velocity_kernel_us": round(t_vel * 1e3, 2), "e2d_stress_kernel_us": round(t_str * 1e3, 2)}


_T0 = time.perf_counter()


def _pow2_floor(n):
    return 1 << (max(int(n), 1).bit_length() - 1)


def _pow2_split(k, chunk):
    """The replays rsg_modelling_plan_run makes for k steps, as text."""
    top, parts = _pow2_floor(chunk), []
    n = top
    while n >= 1:
        c, k = divmod(k, n) if n == top else (k // n, k % n)
        if c:
            parts.append(f"{c} x {n}" if c > 1 else f"{n}")
        n //= 2
    return " + ".join(parts) if parts else "0"


def _progress(msg):
    """One progress line per finished stage, on stderr (stdout holds the JSON)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def build_workload():
    import numpy as np
    from rockseis_amd import core, synth

    nx, nz, lpml, dx, dt, nt, order, f0 = 2001, 1001, 40, 5.0, 5e-4, 6000, 8, 20.0
    vp = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1500.0, 4000.0, nanom=20, seed=2024)
    vs = (vp / 1.73).astype(np.float32)
    rho = synth.gardner(vp)
    m = core.ModelElastic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, False)
    m.Vp[:], m.Vs[:], m.R[:] = vp, vs, rho
    src = core.Data2D(1, nt, dt, 0.0, "P")
    src.scoords[:] = np.float32([[nx // 2 * dx, 10.0]])
    src.data[:] = synth.ricker(nt, dt, f0)
    src.makeMap(m, core.SMAP)
    recs = []
    for f in ("VX", "VZ"):
        r = core.Data2D(nx, nt, dt, 0.0, f)
        r.gcoords[:] = synth.surface_line(nx, dx, 10.0)
        r.makeMap(m, core.GMAP)
        recs.append(r)
    m.staggerModels()
    mod = core.ModellingElastic2D(m, src, order, 1)
    mod.setRecVx(recs[0])
    mod.setRecVz(recs[1])
    cfg = dict(nx=nx, nz=nz, lpml=lpml, dx=dx, dt=dt, nt=nt, order=order, f0=f0,
               nx_pml=nx + 2 * lpml, nz_pml=nz + 2 * lpml)
    return mod, cfg, (vp, vs, rho, src, recs)


def reset(mod):
    w = mod.waves
    for f in ("Sxx", "Szz", "Sxz", "Vx", "Vz"):
        getattr(w, f).zero_()
    for t in w.psi.values():
        t.zero_()
    for t in mod._rec_dev.values():
        t.zero_()


def kernel_time_ms(mod, which, reps):
    """Average duration of one half-step kernel, from HIP events around
    `reps` back-to-back launches on the stream the kernels run on."""
    import torch
    from rockseis_amd import _lib as R
    g, s = mod._args[0], mod._args[1]
    st = torch.cuda.current_stream()
    lib = R.lib()
    R.check(lib.rsg_repeat_step(R.KIND_E2D, which, C.byref(g), C.byref(s), 5, C.c_void_p(st.cuda_stream)))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    R.check(lib.rsg_repeat_step(R.KIND_E2D, which, C.byref(g), C.byref(s), reps, C.c_void_p(st.cuda_stream)))
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def _host_cpu():
    """The host CPU's model name and the cores this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def _e2d_state(mod):
    """The elastic state of a modelling run as one flat host array: Vx, Vz,
    Sxx, Szz, Sxz, then the 16 PML strips in the reference's member order."""
    import numpy as np
    w = mod.waves
    parts = [getattr(w, f).reshape(-1) for f in ("Vx", "Vz", "Sxx", "Szz", "Sxz")]
    parts += [w.psi[n].reshape(-1) for n in w.strips]
    import torch
    return torch.cat(parts).cpu().numpy().astype(np.float32)


def cpu_baseline(cfg, inputs, state, it0, nsteps, max_threads):
    """The reference WavesElastic2D (oracle/_ref, compiled from the reference
    sources) on this host, on a bounded sample of the same workload: time
    steps it0 .. it0+nsteps-1 of the run, started from the state the GPU run
    reached at step it0 (late in the run the wavefield covers most of the
    grid; the first steps of a run, mostly exact zeros, run faster on the
    CPU).  value: one thread.  shot_parallel: min(cores, max_threads) threads
    each running the same window as an independent shot at once (the
    reference's MPI task farm runs one shot per rank,
    parallel.cpp:690-732), summed.  Where the reference build is absent,
    the oracle's C restatement runs the first nsteps steps instead."""
    import concurrent.futures as cf
    import oracle
    from oracle import runs
    vp, vs, rho, src, recs = inputs
    cells = cfg["nx_pml"] * cfg["nz_pml"]
    cpu_model, cores = _host_cpu()
    args = (vp, vs, rho, cfg["lpml"], cfg["dx"], cfg["dx"], cfg["dt"], cfg["nt"], cfg["order"], False,
            src.scoords, src.data, recs[0].gcoords, it0, nsteps, state)
    if not oracle.ref_available():
        t0 = time.perf_counter()
        runs.e2d_run(vp, vs, rho, cfg["lpml"], cfg["dx"], cfg["dx"], cfg["dt"], nsteps, cfg["order"], False, "P",
                     src.scoords, src.data[:, :nsteps], "VX", recs[0].gcoords)
        t = time.perf_counter() - t0
        return {"value": round(cells * nsteps / t / 1e9, 6), "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
                "cpu_model": cpu_model, "host_cores": cores,
                "sample": f"first {nsteps} steps of the workload, oracle C restatement, one thread (set-up included)"}
    t1 = runs.ref_e2d_time_window(*args)
    nthr = max(1, min(cores, max_threads))
    with cf.ThreadPoolExecutor(nthr) as ex:
        ts = list(ex.map(lambda _: runs.ref_e2d_time_window(*args), range(nthr)))
    par = sum(cells * nsteps / t for t in ts)
    return {"value": round(cells * nsteps / t1 / 1e9, 6), "unit": "Gcell-updates/s", "cores": 1,
            "kind": "reference", "cpu_model": cpu_model, "host_cores": cores,
            "seconds_per_step": round(t1 / nsteps, 4),
            "shot_parallel": {"value": round(par / 1e9, 6), "unit": "Gcell-updates/s", "threads": nthr,
                              "seconds_per_step_per_shot": round(max(ts) / nsteps, 4)},
            "sample": (f"steps {it0}..{it0 + nsteps - 1} of the same elastic 2D workload (nt {cfg['nt']}) on the "
                       f"full {cfg['nx_pml']}x{cfg['nz_pml']} padded grid, order {cfg['order']}, started from the "
                       "GPU run's state at that step; reference WavesElastic2D built from its sources, "
                       "velocity + stress + source + 2 receiver sets per step, set-up excluded")}


def cpu_configs(n_window):
    """Single-thread reference step times on configs 3 and 5 (late windows
    from GPU states), and the per-shot FWI gradient times they imply:
    (numforw(nt, 50) + nt) steps of the Revolve schedule (the imaging
    condition and checkpoint copies not counted).  Config 4's 440^3 grid
    takes about a minute per reference step on one core and is left out of
    the default run."""
    import numpy as np
    import torch
    from oracle import runs
    from rockseis_amd import core, synth
    out = {}
    # config 3: acoustic 2D 801 x 401 at 10 m (861 x 461 padded), order 8, smoothed model
    nx, nz, dx, lpml, nt, dt, order = 801, 401, 10.0, 30, 3000, 1e-3, 8
    vt = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1500.0, 4000.0, nanom=20, seed=7)
    vp = synth.gaussian_smooth(vt, 250.0, dx, dx)
    rho = synth.gardner(vp)
    m = core.ModelAcoustic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, False)
    m.Vp[:], m.R[:] = vp, rho
    src = core.Data2D(1, nt, dt, 0.0, "P")
    src.scoords[:] = np.float32([[125.0 + 250.0 * 16, 10.0]])
    src.data[:] = synth.ricker(nt, dt, 10.0)
    src.makeMap(m, core.SMAP)
    rec = core.Data2D(nx, nt, dt, 0.0, "P")
    rec.gcoords[:] = synth.surface_line(nx, dx, 10.0)
    rec.makeMap(m, core.GMAP)
    m.staggerModels()
    mod = core.ModellingAcoustic2D(m, src, order, 1)
    mod.setRecP(rec)
    mod.prepare()
    it0 = nt // 2
    mod.launch(0, it0)
    torch.cuda.synchronize()
    w = mod.waves
    st = torch.cat([w.P.reshape(-1), w.Vx.reshape(-1), w.Vz.reshape(-1)] +
                   [w.psi[n].reshape(-1) for n in w.strips]).cpu().numpy()
    t = runs.ref_a2d_time_window(vp, rho, lpml, dx, dx, dt, nt, order, src.scoords, src.data, rec.gcoords, it0,
                                 n_window, st)
    steps = _numforw(nt, 50) + nt
    out["config3"] = {"seconds_per_step": round(t / n_window, 4),
                      "fwi_gradient_s_per_shot_estimate": round(t / n_window * steps, 1),
                      "sample": (f"acoustic 2D 861x461 padded, order 8: steps {it0}..{it0 + n_window - 1} of a "
                                 f"shot from the GPU state, one thread; gradient = {steps} steps")}
    del mod
    # config 5: one elastic window of 1300 x 600 cells (+lpml 30) at 5 m, order 8
    nxw, nz, dx, lpml, nt, dt = 1300, 600, 5.0, 30, 4000, 4.5e-4
    vp = np.clip(synth.gradient_anomaly_vp(nxw, nz, dx, dx, 1500.0, 4500.0, nanom=12, seed=5), 1500.0, 4500.0)
    vp = vp.astype(np.float32)
    vs, rho = (vp / 1.8).astype(np.float32), synth.gardner(vp)
    m = core.ModelElastic2D(nxw, nz, lpml, dx, dx, 0.0, 0.0, False)
    m.Vp[:], m.Vs[:], m.R[:] = vp, vs, rho
    src = core.Data2D(1, nt, dt, 0.0, "P")
    src.scoords[:] = np.float32([[nxw // 2 * dx, 10.0]])
    src.data[:] = synth.ricker(nt, dt, 15.0)
    src.makeMap(m, core.SMAP)
    recs = []
    for f in ("VX", "VZ"):
        r = core.Data2D(nxw, nt, dt, 0.0, f)
        r.gcoords[:] = synth.surface_line(nxw, dx, 10.0)
        r.makeMap(m, core.GMAP)
        recs.append(r)
    m.staggerModels()
    mod = core.ModellingElastic2D(m, src, 8, 1)
    mod.setRecVx(recs[0])
    mod.setRecVz(recs[1])
    mod.prepare()
    it0 = nt // 2
    mod.launch(0, it0)
    torch.cuda.synchronize()
    st = _e2d_state(mod)
    t = runs.ref_e2d_time_window(vp, vs, rho, lpml, dx, dx, dt, nt, 8, False, src.scoords, src.data,
                                 recs[0].gcoords, it0, n_window, st)
    steps = _numforw(nt, 50) + nt
    out["config5"] = {"seconds_per_step": round(t / n_window, 4),
                      "fwi_gradient_s_per_window_estimate": round(t / n_window * steps, 1),
                      "sample": (f"one elastic window of 1300 x 600 cells (1360 x 660 padded), order 8: steps "
                                 f"{it0}..{it0 + n_window - 1} from the GPU state, one thread; gradient = {steps} "
                                 "steps at the forward step's cost")}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): this path runs "
              "one shot window per GPU and needs no inter-GPU traffic", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if args.pmc_a3d:
        pmc_a3d()
        return
    if args.pmc_dispatch:
        pmc_dispatch()
        return
    if args.roofline_only:
        print(json.dumps(roofline_only(args.kernel_reps)), flush=True)
        return
    mod, cfg, inputs = build_workload()
    K = min(args.steps, cfg["nt"])
    W = min(args.warmup, cfg["nt"])
    mod.setGraphChunk(args.graph_chunk)
    mod.prepare()
    # launch shape picked by the autotuner (timed on a scratch state of this grid)
    from rockseis_amd import autotune
    tuned = autotune.tune_waves(mod.waves, cfg["order"], nshot=1)
    _progress(f"modelling workload built, tile rows per thread {tuned['rz2d']}, PML tile cost {tuned.get('pml_cost')}")
    mod.capture()  # the time loop's graphs, with the tuned launch shape
    if W:
        mod.launch(0, W)
    torch.cuda.synchronize()
    reset(mod)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mod.launch(0, K)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _progress(f"timed {K} modelling steps: {wall:.2f} s")

    cells = cfg["nx_pml"] * cfg["nz_pml"]
    value = cells * K / wall / 1e9
    # algorithmic bytes (fp32 two-pass leapfrog): velocity 7R+2W, stress 8R+3W per cell;
    # PML memory variables: 4 per half step on x-side strips and 4 on z-side, RMW
    strip_cells = cfg["lpml"] * (cfg["nx_pml"] + cfg["nz_pml"]) * 2
    b_vel = 36 * cells + 2 * 8 * strip_cells
    b_str = 44 * cells + 2 * 8 * strip_cells
    t_vel = kernel_time_ms(mod, 0, args.kernel_reps)
    t_str = kernel_time_ms(mod, 1, args.kernel_reps)
    # the CPU sample starts from the GPU run's state at step cpu_it0
    cpu_it0 = max(0, min(args.cpu_start, cfg["nt"] - args.cpu_steps))
    cpu_state = None
    if not args.no_cpu:
        reset(mod)
        mod.launch(0, cpu_it0)
        torch.cuda.synchronize()
        cpu_state = _e2d_state(mod)
    step_bytes = b_vel + b_str
    roof = {"bound": "hbm", "kernel": "e2d_stress_kernel<8>", "achieved": round(b_str / (t_str * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(b_str / (t_str * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": _traffic("e2d_stress"),
            "bytes_per_launch": b_str, "avg_launch_us": round(t_str * 1e3, 2),
            "velocity_kernel": {"achieved": round(b_vel / (t_vel * 1e-3) / 1e9, 1),
                                "avg_launch_us": round(t_vel * 1e3, 2), "bytes_per_launch": b_vel}}
    step_gbs = step_bytes * K / wall / 1e9
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(wall / K * 1e3, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain elastic 2D modelling (ModellingElastic2D::run)",
                   "grid": f"{cfg['nx']}x{cfg['nz']} + lpml {cfg['lpml']} -> {cfg['nx_pml']}x{cfg['nz_pml']}",
                   "dx_m": cfg["dx"], "order": cfg["order"], "dt_s": cfg["dt"], "nt": cfg["nt"],
                   "source": f"explosive {cfg['f0']:g} Hz Ricker", "receivers": "2001 x (vx, vz)",
                   "time_loop": (f"hipGraph replays only: graphs of 1, 2, 4, ..., {_pow2_floor(args.graph_chunk)} "
                                 f"steps captured before the timed region, {K} steps = "
                                 f"{_pow2_split(K, args.graph_chunk)}; source injection and both receiver "
                                 "gathers fused into the stress kernel (2 kernels per step)")
                   if args.graph_chunk > 0 else "eager launches, one step at a time",
                   "graph_chunk": args.graph_chunk, "rows_per_thread": tuned["rz2d"],
                   "pml_tile_cost": tuned.get("pml_cost"),
                   "parallelism": "single GPU"},
        "hbm_frac_whole_step": round(step_gbs / HBM_PEAK_GBS, 4),
        "roofline": roof,
    }
    if not args.no_fwi:
        del mod
        torch.cuda.empty_cache()
        out["fwi"] = fwi_leg(args.fwi_plain_shots)
    legs = [x for x in args.legs.split(",") if x]
    if "a3d" in legs:
        torch.cuda.empty_cache()
        out["a3d"] = a3d_leg()
    if "e2dfwi" in legs:
        torch.cuda.empty_cache()
        out["e2dfwi"] = e2dfwi_leg(args.e2dfwi_shots, eager=args.e2dfwi_eager)
    if "e3d" in legs:
        torch.cuda.empty_cache()
        out["e3d"] = e3d_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg, inputs, cpu_state, cpu_it0, args.cpu_steps, args.cpu_threads)
        _progress("cpu baseline timed")
        if not args.no_fwi:
            out["cpu_baseline"]["per_config"] = cpu_configs(args.cpu_config_steps)
            _progress("cpu per-config windows timed")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def fwi_leg(n_plain):
    """FWI gradient shots/s: acoustic 2D, 801x401 at 10 m (lpml 30), order 8,
    32 shots every 250 m, 801 receivers, 10 Hz Ricker, nt 3000, dt 1 ms;
    observed data from the seeded true model, gradient on the model smoothed
    with a 250 m Gaussian.  Batched = all 32 shots per launch, forward
    pressure of every step kept in HBM (and, for the like-for-like ratio,
    the same batch with 50 Revolve checkpoints); plain = one shot at a time
    through run_optimal with 50 Revolve checkpoints (per-step launches)."""
    import numpy as np
    import torch
    from rockseis_amd import fwi, synth
    nx, nz, dx, lpml, nt, dt, f0, order, nshot = 801, 401, 10.0, 30, 3000, 1e-3, 10.0, 8, 32
    vtrue = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1500.0, 4000.0, nanom=20, seed=7)
    rtrue = synth.gardner(vtrue)
    vp0 = synth.gaussian_smooth(vtrue, 250.0, dx, dx)
    rho0 = synth.gardner(vp0)
    wav = synth.ricker(nt, dt, f0)[None]
    gc = synth.surface_line(nx, dx, 10.0)
    zero = np.zeros((nx, nt), np.float32)
    shots = [fwi.Shot(np.float32([[125.0 + 250.0 * k, 10.0]]), wav, gc, zero) for k in range(nshot)]
    sv = fwi.Survey(shots, lpml, dx, dx, dt, order=order, ncheck=50, mode="store")
    _, _, _, dobs = fwi.batch_gradient(vtrue, rtrue, sv)         # observed data
    _progress("fwi: observed data modelled")
    dobs = dobs.cpu().numpy()
    for k, sh in enumerate(shots):
        sh.dobs = np.ascontiguousarray(dobs[k])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    misfit_b, vg, rg, _ = fwi.batch_gradient(vp0, rho0, sv)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    _progress(f"fwi: batched gradient of {nshot} shots {tb:.2f} s")
    mode_b = sv.mode_used
    # the same batch on the plain path's schedule: Revolve with 50 checkpoints
    sv_r = fwi.Survey(shots, lpml, dx, dx, dt, order=order, ncheck=50, mode="revolve")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    misfit_r, vg_r, _, _ = fwi.batch_gradient(vp0, rho0, sv_r)
    torch.cuda.synchronize()
    tr = time.perf_counter() - t0
    _progress(f"fwi: batched gradient of {nshot} shots, Revolve schedule {tr:.2f} s")
    same_r = bool(np.array_equal(vg_r.cpu().numpy(), vg.cpu().numpy()))
    del vg_r
    plain = fwi.Survey(shots, lpml, dx, dx, dt, order=order, ncheck=50, mode="revolve")
    n_plain = max(1, min(n_plain, nshot))
    t0 = time.perf_counter()
    mis_p = 0.0
    gv_p = np.zeros((nz, nx), np.float32)
    for k in range(n_plain):
        m1, a, _ = fwi.shot_gradient(vp0, rho0, plain, shots[k])
        mis_p += m1
        gv_p += a
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    _progress(f"fwi: plain gradient of {n_plain} shots {tp:.2f} s")
    # the batched gradient of the first n_plain shots must equal the plain one
    check = fwi.Survey(shots[:n_plain], lpml, dx, dx, dt, order=order, ncheck=50, mode="store")
    m_chk, vg_chk, _, _ = fwi.batch_gradient(vp0, rho0, check)
    err = float(np.linalg.norm(vg_chk.cpu().numpy() - gv_p) / np.linalg.norm(gv_p))
    del vg_chk
    # RtmAcoustic2D::run_optimal of shot 0 on the same model with its observed
    # records: zero offset, then 21 horizontal subsurface offsets
    rtm_s = {}
    for nhx in (1, 21):
        rtm_s[nhx] = _rtm_shot(vp0, rho0, lpml, dx, dt, nt, order, shots[0], nhx)
        _progress(f"fwi: RTM of one shot, {nhx} subsurface offsets {rtm_s[nhx]:.2f} s")
    from rockseis_amd import rtm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    img = rtm.batch_image(vp0, rho0, sv_r)
    torch.cuda.synchronize()
    rtm_b = time.perf_counter() - t0
    del img
    _progress(f"fwi: RTM of {nshot} shots as one batch {rtm_b:.2f} s")
    return {"workload": "acoustic 2D FWI gradient, 801x401 at 10 m (861x461 padded), order 8, 32 shots, "
                        "801 receivers, 10 Hz Ricker, nt 3000",
            "batched_shots_per_s": round(nshot / tb, 3), "batched_s": round(tb, 3),
            "batched_history": mode_b,
            "plain_shots_per_s": round(n_plain / tp, 3), "plain_s": round(tp, 3),
            "plain_shots_timed": n_plain, "plain_checkpointing": "Revolve, 50 snapshots in HBM",
            "speedup_batched_vs_plain": round((nshot / tb) / (n_plain / tp), 2),
            "batched_revolve_shots_per_s": round(nshot / tr, 3), "batched_revolve_s": round(tr, 3),
            "speedup_batched_revolve_vs_plain": round((nshot / tr) / (n_plain / tp), 2),
            "batched_revolve_gradient_equals_store": same_r,
            "misfit_32_shots": misfit_b,
            "batched_vs_plain_gradient_rel_l2": err,
            "rtm_shot_s": round(rtm_s[1], 3), "rtm_shot_21_offsets_s": round(rtm_s[21], 3),
            "rtm_batch_s": round(rtm_b, 3), "rtm_batch_shots_per_s": round(nshot / rtm_b, 3),
            "speedup_rtm_batch_vs_single": round((nshot / rtm_b) / (1.0 / rtm_s[1]), 2),
            "rtm": "RtmAcoustic2D::run_optimal of one shot, 50 Revolve checkpoints in HBM, image with nhx 1 / 21; "
                   "batch: the 32 shots stacked in one set of launches per step (rsg_rtm_a2d_batch), same schedule"}


def _rtm_shot(vp0, rho0, lpml, dx, dt, nt, order, shot, nhx):
    """Wall time of RtmAcoustic2D::run_optimal for one shot of the FWI leg
    (the image returned to the host included)."""
    import torch
    from rockseis_amd import core, fwi, rtm
    nz, nx = vp0.shape
    m = core.ModelAcoustic2D(nx, nz, lpml, dx, dx, 0.0, 0.0, False)
    m.Vp[:], m.R[:] = vp0, rho0
    src = core.Data2D(1, nt, dt, 0.0, "P")
    src.scoords[:] = shot.scoords
    src.data[:] = shot.wavelet
    src.makeMap(m, core.SMAP)
    dat = core.Data2D(shot.gcoords.shape[0], nt, dt, 0.0, "P")
    dat.gcoords[:] = shot.gcoords
    dat.data[:] = shot.dobs
    dat.makeMap(m, core.GMAP)
    m.staggerModels()
    img = fwi.Image2D(m, nhx, 1)
    r = rtm.RtmAcoustic2D(m, img, src, dat, order, 1)
    r.setNcheck(50)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r.run_optimal()
    return time.perf_counter() - t0


def e3d_leg():
    """Elastic 3D: the four half-step kernels (forward and adjoint-state
    velocity / stress) at 200^3 (240^3 padded, lpml 20), order 8, free
    surface, timed with HIP events over 20 back-to-back launches on the
    launch stream; then one FwiElastic3D gradient of 120^3 (160^3 padded),
    order 8, nt 500, explosive 15 Hz Ricker, 22 x 22 three-component
    receivers, 20 Revolve checkpoints in 16 bits (S16), after a warm-up
    gradient.  Algorithmic bytes per cell (each array touched once):
    forward velocity 60 (six stresses, three buoyancies, three velocities
    read; three velocities written), forward stress 80 (three velocities,
    six stresses, five moduli read; six stresses written), adjoint velocity
    80 (the forward velocity's arrays and the five moduli; also for the
    workspace entry, whose premultiplied stresses add 68 B per cell of
    their own), adjoint stress 60 (three velocities and six stresses read;
    six stresses written)."""
    import numpy as np
    import torch
    from rockseis_amd import _lib as R, core, fwi, synth
    n, L = 200, 20
    m = core.ModelElastic3D(n, n, n, L, 10.0, 10.0, 10.0, fs=True)
    m.Vp[:], m.Vs[:], m.R[:] = 3000.0, 1700.0, 2000.0
    m.staggerModels()
    w = core.WavesElastic3D(m, 10, 1e-3)
    w.getPml().setSmax(core.SMAX)
    w.getPml().computeABC()
    g, st = w._grid(8), w._state(m)
    s_ = torch.cuda.current_stream()
    sp = C.c_void_p(s_.cuda_stream)
    cells = w.nx_pml * w.ny_pml * w.nz_pml
    lib = R.lib()
    kt = {}
    wsp = torch.empty(6 * cells, dtype=torch.float32, device="cuda")
    bw_ws = lambda g_, s_, p_: lib.rsg_e3d_backward_velocity_ws(g_, s_, C.c_void_p(wsp.data_ptr()), p_)  # noqa: E731
    for name, f, b in (("velocity", lib.rsg_e3d_forward_velocity, 60), ("stress", lib.rsg_e3d_forward_stress, 80),
                       ("velocity_adjoint_state", lib.rsg_e3d_backward_velocity, 80),
                       ("velocity_adjoint_state_workspace", bw_ws, 80),
                       ("stress_adjoint_state", lib.rsg_e3d_backward_stress, 60)):
        for _ in range(3):
            R.check(f(C.byref(g), C.byref(st), sp), name)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s_)
        for _ in range(20):
            f(C.byref(g), C.byref(st), sp)
        e1.record(s_)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / 20 * 1e3
        gbs = cells * b / us / 1e3
        kt[name] = {"avg_launch_us": round(us, 1), "gcell_per_s": round(cells / us / 1e3, 2),
                    "bytes_per_cell": b, "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    _progress("e3d: half-step kernels timed")
    del w, st, wsp
    n, nt, dt = 120, 500, 1e-3
    m = core.ModelElastic3D(n, n, n, L, 10.0, 10.0, 10.0, fs=True)
    m.Vp[:], m.Vs[:], m.R[:] = 3000.0, 1700.0, 2000.0

    def d3(coords, field, data=None):
        c = np.asarray(coords, np.float32).reshape(-1, 3)
        d = core.Data3D(c.shape[0], nt, dt, 0.0, field)
        d.scoords[:] = c
        d.gcoords[:] = c
        if data is not None:
            d.data[:] = data
        return d

    src = d3([[600.0, 600.0, 50.0]], "P", synth.ricker(nt, dt, 15.0)[None] * 1e8)
    src.makeMap(m, core.SMAP)
    xy = np.stack(np.meshgrid(np.arange(50, 1150, 50.0), np.arange(50, 1150, 50.0), indexing="ij"), -1).reshape(-1, 2)
    gc = np.concatenate([xy, np.full((xy.shape[0], 1), 20.0)], 1)
    rng = np.random.default_rng(0)
    obs = []
    for f in ("VX", "VY", "VZ"):
        d = d3(gc, f, (rng.standard_normal((gc.shape[0], nt)) * 1e-3).astype(np.float32))
        d.makeMap(m, core.GMAP)
        obs.append(d)
    m.staggerModels()
    fw = fwi.FwiElastic3D(m, src, *obs, 8, 1)
    imgs = [fwi.Image3D(m) for _ in range(3)]
    fw.setVpgrad(imgs[0])
    fw.setVsgrad(imgs[1])
    fw.setRhograd(imgs[2])
    fw.setNcheck(20)
    fw.setFp16Storage("s16")
    fw.run_optimal()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fw.run_optimal()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _progress(f"e3d: FWI gradient {wall:.2f} s")
    return {"workload": "elastic 3D, 200^3 at 10 m (240^3 padded), order 8, free surface: half-step kernels; "
                        "FwiElastic3D::run_optimal of 120^3 (160^3 padded), order 8, nt 500, 484 x 3 receivers, "
                        "20 S16 Revolve checkpoints",
            "kernels": kt, "fwi_gradient_s": round(wall, 3), "fwi_misfit": fw.misfit64}


def a3d_leg():
    """Acoustic 3D modelling and single-shot FWI gradient: 400^3 at 10 m
    (440^3 padded, lpml 20), order 4, layered Vp (1500-3300 m/s, 10 seeded
    anomalies of +-8%), Gardner rho, 12 Hz Ricker at the surface centre,
    200 x 200 receivers every 20 m at z = 10 m, dt 1 ms, nt 2500.  The
    modelled records of the true model are the observed data of the gradient,
    which runs on the layered background (the true model without its
    anomalies), with 50 Revolve checkpoints in 16 bits (S16)."""
    import torch
    from rockseis_amd import _lib as R, core, fwi, synth
    mod, m, src, rec, gc, (n, lpml, dx, nt, dt, order) = _a3d_model()
    mod.setGraphChunk(64)
    mod.prepare()
    mod.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mod.launch(0, nt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    _progress(f"a3d: {nt} modelling steps {wall:.2f} s")
    mod.finish()
    N = m.getNx_pml() * m.getNy_pml() * m.getNz_pml()
    g, st = mod._args[0], mod._args[1]
    strips = lpml * (m.getNy_pml() * m.getNz_pml() + m.getNx_pml() * m.getNz_pml() + m.getNx_pml() * m.getNy_pml())
    kt = {}
    for which, name, b in ((0, "velocity", 40 * N + 16 * strips), (1, "stress", 24 * N + 16 * strips)):
        s_ = torch.cuda.current_stream()
        R.check(R.lib().rsg_repeat_step(R.KIND_A3D, which, C.byref(g), C.byref(st), 3, C.c_void_p(s_.cuda_stream)))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s_)
        R.check(R.lib().rsg_repeat_step(R.KIND_A3D, which, C.byref(g), C.byref(st), 20, C.c_void_p(s_.cuda_stream)))
        e1.record(s_)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / 20
        kt[name] = {"avg_launch_us": round(ms * 1e3, 1), "bytes_per_launch": b,
                    "achieved_GBs": round(b / (ms * 1e-3) / 1e9, 1),
                    "frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    dobs = rec.data.copy()
    del mod
    torch.cuda.empty_cache()
    v0 = synth.layered_anomaly_vp3d(n, n, n, dx, dx, dx, nanom=0, seed=44)
    m0 = core.ModelAcoustic3D(n, n, n, lpml, dx, dx, dx)
    m0.Vp[:], m0.R[:] = v0, synth.gardner(v0)
    src.makeMap(m0, core.SMAP)
    obs = core.Data3D(gc.shape[0], nt, dt, 0.0, "P")
    obs.gcoords[:] = gc
    obs.data[:] = dobs
    obs.makeMap(m0, core.GMAP)
    m0.staggerModels()
    f = fwi.FwiAcoustic3D(m0, src, obs, order, 1)
    vg, rg = fwi.Image3D(m0), fwi.Image3D(m0)
    f.setVpgrad(vg)
    f.setRhograd(rg)
    f.setNcheck(50)
    f.setFp16Storage("s16")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f.run_optimal()
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    _progress(f"a3d: FWI gradient {tg:.2f} s")
    return {"workload": "acoustic 3D, 400^3 at 10 m (440^3 padded), order 4, nt 2500, 200x200 receivers",
            "modelling_gcell_per_s": round(N * nt / wall / 1e9, 3), "modelling_s": round(wall, 3),
            "kernels": kt, "fwi_gradient_s": round(tg, 2),
            "fwi": ("FwiAcoustic3D::run_optimal, 50 Revolve checkpoints in 16 bits (S16: int16 mantissas, one "
                    "power-of-two exponent per 256 values), layered background model"),
            "misfit": f.misfit64}


def e2dfwi_leg(nshot, n_plain=4, eager=False):
    """Batched elastic FWI gradients over ragged windows: a 6000 x 600 model
    at 5 m (Vp 1500-4500 m/s, Vs = Vp/1.8, Gardner rho), nshot windows of
    seeded widths 600..2000 x 600 cells, order 8, lpml 30, 15 Hz Ricker
    explosive source at the window centre (z = 10 m), receivers at every cell
    of the window at z = 10 m (P, Ux, Uz), dt 0.45 ms, nt 4000, 50 Revolve
    checkpoints in 16 bits (S16); lambda / mu / rho kernels.  The observed records are
    zero (the timed work does not depend on their values).  Plain = the first
    n_plain windows one at a time."""
    import numpy as np
    import torch
    from rockseis_amd import fwi, synth
    nx, nz, dx, lpml, nt, dt, order = 6000, 600, 5.0, 30, 4000, 4.5e-4, 8
    rng = np.random.default_rng(5)
    vp = synth.gradient_anomaly_vp(nx, nz, dx, dx, 1500.0, 4500.0, nanom=60, seed=5)
    vp = np.clip(vp, 1500.0, 4500.0).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    rho = synth.gardner(vp)
    wav = synth.ricker(nt, dt, 15.0)[None]
    shots = []
    cells = 0
    for k in range(nshot):
        w = int(rng.integers(600, 2001))
        x0 = int(rng.integers(0, nx - w + 1))
        gc = synth.surface_line(w, dx, 10.0, x0 * dx)
        sc = np.float32([[(x0 + w // 2) * dx, 10.0]])
        shots.append(fwi.ElasticShot(sc, wav, gc, np.zeros((3, w, nt), np.float32), x0, w, "P"))
        cells += (w + 2 * lpml) * (nz + 2 * lpml)
    sv = fwi.ElasticSurvey(shots, lpml, dx, dx, dt, order, False, 50, fp16="s16")
    _progress(f"e2dfwi: {nshot} windows set up ({cells} padded cells)")
    torch.cuda.synchronize()
    te = mis = None
    if eager:
        t0 = time.perf_counter()
        mis, l, u, r, _ = fwi.elastic_batch_gradient(vp, vs, rho, sv, scale=False)
        torch.cuda.synchronize()
        te = time.perf_counter() - t0
        _progress(f"e2dfwi: batched gradients, eager launches {te:.2f} s")
        del l, u, r
        torch.cuda.empty_cache()
    # the persistent plan: built once (the Revolve sweep captured as two
    # hipGraphs), then one replay = one gradient of all windows
    t0 = time.perf_counter()
    plan = fwi.ElasticFwiPlan(vp, vs, rho, sv, scale=False)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    rz_ragged = plan.tuned["rz2d_ragged"]
    _progress(f"e2dfwi: plan built {t_build:.2f} s (autotuned rows per thread {rz_ragged})")
    t0 = time.perf_counter()
    mis_p = plan.run()[0]
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    plan.close()
    del plan
    torch.cuda.empty_cache()
    _progress(f"e2dfwi: batched gradients, plan replay {tb:.2f} s")
    n_plain = max(1, min(n_plain, nshot))
    t0 = time.perf_counter()
    for sh in shots[:n_plain]:
        one = fwi.ElasticSurvey([sh], lpml, dx, dx, dt, order, False, 50, fp16="s16")
        fwi.elastic_batch_gradient(vp, vs, rho, one, scale=False)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    _progress(f"e2dfwi: plain gradients of {n_plain} windows {tp:.2f} s")
    steps = nt + _numforw(nt, 50)  # forward steps incl. recomputation + adjoint steps
    return {"workload": f"{nshot} elastic windows of 600-2000 x 600 cells (+lpml 30) from a 6000x600 model at "
                        "5 m, order 8, nt 4000, 16-bit (S16) Revolve checkpoints (50)",
            "padded_cells_per_step": cells, "batched_s": round(tb, 2),
            "batched": "persistent plan: the Revolve sweep as two hipGraphs replayed (rsg_fwi_e2d_plan_run)",
            "plan_build_s": round(t_build, 2), "rows_per_thread": rz_ragged,
            "batched_eager_s": None if te is None else round(te, 2),
            "plan_misfit_equals_eager": None if mis is None else bool(abs(mis_p - mis) <= 1e-9 * abs(mis)),
            "batched_shots_per_s": round(nshot / tb, 3),
            "cell_updates_per_s_G": round(cells * steps / tb / 1e9, 2),
            "plain_shots_timed": n_plain, "plain_s": round(tp, 2),
            "plain_shots_per_s": round(n_plain / tp, 3),
            "speedup_batched_vs_plain": round((nshot / tb) / (n_plain / tp), 2)}


def _numforw(nt, snaps):
    from rockseis_amd import _lib as R
    return R.lib().rsg_revolve_numforw(nt, snaps)


def _traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE
    for this kernel at this workload (profiles/<name>_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


if __name__ == "__main__":
    main()
