"""Launch-shape autotuner for the FD step kernels.

The fused 2D kernels give each thread ``rz2d`` rows of a 64-wide tile
(1, 2, 4 or 8) and split the tiles into eight contiguous XCD ranges of equal
estimated cost, a PML / edge tile counting ``pml_cost`` percent of an
interior one; the 3D kernels march ``zc3d`` z planes per block (0 = chosen
from the grid so that a launch has >= 2048 blocks).  Results do not depend on
any knob (every cell is computed by the same arithmetic), only the time
does, so the tuner times a velocity + stress half-step pair for every
candidate on a zeroed scratch state of the caller's shape, keeps the fastest
and sets it with ``rsg_set_tuning`` (2D: rz2d first, then pml_cost at that
rz2d).  Choices are cached per
(kind, padded grid, order, shots) in-process and, if ``RSG_TUNE_CACHE`` names
a JSON file, across processes.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import torch

from . import _lib as R

CANDIDATES = {"rz2d": (1, 2, 4, 8), "zc3d": (0, 2, 4, 8, 16, 32), "pml_cost": (100, 150, 200, 250, 300)}
_cache: dict = {}


def _key(kind, g, nshot):
    return f"{kind}:{g.nx}x{g.ny}x{g.nz}:lpml{g.lpml}:o{g.order}:s{nshot}"


def _load_file():
    path = os.environ.get("RSG_TUNE_CACHE")
    if path and os.path.exists(path):
        with open(path) as fh:
            _cache.update(json.load(fh))


def _save_file():
    path = os.environ.get("RSG_TUNE_CACHE")
    if path:
        with open(path, "w") as fh:
            json.dump(_cache, fh, indent=1, sort_keys=True)


def _strip_sizes(kind, g):
    nx, ny, nz, L = g.nx, g.ny, g.nz, g.lpml
    if kind == R.KIND_A2D:
        x, z = L * nz, nx * L
        return [x, x, z, z, x, x, z, z]
    if kind == R.KIND_E2D:
        x, z = L * nz, nx * L
        return [x, x, x, x, z, z, z, z, x, x, x, x, z, z, z, z]
    lr, fb, tb = nz * ny * L, nz * L * nx, L * ny * nx
    return [lr, lr, fb, fb, tb, tb, lr, lr, fb, fb, tb, tb]


class _Scratch:
    """Zeroed fields and strips of `nshot` shots, plus unit model arrays:
    the timing of a launch does not depend on the values."""

    def __init__(self, kind, g, tab, nshot):
        dev = torch.device("cuda", torch.cuda.current_device())
        n = g.nx * g.ny * g.nz
        self.keep = []

        def buf(m, val=0.0):
            t = torch.full((m,), val, dtype=torch.float32, device=dev)
            self.keep.append(t)
            return t.data_ptr()

        s = R.State()
        fields = {R.KIND_A2D: ("P", "Vx", "Vz"), R.KIND_E2D: ("Sxx", "Szz", "Sxz", "Vx", "Vz"),
                  R.KIND_A3D: ("P", "Vx", "Vy", "Vz")}[kind]
        for f in fields:
            setattr(s, f, buf(n * nshot))
        models = {R.KIND_A2D: ("L", "Rx", "Rz"), R.KIND_E2D: ("L", "L2M", "M", "Rx", "Rz"),
                  R.KIND_A3D: ("L", "Rx", "Ry", "Rz")}[kind]
        for f in models:
            setattr(s, f, buf(n, 1.0))
        s.tab = tab
        sizes = _strip_sizes(kind, g)
        for i, m in enumerate(sizes):
            s.psi[i] = buf(m * nshot)
        self.state = s
        self.batch = R.Batch(nshot=nshot, field_stride=n, xstrip_stride=g.lpml * g.nz,
                             zstrip_stride=g.nx * g.lpml)


def _launch_pair(kind, g, sc, nshot, stream):
    lib = R.lib()
    gs, ss = C.byref(g), C.byref(sc.state)
    if nshot == 1:
        for which in (0, 1):
            R.check(lib.rsg_repeat_step(kind, which, gs, ss, 1, stream), "rsg_repeat_step")
        return
    b = C.byref(sc.batch)
    if kind == R.KIND_A2D:
        R.check(lib.rsg_a2d_forward_velocity_batch(gs, ss, b, stream), "velocity_batch")
        R.check(lib.rsg_a2d_forward_stress_batch(gs, ss, b, stream), "stress_batch")
    else:
        R.check(lib.rsg_e2d_velocity_batch(gs, ss, b, 0, stream), "velocity_batch")
        R.check(lib.rsg_e2d_stress_batch(gs, ss, b, 0, stream), "stress_batch")


def time_pair_ms(kind, g, sc, nshot=1, reps=20):
    """Average time of one velocity + stress pair, from HIP events on the
    launch stream."""
    st = torch.cuda.current_stream()
    stream = C.c_void_p(st.cuda_stream)
    for _ in range(3):
        _launch_pair(kind, g, sc, nshot, stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        _launch_pair(kind, g, sc, nshot, stream)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(kind, g, tab_ptr, nshot=1, reps=20, force=False):
    """Pick the launch shape for one propagator kind on grid `g` (an
    ``_lib.Grid``; PML tables at device pointer `tab_ptr`), set it, and
    return ``{knob: value, "ms": {value: ms}}``.  3D batches are not
    supported (the 3D kernels take one shot)."""
    if kind == R.KIND_A3D and nshot != 1:
        raise R.RsgError("autotune: 3D kernels take one shot")
    knob = "zc3d" if kind == R.KIND_A3D else "rz2d"
    if not _cache:
        _load_file()
    key = _key(kind, g, nshot)
    if key in _cache and not force:
        for k in (knob, "pml_cost"):
            if k in _cache[key]:
                R.set_tuning(k, _cache[key][k])
        return _cache[key]
    sc = _Scratch(kind, g, tab_ptr, nshot)

    def sweep(name):
        # three rounds over the candidates, interleaved, keeping each one's
        # fastest: a single pass in order lets clock ramp-up or a stray
        # delay decide between candidates a few percent apart
        times = {}
        old = R.get_tuning(name)
        try:
            for _ in range(3):
                for v in CANDIDATES[name]:
                    R.set_tuning(name, v)
                    t = time_pair_ms(kind, g, sc, nshot, reps)
                    times[v] = min(times.get(v, t), t)
        except Exception:
            R.set_tuning(name, old)
            raise
        best = min(times, key=times.get)
        R.set_tuning(name, best)
        return best, {str(k): round(t, 5) for k, t in times.items()}

    best, ms = sweep(knob)
    res = {knob: best, "ms": ms}
    if kind != R.KIND_A3D:
        res["pml_cost"], res["ms_pml_cost"] = sweep("pml_cost")
    _cache[key] = res
    _save_file()
    return res


def tune_waves(waves, order, nshot=1, reps=20, force=False):
    """Tune for a ``Waves*`` object of rockseis_amd.core (its grid and PML)."""
    from .core import Der
    g = waves._grid(Der(1, 1, 1, 1, 1, 1, order))
    if waves.Pml.tab is None:
        waves.Pml.computeABC()
    return tune(waves.kind, g, waves.Pml.tab.data_ptr(), nshot, reps, force)


def tune_ragged(kind, grids, tab_ptr, reps=5, force=False):
    """Pick ``rz2d_ragged`` (rows per thread of the ragged-batch launches)
    for shots on the windows ``grids`` (``_lib.Grid`` each, one stencil
    order; PML tables at device pointer `tab_ptr`): every candidate times the
    half steps of one ragged batch (elastic: forward and adjoint-state
    velocity and stress) on zeroed scratch states of those windows, three
    interleaved rounds, fastest kept.  Returns ``{"rz2d_ragged": value,
    "ms": {value: ms}}``."""
    if kind not in (R.KIND_A2D, R.KIND_E2D):
        raise R.RsgError("tune_ragged: 2D kinds only")
    key = f"ragged{kind}:" + ",".join(f"{g.nx}x{g.nz}" for g in grids) + f":lpml{grids[0].lpml}:o{grids[0].order}"
    if not _cache:
        _load_file()
    if key in _cache and not force:
        R.set_tuning("rz2d_ragged", _cache[key]["rz2d_ragged"])
        return _cache[key]
    scratch = [_Scratch(kind, g, tab_ptr, 1) for g in grids]
    n = len(grids)
    garr = (R.Grid * n)(*grids)
    sarr = (R.State * n)(*[sc.state for sc in scratch])
    plan = C.c_void_p()
    lib = R.lib()
    R.check(lib.rsg_ragged2d_create(kind, n, garr, sarr, C.byref(plan)), "rsg_ragged2d_create")
    st = torch.cuda.current_stream()
    stream = C.c_void_p(st.cuda_stream)
    whichs = (0, 1, 2, 3) if kind == R.KIND_E2D else (0, 1)

    def one():
        for w in whichs:
            R.check(lib.rsg_ragged2d_step(plan, w, stream), "rsg_ragged2d_step")

    times, old = {}, R.get_tuning("rz2d_ragged")
    try:
        for _ in range(3):
            for v in CANDIDATES["rz2d"]:
                R.set_tuning("rz2d_ragged", v)
                one()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(st)
                for _ in range(reps):
                    one()
                e1.record(st)
                e1.synchronize()
                t = e0.elapsed_time(e1) / reps
                times[v] = min(times.get(v, t), t)
    except Exception:
        R.set_tuning("rz2d_ragged", old)
        raise
    finally:
        lib.rsg_ragged2d_destroy(plan)
    best = min(times, key=times.get)
    R.set_tuning("rz2d_ragged", best)
    _cache[key] = {"rz2d_ragged": best, "ms": {str(k): round(t, 5) for k, t in times.items()}}
    _save_file()
    return _cache[key]
