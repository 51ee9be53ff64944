"""The C-ABI library: it loads, exports every function include/rsgpu.h
declares, and its host-side set-up (PML tables, Der coefficients, model
staggering, coordinate maps) is bit-exact with the reference fixtures.
No GPU is needed for any of these."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from rockseis_amd import _lib as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "rsgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rsg_[a-z0-9_]+)\s*\(", txt)))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def test_exports_every_declared_symbol():
    L = R.lib()
    names = _declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(R.EXPORTS)


def test_error_path_reports_message():
    L = R.lib()
    g = R.Grid(nx=10, ny=1, nz=10, lpml=3, order=9, fs=0, dx=1, dy=1, dz=1, dt=1e-3)
    s = R.State()
    rc = L.rsg_a2d_forward_velocity(C.byref(g), C.byref(s), None)
    assert rc == -1
    assert b"order" in L.rsg_last_error()


def test_pml_tables_match_reference(golden):
    g = golden("pml")
    for k in g.files:
        lpml = int(k[1:].split("_")[0])
        dt = float(k.split("_dt")[1].split("_")[0])
        sign = int(k.split("_s")[1])
        t = np.zeros(12 * lpml, np.float32)
        R.check(R.lib().rsg_pml_tables(lpml, dt, 1200.0, sign, _p(t)))
        assert np.array_equal(t, g[k]), k


def test_der_coeffs():
    import oracle
    c = np.zeros(36, np.float32)
    R.check(R.lib().rsg_der_coeffs(_p(c)))
    o = np.zeros(36, np.float32)
    oracle.fn("der_coeffs")(o)
    assert np.array_equal(c, o)


def test_stagger_match_reference(golden):
    g = golden("stagger")
    lpml = int(g["lpml"])
    nz, nx = g["vp"].shape
    nz3, ny3, nx3 = g["vp3"].shape
    L = R.lib()
    for fs in (0, 1):
        shp = (nz + 2 * lpml, nx + 2 * lpml)
        out = [np.zeros(shp, np.float32) for _ in range(3)]
        R.check(L.rsg_stagger_a2d(_p(g["vp"]), _p(g["rho"]), nx, nz, lpml, fs, *map(_p, out)))
        for o, k in zip(out, ("L", "Rx", "Rz")):
            assert np.array_equal(o, g[f"a2d_fs{fs}_{k}"]), k
        out = [np.zeros(shp, np.float32) for _ in range(5)]
        R.check(L.rsg_stagger_e2d(_p(g["vp"]), _p(g["vs"]), _p(g["rho"]), nx, nz, lpml, fs, *map(_p, out)))
        for o, k in zip(out, ("L", "L2M", "M", "Rx", "Rz")):
            assert np.array_equal(o, g[f"e2d_fs{fs}_{k}"]), k
        shp3 = (nz3 + 2 * lpml, ny3 + 2 * lpml, nx3 + 2 * lpml)
        out = [np.zeros(shp3, np.float32) for _ in range(4)]
        R.check(L.rsg_stagger_a3d(_p(g["vp3"]), _p(g["rho3"]), nx3, ny3, nz3, lpml, fs, *map(_p, out)))
        for o, k in zip(out, ("L", "Rx", "Ry", "Rz")):
            assert np.array_equal(o, g[f"a3d_fs{fs}_{k}"]), k


def test_zero_density_is_an_error():
    vp = np.full((4, 5), 2000, np.float32)
    rho = np.ones((4, 5), np.float32)
    rho[2, :] = 0.0  # staggered averages vanish along this row
    out = [np.zeros((8, 9), np.float32) for _ in range(3)]
    rc = R.lib().rsg_stagger_a2d(_p(vp), _p(rho), 5, 4, 2, 0, *map(_p, out))
    assert rc != 0 and b"Zero density" in R.lib().rsg_last_error()


def test_makemap_match_reference(golden):
    g = golden("makemap")
    coords = g["coords"]
    n = coords.shape[0]
    for field in (0, 1, 3):
        m = np.full((n, 2), -1, np.int32)
        s = np.zeros((n, 2), np.float32)
        R.check(R.lib().rsg_makemap2d(_p(coords), n, 40, 30, 10.0, 10.0, 0.0, 0.0, field, _p(m), _p(s)))
        assert np.array_equal(m, g[f"map_f{field}"])
        ok = m >= 0
        assert np.array_equal(s[ok], g[f"shift_f{field}"][ok])


def test_makemap3d_against_oracle():
    from oracle import runs
    rng = np.random.default_rng(4)
    c = np.stack([rng.uniform(-20, 230, 200), rng.uniform(-20, 250, 200), rng.uniform(-20, 210, 200)],
                 1).astype(np.float32)
    c[:4] = np.float32([[0, 0, 0], [10, 20, 30], [220, 230, 200], [5, 5, 5]])
    for name, f in (("P", 0), ("VX", 1), ("VY", 2), ("VZ", 3)):
        m = np.full((200, 3), -1, np.int32)
        s = np.zeros((200, 3), np.float32)
        R.check(R.lib().rsg_makemap3d(_p(c), 200, 22, 24, 20, 10.0, 10.0, 10.0, 0.0, 0.0, 0.0, f, _p(m), _p(s)))
        mo, so = runs.makemap3d(c, (22, 24, 20), (10.0, 10.0, 10.0), name)
        assert np.array_equal(m, mo)
        assert np.array_equal(s[m >= 0], so[mo >= 0])


def test_revolve_schedule_matches_reference(golden):
    from tests.golden.make_golden import REVOLVE_CASES
    g = golden("revolve")
    L = R.lib()
    for nt, sn in REVOLVE_CASES:
        ref = g[f"n{nt}_s{sn}"]
        a = np.zeros(ref.shape[0] + 8, np.int32)
        c = a.copy()
        k = a.copy()
        n = C.c_int()
        R.check(L.rsg_revolve_schedule(nt, sn, _p(a), _p(c), _p(k), a.size, C.byref(n)))
        got = np.stack([a[:n.value], c[:n.value], k[:n.value]], 1)
        assert np.array_equal(got, ref), (nt, sn)
        assert L.rsg_revolve_numforw(nt, sn) == int(g[f"n{nt}_s{sn}_numforw"])


def test_snap_cadence_matches_oracle():
    from oracle.fwi import snap_cadence
    for nt in (1, 300, 2999, 3000, 6000):
        for dt in (1e-3, 4.5e-4):
            for snapinc in (1, 3, 7, 50):
                a, b = C.c_int(), C.c_int()
                R.check(R.lib().rsg_snap_cadence(nt, dt, snapinc, C.byref(a), C.byref(b)))
                assert (a.value, b.value) == snap_cadence(nt, dt, snapinc)
    a, b = C.c_int(), C.c_int()
    assert R.lib().rsg_snap_cadence(10, 1e-3, 0, C.byref(a), C.byref(b)) == -1  # RSG_EINVAL


def test_struct_layouts_match_header(tmp_path):
    """Every struct of include/rsgpu.h has the size and member offsets that
    the ctypes mirrors in rockseis_amd._lib assume (compiled with gcc)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    structs = {"rsg_grid": R.Grid, "rsg_state": R.State, "rsg_traces": R.Traces, "rsg_batch": R.Batch,
               "rsg_fwi_a2d": R.Fwi, "rsg_fwi_e2d_shot": R.FwiE2DShot, "rsg_fwi_e2d": R.FwiE2D,
               "rsg_fwi_a3d": R.FwiA3D, "rsg_e2d_records": R.E2DRecords, "rsg_rtm_a2d": R.RtmA2D,
               "rsg_rtm_a3d": R.RtmA3D, "rsg_rss_header": R.RssHeader, "rsg_image2d": R.Image2DDesc,
               "rsg_state_e3d": R.StateE3D, "rsg_fwi_e3d": R.FwiE3D}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rsgpu.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'    printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'    printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("    return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                              text=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)


def _kernel_private_sizes():
    """{kernel: private segment bytes per lane} for every gfx950 kernel of
    librsgpu.so, read from the metadata notes of its code objects (the
    .hip_fatbin section holds one offload bundle per source file)."""
    import struct
    import subprocess
    import tempfile
    llvm = "/opt/rocm/llvm/bin"
    if not os.path.exists(os.path.join(llvm, "llvm-readelf")):
        pytest.skip("llvm-readelf not installed")
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        fb = os.path.join(tmp, "fatbin")
        subprocess.run([os.path.join(llvm, "llvm-objcopy"), "--dump-section", f".hip_fatbin={fb}", R.LIBPATH,
                        os.path.join(tmp, "lib.so")], check=True, capture_output=True)
        data = open(fb, "rb").read()
        start = data.find(magic)
        assert start >= 0
        while start >= 0:
            n, = struct.unpack_from("<Q", data, start + 24)
            pos = start + 32
            for _ in range(n):
                off, size, tl = struct.unpack_from("<QQQ", data, pos)
                triple = data[pos + 24:pos + 24 + tl].decode()
                pos += 24 + tl
                if "gfx950" in triple and size:
                    co = os.path.join(tmp, "co.o")
                    open(co, "wb").write(data[start + off:start + off + size])
                    notes = subprocess.run([os.path.join(llvm, "llvm-readelf"), "--notes", co], check=True,
                                           capture_output=True, text=True).stdout
                    name = None
                    for line in notes.splitlines():
                        line = line.strip()
                        if line.startswith(".name:"):
                            name = line.split(":", 1)[1].strip()
                        elif line.startswith(".private_segment_fixed_size:") and name:
                            out[name] = int(line.split(":", 1)[1])
            start = data.find(magic, start + 1)
    return out


def test_no_kernel_uses_scratch():
    """Register spills or a private copy of a kernel's argument struct put
    per-lane data in scratch memory, which costs an order of magnitude in
    the stencil kernels: every kernel of the library must use none."""
    sizes = _kernel_private_sizes()
    assert len(sizes) > 100
    assert any("e2d_stress_ragged_kernel" in k for k in sizes)
    bad = {k: v for k, v in sizes.items() if v}
    assert not bad, bad


def test_fwi_e2d_plan_rejects_bad_arguments():
    """The plan entry points validate their arguments before any device
    work: null plan pointer, null settings, no shots, a bad misfit type;
    destroy takes null."""
    L = R.lib()
    h = C.c_void_p()
    assert L.rsg_fwi_e2d_plan_create(None, None, 1, C.byref(h)) != 0
    assert L.rsg_fwi_e2d_plan_create(None, None, 1, None) != 0
    f = R.FwiE2D()
    shots = (R.FwiE2DShot * 1)()
    assert L.rsg_fwi_e2d_plan_create(C.byref(f), shots, 0, C.byref(h)) != 0
    assert L.rsg_fwi_e2d_plan_create(C.byref(f), shots, 1, C.byref(h)) != 0  # null tables / images
    assert b"null" in L.rsg_last_error()
    f.tab, f.nx, f.nz, f.misfit_type = 1, 10, 10, 7
    for c in range(3):
        f.grad[c] = 1
    assert L.rsg_fwi_e2d_plan_create(C.byref(f), shots, 1, C.byref(h)) != 0
    assert b"misfit type" in L.rsg_last_error()
    assert h.value is None
    assert L.rsg_fwi_e2d_plan_run(None, None, None) != 0
    assert L.rsg_fwi_e2d_plan_destroy(None) == 0


def test_modelling_plan_rejects_bad_arguments():
    """rsg_modelling_plan_create / _run validate their arguments before any
    device work: null output, null grid or state, a chunk outside 1..4096,
    an unknown kind, a snapshot set without a field; run and destroy take
    a null plan."""
    L = R.lib()
    h = C.c_void_p()
    g, s = R.Grid(), R.State()
    assert L.rsg_modelling_plan_create(R.KIND_A2D, C.byref(g), C.byref(s), None, 0, None, 0, None, 8,
                                       None, None) != 0
    assert L.rsg_modelling_plan_create(R.KIND_A2D, None, C.byref(s), None, 0, None, 0, None, 8,
                                       None, C.byref(h)) != 0
    for chunk in (0, 5000):
        assert L.rsg_modelling_plan_create(R.KIND_A2D, C.byref(g), C.byref(s), None, 0, None, 0, None,
                                           chunk, None, C.byref(h)) != 0
        assert b"max_chunk" in L.rsg_last_error()
    assert L.rsg_modelling_plan_create(9, C.byref(g), C.byref(s), None, 0, None, 0, None, 8,
                                       None, C.byref(h)) != 0
    sn = (R.SnapSet * 1)()
    sn[0].snapinc = 1
    sn[0].out = 1
    assert L.rsg_modelling_plan_create(R.KIND_A2D, C.byref(g), C.byref(s), None, 0, None, 1, sn, 8,
                                       None, C.byref(h)) != 0
    assert b"snapshot" in L.rsg_last_error()
    assert h.value is None
    assert L.rsg_modelling_plan_run(None, 0, 1, None) != 0
    assert L.rsg_modelling_plan_destroy(None) == 0


def test_fwi_e3d_rejects_bad_arguments():
    """rsg_fwi_e3d_gradient and the elastic 3D adjoint-state entry points
    validate their arguments before any device work: null arguments, a
    source of another field, Ux / Uy / Uz records of the wrong fields or
    trace counts, a domain model, an image of the wrong size, an unknown
    storage kind or misfit type."""
    L = R.lib()
    g, f = R.Grid(), R.FwiE3D()
    assert L.rsg_fwi_e3d_gradient(None, None, None) != 0
    assert b"null" in L.rsg_last_error()
    for n in ("L", "L2M", "Mxz", "Myz", "Mxy", "Rx", "Ry", "Rz", "tab", "Vp", "Vs", "Rho", "vpgrad", "vsgrad",
              "rhograd"):
        setattr(f, n, 64)
    f.src.data, f.src.map, f.src.shift = 64, 64, 64
    f.src.ntrace, f.src.nt, f.src.field = 1, 50, 5
    assert L.rsg_fwi_e3d_gradient(C.byref(g), C.byref(f), None) != 0
    assert b"source field" in L.rsg_last_error()
    f.src.field = R.PRESSURE
    for c, fld in enumerate((R.VX, R.VY, R.VZ)):
        f.obs[c].data, f.obs[c].map, f.obs[c].shift = 64, 64, 64
        f.obs[c].ntrace, f.obs[c].nt, f.obs[c].field = 4, 50, fld
        f.dmod[c], f.dres[c] = 64, 64
    f.obs[1].field = R.VX
    assert L.rsg_fwi_e3d_gradient(C.byref(g), C.byref(f), None) != 0
    assert b"VX, VY, VZ" in L.rsg_last_error()
    f.obs[1].field, f.obs[2].ntrace = R.VY, 3
    assert L.rsg_fwi_e3d_gradient(C.byref(g), C.byref(f), None) != 0
    assert b"equal count" in L.rsg_last_error()
    f.obs[2].ntrace = 4
    g.nx, g.ny, g.nz, g.lpml, g.domdec = 20, 20, 20, 4, 1
    assert L.rsg_fwi_e3d_gradient(C.byref(g), C.byref(f), None) != 0
    assert b"domain" in L.rsg_last_error()
    g.domdec = 0
    f.nx, f.ny, f.nz = 12, 12, 11
    assert L.rsg_fwi_e3d_gradient(C.byref(g), C.byref(f), None) != 0
    assert b"unpadded" in L.rsg_last_error()
    f.nz, f.fp16 = 12, 7
    assert L.rsg_fwi_e3d_gradient(C.byref(g), C.byref(f), None) != 0
    assert b"storage" in L.rsg_last_error()
    f.fp16, f.misfit_type = 0, 9
    assert L.rsg_fwi_e3d_gradient(C.byref(g), C.byref(f), None) != 0
    assert b"misfit type" in L.rsg_last_error()
    for fn in (L.rsg_e3d_backward_velocity, L.rsg_e3d_backward_stress):
        assert fn(None, None, None) != 0
    assert L.rsg_e3d_backward_velocity_ws(None, None, None, None) != 0
