"""The exact-adjoint oracle (oracle/adjoint.py, oracle/rsoracle_impl.h
orc_deriv_t / orc_a2d_adj_*) pinned by what the mathematics fixes: the
transposed derivative against the forward one (<D u, v> = <u, D^T v>, every
order, direction and axis, edge stencils included); the transposed model
staggering against finite differences of the staggering itself; the
adjoint dot-product test of the whole time loop (<F w, r> = <w, F^T r>,
wavelet to records); and the gradient against central finite differences of
the FP64 misfit at 16 random model points, with and without a free surface.
A plausible mistake (a dropped PML term, a wrong edge order, a missed
free-surface factor) fails the last two."""
import ctypes as C

import numpy as np
import pytest

import oracle
from oracle.adjoint import a2d_misfit, a2d_model_grad, fwi_a2d_exact
from oracle.runs import stagger_a2d
from tests.exact_util import exact_case, observed, oracle_args


@pytest.mark.parametrize("axis", [0, 2])
@pytest.mark.parametrize("fw", [0, 1])
def test_deriv_t_is_transpose(axis, fw):
    rng = np.random.default_rng(axis * 10 + fw)
    nz, nx = 23, 29
    for order in range(1, 9):
        u = rng.standard_normal((nz, nx))
        v = rng.standard_normal((nz, nx))
        du = np.zeros_like(u)
        oracle.fn("deriv", np.float64)(u.reshape(-1), du.reshape(-1), nx, 1, nz, axis, fw, order, 7.0)
        dtv = np.zeros_like(v)
        oracle.fn("deriv_t", np.float64)(v.reshape(-1), dtv.reshape(-1), nx, 1, nz, axis, fw, order, 7.0)
        a, b = np.sum(du * v), np.sum(u * dtv)
        assert abs(a - b) <= 1e-13 * (abs(a) + np.abs(du).sum()), (order, a, b)


@pytest.mark.parametrize("fs", [False, True])
def test_model_grad_is_staggering_transpose(fs):
    """J = <gL, L> + <gRx, Rx> + <gRz, Rz> through orc_stagger_a2d (FP64):
    a2d_model_grad against central differences at 12 random model cells
    (edge cells, whose padding replicates them, included)."""
    rng = np.random.default_rng(5 + fs)
    nz, nx, lpml = 9, 11, 3
    vp = 2000 + 500 * rng.random((nz, nx))
    rho = 1800 + 400 * rng.random((nz, nx))
    shp = (nz + 2 * lpml, nx + 2 * lpml)
    gL, gRx, gRz = (rng.standard_normal(shp) for _ in range(3))
    gRx *= 1e10
    gRz *= 1e10
    gv, gr = a2d_model_grad(gL, gRx, gRz, vp, rho, lpml, fs)

    def J(v, r):
        L, Rx, Rz = stagger_a2d(v, r, lpml, fs, np.float64)
        return np.sum(gL * L) + np.sum(gRx * Rx) + np.sum(gRz * Rz)
    pts = [(0, 0), (nz - 1, nx - 1), (0, nx - 1)] + [tuple(rng.integers(0, [nz, nx])) for _ in range(9)]
    for iz, ix in pts:
        for name, arr, g in (("vp", vp, gv), ("rho", rho, gr)):
            h = 1e-4 * arr[iz, ix]
            a, b = arr.copy(), arr.copy()
            a[iz, ix] += h
            b[iz, ix] -= h
            fd = (J(a, rho) - J(b, rho)) / (2 * h) if name == "vp" else (J(vp, a) - J(vp, b)) / (2 * h)
            assert abs(fd - g[iz, ix]) <= 1e-6 * np.abs(g).max(), (name, iz, ix, fd, g[iz, ix])


@pytest.mark.parametrize("order,fs", [(4, False), (8, True), (2, True)])
def test_exact_dot_product(order, fs):
    c = exact_case(seed=order, order=order, fs=fs, nt=160)
    dobs = observed(c)
    r = fwi_a2d_exact(*oracle_args(c, dobs))
    lhs = float(np.sum(r["dmod"] * r["dres"]))
    rhs = float(np.sum(c["wav"] * r["wavgrad"]))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("order,fs", [(4, False), (8, True)])
def test_exact_gradient_vs_finite_differences(order, fs):
    c = exact_case(seed=10 + order, order=order, fs=fs, nt=160)
    dobs = observed(c)
    r = fwi_a2d_exact(*oracle_args(c, dobs))
    nz, nx = c["vp"].shape
    rng = np.random.default_rng(99)
    pts = [tuple(rng.integers(0, [nz, nx])) for _ in range(16)]
    args = list(oracle_args(c, dobs))
    for k, (iz, ix) in enumerate(pts):
        name = ("vp", "rho")[k % 2]
        idx = 0 if name == "vp" else 1
        g = r["vpgrad" if name == "vp" else "rhograd"]
        arr = args[idx]
        h = 1e-4 * arr[iz, ix]
        a, b = arr.copy(), arr.copy()
        a[iz, ix] += h
        b[iz, ix] -= h
        args[idx] = a
        jp = a2d_misfit(*args)
        args[idx] = b
        jm = a2d_misfit(*args)
        args[idx] = arr
        fd = (jp - jm) / (2 * h)
        assert abs(fd - g[iz, ix]) <= 1e-6 * np.abs(g).max(), (name, iz, ix, fd, g[iz, ix])


def test_exact_differs_from_reference_adjoint():
    """The reference's adjoint (forward operator + crossCorr) is not the
    gradient of the discrete misfit: the two gradients differ well beyond
    rounding, so the exact mode is a separate quantity."""
    from oracle.fwi import fwi_a2d
    c = exact_case(seed=3, order=4, nt=160)
    dobs = observed(c)
    r = fwi_a2d_exact(*oracle_args(c, dobs))
    q = fwi_a2d(c["vp"], c["rho"], c["lpml"], c["dx"], c["dx"], c["dt"], c["nt"], c["order"], c["fs"], 8,
                c["scoords"], c["wav"], c["gcoords"], dobs, real=np.float64)
    e = np.linalg.norm(r["vpgrad"] - q["vpgrad"]) / np.linalg.norm(r["vpgrad"])
    assert e > 1e-3
