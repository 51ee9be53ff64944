"""The acoustic 3D exact-adjoint oracle (oracle/adjoint.py fwi_a3d_exact,
oracle/rsoracle_impl.h orc_a3d_adj_*) pinned by what the mathematics fixes:
the transposed 3D model staggering against finite differences of the
staggering itself, the adjoint dot-product test of the whole time loop
(wavelet to records) and the Vp / rho gradient against central finite
differences of the FP64 misfit, with and without a free surface.  A dropped
strip term on one of the three axes, a wrong y stride or a missed
free-surface factor fails one of them."""
import numpy as np
import pytest

from oracle.adjoint import a3d_misfit, a3d_model_grad, fwi_a3d_exact
from oracle.runs import stagger_a3d
from tests.exact_util import exact_case_a3d, observed_a3d, oracle_args_a3d


@pytest.mark.parametrize("fs", [False, True])
def test_a3d_model_grad_is_staggering_transpose(fs):
    rng = np.random.default_rng(31 + fs)
    nz, ny, nx, lpml = 5, 6, 7, 2
    vp = 2000 + 500 * rng.random((nz, ny, nx))
    rho = 1800 + 400 * rng.random((nz, ny, nx))
    shp = (nz + 2 * lpml, ny + 2 * lpml, nx + 2 * lpml)
    gs = [rng.standard_normal(shp) for _ in range(4)]
    for g in gs[1:]:
        g *= 1e10
    gv, gr = a3d_model_grad(*gs, vp, rho, lpml, fs)

    def J(v, r):
        return sum(np.sum(g * a) for g, a in zip(gs, stagger_a3d(v, r, lpml, fs, np.float64)))
    pts = [(0, 0, 0), (nz - 1, ny - 1, nx - 1), (0, ny - 1, 0)] + [tuple(rng.integers(0, [nz, ny, nx]))
                                                                 for _ in range(9)]
    for p in pts:
        for k, g in enumerate((gv, gr)):
            m = [vp, rho]
            h = 1e-4 * m[k][p]
            a, b = m[k].copy(), m[k].copy()
            a[p] += h
            b[p] -= h
            ma, mb = list(m), list(m)
            ma[k], mb[k] = a, b
            fd = (J(*ma) - J(*mb)) / (2 * h)
            assert abs(fd - g[p]) <= 1e-6 * np.abs(g).max(), (k, p, fd, g[p])


@pytest.mark.parametrize("order,fs", [(4, False), (8, True), (2, True)])
def test_a3d_exact_dot_product(order, fs):
    c = exact_case_a3d(seed=order, order=order, fs=fs)
    dobs = observed_a3d(c)
    r = fwi_a3d_exact(*oracle_args_a3d(c, dobs))
    lhs = float(np.sum(r["dmod"] * r["dres"]))
    rhs = float(np.sum(c["wav"] * r["wavgrad"]))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("order,fs", [(4, False), (8, True)])
def test_a3d_exact_gradient_vs_finite_differences(order, fs):
    c = exact_case_a3d(seed=10 + order, order=order, fs=fs, nt=90)
    dobs = observed_a3d(c)
    r = fwi_a3d_exact(*oracle_args_a3d(c, dobs))
    rng = np.random.default_rng(5)
    args = list(oracle_args_a3d(c, dobs))
    pts = [tuple(rng.integers(0, c["vp"].shape)) for _ in range(15)] + [(0, 3, 4)]
    for k, p in enumerate(pts):
        idx = k % 2
        g = r[("vpgrad", "rhograd")[idx]]
        arr = args[idx]
        h = 1e-4 * arr[p]
        a, b = arr.copy(), arr.copy()
        a[p] += h
        b[p] -= h
        args[idx] = a
        jp = a3d_misfit(*args)
        args[idx] = b
        jm = a3d_misfit(*args)
        args[idx] = arr
        fd = (jp - jm) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * np.abs(g).max(), (idx, p, fd, g[p])
