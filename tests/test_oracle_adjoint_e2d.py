"""The elastic exact-adjoint oracle (oracle/adjoint.py fwi_e2d_exact,
oracle/rsoracle_impl.h orc_e2d_adj_*) pinned by what the mathematics fixes:
the transposed elastic model staggering against finite differences of the
staggering itself; the adjoint dot-product test of the whole time loop
(<F w, r> = <w, F^T r>, wavelet to the P, Ux and Uz records) for explosive
and force sources; and the Vp / Vs / rho gradient against central finite
differences of the FP64 misfit at 16 random model points, with and without a
free surface.  A dropped strip term, a wrong modulus pairing (L against
L2M), a missed free-surface row or a wrong staggering order fails one of
them."""
import numpy as np
import pytest

from oracle.adjoint import e2d_misfit, e2d_model_grad, fwi_e2d_exact
from oracle.runs import stagger_e2d
from tests.exact_util import exact_case_e2d, observed_e2d, oracle_args_e2d


def _lame_to_v(lam, mu, rho):
    return np.sqrt((lam + 2 * mu) / rho), np.sqrt(mu / rho), rho


@pytest.mark.parametrize("fs,lame", [(False, False), (True, False), (False, True), (True, True)])
def test_e2d_model_grad_is_staggering_transpose(fs, lame):
    """With lame the parameters are (lambda, mu, rho), mapped to (vp, vs,
    rho) before the staggering."""
    rng = np.random.default_rng(21 + fs)
    nz, nx, lpml = 9, 11, 3
    vp = 2500 + 500 * rng.random((nz, nx))
    vs = 1200 + 300 * rng.random((nz, nx))
    rho = 1800 + 400 * rng.random((nz, nx))
    shp = (nz + 2 * lpml, nx + 2 * lpml)
    gs = [rng.standard_normal(shp) for _ in range(5)]
    gs[3] *= 1e10
    gs[4] *= 1e10
    gv, gw, gr = e2d_model_grad(*gs, vp, vs, rho, lpml, fs, lame=lame)

    def J(a, b, r):
        v, w, r = _lame_to_v(a, b, r) if lame else (a, b, r)
        return sum(np.sum(g * a) for g, a in zip(gs, stagger_e2d(v, w, r, lpml, fs, np.float64)))
    pts = [(0, 0), (nz - 1, nx - 1), (0, nx - 1)] + [tuple(rng.integers(0, [nz, nx])) for _ in range(9)]
    base = [rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2, rho] if lame else [vp, vs, rho]
    for iz, ix in pts:
        for k, g in enumerate((gv, gw, gr)):
            m = list(base)
            h = 1e-4 * m[k][iz, ix]
            a, b = m[k].copy(), m[k].copy()
            a[iz, ix] += h
            b[iz, ix] -= h
            ma, mb = list(m), list(m)
            ma[k], mb[k] = a, b
            fd = (J(*ma) - J(*mb)) / (2 * h)
            assert abs(fd - g[iz, ix]) <= 1e-6 * np.abs(g).max(), (k, iz, ix, fd, g[iz, ix])


@pytest.mark.parametrize("order,fs,src", [(4, False, "P"), (8, True, "P"), (2, True, "VX"), (4, False, "VZ")])
def test_e2d_exact_dot_product(order, fs, src):
    c = exact_case_e2d(seed=order, order=order, fs=fs, srcfield=src)
    dobs = observed_e2d(c)
    r = fwi_e2d_exact(*oracle_args_e2d(c, dobs))
    lhs = sum(float(np.sum(r["dmod"][f] * r["dres"][f])) for f in ("P", "VX", "VZ"))
    rhs = float(np.sum(c["wav"] * r["wavgrad"]))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("order,fs,src", [(4, False, "P"), (8, True, "P"), (4, True, "VZ")])
def test_e2d_exact_gradient_vs_finite_differences(order, fs, src):
    c = exact_case_e2d(seed=30 + order, order=order, fs=fs, srcfield=src)
    dobs = observed_e2d(c)
    r = fwi_e2d_exact(*oracle_args_e2d(c, dobs))
    nz, nx = c["vp"].shape
    rng = np.random.default_rng(77)
    pts = [tuple(rng.integers(0, [nz, nx])) for _ in range(15)] + [(0, int(rng.integers(0, nx)))]
    args = list(oracle_args_e2d(c, dobs))
    for k, (iz, ix) in enumerate(pts):
        idx = k % 3
        g = r[("vpgrad", "vsgrad", "rhograd")[idx]]
        arr = args[idx]
        h = 1e-4 * arr[iz, ix]
        a, b = arr.copy(), arr.copy()
        a[iz, ix] += h
        b[iz, ix] -= h
        args[idx] = a
        jp = e2d_misfit(*args)
        args[idx] = b
        jm = e2d_misfit(*args)
        args[idx] = arr
        fd = (jp - jm) / (2 * h)
        assert abs(fd - g[iz, ix]) <= 1e-6 * np.abs(g).max(), (idx, iz, ix, fd, g[iz, ix])
