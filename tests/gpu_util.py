"""Helpers for the GPU parity tests: build the rockseis_amd objects of a
golden case and compare with a relative L2 tolerance."""
import numpy as np

from rockseis_amd import core

RTOL_L2 = 1e-5  # receiver gathers and snapshots vs the reference fp32 path


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def data2d(coords, nt, dt, field, wav=None):
    coords = np.asarray(coords, np.float32).reshape(-1, 2)
    d = core.Data2D(coords.shape[0], nt, dt, 0.0, field)
    d.scoords[:] = coords
    d.gcoords[:] = coords
    if wav is not None:
        d.data[:] = np.asarray(wav, np.float32).reshape(d.data.shape)
    return d


def data3d(coords, nt, dt, field, wav=None):
    coords = np.asarray(coords, np.float32).reshape(-1, 3)
    d = core.Data3D(coords.shape[0], nt, dt, 0.0, field)
    d.scoords[:] = coords
    d.gcoords[:] = coords
    if wav is not None:
        d.data[:] = np.asarray(wav, np.float32).reshape(d.data.shape)
    return d
