"""Pins of the amplitude-interpolation oracle (oracle/few_oracle_amp.c, the
CPU branch of interp2D, src/few/cutils/AmpInterp2D.cu:1-208) and of the
kernel's per-point basis / per-set sum against it (host twin).

FITPACK's bispev (scipy.interpolate.bisplev) evaluates the same
tensor-product B-spline with the same routines (fpbspl, fpbisp): the oracle
must agree with it to the last bit or nearly, on random knots, coefficients
and points, including points outside the knot range (clamped) and on knots.
"""

import ctypes

import numpy as np
import pytest
from scipy.interpolate import bisplev

import oracle

P = ctypes.c_void_p


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _case(seed, nxi=9, nyi=12, kx=3, ky=3, sets=5, npts=400):
    rng = np.random.default_rng(seed)
    ix = np.sort(rng.uniform(0.0, 1.0, nxi))
    iy = np.sort(rng.uniform(-1.0, 2.0, nyi))
    tx = np.concatenate([[0.0] * (kx + 1), ix, [1.0] * (kx + 1)])
    ty = np.concatenate([[-1.0] * (ky + 1), iy, [2.0] * (ky + 1)])
    ncoef = (len(tx) - kx - 1) * (len(ty) - ky - 1)
    c = rng.normal(size=(sets, ncoef))
    x = np.concatenate([rng.uniform(-0.1, 1.1, npts - len(ix)), ix])
    y = np.concatenate([rng.uniform(-1.2, 2.2, npts - len(iy)), iy])[:len(x)]
    return tx, ty, c, x, y


@pytest.mark.parametrize("seed,kx,ky", [(1, 3, 3), (2, 3, 3), (3, 1, 2), (4, 5, 4), (5, 2, 1)])
def test_interp2d_vs_fitpack(seed, kx, ky):
    tx, ty, c, x, y = _case(seed, kx=kx, ky=ky)
    z = oracle.interp2d(tx, ty, c, x, y, kx, ky)
    xc = np.clip(x, tx[kx], tx[-kx - 1])
    yc = np.clip(y, ty[ky], ty[-ky - 1])
    for s in range(c.shape[0]):
        ref = np.array([bisplev(a, b, (tx, ty, c[s], kx, ky)) for a, b in zip(xc, yc)])
        assert np.max(np.abs(z[s] - ref)) <= 4e-16 * max(1.0, np.abs(ref).max()), s


def test_interp2d_reproduces_bicubic_polynomial():
    """A tensor-product spline fitted exactly to a bicubic polynomial returns it."""
    from scipy.interpolate import RectBivariateSpline
    gx = np.linspace(0, 1, 15)
    gy = np.linspace(-1, 1, 11)
    f = lambda a, b: 1 + 2 * a - a ** 2 * b + 0.5 * a ** 3 * b ** 3 - b ** 2
    spl = RectBivariateSpline(gx, gy, f(gx[:, None], gy[None, :]), kx=3, ky=3, s=0)
    tx, ty, cc = spl.tck
    rng = np.random.default_rng(0)
    x, y = rng.uniform(0, 1, 300), rng.uniform(-1, 1, 300)
    z = oracle.interp2d(tx, ty, cc, x, y)
    assert np.max(np.abs(z[0] - f(x, y))) < 1e-12


def test_interp2d_mismatched_lengths_raise():
    tx, ty, c, x, y = _case(1)
    with pytest.raises(ValueError):
        oracle.interp2d(tx, ty, c, x, y[:-1])


def test_twin_interp2d_bitwise(twin):
    twin.twin_interp2d.argtypes = [P, P, ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int,
                                   ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int, ctypes.c_int]
    for seed, kx, ky in [(1, 3, 3), (4, 5, 4), (3, 1, 2)]:
        tx, ty, c, x, y = _case(seed, kx=kx, ky=ky)
        ref = oracle.interp2d(tx, ty, c, x, y, kx, ky)
        z = np.zeros_like(ref)
        assert twin.twin_interp2d(_p(z), _p(tx), len(tx), _p(ty), len(ty), _p(c), kx, ky, _p(x),
                                  len(x), _p(y), c.shape[0], c.shape[1]) == 0
        assert np.array_equal(z, ref)
