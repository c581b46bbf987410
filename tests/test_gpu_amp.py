"""GPU parity of the amplitude interpolation (csrc/amp2d.hip, few_interp2d_async,
few_interp2D_wrap) against the oracle: bit-identical (same fpbspl / fpbisp
operations in the same order, contraction off, no library calls)."""

import numpy as np
import pytest

import oracle
from test_oracle_amp import _case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,kx,ky,sets,npts", [(1, 3, 3, 5, 400), (2, 3, 3, 700, 130),
                                                  (4, 5, 4, 3, 70), (3, 1, 2, 9, 1)])
def test_interp2d_wrap_bitwise(gpu, seed, kx, ky, sets, npts):
    """The reference-shaped backend method on numpy arrays (pyampinterp2D.pyx:15-29)."""
    from fastemriwaveforms_amd import pyinterp
    tx, ty, c, x, y = _case(seed, kx=kx, ky=ky, sets=sets, npts=max(npts, 30))
    x, y = x[:npts], y[:npts]
    ref = oracle.interp2d(tx, ty, c, x, y, kx, ky)
    z = np.zeros(sets * npts)
    pyinterp.interp2D(z, tx, len(tx), ty, len(ty), c.flatten(), kx, ky, x, npts, y, npts, sets,
                      c.shape[1])
    assert np.array_equal(z.reshape(sets, npts), ref)


def test_interp2d_errors(gpu):
    from fastemriwaveforms_amd import pyinterp
    from fastemriwaveforms_amd.exceptions import FewRocmError
    tx, ty, c, x, y = _case(1)
    z = np.zeros(c.shape[0] * len(x))
    with pytest.raises(ValueError):
        pyinterp.interp2D(z, tx, len(tx), ty, len(ty), c.flatten(), 3, 3, x, len(x), y,
                          len(y) - 1, c.shape[0], c.shape[1])
    with pytest.raises(FewRocmError):   # fewer coefficients per set than the knots imply
        pyinterp.interp2D(z, tx, len(tx), ty, len(ty), c.flatten(), 3, 3, x, len(x), y, len(y),
                          c.shape[0], c.shape[1] - 1)


def test_amp_interp2d_class(gpu):
    """AmpInterp2D(w_knots, u_knots, coefficients, l, m, k, n)(w, u) against the
    oracle, all modes and a subset (ampinterp2d.py:96-190)."""
    from fastemriwaveforms_amd.amplitude import AmpInterp2D
    rng = np.random.default_rng(7)
    tx, ty, _, x, y = _case(2, npts=257)
    nc = (len(tx) - 4) * (len(ty) - 4)
    nmodes = 37
    coeff = rng.normal(size=(nmodes, 2, nc))
    amp = AmpInterp2D(tx, ty, coeff, *(np.arange(nmodes),) * 4)
    got = amp(x, y).cpu().numpy()
    ref = oracle.interp2d(tx, ty, coeff.reshape(2 * nmodes, nc), x, y).reshape(nmodes, 2, -1)
    assert got.shape == (len(x), nmodes)
    assert np.array_equal(got, (ref[:, 0] + 1j * ref[:, 1]).T)
    sub = np.array([3, 0, 36, 11])
    got_sub = amp(x, y, mode_indexes=sub).cpu().numpy()
    assert np.array_equal(got_sub, got[:, sub])
