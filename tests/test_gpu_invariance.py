"""Physical invariances of the GPU mode sum (tests/test_physics.py:26-93,212-233).

The reference checks them on whole waveforms through the mismatch; restated for
the spline-plus-sum path they are:

* sampling: sample i sits at t = i dt on an absolute grid, so dt = 10 and
  dt = 5 [::2] give the same samples (test_physics.py:26-51);
* duration: a longer trajectory with the same leading knots gives the same
  leading samples, up to the not-a-knot end condition, whose influence decays
  geometrically along the knots (test_physics.py:53-63);
* time rescaling: knots and dt doubled with the phases kept is the same
  waveform (the mode-sum part of test_mass_invariance, test_physics.py:65-93);
  doubling is exact in binary floating point, so every spline coefficient
  scales exactly and the FP64 samples agree bit for bit;
* amplitude scaling: the sum is linear in the amplitudes, as the 1/distance
  scaling relies on (test_physics.py:212-233); tiny physical amplitudes
  (1e-22) must not underflow the FP16 tiles of the mixed path.

FP64 must meet the reference's mismatch < 1e-14; the mixed paths the rel-L2
< 1e-6 bar of the FP64 oracle comparison.
"""

import dataclasses

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

PRECS = ["fp64", "mixed", "mixed16"]
MIXED_TOL = 1e-6


def _run(inp, precision):
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    return ModeSumProblem(inp).run(precision).cpu().numpy()


def _check(prec, a, b, what):
    mm = oracle.mismatch(a, b)
    err = oracle.rel_l2(a, b)
    print(f"{what} {prec}: mismatch {mm:.3e} rel-L2 {err:.3e} "
          f"bitwise {np.array_equal(a, b)}")
    if prec == "fp64":
        assert mm < 1e-14 and err < 1e-12
    else:
        assert err < MIXED_TOL


@pytest.mark.parametrize("prec", PRECS)
def test_sampling_rate(gpu, prec):
    from fastemriwaveforms_amd import synthetic
    i10 = synthetic.make_input(0.01, 10.0, 40, "m_nonneg_30", 300, 4101)
    i5 = synthetic.make_input(0.01, 5.0, 40, "m_nonneg_30", 300, 4101)
    np.testing.assert_array_equal(i10.t, i5.t)
    h10, h5 = _run(i10, prec), _run(i5, prec)
    n = min(len(h10), len(h5[::2]))
    assert n > 30000
    _check(prec, h10[:n], h5[::2][:n], "dt 10 vs 5[::2]")
    if prec == "fp64":
        np.testing.assert_array_equal(h10[:n], h5[::2][:n])


@pytest.mark.parametrize("prec", PRECS)
def test_duration_prefix(gpu, prec):
    from fastemriwaveforms_amd import synthetic
    long = synthetic.make_input(0.02, 10.0, 81, "m_nonneg_30", 300, 4102)
    cut = 41
    short = dataclasses.replace(
        long, t=long.t[:cut].copy(), teuk=np.ascontiguousarray(long.teuk[:cut]),
        phase_knots=np.ascontiguousarray(long.phase_knots[:, :cut]),
        freqs=np.ascontiguousarray(long.freqs[:, :cut]),
        num_pts=int(long.t[cut - 1] / long.dt) + 1)
    hl, hs = _run(long, prec), _run(short, prec)
    # samples more than 20 knots before the short trajectory's end
    n = int(long.t[cut - 21] / long.dt)
    assert n > 1000
    _check(prec, hs[:n], hl[:n], "T vs 2T prefix")


@pytest.mark.parametrize("prec", PRECS)
def test_time_rescaling(gpu, prec):
    from fastemriwaveforms_amd import synthetic
    a = synthetic.make_input(0.01, 10.0, 60, "m_nonneg_30", 400, 4103)
    b = dataclasses.replace(a, t=2.0 * a.t, dt=2.0 * a.dt, freqs=0.5 * a.freqs)
    ha, hb = _run(a, prec), _run(b, prec)
    assert len(ha) == len(hb)
    _check(prec, ha, hb, "t, dt doubled")
    if prec == "fp64":
        np.testing.assert_array_equal(ha, hb)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("factor", [0.37, 3.7e-23])
def test_amplitude_scaling(gpu, prec, factor):
    from fastemriwaveforms_amd import synthetic
    a = synthetic.make_input(0.01, 10.0, 60, "m_nonneg_30", 400, 4104)
    b = dataclasses.replace(a, teuk=np.ascontiguousarray(a.teuk * factor))
    ha, hb = _run(a, prec), _run(b, prec)
    assert np.all(np.isfinite(hb))
    _check(prec, ha * factor, hb, f"amplitudes x {factor}")
