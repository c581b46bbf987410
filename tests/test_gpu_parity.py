"""GPU parity of the HIP spline build and the plain FP64 mode sum against the oracle.

Bars: spline coefficients bit-identical to the oracle (same recurrence, same
rounding); FP64 waveforms within rel-L2 1e-12 of the oracle (the only
difference allowed is the last bit of sin/cos and FMA in the amplitude /
complex products -- the phase argument is bit-identical).
"""

import numpy as np
import pytest

import oracle
from helpers import oracle_coefficients, oracle_waveform, spread_indices

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _problem(inp):
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    return ModeSumProblem(inp)


def test_interpolate_arrays_dropin_bitwise(gpu):
    """few_interpolate_arrays through the reference-shaped wrapper, numpy in/out."""
    from fastemriwaveforms_amd import pyinterp
    rng = np.random.default_rng(3)
    for ninterps, length in [(1, 3), (5, 4), (70, 120), (129, 37)]:
        t = np.sort(rng.uniform(0, 1e6, (ninterps, length)), axis=1)
        y = rng.normal(size=(ninterps, length))
        interp = np.zeros((4, ninterps, length))
        interp[0] = y
        interp = interp.flatten()
        B, ud, dg, ld = (np.zeros(ninterps * length) for _ in range(4))
        pyinterp.interpolate_arrays_wrap(t.flatten(), interp, ninterps, length, B, ud, dg, ld)
        ref, _, info = oracle.interpolate_arrays(t, y)
        assert info == (3 if length == 3 else 0)  # 3 knots: singular, LAPACK's early return
        got = interp.reshape(4, ninterps, length)
        assert np.array_equal(got, ref), (ninterps, length, np.max(np.abs(got - ref)))


def test_cubic_spline_native(gpu):
    """Restates tests/test_mode_summation.py:23-52 on the native class."""
    from scipy.interpolate import CubicSpline
    from fastemriwaveforms_amd.summation import CubicSplineInterpolant
    t_test = np.linspace(0, 100, 1000)
    f_test = np.vstack((np.cos(t_test), np.sin(t_test)))
    few_spl = CubicSplineInterpolant(t_test, f_test)
    scipy_spl = CubicSpline(t_test, f_test, axis=-1)
    np.testing.assert_allclose(few_spl(t_test).cpu().numpy(), f_test, rtol=1e-10)
    for d in range(1, 4):
        np.testing.assert_allclose(few_spl(t_test, deriv_order=d).cpu().numpy(),
                                   scipy_spl(t_test, nu=d), rtol=1e-10)
    t_eval = np.linspace(0, 100, 1852)
    for d in range(4):
        np.testing.assert_allclose(few_spl(t_eval, deriv_order=d).cpu().numpy(),
                                   scipy_spl(t_eval, nu=d), rtol=1e-10)


def test_prepare_modes_bitwise(gpu):
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.make_workload("cpu0")
    p = _problem(inp)
    p.build_splines()
    interp_tl, pcoef = oracle_coefficients(inp)
    assert np.array_equal(p.interp.cpu().numpy(), interp_tl)
    assert np.array_equal(p.pcoef.cpu().numpy(), pcoef)


def test_direct_fp64_cpu0_full(gpu):
    """Workload cpu0 (0.1 yr, 500 modes, 100 knots): every sample against the oracle."""
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.make_workload("cpu0")
    h = _problem(inp).run("fp64").cpu().numpy()
    ref = oracle_waveform(inp)
    err = oracle.rel_l2(h, ref)
    print(f"cpu0 fp64 rel-L2 {err:.3e}")
    assert err <= FP64_TOL


def test_direct_fp64_full1_samples(gpu):
    """Workload full1 (1 yr, 3843 modes, 120 knots): 2000+ spread samples."""
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.make_workload("full1")
    h = _problem(inp).run("fp64").cpu().numpy()
    idx = spread_indices(inp.num_pts, count=2000)
    ref = oracle_waveform(inp, idx)
    err = oracle.rel_l2(h[idx], ref)
    print(f"full1 fp64 rel-L2 {err:.3e} over {len(idx)} samples")
    assert err <= FP64_TOL


def test_dropin_get_waveform_numpy(gpu):
    """The reference-shaped wrappers driven exactly as InterpolatedModeSum.sum drives them
    (interpolatedmodesum.py:326-375), with numpy arrays (Feature.NUMPY drop-in)."""
    from fastemriwaveforms_amd import pyinterp, synthetic
    inp = synthetic.small_input(seed=11, L=15, K=40, T_years=0.003)
    L, K = inp.L, inp.K
    y_all = np.concatenate([inp.teuk.T.real, inp.teuk.T.imag])
    interp = np.zeros((4, 2 * K, L))
    interp[0] = y_all
    interp = interp.flatten()
    tt = np.tile(inp.t, (2 * K, 1)).flatten()
    B, ud, dg, ld = (np.zeros(2 * K * L) for _ in range(4))
    pyinterp.interpolate_arrays_wrap(tt, interp, 2 * K, L, B, ud, dg, ld)
    interp_tl = np.transpose(interp.reshape(4, 2 * K, L), [0, 2, 1]).flatten()
    _, pcoef = oracle_coefficients(inp)
    wave = np.zeros(inp.num_pts, dtype=np.complex128)
    pyinterp.get_waveform_wrap(wave, interp_tl, inp.t, pcoef.flatten(), inp.m, inp.k, inp.n, L,
                               inp.num_pts, K, inp.ylms, inp.dt, inp.t, 0)
    ref = oracle_waveform(inp)
    assert oracle.rel_l2(wave, ref) <= FP64_TOL


@pytest.mark.parametrize("precision", ["mixed16", "mixed", "fact64"])
def test_dropin_get_waveform_numpy_mixed(gpu, precision):
    """The reference-shaped wrappers driven exactly as InterpolatedModeSum.sum drives
    them (interpolatedmodesum.py:326-375), numpy arrays, workload cpu0, with the
    backend's precision set to a factorised sum: rel-L2 < 1e-6 against the oracle
    (1e-8 for the FP64 factorised sum); the mode plan is built once and reused by
    a second call."""
    from fastemriwaveforms_amd import pyinterp, synthetic
    from fastemriwaveforms_amd.backend import RocmBackend
    inp = synthetic.make_workload("cpu0")
    L, K = inp.L, inp.K
    y_all = np.concatenate([inp.teuk.T.real, inp.teuk.T.imag])
    interp = np.zeros((4, 2 * K, L))
    interp[0] = y_all
    interp = interp.flatten()
    tt = np.tile(inp.t, (2 * K, 1)).flatten()
    B, ud, dg, ld = (np.zeros(2 * K * L) for _ in range(4))
    pyinterp.interpolate_arrays_wrap(tt, interp, 2 * K, L, B, ud, dg, ld)
    interp_tl = np.transpose(interp.reshape(4, 2 * K, L), [0, 2, 1]).flatten()
    _, pcoef = oracle_coefficients(inp)
    ref = oracle_waveform(inp)
    backend = RocmBackend()
    before = backend.precision
    backend.precision = precision
    try:
        pyinterp._plans.clear()
        for _ in range(2):
            wave = np.zeros(inp.num_pts, dtype=np.complex128)
            pyinterp.get_waveform_wrap(wave, interp_tl, inp.t, pcoef.flatten(), inp.m, inp.k,
                                       inp.n, L, inp.num_pts, K, inp.ylms, inp.dt, inp.t, 0)
            err = oracle.rel_l2(wave, ref)
            print(f"drop-in {precision}: rel-L2 vs oracle {err:.3e}")
            assert err < (1e-8 if precision == "fact64" else 1e-6)
        assert len(pyinterp._plans) == 1
    finally:
        backend.precision = before


def test_native_interpolated_mode_sum(gpu):
    """Native InterpolatedModeSum.__call__ (summation/base.py:66 + interpolatedmodesum.py:283)."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.summation import InterpolatedModeSum
    inp = synthetic.small_input(seed=5, L=20, K=60, T_years=0.004)
    _, pcoef = oracle_coefficients(inp)
    dopr = np.transpose(pcoef, (1, 2, 0))  # (L-1, 3, 8): the trajectory's layout
    ylms = inp.ylms.copy()
    ylms[inp.K:] *= (-1.0) ** inp.l  # sum() applies (-1)^l again (interpolatedmodesum.py:356-358)
    summ = InterpolatedModeSum()
    h = summ(inp.t, inp.teuk, ylms, inp.t, dopr, inp.l, inp.m, inp.k, inp.n,
             T=inp.T_years, dt=inp.dt).cpu().numpy()
    assert summ.num_pts == inp.num_pts
    ref = oracle_waveform(inp)
    assert oracle.rel_l2(h, ref) <= FP64_TOL


@pytest.mark.parametrize("precision", ["mixed", "mixed16"])
def test_native_interpolated_mode_sum_mixed(gpu, precision):
    """The native class in the factorised precisions: same call, rel-L2 < 1e-6,
    and the cached mode plan is reused on a second call."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.summation import InterpolatedModeSum
    inp = synthetic.small_input(seed=6, L=20, K=300, T_years=0.004)
    _, pcoef = oracle_coefficients(inp)
    dopr = np.transpose(pcoef, (1, 2, 0))
    ylms = inp.ylms.copy()
    ylms[inp.K:] *= (-1.0) ** inp.l
    summ = InterpolatedModeSum(precision=precision)
    ref = oracle_waveform(inp)
    for _ in range(2):
        h = summ(inp.t, inp.teuk, ylms, inp.t, dopr, inp.l, inp.m, inp.k, inp.n,
                 T=inp.T_years, dt=inp.dt).cpu().numpy()
        assert oracle.rel_l2(h, ref) < 1e-6
    assert len(summ._plans) == 1


def test_brute_force_small(gpu):
    """Direct numpy sum over modes (directmodesum.py:91-126 pattern) on dense
    spline evaluations: pins the whole FP64 pipeline on a tiny case."""
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.small_input(seed=2, L=10, K=12, T_years=0.001)
    h = _problem(inp).run("fp64").cpu().numpy()
    interp_tl, pcoef = oracle_coefficients(inp)
    N = inp.num_pts
    tt = np.arange(N) * inp.dt
    j = np.clip(np.searchsorted(inp.t, tt, side="right") - 1, 0, inp.L - 2)
    x = tt - inp.t[j]
    K = inp.K
    A = np.zeros((N, K), dtype=np.complex128)
    for p in range(4):
        A += (interp_tl[p][j, :K] + 1j * interp_tl[p][j, K:]) * x[:, None] ** p
    s = (tt - inp.t[j]) / (inp.t[j + 1] - inp.t[j])
    ph = []
    for q in range(3):
        r = pcoef[:, j, q]
        ph.append(r[0] + s * (r[1] + (1 - s) * (r[2] + s * r[3])))
    phase = inp.m[None] * ph[0][:, None] + inp.k[None] * ph[1][:, None] + inp.n[None] * ph[2][:, None]
    P = A * np.exp(-1j * phase)
    yp, ym = inp.ylms[:K], inp.ylms[K:]
    ref = (P * yp).sum(1) + ((inp.m != 0) * np.conj(P) * ym).sum(1)
    assert oracle.rel_l2(h, ref) < 1e-10


@pytest.mark.parametrize("case", ["many_modes", "knots_on_grid", "knot_past_end", "accumulate"])
def test_direct_fp64_edges(gpu, case):
    import torch
    from fastemriwaveforms_amd import synthetic
    if case == "many_modes":
        # > 5000 modes: exercises the 5000-mode flush of the CPU chunking
        inp = synthetic.make_input(0.0015, 10.0, 8, "m_all_55", 5500, 17, name=case)
    else:
        inp = synthetic.small_input(seed=23, L=14, K=30, T_years=0.002)
    if case == "knots_on_grid":
        inp.t[1:-1] = np.round(inp.t[1:-1] / inp.dt) * inp.dt  # samples exactly on knots
    if case == "knot_past_end":
        inp.num_pts = inp.num_pts - 1000  # a middle knot lies past (N-1)*dt: find_start_inds breaks
    p = _problem(inp)
    if case == "accumulate":
        init = (torch.arange(inp.num_pts, dtype=torch.float64, device=p.wave.device) * 1e-3).to(torch.complex128)
        p.wave.copy_(init)
        p.build_splines()
        p.mode_sum("fp64")
        h = (p.wave - init).cpu().numpy()
    else:
        h = p.run("fp64").cpu().numpy()
    ref = oracle_waveform(inp)
    err = oracle.rel_l2(h, ref)
    assert err <= FP64_TOL, (case, err)


def test_spline_eval_kernel_bitwise(gpu):
    """CubicSplineInterpolant.__call__ (one HIP launch) equals the oracle's restatement of
    interpolatedmodesum.py:161-257 bit for bit: values and derivatives, knots hit
    exactly, the last knot, shared and per-array tnew; edge fill + ValueError outside."""
    import torch
    from fastemriwaveforms_amd.summation import CubicSplineInterpolant
    rng = np.random.default_rng(12)
    L, n = 37, 9
    t = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, L - 2)), [1.0]]) * 1e5
    y = rng.normal(size=(n, L))
    spl = CubicSplineInterpolant(t, y)
    interp_ref, _, info = oracle.interpolate_arrays(t, y)
    assert info == 0
    tk = np.tile(t, (n, 1))
    tnew = np.concatenate([rng.uniform(0, 1e5, 500), t, [t[-1]]])
    for d in range(4):
        got = spl(tnew, deriv_order=d).cpu().numpy()
        ref = oracle.spline_eval(tk, interp_ref, tnew, d)
        assert np.array_equal(got, ref), d
    t2 = rng.uniform(0, 1e5, (n, 64))
    assert np.array_equal(spl(t2).cpu().numpy(), oracle.spline_eval(tk, interp_ref, t2))
    out = spl(np.array([-5.0, 2e5])).cpu().numpy()
    assert np.array_equal(out[:, 0], y[:, 0]) and np.array_equal(out[:, 1], y[:, -1])
    with pytest.raises(ValueError):
        spl(np.array([2e5]), deriv_order=1)
    with pytest.raises(ValueError):
        spl(np.array([1.0]), deriv_order=4)
    # knots per array (2-D t)
    t2k = np.sort(rng.uniform(0, 1, (n, L)), axis=1) * 1e3
    spl2 = CubicSplineInterpolant(t2k, y)
    interp2, _, info = oracle.interpolate_arrays(t2k, y)
    assert info == 0
    te = rng.uniform(t2k[:, :1].max(), t2k[:, -1:].min(), 100)
    assert np.array_equal(spl2(te, 1).cpu().numpy(), oracle.spline_eval(t2k, interp2, te, 1))


def test_direct_mode_sum(gpu):
    """DirectModeSum (directmodesum.py:70-126) on dense inputs against a numpy restatement."""
    from fastemriwaveforms_amd.summation import DirectModeSum
    rng = np.random.default_rng(4)
    N, K = 3000, 70
    l = rng.integers(2, 11, K).astype(np.int32)
    m = np.array([rng.integers(0, li + 1) for li in l], np.int32)
    k = rng.integers(-2, 3, K).astype(np.int32)
    n = rng.integers(-30, 31, K).astype(np.int32)
    teuk = (rng.normal(size=(N, K)) + 1j * rng.normal(size=(N, K))) * 0.1
    ylms = np.exp(1j * rng.uniform(0, 2 * np.pi, 2 * K))
    t = np.arange(N) * 10.0
    Pp, Pt, Pr = (np.cumsum(rng.uniform(0.001, 0.02, N)) * 1e3 for _ in range(3))
    dm = DirectModeSum()
    h = dm(t, teuk, ylms, t, (Pp, Pt, Pr), l, m, k, n, T=N * 10.0 / oracle.YRSID_SI, dt=10.0)
    h = h.cpu().numpy()
    phase = m[None] * Pp[:, None] + k[None] * Pt[:, None] + n[None] * Pr[:, None]
    w1 = np.sum(ylms[None, :K] * teuk * np.exp(-1j * phase), axis=1)
    sel = m > 0
    w2 = np.sum((-1.0) ** l[None, sel] * ylms[None, K:][:, sel] * np.conj(teuk[:, sel])
                * np.exp(1j * phase[:, sel]), axis=1)
    ref = w1 + w2
    assert len(h) == N
    assert oracle.rel_l2(h, ref) < 1e-13


@pytest.mark.parametrize("length", [3, 4, 5, 9, 33, 120, 150, 151])
def test_spline_shared_knots_bitwise(gpu, length):
    """The shared-knot block path of the spline build (t_sys = 0: the matrix
    eliminated once per block, 32 systems per block, coefficients spread over
    four waves) at small, odd and limit lengths (150 is the last LDS length,
    151 takes the per-lane path) and system counts that leave partial blocks:
    bit for bit the oracle's interpolate_arrays on the tiled knots.  3 knots
    is singular (LAPACK's early return; the block falls back per system)."""
    import torch
    from fastemriwaveforms_amd._lib import lib
    from fastemriwaveforms_amd.pipeline import _stream
    rng = np.random.default_rng(length)
    for nsys in (1, 31, 33, 70):
        h = rng.uniform(0.2, 1.0, length - 1)
        h[rng.random(length - 1) < 0.2] *= 40.0  # uneven spacing: row interchanges
        t = np.concatenate([[0.0], np.cumsum(h)]) * 1e4
        y = rng.normal(size=(nsys, length))
        dev = torch.device("cuda", 0)
        t_d = torch.from_numpy(t).to(dev)
        y_d = torch.from_numpy(y).to(dev)
        out = torch.zeros((4, nsys, length), dtype=torch.float64, device=dev)
        scratch = torch.zeros(4 * length * nsys, dtype=torch.float64, device=dev)
        rc = lib().few_spline_build_async(y_d.data_ptr(), length, 1, t_d.data_ptr(), 0, 1,
                                          out.data_ptr(), nsys * length, length, 1, nsys, length,
                                          0, scratch.data_ptr(), _stream())
        assert rc == 0, lib().few_last_error()
        torch.cuda.synchronize()
        ref, _, info = oracle.interpolate_arrays(np.tile(t, (nsys, 1)), y)
        assert info == (3 if length == 3 else 0)
        got = out.cpu().numpy()
        assert np.array_equal(got, ref), (nsys, length)


@pytest.mark.parametrize("precision", ["fp64", "mixed16"])
def test_native_pad_output_odd_len(gpu, precision):
    """SummationBase.__call__'s output sizing (summation/base.py:83-112): with
    pad_output the waveform runs to T (zeros after the trajectory), odd_len makes
    the length odd; the summed samples equal the oracle's."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.summation import InterpolatedModeSum
    inp = synthetic.small_input(seed=15, L=16, K=80, T_years=0.003)
    _, pcoef = oracle_coefficients(inp)
    dopr = np.transpose(pcoef, (1, 2, 0))
    ylms = inp.ylms.copy()
    ylms[inp.K:] *= (-1.0) ** inp.l
    ref = oracle_waveform(inp)
    T = 1.7 * inp.T_years  # observation longer than the trajectory
    for pad, odd in [(True, False), (True, True), (False, True)]:
        summ = InterpolatedModeSum(precision=precision, pad_output=pad, odd_len=odd)
        h = summ(inp.t, inp.teuk, ylms, inp.t, dopr, inp.l, inp.m, inp.k, inp.n, T=T,
                 dt=inp.dt).cpu().numpy()
        npts, npad = oracle.num_points(inp.t[0], inp.t[-1], T, inp.dt, pad_output=pad,
                                       odd_len=odd)
        assert npts == inp.num_pts and summ.num_pts == inp.num_pts
        assert len(h) == npts + npad
        if odd:
            assert len(h) % 2 == 1
        if pad:
            assert len(h) > inp.num_pts + 1000
        err = oracle.rel_l2(h[:inp.num_pts], ref)
        assert err <= (FP64_TOL if precision == "fp64" else 1e-6), (pad, odd, err)
        assert np.count_nonzero(h[inp.num_pts:]) == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed16"])
def test_native_integrate_backwards(gpu, precision):
    """integrate_backwards (interpolatedmodesum.py:341-345): amplitude and phase knots
    shifted by t[-1] - int(t[-1] / dt) dt before the sum, against the oracle on the
    same shifted knots."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.summation import InterpolatedModeSum
    inp = synthetic.small_input(seed=17, L=14, K=60, T_years=0.003)
    interp_tl, pcoef = oracle_coefficients(inp)
    dopr = np.transpose(pcoef, (1, 2, 0))
    ylms = inp.ylms.copy()
    ylms[inp.K:] *= (-1.0) ** inp.l
    offset = inp.t[-1] - int(inp.t[-1] / inp.dt) * inp.dt
    assert offset > 0
    ts = inp.t - offset
    ref = oracle.get_waveform(interp_tl, ts, pcoef, inp.m, inp.k, inp.n, inp.L, inp.num_pts, inp.K,
                              inp.ylms, inp.dt, ts)
    summ = InterpolatedModeSum(precision=precision)
    h = summ(inp.t, inp.teuk, ylms, inp.t, dopr, inp.l, inp.m, inp.k, inp.n, T=inp.T_years,
             dt=inp.dt, integrate_backwards=True).cpu().numpy()
    err = oracle.rel_l2(h, ref)
    print(f"integrate_backwards {precision}: rel-L2 {err:.3e}")
    assert err <= (FP64_TOL if precision == "fp64" else 1e-6)
    # the unshifted sum differs: the shift matters
    plain = oracle_waveform(inp)
    assert oracle.rel_l2(plain, ref) > 1e-6
