"""Pins of the CPU oracle (no GPU).

Each oracle function is checked against something that fixes it
independently: scipy's LAPACK dgtsv and CubicSpline (the same routine and
boundary condition the reference uses), closed forms, brute-force numpy
sums, and the reference's own known-answer tests restated.
"""

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.linalg import lapack

import oracle
from helpers import dense_basis_monomials, dense_eval_monomial, oracle_coefficients, oracle_waveform


def test_dgtsv_bitwise_vs_lapack():
    """orc_dgtsv restates LAPACK dgtsv: identical INFO and bit-identical solutions,
    including systems that need row interchanges."""
    rng = np.random.default_rng(0)
    swaps = 0
    for _ in range(300):
        n = int(rng.integers(2, 200))
        dl, d, du, b = rng.normal(size=n - 1), rng.normal(size=n), rng.normal(size=n - 1), rng.normal(size=n)
        x, info = oracle.dgtsv(dl, d, du, b)
        _, _, _, xs, infos = lapack.dgtsv(dl.copy(), d.copy(), du.copy(), b.copy())
        assert info == infos
        assert np.array_equal(x, xs)
        swaps += int(np.any(np.abs(d[:-1]) < np.abs(dl)))
    assert swaps > 50  # the pivoting branch is exercised


def test_dgtsv_singular_info():
    dl = np.array([0.0, 1.0])
    d = np.array([0.0, 1.0, 1.0])
    du = np.array([1.0, 1.0])
    b = np.ones(3)
    _, info = oracle.dgtsv(dl, d, du, b)
    *_, infos = lapack.dgtsv(dl.copy(), d.copy(), du.copy(), b.copy())
    assert info == infos == 1


def test_three_knots_singular():
    """With 3 knots both not-a-knot end rows describe the same constraint and the
    reference's system is singular (scipy special-cases n = 3 as a parabola);
    LAPACK reports INFO = 3 and the reference ignores it (interpolate.cu:235).
    The oracle keeps that behaviour (and the GPU kernel reproduces it bit for bit)."""
    t = np.array([0.0, 1.0, 3.0])
    y = np.array([[1.0, 2.0, 0.5]])
    _, _, info = oracle.interpolate_arrays(t, y)
    assert info == 3


@pytest.mark.parametrize("L", [4, 5, 17, 120, 200])
def test_spline_vs_scipy_not_a_knot(L):
    rng = np.random.default_rng(L)
    t = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, L - 2)), [1.0]]) * 3.2e7
    y = rng.normal(size=(7, L))
    interp, _, info = oracle.interpolate_arrays(t, y)
    assert info == 0
    cs = CubicSpline(t, y, axis=-1, bc_type="not-a-knot")
    for p in range(4):
        np.testing.assert_allclose(interp[p][:, :-1], cs.c[3 - p].T, rtol=1e-10, atol=1e-12 * np.abs(cs.c[3 - p]).max())
    # final-knot re-expansion (interpolate.cu:242-280): derivatives of the last piece at t[-1]
    x = t[-1] - t[-2]
    np.testing.assert_allclose(interp[1][:, -1], cs(t[-1], nu=1), rtol=1e-9)
    np.testing.assert_allclose(interp[2][:, -1], cs(t[-1], nu=2) / 2, rtol=1e-9)
    np.testing.assert_allclose(interp[3][:, -1], interp[3][:, -2], rtol=0)
    assert x > 0


def test_spline_reproduces_cubics():
    """Not-a-knot splines reproduce any cubic exactly (closed form)."""
    rng = np.random.default_rng(1)
    t = np.sort(rng.uniform(0, 10, 30))
    a = rng.normal(size=4)
    y = a[0] + a[1] * t + a[2] * t ** 2 + a[3] * t ** 3
    interp, _, _ = oracle.interpolate_arrays(t, y)
    te = np.linspace(t[0], t[-1], 500)
    got = oracle.spline_eval(np.atleast_2d(t), interp, te)
    np.testing.assert_allclose(got, a[0] + a[1] * te + a[2] * te ** 2 + a[3] * te ** 3, rtol=1e-11, atol=1e-11)


def test_cubic_spline_known_answer():
    """tests/test_mode_summation.py:23-52 (test_cubic_spline), restated on the oracle."""
    t_test = np.linspace(0, 100, 1000)
    f_test = np.vstack((np.cos(t_test), np.sin(t_test)))
    interp, _, _ = oracle.interpolate_arrays(t_test, f_test)
    tk = np.tile(t_test, (2, 1))
    sp = CubicSpline(t_test, f_test, axis=-1)
    np.testing.assert_allclose(oracle.spline_eval(tk, interp, t_test), f_test, rtol=1e-10)
    for d in range(1, 4):
        np.testing.assert_allclose(oracle.spline_eval(tk, interp, t_test, d), sp(t_test, nu=d), rtol=1e-10)
    te = np.linspace(0, 100, 1852)
    for d in range(4):
        np.testing.assert_allclose(oracle.spline_eval(tk, interp, te, d), sp(te, nu=d), rtol=1e-10)


def test_spline_eval_edges():
    """Edge fill outside the range; derivatives outside raise (interpolatedmodesum.py:231-255)."""
    t = np.linspace(0, 10, 11)
    y = np.atleast_2d(np.sin(t))
    interp, _, _ = oracle.interpolate_arrays(t, y)
    out = oracle.spline_eval(np.atleast_2d(t), interp, np.array([-1.0, 11.0]))
    np.testing.assert_array_equal(out, [y[0, 0], y[0, -1]])
    with pytest.raises(ValueError):
        oracle.spline_eval(np.atleast_2d(t), interp, np.array([11.0]), deriv_order=1)
    with pytest.raises(ValueError):
        oracle.spline_eval(np.atleast_2d(t), interp, np.array([1.0]), deriv_order=4)


def test_find_start_inds_known_answers():
    dt = 10.0
    # knots 0, 25, 50 (on the grid), 95; out_len 10 -> T = 90 -> knot 95 >= T
    h_t = np.array([0.0, 25.0, 50.0, 95.0])
    st, used = oracle.find_start_inds(h_t, dt, 10)
    assert used == 4
    np.testing.assert_array_equal(st, [0, 3, 5, 10])  # 50 exactly on a sample -> later interval
    # a middle knot past (N-1)*dt: the rule stops there
    h_t = np.array([0.0, 25.0, 95.0, 120.0, 150.0])
    st, used = oracle.find_start_inds(h_t, dt, 10)
    assert used == 3
    np.testing.assert_array_equal(st, [0, 3, 10])
    # empty interval (two knots inside one sample gap)
    h_t = np.array([0.0, 21.0, 22.0, 100.0])
    st, used = oracle.find_start_inds(h_t, dt, 10)
    np.testing.assert_array_equal(st, [0, 3, 3, 10])


def test_dense_output_identity():
    """cubic piece -> 8-coefficient form: the nested form of interpolate.cu:642 with
    r5..r8 = 0 equals the cubic at random s (polynomial identity)."""
    rng = np.random.default_rng(4)
    t = np.sort(rng.uniform(0, 1e7, 40))
    y = rng.normal(size=(3, 40)) * 1e5
    coeffs, _, _ = oracle.interpolate_arrays(t, y)
    dense = oracle.cubic_to_dense(coeffs, t)
    for j in [0, 7, 38]:
        h = t[j + 1] - t[j]
        s = rng.uniform(0, 1, 50)
        s1 = 1 - s
        for p in range(3):
            r = dense[:, j, p]
            P = r[0] + s * (r[1] + s1 * (r[2] + s * (r[3] + s1 * (r[4] + s * (r[5] + s1 * (r[6] + s * r[7]))))))
            x = s * h
            cub = coeffs[0, p, j] + coeffs[1, p, j] * x + coeffs[2, p, j] * x ** 2 + coeffs[3, p, j] * x ** 3
            np.testing.assert_allclose(P, cub, rtol=1e-12, atol=1e-9)
        assert np.all(dense[4:, j] == 0)


def _brute(inp):
    """Dense per-sample sum with every phase from the monomial expansion of the
    8-coefficient dense output (all eight terms)."""
    interp_tl, pcoef = oracle_coefficients(inp)
    N, K = inp.num_pts, inp.K
    tt = np.arange(N) * inp.dt
    j = np.clip(np.searchsorted(inp.t, tt, side="right") - 1, 0, inp.L - 2)
    x = tt - inp.t[j]
    A = sum((interp_tl[p][j, :K] + 1j * interp_tl[p][j, K:]) * x[:, None] ** p for p in range(4))
    s = (tt - inp.t[j]) / (inp.t[j + 1] - inp.t[j])
    ph = [dense_eval_monomial(pcoef[:, j, q], s) for q in range(3)]
    phase = inp.m[None] * ph[0][:, None] + inp.k[None] * ph[1][:, None] + inp.n[None] * ph[2][:, None]
    P = A * np.exp(-1j * phase)
    return (P * inp.ylms[:K]).sum(1) + ((inp.m != 0) * np.conj(P) * inp.ylms[K:]).sum(1)


@pytest.mark.parametrize("phases,modes", [("spline", "m_nonneg_30"), ("dopr", "m_nonneg_30"),
                                          ("dopr", "k1_20")])
def test_mode_sum_vs_brute_force(phases, modes):
    """DirectModeSum pattern (directmodesum.py:91-126) on dense spline values; with
    DOPR853-style phases (non-zero rcont5..8, Phi_theta != 0) and k != 0 modes."""
    from fastemriwaveforms_amd import synthetic
    for seed in (2, 9):
        inp = synthetic.make_input(0.001, 10.0, 10, modes, 15, seed, phases=phases)
        if phases == "dopr":
            assert np.all(np.abs(inp.dense[4:]) > 0) and np.any(inp.phase_knots[1] != 0)
        if modes == "k1_20":
            assert np.any(inp.k != 0)
        np.testing.assert_allclose(oracle_waveform(inp), _brute(inp), rtol=1e-10, atol=1e-13)


def test_dense_output_form_pinned():
    """The oracle's 8-coefficient phase evaluation (interpolate.cu:642) against the
    monomial expansion of its basis (polynomial products, no nested form), with
    random non-zero r1..r8 on all three phases; each phase reaches the sum
    through a mode of its own (m, k or n = 1), so a dropped, swapped or
    mis-indexed coefficient or phase shows as a phase error."""
    rng = np.random.default_rng(12)
    L, N, dt = 5, 100, 10.0
    t = np.array([0.0, 137.0, 400.0, 655.0, 1000.0])
    scale = np.array([30.0, 20.0, 5.0, 3.0, 2.0, 1.0, 0.5, 0.3])
    dense = np.ascontiguousarray(rng.normal(size=(8, L - 1, 3)) * scale[:, None, None])
    teuk = np.ones((L, 1), np.complex128)
    interp, _, _ = oracle.interpolate_arrays(t, np.concatenate([teuk.T.real, teuk.T.imag]))
    interp_tl = np.ascontiguousarray(np.transpose(interp, (0, 2, 1)))
    tt = np.arange(N) * dt
    j = np.searchsorted(t, tt, side="right") - 1
    s = (tt - t[j]) / (t[j + 1] - t[j])
    for q, (m, k, n) in enumerate([(1, 0, 0), (0, 1, 0), (0, 0, 1)]):
        h = oracle.get_waveform(interp_tl, t, dense, [m], [k], [n], L, N, 1,
                                np.array([1.0 + 0j, 0.0]), dt, t)
        P = dense_eval_monomial(dense[:, j, q], s)
        np.testing.assert_allclose(np.abs(h), 1.0, atol=1e-14)
        np.testing.assert_allclose(np.angle(h * np.exp(1j * P)), 0.0, atol=1e-12)
    # the expansion itself: the DOP853 end conditions (dopr853.py:631-642)
    from numpy.polynomial import polynomial as Pn
    r = rng.normal(size=8)
    mono = dense_basis_monomials().T @ r
    d1 = Pn.polyder(mono)
    np.testing.assert_allclose([Pn.polyval(0.0, mono), Pn.polyval(1.0, mono)],
                               [r[0], r[0] + r[1]], rtol=1e-14)
    np.testing.assert_allclose([Pn.polyval(0.0, d1), Pn.polyval(1.0, d1)],
                               [r[1] + r[2], r[1] - r[2] - r[3]], rtol=1e-13, atol=1e-14)


def test_synthetic_dense_output():
    """The DOPR853-style generator: the first four coefficients are the Hermite
    cubic (end values and slopes exact), the eight together follow the
    non-polynomial inspiral phase to 1e-11 of its value inside every interval
    of a 120-knot year, and rcont5..8 are far from zero."""
    from fastemriwaveforms_amd import synthetic
    rng = np.random.default_rng(0)
    Ts = synthetic.YRSID_SI
    t = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, 118)) * Ts, [Ts]])
    h = np.diff(t)
    for terms in synthetic.inspiral_phase_terms(0.87, 0.95):
        fn = lambda x: synthetic.inspiral_phase(x, Ts, terms)  # noqa: E731
        r = synthetic.dense_output(t, fn)
        y, dy = fn(t)
        np.testing.assert_allclose(r[0] + r[1], y[1:], rtol=1e-15)
        np.testing.assert_allclose(r[1] + r[2], h * dy[:-1], rtol=1e-12)
        np.testing.assert_allclose(r[1] - r[2] - r[3], h * dy[1:], rtol=1e-12)
        assert np.median(np.abs(r[4])) > 1e-6 and np.median(np.abs(r[7])) > 1e-12
        j = rng.integers(0, len(h), 400)
        sv = rng.uniform(0, 1, 400)
        P = dense_eval_monomial(r[:, j], sv)
        exact, _ = fn(t[j] + sv * h[j])
        np.testing.assert_allclose(P, exact, rtol=1e-11)


def test_mode_sum_closed_form_geometric():
    """Constant amplitude, linear phase, one mode: h_i = A Y exp(-i m w dt i),
    a geometric series in closed form; Y- = 0 disables the conjugate term and
    m = 0 drops it by rule."""
    L, N, dt = 6, 400, 10.0
    t = np.linspace(0, (N - 1) * dt + 7.0, L)
    w = 1.3e-3
    A = 0.7 - 0.2j
    for m, ym in [(3, 0.0), (0, 1.0 + 0.5j)]:
        teuk = np.full((L, 1), A)
        interp, _, _ = oracle.interpolate_arrays(t, np.concatenate([teuk.T.real, teuk.T.imag]))
        interp_tl = np.transpose(interp, (0, 2, 1)).copy()
        # Phi_phi = w t exactly representable as a cubic -> dense form
        pk = np.stack([w * t, np.zeros(L), np.zeros(L)])
        pc, _, _ = oracle.interpolate_arrays(t, pk)
        pcoef = oracle.cubic_to_dense(pc, t)
        ylms = np.array([1.0 + 0j, ym])
        h = oracle.get_waveform(interp_tl, t, pcoef, [m], [0], [0], L, N, 1, ylms, dt, t)
        i = np.arange(N)
        exact = A * np.exp(-1j * m * w * dt * i)
        np.testing.assert_allclose(h, exact, rtol=1e-9, atol=1e-12)
        # geometric series: sum_i h_i = A (1 - q^N) / (1 - q)
        if m != 0:
            q = np.exp(-1j * m * w * dt)
            np.testing.assert_allclose(h.sum(), A * (1 - q ** N) / (1 - q), rtol=1e-8)


def test_minus_m_convention():
    """ylms[K:] carries (-1)^l (interpolatedmodesum.py:356-358 == directmodesum.py:108-122):
    the conjugate term equals an explicit -m mode with amplitude (-1)^l conj(A)."""
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.small_input(seed=3, L=9, K=6, T_years=0.001)
    inp.m[:] = np.abs(inp.m) + 1
    h = oracle_waveform(inp)
    K = inp.K
    # explicit form: modes (m, n) with Y+ and (-m, -n) with conj amplitude and Y-
    inp2 = synthetic.small_input(seed=3, L=9, K=6, T_years=0.001)
    inp2.m[:] = inp.m
    teuk2 = np.concatenate([inp2.teuk, np.conj(inp2.teuk)], axis=1)
    yl = np.concatenate([inp2.ylms[:K], inp2.ylms[K:], np.zeros(2 * K)])
    inp2.teuk = teuk2
    inp2.ylms = yl
    inp2.m = np.concatenate([inp.m, -inp.m]).astype(np.int32)
    inp2.k = np.zeros(2 * K, np.int32)
    inp2.n = np.concatenate([inp.n, -inp.n]).astype(np.int32)
    inp2.l = np.concatenate([inp.l, inp.l])
    h2 = oracle_waveform(inp2)
    np.testing.assert_allclose(h, h2, rtol=1e-9, atol=1e-12)


def test_sampling_invariance():
    """tests/test_physics.py:26-63: sample i is evaluated at t = i*dt on an absolute
    grid, so dt vs dt/2 [::2] and T vs 2T prefixes agree."""
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.small_input(seed=8, L=11, K=10, T_years=0.002, dt=10.0)
    h10 = oracle_waveform(inp)
    inp5 = synthetic.small_input(seed=8, L=11, K=10, T_years=0.002, dt=5.0)
    h5 = oracle_waveform(inp5)
    n = min(len(h10), len(h5[::2]))
    assert oracle.mismatch(h10[:n], h5[::2][:n]) < 1e-14
    np.testing.assert_array_equal(h10[:n], h5[::2][:n])


def test_num_points():
    """summation/base.py:83-102"""
    Y = oracle.YRSID_SI
    assert oracle.num_points(0.0, Y, 1.0, 10.0) == (int(Y / 10.0) + 1, 0)
    # trajectory shorter than T: samples stop at the last knot; padding on request
    n, pad = oracle.num_points(0.0, 1000.0, 1.0, 10.0, pad_output=True)
    assert n == 101 and pad == int(int(Y / 10.0) * 10.0 / 10.0) + 1 - 101
    n, pad = oracle.num_points(0.0, 1000.0, 1.0, 10.0, odd_len=True)
    assert (n + pad) % 2 == 1
