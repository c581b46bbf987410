"""Pins of the AAK oracle (oracle/few_oracle_aak.c, the CPU branch of
gpuAAK.cu's get_waveform) and of the kernel's per-sample code against it
(host twin).

* the Numerical Recipes Bessel functions against scipy's jv, bounded per
  order by the routines' own accuracy (Miller's recurrence started at
  n + sqrt(40 n) is not fully converged for the lowest orders);
* the rotation coefficients against vector identities computed with numpy
  (cos = (L.S - (n.L)(n.S)) / (|n x L||n x S|), sin = 2 n.(S x L) / ...);
* the segment plan against known answers (gpuAAK.cu:458-483);
* the eccentric harmonics' Bessel brackets against the Peters-Mathews power
  per harmonic (scipy's jv) and summed over harmonics (closed form f(e));
* the harmonic sum in the circular limit (e = 0): only the n = 2 harmonic is
  non-zero, so 2 harmonics and 12 give the same waveform; the amplitude
  scales as mu / dist and as M^(2/3).
"""

import ctypes

import numpy as np
import pytest
import scipy.special as sp

import oracle
from helpers import aak_input, oracle_aak

P = ctypes.c_void_p


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


BESSEL_BOUND = {0: 1e-8, 1: 1e-8, 2: 3e-7, 3: 3e-7}


@pytest.mark.parametrize("n", list(range(0, 34)))
def test_bessel_vs_scipy(n):
    """Orders and arguments the AAK sum uses: J_k(n e) with e < 1, k <= n + 2."""
    xs = (np.linspace(1e-3, max(n, 1) * 0.999, 300) if n >= 2
          else np.concatenate([np.linspace(1e-3, 8 * 0.999, 200), np.linspace(8.0, 30.0, 100)]))
    err = max(abs(oracle.bessel_jn(n, x) - sp.jv(n, x)) for x in xs)
    bound = BESSEL_BOUND.get(n, 1e-8 if n < 10 else 1e-11)
    assert err < bound, (n, err)


def test_bessel_special_values():
    assert oracle.bessel_jn(5, 0.0) == 0.0
    assert abs(oracle.bessel_jn(-1, 0.3) + sp.jv(1, 0.3)) < 1e-8   # gpuAAK's J0 at n = 1
    assert abs(oracle.bessel_jn(3, -0.7) + oracle.bessel_jn(3, 0.7)) == 0.0   # odd order, odd


@pytest.mark.parametrize("e", [0.05, 0.3, 0.5, 0.7])
def test_harmonics_peters_mathews_power(e):
    """The eccentric harmonic brackets the waveform is built from carry the
    Peters-Mathews power: with alpha = n X, beta = n sqrt(1-e^2) Y, gamma = 2 J_n,
    g(n, e) = n^2/32 (alpha^2 + beta^2 + gamma^2/3) (Peters & Mathews 1963), each
    g(n, e) equals the closed form in scipy's Bessel functions, and the sum over
    harmonics is f(e) = (1 + 73/24 e^2 + 37/96 e^4) / (1 - e^2)^(7/2)."""
    total = 0.0
    for n in range(1, 121):
        X, Y, Jn = oracle.aak_harmonic(n, e)
        g = n * n / 32.0 * ((n * X) ** 2 + n * n * (1 - e * e) * Y ** 2 + (2 * Jn) ** 2 / 3.0)
        J = lambda k: sp.jv(k, n * e)
        g_ref = n ** 4 / 32.0 * ((J(n - 2) - 2 * e * J(n - 1) + 2.0 / n * J(n) + 2 * e * J(n + 1)
                                  - J(n + 2)) ** 2
                                 + (1 - e * e) * (J(n - 2) - 2 * J(n) + J(n + 2)) ** 2
                                 + 4.0 / (3 * n * n) * J(n) ** 2)
        # the routines' J_2, J_3 carry up to 3e-7 (Miller's start is not
        # converged for the lowest orders), which n^2 X^2 turns into ~1e-6
        assert abs(g - g_ref) <= 5e-6 * max(g_ref, 1e-12), (n, g, g_ref)
        total += g
    f = (1 + 73 / 24 * e ** 2 + 37 / 96 * e ** 4) / (1 - e * e) ** 3.5
    assert total == pytest.approx(f, rel=2e-6)


def test_harmonics_circular_limit():
    """e = 0: only n = 2 radiates; its brackets are X = Y = J0(0) = 1 (to the
    rational fit's 3e-9 at 0) and J_2(0) = 0, so g(2, 0) = 1 = f(0)."""
    X, Y, Jn = oracle.aak_harmonic(2, 0.0)
    assert X == Y and abs(X - 1.0) < 5e-9 and Jn == 0.0
    for n in (1, 3, 4, 7):
        assert oracle.aak_harmonic(n, 0.0) == (0.0, 0.0, 0.0)


@pytest.mark.parametrize("seed", range(6))
def test_rot_coeff_identities(seed):
    rng = np.random.default_rng(seed)
    iota, tS, pS, tK, pK, alp = rng.uniform(0.1, 3.0, 6)
    rot, n, L, S, nxL, nxS = oracle.aak_rot_coeff(iota, tS, pS, tK, pK, alp)
    assert np.allclose([np.linalg.norm(v) for v in (n, L, S)], 1.0, atol=1e-14)
    norm = np.linalg.norm(np.cross(n, L)) * np.linalg.norm(np.cross(n, S))
    cosrot = (L @ S - (n @ L) * (n @ S)) / norm
    sinrot = 2 * (n @ np.cross(S, L)) / norm
    assert rot[0] == pytest.approx(2 * cosrot ** 2 - 1, abs=1e-13)
    assert rot[1] == pytest.approx(cosrot * sinrot, abs=1e-13)
    assert rot[2] == -rot[1] and rot[3] == rot[0]


def test_start_inds_known_answers():
    dt = 10.0
    h_t = np.array([0.0, 25.0, 30.0, 99.0, 1000.0])
    st, npts = oracle.aak_start_inds(h_t, dt, 11)   # T = 100 s: knot 4 reaches the end
    assert npts == 5 and list(st) == [0, 3, 3, 10, 11]
    st, npts = oracle.aak_start_inds(h_t, dt, 8)    # T = 70: knot 3 (99 s) reaches it
    assert npts == 4 and list(st) == [0, 3, 3, 8]
    st, npts = oracle.aak_start_inds(np.array([0.0, 5.0, 7.0]), dt, 100)   # never reached
    assert npts == 3 and list(st) == [0, 1, 1]


def test_circular_limit_single_harmonic():
    t, c, prm, n_out = aak_input(T_years=0.01, L=12, e0=0.0, e1=0.0)
    h2 = oracle_aak(t, c, prm, n_out, nmodes=2)
    h12 = oracle_aak(t, c, prm, n_out, nmodes=12)
    assert np.abs(h2).max() > 0
    assert np.array_equal(h2, h12)
    h1 = oracle_aak(t, c, prm, n_out, nmodes=1)
    assert np.abs(h1).max() < 1e-12 * np.abs(h2).max()


def test_amplitude_scaling():
    t, c, prm, n_out = aak_input(T_years=0.01, L=12)
    h = oracle_aak(t, c, prm, n_out, nmodes=8)
    h_far = oracle_aak(t, c, dict(prm, dist=4.0, mu=20.0), n_out, nmodes=8)
    np.testing.assert_allclose(h_far, h * (20.0 / 10.0) / 4.0, rtol=1e-12, atol=0)
    h_heavy = oracle_aak(t, c, dict(prm, M=8e6), n_out, nmodes=8)
    np.testing.assert_allclose(h_heavy, h * 4.0, rtol=1e-12, atol=0)   # (8)^(2/3) = 4


def test_dense_rate_matches_finite_difference():
    """Omega from the dense-output derivative formula (gpuAAK.cu:330-332) is the
    time derivative of the dense phase: checked through the amplitude, which is
    (|Omega_phi| M)^(2/3): a constant-e, S = 0, e = 0 waveform's envelope follows it."""
    t, c, prm, n_out = aak_input(T_years=0.01, L=12, e0=0.0, e1=0.0)
    j = 5
    w = t[j + 1] - t[j]
    r = c[j, 3]                                    # Phi_phi dense coefficients
    def dense(s):
        s1 = 1 - s
        return r[0] + s * (r[1] + s1 * (r[2] + s * (r[3] + s1 * (r[4] + s * (r[5] + s1 * (r[6] + s * r[7]))))))
    s = 0.37
    s2, s3, s4, s5, s6 = s ** 2, s ** 3, s ** 4, s ** 5, s ** 6
    rate = (r[1] + r[2] * (1. - 2. * s) + r[3] * (2. * s - 3. * s2) + r[4] * (2. * s - 6. * s2 + 4. * s3)
            + r[5] * (3. * s2 - 8. * s3 + 5. * s4) + r[6] * (3. * s2 - 12. * s3 + 15 * s4 - 6. * s5)
            + r[7] * (4. * s3 - 15. * s4 + 18. * s5 - 7. * s6)) / w
    hstep = 1e-5
    fd = (dense(s + hstep) - dense(s - hstep)) / (2 * hstep) / w
    assert rate == pytest.approx(fd, rel=1e-7)


def test_twin_bessel_bitwise(twin):
    """The kernel's Bessel routines (nr_jn and the lockstep nr_jn5), run on the
    host, equal the oracle's bit for bit over the orders and arguments the sum
    uses (both recurrence branches, orders 0 and 1, x = 0)."""
    twin.twin_bessel.argtypes = [P, P, ctypes.c_long, P, P]
    rng = np.random.default_rng(7)
    orders, xs = [], []
    for n in range(1, 60):
        for x in np.concatenate([[0.0], rng.uniform(0, 0.99, 6) * n, [n * 0.999, n + 0.5]]):
            orders.append(n - 2)
            xs.append(x)
    orders = np.array(orders, dtype=np.int32)
    xs = np.array(xs)
    out1 = np.zeros(len(xs))
    out5 = np.zeros((len(xs), 5))
    twin.twin_bessel(_p(orders), _p(xs), len(xs), _p(out1), _p(out5))
    for i, (o, x) in enumerate(zip(orders, xs)):
        ref = [oracle.bessel_jn(o + j, x) for j in range(5)]
        if o >= 0:
            assert out1[i] == ref[0], (o, x)
        assert list(out5[i][1:]) == ref[1:], (o, x)
        if o >= 0:
            assert out5[i][0] == ref[0], (o, x)


def test_twin_aak_bitwise(twin):
    """The kernel's per-sample routine, run on the host over the oracle's plan,
    equals the oracle bit for bit (host libm on both sides)."""
    twin.twin_aak_waveform.argtypes = [P, P] + [ctypes.c_double] * 8 + [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, P]
    for mich in (False, True):
        t, c, prm, n_out = aak_input(T_years=0.005, L=10, seed=4)
        ref = oracle_aak(t, c, prm, n_out, nmodes=15, mich=mich)
        dense = np.ascontiguousarray(np.transpose(c, (2, 0, 1)))
        got = np.zeros(n_out, dtype=np.complex128)
        assert twin.twin_aak_waveform(_p(got), _p(dense), prm["M"], prm["S"], prm["mu"], prm["qS"],
                                      prm["phiS"], prm["qK"], prm["phiK"], prm["dist"], 15,
                                      int(mich), len(t), n_out, 10.0, _p(t)) == 0
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref), mich
