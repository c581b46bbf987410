"""The HIP kernels' scalar code, compiled for the host, against the oracle (no GPU).

tests/host_twin.hip includes the product's kernel sources and runs their
host-callable routines (the fused spline recurrence, the dense-output phase,
the per-mode FP64 term, the interval planning rule) on the CPU.
"""

import ctypes

import numpy as np
import pytest

import oracle
from helpers import oracle_coefficients, oracle_waveform, spread_indices

P = ctypes.c_void_p


def _p(a):
    return a.ctypes.data


@pytest.fixture(scope="module")
def tw(twin):
    twin.twin_spline_build.argtypes = [P, ctypes.c_long, ctypes.c_long, P, ctypes.c_long,
                                       ctypes.c_long, P, ctypes.c_long, ctypes.c_long,
                                       ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P]
    twin.twin_prepare_modes.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, P, P, P]
    twin.twin_plan.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_int, P, P]
    twin.twin_direct_samples.argtypes = [P, P, ctypes.c_long, P, P, P, P, P, P, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, P, ctypes.c_double, P]
    twin.twin_spline_adjoint.argtypes = [P, P, ctypes.c_int, ctypes.c_int, P]
    return twin


@pytest.mark.parametrize("nsys,length", [(3, 4), (2, 5), (4, 9), (5, 13), (7, 120), (3, 200)])
def test_twin_spline_adjoint_bitwise(tw, nsys, length):
    """The kernels' transposed spline build equals the oracle's bit for bit
    (same operation order, contraction off), including pivoting rows."""
    rng = np.random.default_rng(100 + length)
    t = np.concatenate([[0.0], np.sort(rng.uniform(0, 1, length - 2)), [1.0]]) * 3e7
    G = np.ascontiguousarray(rng.normal(size=(4, nsys, length)))
    gy_ref, info = oracle.spline_adjoint(t, G)
    assert info == 0
    gy = np.zeros((nsys, length))
    assert tw.twin_spline_adjoint(_p(t), _p(G), nsys, length, _p(gy)) == 0
    assert np.array_equal(gy, gy_ref)


@pytest.mark.parametrize("ninterps,length", [(1, 3), (3, 4), (2, 9), (2, 10), (2, 11), (2, 12), (5, 13), (33, 120), (7, 200)])
def test_twin_spline_reference_layout_bitwise(tw, ninterps, length):
    rng = np.random.default_rng(length)
    t = np.sort(rng.uniform(0, 1e7, (ninterps, length)), axis=1)
    y = rng.normal(size=(ninterps, length))
    interp = np.zeros((4, ninterps, length))
    interp[0] = y
    scratch = np.zeros(4 * ninterps * length)
    info = np.zeros(1, np.int32)
    tw.twin_spline_build(_p(interp), length, 1, _p(t), length, 1, _p(interp), ninterps * length,
                         length, 1, ninterps, length, 0, _p(scratch), _p(info))
    ref, _, oinfo = oracle.interpolate_arrays(t, y)
    # 3 knots: both not-a-knot end rows coincide and the system is singular;
    # LAPACK returns INFO = 3 (ignored by the reference) before back substitution
    assert oinfo == (3 if length == 3 else 0)
    assert info[0] == (1 if length == 3 else 0)
    assert np.array_equal(interp, ref)


def test_twin_spline_pivoting_bitwise(tw):
    """Knot spacings with large ratios force dgtsv row interchanges."""
    rng = np.random.default_rng(5)
    L = 60
    gaps = np.exp(rng.uniform(-6, 6, L - 1))
    t = np.concatenate([[0.0], np.cumsum(gaps)])
    y = rng.normal(size=(4, L))
    interp = np.zeros((4, 4, L))
    interp[0] = y
    scratch = np.zeros(4 * 4 * L)
    tw.twin_spline_build(_p(interp), L, 1, _p(t), 0, 1, _p(interp), 4 * L, L, 1, 4, L, 0,
                         _p(scratch), None)
    ref, _, _ = oracle.interpolate_arrays(t, y)
    assert np.array_equal(interp, ref)


@pytest.mark.parametrize("L,K", [(4, 3), (9, 20), (120, 50), (150, 7)])
def test_twin_spline_shared_bitwise(tw, L, K):
    """The shared-knot (LDS) path of the spline build -- one elimination of the
    common matrix, then right-hand-side sweeps -- equals the oracle bit for bit."""
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.make_input(0.01, 10.0, L, "m_nonneg_30", K, 100 + L, name="sh")
    tw.twin_prepare_modes_shared.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, P, P]
    interp = np.zeros((4, L, 2 * K))
    pcoef = np.zeros((8, L - 1, 3))
    teuk = np.ascontiguousarray(inp.teuk)
    pk = np.ascontiguousarray(inp.phase_knots)
    assert tw.twin_prepare_modes_shared(_p(teuk), _p(pk), _p(inp.t), L, K, _p(interp),
                                        _p(pcoef)) == 0
    interp_ref, pcoef_ref = oracle_coefficients(inp)
    assert np.array_equal(interp, interp_ref)
    assert np.array_equal(pcoef, pcoef_ref)


def test_twin_prepare_modes_bitwise(tw):
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.make_workload("cpu0")
    L, K = inp.L, inp.K
    interp = np.zeros((4, L, 2 * K))
    pcoef = np.zeros((8, L - 1, 3))
    scratch = np.zeros(4 * L * (2 * K + 3))
    teuk = np.ascontiguousarray(inp.teuk)
    tw.twin_prepare_modes(_p(teuk), _p(inp.phase_knots), _p(inp.t), L, K, _p(interp), _p(pcoef),
                          _p(scratch))
    ref_i, ref_p = oracle_coefficients(inp)
    assert np.array_equal(interp, ref_i)
    assert np.array_equal(pcoef, ref_p)


@pytest.mark.parametrize("out_delta", [0, -1000, -4000])
def test_twin_plan_matches_find_start_inds(tw, out_delta):
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.small_input(seed=23, L=14, K=4, T_years=0.002)
    N = inp.num_pts + out_delta
    st = np.zeros(inp.L, np.int32)
    en = np.zeros(inp.L, np.int32)
    used = tw.twin_plan(_p(inp.t), inp.L, inp.dt, N, _p(st), _p(en))
    ref, rused = oracle.find_start_inds(inp.t, inp.dt, N)
    assert used == rused
    for j in range(rused - 1):
        assert (st[j], en[j]) == (ref[j], ref[j + 1])
    # intervals after the stop own no samples and sit at out_len, so the last
    # interval's end is where the covered samples end
    assert np.all(st[rused - 1:inp.L - 1] == N) and np.all(en[rused - 1:inp.L - 1] == N)
    assert en[inp.L - 2] == (N if rused < inp.L else ref[inp.L - 1])


def test_twin_direct_fp64_vs_oracle(tw):
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.make_workload("cpu0")
    interp, pcoef = oracle_coefficients(inp)
    idx = spread_indices(inp.num_pts, count=600)
    out = np.zeros(2 * len(idx))
    m, k, n = (np.ascontiguousarray(a, np.int32) for a in (inp.m, inp.k, inp.n))
    yl = np.ascontiguousarray(inp.ylms)
    tw.twin_direct_samples(_p(out), _p(idx), len(idx), _p(interp), _p(inp.t), _p(pcoef), _p(m),
                           _p(k), _p(n), inp.L, inp.num_pts, inp.K, _p(yl), inp.dt, _p(inp.t))
    ref = oracle_waveform(inp, idx)
    assert oracle.rel_l2(out.view(np.complex128), ref) < 1e-14


def test_twin_sincos_turns(tw):
    """FP64 sin/cos of 2*pi*f (few_math.hpp) against numpy, over reduced and
    unreduced turns up to 1e6 (phases of a two-year inspiral)."""
    tw.twin_sincos_turns.argtypes = [P, P, P, ctypes.c_long]
    rng = np.random.default_rng(0)
    f = np.concatenate([np.linspace(-0.5, 0.5, 20001), rng.uniform(-1e6, 1e6, 20000),
                        np.array([0.0, 0.25, -0.25, 0.125, 0.5, -0.5, 1e-300])])
    s = np.zeros_like(f)
    c = np.zeros_like(f)
    tw.twin_sincos_turns(_p(f), _p(s), _p(c), len(f))
    r = f - np.rint(f)  # the reduction the kernels apply
    np.testing.assert_allclose(s, np.sin(2 * np.pi * r), rtol=0, atol=4e-16)
    np.testing.assert_allclose(c, np.cos(2 * np.pi * r), rtol=0, atol=4e-16)


def test_twin_sincos_circle(tw):
    """The FP16-split paths' sin/cos of 2*pi*f without a quadrant reduction
    (degree-8 polynomials in x^2 over the whole circle) against numpy: the
    bound its fit states (3.1e-13 for sin, 2.6e-12 for cos), at the ends of
    the range, the quadrant boundaries and zero included."""
    tw.twin_sincos_circle.argtypes = [P, P, P, ctypes.c_long]
    rng = np.random.default_rng(1)
    f = np.concatenate([np.linspace(-0.5, 0.5, 400001), rng.uniform(-1e6, 1e6, 20000),
                        np.array([0.0, 0.25, -0.25, 0.125, 0.5, -0.5, 0.375, 1e-300])])
    s = np.zeros_like(f)
    c = np.zeros_like(f)
    tw.twin_sincos_circle(_p(f), _p(s), _p(c), len(f))
    r = f - np.rint(f)
    np.testing.assert_allclose(s, np.sin(2 * np.pi * r), rtol=0, atol=3.5e-13)
    np.testing.assert_allclose(c, np.cos(2 * np.pi * r), rtol=0, atol=3e-12)
    # a dropped or mistyped coefficient fails this by orders of magnitude
    assert np.max(np.abs(s * s + c * c - 1.0)) < 6e-12


def _sincos_ref(x):
    """sin and cos of a double to ~1e-30 via an 80-digit reduction (Machin's pi)."""
    from decimal import Decimal, getcontext
    getcontext().prec = 80

    def atan_inv(n):
        n = Decimal(n)
        term = 1 / n
        s, k, sign = term, 1, -1
        while True:
            term /= n * n
            k += 2
            t = term / k
            if t < Decimal(10) ** -78:
                return s
            s += sign * t
            sign = -sign

    pi = 16 * atan_inv(5) - 4 * atan_inv(239)
    out_s, out_c = [], []
    for v in x:
        d = Decimal(float(v))
        k = (d / (pi / 2)).to_integral_value()
        r = d - k * (pi / 2)
        s, c, term, n = Decimal(0), Decimal(0), Decimal(1), 0
        while True:  # Taylor at |r| <= pi/4
            if n % 2 == 0:
                c += term if (n // 2) % 2 == 0 else -term
            else:
                s += term if (n // 2) % 2 == 0 else -term
            n += 1
            term = term * r / n
            if abs(term) < Decimal(10) ** -40:
                break
        q = int(k) % 4
        s, c = [(s, c), (c, -s), (-s, -c), (-c, s)][q]
        out_s.append(float(s))
        out_c.append(float(c))
    return np.array(out_s), np.array(out_c)


def test_twin_sincos_cw(tw):
    """The FP64 mode-sum sincos (Cody-Waite reduction + Taylor, few_math.hpp)
    against an 80-digit reference up to 1e8 rad: within 2 ulp of 1; and the
    fallback signal beyond 2^26 quadrants."""
    tw.twin_sincos_cw.argtypes = [P, P, P, P, ctypes.c_long]
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-1e8, 1e8, 400), rng.uniform(-1e4, 1e4, 200),
                        rng.uniform(-4, 4, 200), np.array([0.0, -0.0, 1e-300, np.pi / 2, -np.pi,
                                                           1.0e8, -1.0e8, 355.0, 103993.0])])
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    ok = np.zeros(len(x), np.int32)
    tw.twin_sincos_cw(_p(x), _p(s), _p(c), _p(ok), len(x))
    assert ok.all()
    rs, rc = _sincos_ref(x)
    ulp = np.spacing(1.0)
    assert np.max(np.abs(s - rs)) <= 2 * ulp, np.max(np.abs(s - rs))
    assert np.max(np.abs(c - rc)) <= 2 * ulp, np.max(np.abs(c - rc))
    big = np.array([2.0e8, -3.0e9, np.inf, np.nan])
    ok2 = np.zeros(len(big), np.int32)
    s2, c2 = np.zeros_like(big), np.zeros_like(big)
    tw.twin_sincos_cw(_p(big), _p(s2), _p(c2), _p(ok2), len(big))
    assert not ok2.any()


@pytest.mark.parametrize("n", [3, 4, 5, 9, 17, 120, 150])
def test_twin_spline_block_pieces_bitwise(tw, n):
    """The shared-knot block's read-ahead, branch-free elimination and sweeps
    (dgtsv_factor_ahead, dgtsv_sweeps) against dgtsv_factor and the plain
    sweep of spline_system_shared, bit for bit, on knots whose spacings force
    row interchanges at random rows."""
    tw.twin_spline_block_pieces.argtypes = [P, P, ctypes.c_int, P, P, P, P, P]
    rng = np.random.default_rng(n)
    swaps = 0
    for trial in range(20):
        h = rng.uniform(0.05, 1.0, n - 1)
        # occasional very uneven neighbours make |dl| > |d| (interchanges)
        h[rng.random(n - 1) < 0.3] *= rng.uniform(20.0, 200.0)
        x = np.concatenate([[0.0], np.cumsum(h)])
        y = rng.normal(size=n)
        Fp = np.zeros(5 * n)
        Fa = np.zeros(5 * n)
        info = np.zeros(2, np.int32)
        slopes = np.zeros(n)
        c1 = np.zeros(n - 1)
        tw.twin_spline_block_pieces(_p(x), _p(y), n, _p(Fp), _p(Fa), _p(info), _p(slopes), _p(c1))
        assert info[0] == info[1]
        if info[0] != 0:
            continue
        assert np.array_equal(Fp[:2 * n - 1], Fa[:2 * n - 1])          # piv, fact
        assert np.array_equal(Fp[2 * n:3 * n], Fa[2 * n:3 * n])        # d
        assert np.array_equal(Fp[3 * n:4 * n - 1], Fa[3 * n:4 * n - 1])  # du
        assert np.array_equal(Fp[4 * n:5 * n - 2], Fa[4 * n:5 * n - 2])  # dl (fill-in)
        assert np.array_equal(slopes[:n - 1], c1)
        swaps += int(np.sum(Fa[:n - 1] == 1.0))
    assert n < 4 or swaps > 0   # the interchange branch was exercised


def test_twin_factor_ahead_interchanges_and_singular(tw):
    """Hand-built matrices: interchanges at every step, and dgtsv's early
    return on an exactly singular leading block (same INFO, same stored rows)."""
    tw.twin_dgtsv_pair.argtypes = [ctypes.c_int, P, P, P, P, P]
    n = 12
    rng = np.random.default_rng(7)
    d = rng.uniform(0.01, 0.1, n)
    du = rng.uniform(0.5, 1.0, n - 1)
    dl = rng.uniform(5.0, 10.0, n - 1)   # |dl| > |d|: interchange at every step
    out = np.zeros((2, 5 * n))
    info = np.zeros(2, np.int32)
    tw.twin_dgtsv_pair(n, _p(d), _p(du), _p(dl), _p(out), _p(info))
    assert info[0] == info[1] == 0
    assert np.all(out[0, :n - 1] == 1.0)
    assert np.array_equal(out[0], out[1])
    d2 = d.copy()
    d2[0] = 0.0
    dl2 = dl.copy()
    dl2[0] = 0.0   # no interchange possible and a zero pivot: INFO = 1
    tw.twin_dgtsv_pair(n, _p(d2), _p(du), _p(dl2), _p(out), _p(info))
    assert info[0] == info[1] == 1
    assert np.array_equal(out[0], out[1])
