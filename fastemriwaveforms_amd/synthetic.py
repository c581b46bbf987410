"""Seeded synthetic inputs for the interpolated mode sum (host numpy arrays).

Workloads (sizes follow summation/base.py:83-102 with YRSID_SI from
cutils/global.h:23):

* ``"cpu0"``   -- T = 0.1 yr, dt = 10 s, 100 knots, 500 of the 3843-mode set.
* ``"full1"``  -- T = 1 yr, dt = 10 s, 120 knots, all 3843 modes.
* ``"mixed2"`` -- T = 2 yr, dt = 5 s, 200 knots, 2000 of the 12,987-mode set
  (|m| <= l, n in [-55, 55]) with Y- = 0 for every mode (explicit +-m).
* batch    -- :func:`make_batch`: uneven durations / knots / mode counts.

Distributions: knots t0 = 0, t_{L-1} = T, interior sorted uniform draws.
f(t) = 1e-3 + 2e-3 (3 tau^2 - 2 tau^3) Hz with tau = t/T (or, with
chirp="rising", 1e-3 + 1e-3 (tau + tau^2 - tau^3 / 2)) and Phi_phi = 2 pi int f (analytic);
the frequencies at the knots are kept in ``freqs``; Phi_r = kappa Phi_phi with
kappa ~ U(0.8, 0.95); Phi_theta = 0.  Knot amplitudes
CN(0,1) 10^(-|n|/10) 10^(-(l-2)/3) (1 + sum_j a_j tau^j), a_j ~ N(0, 0.1^2).
Y+ and Y- have unit modulus and uniform phase.
"""

from __future__ import annotations

import dataclasses

import numpy as np

YRSID_SI = 31558149.763545595  # cutils/global.h:23

WORKLOADS = {
    "cpu0": dict(T=0.1, dt=10.0, L=100, mode_set="m_nonneg_30", keep=500, seed=1000),
    "full1": dict(T=1.0, dt=10.0, L=120, mode_set="m_nonneg_30", keep=None, seed=1001),
    "mixed2": dict(T=2.0, dt=5.0, L=200, mode_set="m_all_55", keep=2000, seed=1002),
}


@dataclasses.dataclass
class ModeSumInput:
    """One waveform's inputs to InterpolatedModeSum.sum-equivalent code."""

    t: np.ndarray            # (L,) knot times [s]
    teuk: np.ndarray         # (L, K) complex128 knot amplitudes
    ylms: np.ndarray         # (2K,) complex128: Y+ then Y- (as the kernel reads them)
    phase_knots: np.ndarray  # (3, L) Phi_phi, Phi_theta, Phi_r at the knots
    l: np.ndarray
    m: np.ndarray
    k: np.ndarray
    n: np.ndarray            # (K,) int32
    dt: float
    T_years: float
    num_pts: int
    name: str = ""
    freqs: np.ndarray | None = None  # (3, L) f_phi, f_theta, f_r [Hz] at the knots
    # (8, L-1, 3) dense-output phase coefficients handed over as they are (the
    # trajectory's DOPR853 output); None: the phases are cubic splines of
    # phase_knots, converted to that layout with rcont5..8 = 0
    dense: np.ndarray | None = None

    @property
    def L(self) -> int:
        return len(self.t)

    @property
    def K(self) -> int:
        return self.teuk.shape[1]

    @property
    def mode_samples(self) -> int:
        return self.K * self.num_pts


def mode_set(kind: str):
    """(l, m, n) arrays of a mode set (k = 0), or (l, m, k, n) for the k sets ("k1_20")."""
    ls, ms, ks, ns = [], [], [], []
    if kind == "m_nonneg_30":      # SchwarzschildEccentric(lmax=10, nmax=30): 63 * 61 = 3843
        for l in range(2, 11):
            for m in range(0, l + 1):
                for n in range(-30, 31):
                    ls.append(l); ms.append(m); ns.append(n)
    elif kind == "k1_20":          # l = 2..4, m = 0..l, k = -1..1, n = -20..20: 12 * 3 * 41 = 1476
        for l in range(2, 5):
            for m in range(0, l + 1):
                for kk in (-1, 0, 1):
                    for n in range(-20, 21):
                        ls.append(l); ms.append(m); ks.append(kk); ns.append(n)
    elif kind == "m_all_55":       # 117 (l, m) pairs with |m| <= l, times 111 n = 12,987
        for l in range(2, 11):
            for m in range(-l, l + 1):
                for n in range(-55, 56):
                    ls.append(l); ms.append(m); ns.append(n)
    else:
        raise ValueError(kind)
    out = [np.asarray(ls, np.int32), np.asarray(ms, np.int32), np.asarray(ns, np.int32)]
    if ks:
        out.insert(2, np.asarray(ks, np.int32))
    return tuple(out)


def num_points(t_first: float, t_last: float, T_years: float, dt: float) -> int:
    """summation/base.py:83-98 without padding."""
    n_pts = int(T_years * YRSID_SI / dt)
    T = n_pts * dt
    if T < t_last:
        return int((T - t_first) / dt) + 1
    return int((t_last - t_first) / dt) + 1


def phase_phi(t: np.ndarray, Ts: float, chirp: str = "smooth") -> np.ndarray:
    """2 pi int_0^t f with f = 1e-3 + 2e-3 (3 tau^2 - 2 tau^3) ("smooth", f' = 0
    at both ends) or f = 1e-3 + 1e-3 (tau + tau^2 - tau^3 / 2) ("rising": f' > 0
    throughout and a cubic term in every knot interval), tau = t/Ts."""
    tau = t / Ts
    if chirp == "rising":
        return 2.0 * np.pi * (1e-3 * t + 1e-3 * Ts * (tau ** 2 / 2 + tau ** 3 / 3 - tau ** 4 / 8))
    return 2.0 * np.pi * (1e-3 * t + 2e-3 * Ts * (tau ** 3 - 0.5 * tau ** 4))


def freq_phi(t: np.ndarray, Ts: float, chirp: str = "smooth") -> np.ndarray:
    """f = d phase_phi / dt / (2 pi) [Hz]."""
    tau = t / Ts
    if chirp == "rising":
        return 1e-3 + 1e-3 * (tau + tau ** 2 - 0.5 * tau ** 3)
    return 1e-3 + 2e-3 * (3 * tau ** 2 - 2 * tau ** 3)


def inspiral_phase_terms(kappa: float, kappa_theta: float):
    """Frequency terms (f [Hz], a) of a quasi-Newtonian inspiral for the three
    phases: each term contributes f u^(a-1) to the frequency, u = 1 - tau/1.25,
    tau = t/T (f_phi rises from 1e-3 to 1.83e-3 Hz as u^(-3/8), the leading-order
    chirp law).  Phi_theta and Phi_r follow Phi_phi with ratios kappa_theta,
    kappa plus slowly drifting terms, so none of the three is a polynomial."""
    return ([(1e-3, 5.0 / 8.0)],
            [(kappa_theta * 1e-3, 5.0 / 8.0), (2e-5, 7.0 / 8.0)],
            [(kappa * 1e-3, 5.0 / 8.0), (-1e-5, 3.0 / 4.0)])


INSPIRAL_TC = 1.25  # coalescence at 1.25 T: u = 1 - t / (1.25 T) stays >= 0.2


def inspiral_phase(t: np.ndarray, Ts: float, terms) -> tuple[np.ndarray, np.ndarray]:
    """Phase sum_terms 2 pi f Ts tc / a (1 - u^a) and its rate 2 pi sum f u^(a-1)."""
    tc = INSPIRAL_TC
    u = 1.0 - np.asarray(t, np.float64) / (tc * Ts)
    ph = np.zeros_like(u)
    rate = np.zeros_like(u)
    for f, a in terms:
        ph += 2.0 * np.pi * f * Ts * tc / a * (1.0 - u ** a)
        rate += 2.0 * np.pi * f * u ** (a - 1.0)
    return ph, rate


def dense_output(t: np.ndarray, fn, nodes: int = 24) -> np.ndarray:
    """DOPR853-style dense-output coefficients (8, L-1) of a smooth function on
    the knots t, for get_waveform's nested form
    P(s) = r1 + s (r2 + s1 (r3 + s (r4 + s1 (r5 + s (r6 + s1 (r7 + s r8)))))).

    r1..r4 are DOP853's own (trajectory/dopr853.py:631-642): the Hermite cubic
    of the values and slopes at both ends, rcont3 = h y'_0 - (y_1 - y_0),
    rcont4 = (y_1 - y_0) - h y'_1 - rcont3.  DOP853 builds r5..r8 from its
    extra stages; here they are the least-squares fit of the remainder
    y - Hermite = s^2 s1^2 (r5 + s r6 + s s1 r7 + s^2 s1 r8) on Chebyshev
    nodes, which has the same structure (no change of the end values and
    slopes) and order.  fn(t) returns (y, y')."""
    t = np.asarray(t, np.float64)
    h = np.diff(t)
    y, dy = fn(t)
    out = np.zeros((8, len(h)))
    out[0] = y[:-1]
    out[1] = y[1:] - y[:-1]
    out[2] = h * dy[:-1] - out[1]
    out[3] = out[1] - h * dy[1:] - out[2]
    k = np.arange(nodes)
    sn = 0.5 - 0.5 * np.cos(np.pi * (k + 0.5) / nodes)          # (nodes,) in (0, 1)
    s1 = 1.0 - sn
    w = sn ** 2 * s1 ** 2
    basis = np.stack([w, w * sn, w * sn * s1, w * sn ** 2 * s1], axis=1)   # (nodes, 4)
    tt = t[:-1, None] + sn[None, :] * h[:, None]
    yv, _ = fn(tt.ravel())
    r = out[:, :, None]
    herm = r[0] + sn * (r[1] + s1 * (r[2] + sn * r[3]))
    resid = yv.reshape(tt.shape) - herm                                  # (L-1, nodes)
    coef, *_ = np.linalg.lstsq(basis, resid.T, rcond=None)               # (4, L-1)
    out[4:] = coef
    return out


def make_input(T_years: float, dt: float, L: int, mode_kind: str, keep, seed: int,
               minus_m_ylm: bool = True, name: str = "", chirp: str = "smooth",
               phases: str = "spline") -> ModeSumInput:
    """phases="spline": Phi_phi from the chirp law ``chirp`` (a quartic), Phi_r =
    kappa Phi_phi, Phi_theta = 0, all splined by the build (rcont5..8 = 0).
    phases="dopr": the non-polynomial inspiral of :func:`inspiral_phase_terms`
    on all three phases (Phi_theta != 0), handed over as DOPR853-style dense
    output with non-zero rcont5..8 (``dense``)."""
    rng = np.random.default_rng(seed)
    Ts = T_years * YRSID_SI
    interior = np.sort(rng.uniform(0.0, 1.0, L - 2)) * Ts
    t = np.concatenate([[0.0], interior, [Ts]])
    ms_all = mode_set(mode_kind)
    if len(ms_all) == 4:
        l_all, m_all, k_all, n_all = ms_all
    else:
        l_all, m_all, n_all = ms_all
        k_all = np.zeros_like(m_all)
    if keep is not None and keep < len(l_all):
        sel = np.sort(rng.choice(len(l_all), size=keep, replace=False))
    else:
        sel = np.arange(len(l_all))
    l, m, kk, n = l_all[sel], m_all[sel], k_all[sel], n_all[sel]
    K = len(l)
    tau = t / Ts
    base = (rng.normal(size=K) + 1j * rng.normal(size=K)) / np.sqrt(2.0)
    scale = 10.0 ** (-np.abs(n) / 10.0) * 10.0 ** (-(l - 2) / 3.0)
    a = rng.normal(0.0, 0.1, size=(3, K))
    mod = 1.0 + a[0] * tau[:, None] + a[1] * tau[:, None] ** 2 + a[2] * tau[:, None] ** 3
    teuk = (base * scale)[None, :] * mod
    yp = np.exp(1j * rng.uniform(0, 2 * np.pi, K))
    ym = np.exp(1j * rng.uniform(0, 2 * np.pi, K))
    if not minus_m_ylm:
        ym = np.zeros(K, dtype=np.complex128)
    kappa = rng.uniform(0.8, 0.95)
    dense = None
    if phases == "dopr":
        kappa_theta = rng.uniform(0.9, 0.99)
        terms = inspiral_phase_terms(kappa, kappa_theta)
        vals, rates = zip(*(inspiral_phase(t, Ts, tm) for tm in terms))
        phase_knots = np.stack(vals)
        freqs = np.stack(rates) / (2.0 * np.pi)
        dense = np.stack([dense_output(t, lambda x, tm=tm: inspiral_phase(x, Ts, tm))
                          for tm in terms], axis=-1)                     # (8, L-1, 3)
    elif phases == "spline":
        Pphi = phase_phi(t, Ts, chirp)
        phase_knots = np.stack([Pphi, np.zeros_like(Pphi), kappa * Pphi])
        fphi = freq_phi(t, Ts, chirp)
        freqs = np.stack([fphi, np.zeros_like(fphi), kappa * fphi])
    else:
        raise ValueError(phases)
    return ModeSumInput(
        t=t, teuk=np.ascontiguousarray(teuk.astype(np.complex128)),
        ylms=np.concatenate([yp, ym]).astype(np.complex128),
        phase_knots=np.ascontiguousarray(phase_knots), l=l, m=m,
        k=np.ascontiguousarray(kk, np.int32), n=n, dt=float(dt), T_years=float(T_years),
        num_pts=num_points(t[0], t[-1], T_years, dt), name=name, freqs=freqs,
        dense=None if dense is None else np.ascontiguousarray(dense))


def make_workload(name: str, seed: int | None = None) -> ModeSumInput:
    w = WORKLOADS[name]
    return make_input(w["T"], w["dt"], w["L"], w["mode_set"], w["keep"],
                      w["seed"] if seed is None else seed,
                      minus_m_ylm=(w["mode_set"] != "m_all_55"), name=name)


def make_batch(count: int = 256, seed: int = 1003, t_range=(0.1, 2.0), L_range=(50, 200),
               K_range=(50, 3000), dt: float = 10.0) -> list[ModeSumInput]:
    """Uneven batch: durations U(t_range) yr, knots U{L_range}, modes U{K_range} of the 3843 set."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        T = float(rng.uniform(*t_range))
        L = int(rng.integers(L_range[0], L_range[1] + 1))
        K = int(rng.integers(K_range[0], K_range[1] + 1))
        out.append(make_input(T, dt, L, "m_nonneg_30", K, int(rng.integers(1 << 31)),
                              name=f"batch{b}"))
    return out


def small_input(seed: int = 7, L: int = 12, K: int = 20, T_years: float = 0.002,
                dt: float = 10.0) -> ModeSumInput:
    """A tiny case (a few thousand samples) for brute-force checks."""
    return make_input(T_years, dt, L, "m_nonneg_30", K, seed, name="small")
