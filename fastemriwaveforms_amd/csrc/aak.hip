// Augmented Analytic Kludge summation (get_waveform of
// src/few/cutils/gpuAAK.cu:181-577): per output sample, the trajectory's
// dense-output splines (e, Y = cos iota, the three phases and their rates),
// the source-frame geometry (orbital-plane precession, spin-orbit angle beta,
// optional LISA long-wavelength response) and the sum over harmonics n of the
// Peters-Mathews amplitudes with Bessel functions J_{n-2..n+2}(n e),
// rotated to the wave frame.
//
// The Bessel functions are the Numerical Recipes routines the reference's CPU
// branch uses (rational / asymptotic J0, J1; Miller's backward recurrence for
// J_n below its order), restated here so the device result follows the CPU
// reference rather than a library jn.  Everything else is the reference's
// arithmetic in its order with contraction off, on the device's FP64 libm;
// the cos and sin of one argument are taken with one sincos call (as an
// optimising compiler merges them), so the result does not depend on
// whether a compiler merges the pair.
//
// Work layout: a one-thread planning kernel turns the knot times into sample
// ranges with the reference's rule (gpuAAK.cu:458-483); the sum kernel runs
// one lane per output sample over all segments in one launch (the reference
// launches one kernel per segment on its own stream).
#include <math.h>

#include "aak.hpp"
#include "few_common.hpp"

#pragma clang fp contract(off)

namespace few {

#define AAK_YRSID_SI 31558149.763545595      // global.h:23
#define AAK_MTSUN_SI 4.9254909491978065e-06  // global.h:24
#define AAK_GPCINSEC 1.02927125054339e+17    // global.h:26
#define AAK_AUSEC 499.00478383615643         // global.h:27

// Bessel functions of the first kind as Numerical Recipes computes them
// (Press et al., Numerical Recipes in C, 2nd ed., sec. 6.5), the routines of
// gpuAAK.cu:8-118.  The fitted constants and the order of every operation are
// those routines' (bit-for-bit agreement with the CPU reference needs them);
// the code is organised around the three methods: a rational fit below
// |x| = 8, Hankel's asymptotic form above, and recurrences in the order.

// c[0] + y (c[1] + y (... + y c[N-1])): the nested order of the fits.
template <int N>
__host__ __device__ inline double nr_nested(const double (&c)[N], double y)
{
    double acc = c[N - 1];
#pragma unroll
    for (int i = N - 2; i >= 0; i--) acc = c[i] + y * acc;
    return acc;
}

// Hankel's form for |x| >= 8: sqrt(2 / (pi x)) (cos(x - w) P - (8/x) sin(x - w) Q),
// P and Q nested in y = (8/x)^2, w = pi/4 for order 0 and 3 pi/4 for order 1.
template <int NP, int NQ>
__host__ __device__ inline double nr_hankel(double ax, double w, const double (&pc)[NP],
                                            const double (&qc)[NQ])
{
    const double z = 8.0 / ax;
    const double y = z * z;
    const double P = nr_nested(pc, y);
    const double Q = nr_nested(qc, y);
    double sn, cs;
    sincos(ax - w, &sn, &cs);
    return sqrt(0.636619772 / ax) * (cs * P - z * sn * Q);
}

// J0 (gpuAAK.cu:8-32)
__host__ __device__ inline double nr_j0(double x)
{
    const double ax = fabs(x);
    if (ax < 8.0) {
        const double num[6] = {57568490574.0, -13362590354.0, 651619640.7, -11214424.18, 77392.33017,
                               -184.9052456};
        const double den[6] = {57568490411.0, 1029532985.0, 9494680.718, 59272.64853, 267.8532712, 1.0};
        const double y = x * x;
        return nr_nested(num, y) / nr_nested(den, y);
    }
    const double pc[5] = {1.0, -0.1098628627e-2, 0.2734510407e-4, -0.2073370639e-5, 0.2093887211e-6};
    const double qc[5] = {-0.1562499995e-1, 0.1430488765e-3, -0.6911147651e-5, 0.7621095161e-6,
                          -0.934935152e-7};
    return nr_hankel(ax, 0.785398164, pc, qc);
}

// J1 (gpuAAK.cu:34-60), odd in x
__host__ __device__ inline double nr_j1(double x)
{
    const double ax = fabs(x);
    if (ax < 8.0) {
        const double num[6] = {72362614232.0, -7895059235.0, 242396853.1, -2972611.439, 15704.48260,
                               -30.16036606};
        const double den[6] = {144725228442.0, 2300535178.0, 18583304.74, 99447.43394, 376.9991397, 1.0};
        const double y = x * x;
        return x * nr_nested(num, y) / nr_nested(den, y);
    }
    const double pc[5] = {1.0, 0.183105e-2, -0.3516396496e-4, 0.2457520174e-5, -0.240337019e-6};
    const double qc[5] = {0.04687499995, -0.2002690873e-3, 0.8449199096e-5, -0.88228987e-6, 0.105787412e-6};
    const double v = nr_hankel(ax, 2.356194491, pc, qc);
    return x < 0.0 ? -v : v;
}

// J_n(ax), ax > n: the upward recurrence J_{k+1} = (2k/x) J_k - J_{k-1} from
// J0 and J1, stable while the order stays below the argument.
__host__ __device__ inline double nr_jn_upward(int n, double ax)
{
    const double two_over_x = 2.0 / ax;
    double lower = nr_j0(ax), upper = nr_j1(ax);
    for (int k = 1; k < n; k++) {
        const double next = k * two_over_x * upper - lower;
        lower = upper;
        upper = next;
    }
    return upper;
}

// Miller's algorithm: J_n(ax), 0 < ax <= n.  The downward recurrence runs
// from the even start 2 floor((n + sqrt(40 n)) / 2) with arbitrary seeds,
// the terms are rescaled by 1e-10 whenever one exceeds 1e10, and the value
// at order n is normalised by 1 = J0 + 2 (J2 + J4 + ...).
__host__ __device__ inline int nr_miller_start(int n)
{
    return 2 * ((n + (int)sqrt(40.0 * n)) / 2);
}

__host__ __device__ inline double nr_jn_miller(int n, double ax)
{
    const double two_over_x = 2.0 / ax;
    double cur = 1.0, prev = 0.0, even_sum = 0.0, at_n = 0.0;
    bool odd_step = false;  // the start is even, so every second term is an even order
    for (int k = nr_miller_start(n); k > 0; k--) {
        const double below = k * two_over_x * cur - prev;
        prev = cur;
        cur = below;
        if (fabs(cur) > 1.0e10) {
            cur *= 1.0e-10;
            prev *= 1.0e-10;
            at_n *= 1.0e-10;
            even_sum *= 1.0e-10;
        }
        if (odd_step) even_sum += cur;
        odd_step = !odd_step;
        if (k == n) at_n = prev;
    }
    return at_n / (2.0 * even_sum - cur);
}

// J_n(x) (gpuAAK.cu:62-118).  As there, orders 0 and 1 return J(|x|), and
// the sign for negative x is applied only to odd n >= 2 (n % 2 == 1).
__host__ __device__ inline double nr_jn(int n, double x)
{
    const double ax = fabs(x);
    if (n == 0) return nr_j0(ax);
    if (n == 1) return nr_j1(ax);
    if (ax == 0.0) return 0.0;
    const double v = ax > (double)n ? nr_jn_upward(n, ax) : nr_jn_miller(n, ax);
    return x < 0.0 && n % 2 == 1 ? -v : v;
}

// One step k of nr_jn_miller for the chains LO..4 of nr_jn5 (each chain's
// operations in nr_jn_miller's order: the step, the rescaling, the even-order
// sum -- on odd k, as every start is even -- and the capture at k = order,
// which only the CAP steps test).
template <int LO, bool ODD, bool CAP>
__host__ __device__ inline void nr_miller_step(int k, double two_over_x, int nu0, double (&cur)[5],
                                               double (&prev)[5], double (&even_sum)[5],
                                               double (&at_n)[5])
{
    const double kt = k * two_over_x;
    bool big = false;
#pragma unroll
    for (int i = LO; i < 5; i++) {
        const double below = kt * cur[i] - prev[i];
        prev[i] = cur[i];
        cur[i] = below;
        big = big || fabs(below) > 1.0e10;
    }
    if (big) {
#pragma unroll
        for (int i = LO; i < 5; i++) {
            if (fabs(cur[i]) > 1.0e10) {
                cur[i] *= 1.0e-10;
                prev[i] *= 1.0e-10;
                at_n[i] *= 1.0e-10;
                even_sum[i] *= 1.0e-10;
            }
        }
    }
    if (ODD) {
#pragma unroll
        for (int i = LO; i < 5; i++) even_sum[i] += cur[i];
    }
    if (CAP) {
#pragma unroll
        for (int i = LO; i < 5; i++)
            if (k == nu0 + i) at_n[i] = prev[i];
    }
}

// Steps k, k - 1, ..., stop + 1 of the chains LO..4, k and stop even: step
// pairs (even k, odd k - 1), so the even-order sum needs no parity test and
// the two steps share their registers; only the pairs that hold an order
// nu0 .. nu0 + 4 test for the capture.
template <int LO>
__host__ __device__ inline void nr_miller_run(int &k, int stop, double two_over_x, int nu0,
                                              double (&cur)[5], double (&prev)[5],
                                              double (&even_sum)[5], double (&at_n)[5])
{
    for (; k > stop; k -= 2) {
        if (k - 1 <= nu0 + 4 && k >= nu0) {
            nr_miller_step<LO, false, true>(k, two_over_x, nu0, cur, prev, even_sum, at_n);
            nr_miller_step<LO, true, true>(k - 1, two_over_x, nu0, cur, prev, even_sum, at_n);
        } else {
            nr_miller_step<LO, false, false>(k, two_over_x, nu0, cur, prev, even_sum, at_n);
            nr_miller_step<LO, true, false>(k - 1, two_over_x, nu0, cur, prev, even_sum, at_n);
        }
    }
}

// nr_jn(nu0 + i, x), i = 0..4, for x >= 0 (the five Bessel values of one
// harmonic), bit for bit.  The five Miller recurrences run in lockstep
// (independent chains, instruction-level parallelism; the start grows with
// the order, so chain i joins at k = its start); orders 0 and 1, x = 0 and
// x > order take nr_jn's other branches (the lockstep chain of such an order
// runs and is discarded).
__host__ __device__ inline void nr_jn5(int nu0, double x, double (&out)[5])
{
    const double ax = fabs(x);
    const double two_over_x = 2.0 / ax;
    int start[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const int nu = nu0 + i;
        start[i] = nu >= 2 ? nr_miller_start(nu) : 0;
    }
    double cur[5], prev[5], even_sum[5], at_n[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        cur[i] = 1.0;
        prev[i] = at_n[i] = even_sum[i] = 0.0;
    }
    // every start is even (or 0), so every run spans whole step pairs
    int k = start[4];
    nr_miller_run<4>(k, start[3], two_over_x, nu0, cur, prev, even_sum, at_n);
    nr_miller_run<3>(k, start[2], two_over_x, nu0, cur, prev, even_sum, at_n);
    nr_miller_run<2>(k, start[1], two_over_x, nu0, cur, prev, even_sum, at_n);
    nr_miller_run<1>(k, start[0], two_over_x, nu0, cur, prev, even_sum, at_n);
    nr_miller_run<0>(k, 0, two_over_x, nu0, cur, prev, even_sum, at_n);
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const int nu = nu0 + i;
        if (nu >= 2 && ax != 0.0 && !(ax > (double)nu))
            out[i] = at_n[i] / (2.0 * even_sum[i] - cur[i]);
        else
            out[i] = nr_jn(nu, x);
    }
}

// gpuAAK.cu:136-176: rotation coefficients source frame -> wave frame
__host__ __device__ inline void aak_rot_coeff(double *rot, double iota, double theta_S, double phi_S,
                                              double theta_K, double phi_K, double alpha)
{
    double n[3], S[3], L[3], nxL[3], nxS[3];
    double sTS, cTS, sPS, cPS, sTK, cTK, sPK, cPK, sI, cI, sA, cA;
    sincos(theta_S, &sTS, &cTS);
    sincos(phi_S, &sPS, &cPS);
    sincos(theta_K, &sTK, &cTK);
    sincos(phi_K, &sPK, &cPK);
    sincos(iota, &sI, &cI);
    sincos(alpha, &sA, &cA);
    n[0] = sTS * cPS;
    n[1] = sTS * sPS;
    n[2] = cTS;
    S[0] = sTK * cPK;
    S[1] = sTK * sPK;
    S[2] = cTK;
    L[0] = cI * sTK * cPK + sI * (sA * sPK - cA * cTK * cPK);
    L[1] = cI * sTK * sPK - sI * (sA * cPK + cA * cTK * sPK);
    L[2] = cI * cTK + sI * cA * sTK;
    nxL[0] = n[1] * L[2] - n[2] * L[1];
    nxL[1] = n[2] * L[0] - n[0] * L[2];
    nxL[2] = n[0] * L[1] - n[1] * L[0];
    nxS[0] = n[1] * S[2] - n[2] * S[1];
    nxS[1] = n[2] * S[0] - n[0] * S[2];
    nxS[2] = n[0] * S[1] - n[1] * S[0];
    double norm = sqrt(nxL[0] * nxL[0] + nxL[1] * nxL[1] + nxL[2] * nxL[2]) *
                  sqrt(nxS[0] * nxS[0] + nxS[1] * nxS[1] + nxS[2] * nxS[2]);
    double dot = nxL[0] * nxS[0] + nxL[1] * nxS[1] + nxL[2] * nxS[2];
    if (norm < 1e-6) norm = 1e-6;
    const double cosrot = dot / norm;
    dot = L[0] * nxS[0] + L[1] * nxS[1] + L[2] * nxS[2];
    double sinrot = dot;
    dot = S[0] * nxL[0] + S[1] * nxL[1] + S[2] * nxL[2];
    sinrot -= dot;
    sinrot /= norm;
    rot[0] = 2. * cosrot * cosrot - 1.;
    rot[1] = cosrot * sinrot;
    rot[2] = -rot[1];
    rot[3] = rot[0];
}

__host__ __device__ inline double aak_dense(const double *r, double s, double s1)
{
    return r[0] + s * (r[1] + s1 * (r[2] + s * (r[3] + s1 * (r[4] + s * (r[5] + s1 * (r[6] + s * r[7]))))));
}

__host__ __device__ inline double aak_dense_dot(const double *r, double s, double s2, double s3,
                                                double s4, double s5, double s6)
{
    return r[1] + r[2] * (1. - 2. * s) + r[3] * (2. * s - 3. * s2) + r[4] * (2. * s - 6. * s2 + 4. * s3) +
           r[5] * (3. * s2 - 8. * s3 + 5. * s4) + r[6] * (3. * s2 - 12. * s3 + 15 * s4 - 6. * s5) +
           r[7] * (4. * s3 - 15. * s4 + 18. * s5 - 7. * s6);
}

// One output sample i of segment j (gpuAAK.cu:270-451); sc[p][c] are the
// segment's dense coefficients.  Returns hI - i hII.
__host__ __device__ inline double2 aak_sample(const AakArgs &a, const double (*sc)[8], double start_t,
                                              double segwidth, long i)
{
    double qS = a.qS, qK = a.qK;
    const double phiS = a.phiS, phiK = a.phiK;
    const double fill_val = 1e-6;
    if (qS < fill_val) qS = fill_val;
    if (qK < fill_val) qK = fill_val;
    if (qS > M_PI - fill_val) qS = M_PI - fill_val;
    if (qK > M_PI - fill_val) qK = M_PI - fill_val;
    double cosqS, sinqS, cosqK, sinqK, cosphiK, sinphiK, cosphiS, sinphiS, cosdphi, sindphi;
    sincos(qS, &sinqS, &cosqS);
    sincos(qK, &sinqK, &cosqK);
    sincos(phiK, &sinphiK, &cosphiK);
    sincos(phiS, &sinphiS, &cosphiS);
    sincos(phiK - phiS, &sindphi, &cosdphi);
    const double halfsqrt3 = sqrt(3.) / 2.;
    const double mu_sec = a.mu * AAK_MTSUN_SI;
    const double zeta = mu_sec / a.dist / AAK_GPCINSEC;
    const double t = a.dt * i;
    const double s = (t - start_t) / segwidth;
    const double s1 = (1.0 - s);
    const double s2 = s * s, s3 = s * s2, s4 = s * s3, s5 = s * s4, s6 = s * s5;
    const double e = aak_dense(sc[1], s, s1);
    const double Y = aak_dense(sc[2], s, s1);
    const double Phi_phi = aak_dense(sc[3], s, s1);
    const double Phi_theta = aak_dense(sc[4], s, s1);
    const double Phi_r = aak_dense(sc[5], s, s1);
    const double OmegaPhi = aak_dense_dot(sc[3], s, s2, s3, s4, s5, s6) / segwidth;
    const double OmegaTheta = aak_dense_dot(sc[4], s, s2, s3, s4, s5, s6) / segwidth;
    const double OmegaR = aak_dense_dot(sc[5], s, s2, s3, s4, s5, s6) / segwidth;
    const double Phi = Phi_r;
    const double gim = Phi_theta - Phi_r;
    const double alp = Phi_phi - Phi_theta;
    const double nu = OmegaR / (2. * M_PI);
    const double gimdot = OmegaTheta - OmegaR;
    const double lam = acos(Y);
    double coslam, sinlam, cosalp, sinalp;
    sincos(lam, &sinlam, &coslam);
    sincos(alp, &sinalp, &cosalp);
    const double cosqL = cosqK * coslam + sinqK * sinlam * cosalp;
    const double sinqL = sqrt(1. - cosqL * cosqL);
    const double phiLup = sinqK * sinphiK * coslam - cosphiK * sinlam * sinalp - cosqK * sinphiK * sinlam * cosalp;
    const double phiLdown = sinqK * cosphiK * coslam + sinphiK * sinlam * sinalp - cosqK * cosphiK * sinlam * cosalp;
    const double phiL = atan2(phiLup, phiLdown);
    const double Ldotn = cosqL * cosqS + sinqL * sinqS * cos(phiL - phiS);
    const double Ldotn2 = Ldotn * Ldotn;
    const double Sdotn = cosqK * cosqS + sinqK * sinqS * cosdphi;
    double beta;
    if (a.S_phys == 0.0 || lam < 1e-10 || lam > (M_PI - 1e-10)) {
        beta = 0.0;
    } else {
        const double betaup = -Sdotn + coslam * Ldotn;
        const double betadown = sinqS * sindphi * sinlam * cosalp + (cosqK * Sdotn - cosqS) / sinqK * sinlam * sinalp;
        beta = atan2(betaup, betadown);
    }
    const double gam = 2. * (gim + beta);
    double cos2gam, sin2gam;
    sincos(gam, &sin2gam, &cos2gam);
    double cosorbphs = 0, FplusI, FcrosI, FplusII, FcrosII;
    if (a.mich) {
        const double orbphs = 2. * M_PI * t / AAK_YRSID_SI;
        double sinorbphs, sinorb, cosorb;
        sincos(orbphs - phiS, &sinorbphs, &cosorbphs);
        sincos(orbphs, &sinorb, &cosorb);
        const double cosq = .5 * cosqS - halfsqrt3 * sinqS * cosorbphs;
        const double phiw = orbphs + atan2(halfsqrt3 * cosqS + .5 * sinqS * cosorbphs, sinqS * sinorbphs);
        const double psiup = .5 * cosqK - halfsqrt3 * sinqK * cos(orbphs - phiK) - cosq * (cosqK * cosqS + sinqK * sinqS * cosdphi);
        const double psidown = .5 * sinqK * sinqS * sindphi - halfsqrt3 * cosorb * (cosqK * sinqS * sinphiS - cosqS * sinqK * sinphiK) - halfsqrt3 * sinorb * (cosqS * sinqK * cosphiK - cosqK * sinqS * cosphiS);
        const double psi = atan2(psiup, psidown);
        const double cosq1 = .5 * (1. + cosq * cosq);
        double cos2phi, sin2phi, cos2psi, sin2psi;
        sincos(2. * phiw, &sin2phi, &cos2phi);
        sincos(2. * psi, &sin2psi, &cos2psi);
        FplusI = cosq1 * cos2phi * cos2psi - cosq * sin2phi * sin2psi;
        FcrosI = cosq1 * cos2phi * sin2psi + cosq * sin2phi * cos2psi;
        FplusII = cosq1 * sin2phi * cos2psi + cosq * cos2phi * sin2psi;
        FcrosII = cosq1 * sin2phi * sin2psi - cosq * cos2phi * cos2psi;
    } else {
        FplusI = 1.0;
        FcrosI = 0.0;
        FplusII = 0.0;
        FcrosII = 1.0;
    }
    const double Amp = pow(fabs(OmegaPhi) * a.M_phys * AAK_MTSUN_SI, 2. / 3.) * zeta;
    double rot[4];
    aak_rot_coeff(rot, lam, qS, phiS, qK, phiK, alp);
    double hItemp = 0.0, hIItemp = 0.0;
    for (int n = 1; n <= a.nmodes; n++) {
        double nPhi;
        if (a.mich) {
            const double fn = n * nu + gimdot / M_PI;
            const double Doppler = 2. * M_PI * fn * AAK_AUSEC * sinqS * cosorbphs;
            nPhi = n * Phi + Doppler;
        } else {
            nPhi = n * Phi;
        }
        const double ne = n * e;
        double Jv[5];
        nr_jn5(n - 2, ne, Jv);
        const double J0 = (n == 1) ? -1.0 * nr_j1(ne) : Jv[0];
        const double J1 = Jv[1], J2 = Jv[2], J3 = Jv[3], J4 = Jv[4];
        double sn, cn;
        sincos(nPhi, &sn, &cn);
        const double aa = -n * Amp * (J0 - 2. * e * J1 + 2. / n * J2 + 2. * e * J3 - J4) * cn;
        const double bb = -n * Amp * sqrt(1 - e * e) * (J0 - 2. * J2 + J4) * sn;
        const double cc = 2. * Amp * J2 * cn;
        const double Aplus0 = -(1. + Ldotn2) * (aa * cos2gam - bb * sin2gam) + cc * (1 - Ldotn2);
        const double Acros0 = 2. * Ldotn * (bb * cos2gam + aa * sin2gam);
        const double Aplus = Aplus0 * rot[0] + Acros0 * rot[1];
        const double Acros = Aplus0 * rot[2] + Acros0 * rot[3];
        double hnI, hnII;
        if (a.mich) {
            hnI = halfsqrt3 * (FplusI * Aplus + FcrosI * Acros);
            hnII = halfsqrt3 * (FplusII * Aplus + FcrosII * Acros);
        } else {
            hnI = FplusI * Aplus + FcrosI * Acros;
            hnII = FplusII * Aplus + FcrosII * Acros;
        }
        hItemp += hnI;
        hIItemp += hnII;
    }
    return make_double2(hItemp, -hIItemp);
}

// gpuAAK.cu:458-483: start index of every segment and the number of knots
// used; segments stop at the last knot when none reaches the end (the
// reference then reads past its arrays).
__host__ __device__ inline int aak_plan(const double *h_t, int L, double dt, int out_len, int *st)
{
    const double T = (out_len - 1) * dt;
    st[0] = 0;
    int i;
    for (i = 1; i < L; i++) {
        const double t = h_t[i];
        if (t < T) {
            st[i] = (int)ceil(t / dt);
        } else {
            st[i] = out_len;
            break;
        }
    }
    return i < L ? i + 1 : L;
}

__global__ __launch_bounds__(64) void aak_plan_kernel(AakArgs a)
{
    if (threadIdx.x == 0) a.plan[a.L] = aak_plan(a.h_t, a.L, a.dt, a.out_len, a.plan);
}

#define AAK_THREADS 128

// Three waves per SIMD (at most 168 registers, a few spills): the per-sample
// Bessel recurrences and sincos chains are FP64-latency-bound, and the third
// wave hides more of that than the spills cost (full year at 10 s, 30
// harmonics: 7.1 ms against 8.2 ms at two waves; four waves, 7.5 ms).
__global__ __launch_bounds__(AAK_THREADS) __attribute__((amdgpu_waves_per_eu(3, 3)))
void aak_sum_kernel(AakArgs a)
{
    const int npts = a.plan[a.L];
    const int nseg = npts - 1;
    const long i = (long)blockIdx.x * AAK_THREADS + threadIdx.x;
    if (nseg < 1 || i >= a.plan[nseg] || i >= a.out_len) return;
    // segment j: plan[j] <= i < plan[j + 1] (the last such j, as segments
    // later in the loop overwrite earlier ones in the reference)
    int lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.plan[mid] <= i) lo = mid; else hi = mid;
    }
    const int j = lo;
    if (i < a.plan[j]) return;
    double sc[6][8];
    const long np = (long)(a.L - 1) * 6;
    for (int c = 0; c < 8; c++)
        for (int p = 0; p < 6; p++) sc[p][c] = a.interp[c * np + (long)j * 6 + p];
    a.wave[i] = aak_sample(a, sc, a.h_t[j], a.h_t[j + 1] - a.h_t[j], i);
}

size_t aak_workspace_bytes(int init_len) { return sizeof(int) * (size_t)(init_len + 1); }

int launch_aak(AakArgs a, void *workspace, size_t workspace_bytes, hipStream_t stream)
{
    if (a.out_len <= 0) return FEW_OK;
    if (workspace_bytes < aak_workspace_bytes(a.L)) {
        set_error("aak: workspace of %zu bytes, need %zu", workspace_bytes, aak_workspace_bytes(a.L));
        return FEW_ERR_ARG;
    }
    a.plan = static_cast<int *>(workspace);
    hipLaunchKernelGGL(aak_plan_kernel, dim3(1), dim3(64), 0, stream, a);
    FEW_HIP(hipGetLastError());
    const long blocks = ((long)a.out_len + AAK_THREADS - 1) / AAK_THREADS;
    hipLaunchKernelGGL(aak_sum_kernel, dim3((unsigned)blocks), dim3(AAK_THREADS), 0, stream, a);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

}  // namespace few
