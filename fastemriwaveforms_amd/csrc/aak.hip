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

// gpuAAK.cu:8-32
__host__ __device__ inline double aak_j0(double x)
{
    double ax, z, xx, y, ans, ans1, ans2;
    if ((ax = fabs(x)) < 8.0) {
        y = x * x;
        ans1 = 57568490574.0 + y * (-13362590354.0 + y * (651619640.7 + y * (-11214424.18 + y * (77392.33017 + y * (-184.9052456)))));
        ans2 = 57568490411.0 + y * (1029532985.0 + y * (9494680.718 + y * (59272.64853 + y * (267.8532712 + y * 1.0))));
        ans = ans1 / ans2;
    } else {
        z = 8.0 / ax;
        y = z * z;
        xx = ax - 0.785398164;
        ans1 = 1.0 + y * (-0.1098628627e-2 + y * (0.2734510407e-4 + y * (-0.2073370639e-5 + y * 0.2093887211e-6)));
        ans2 = -0.1562499995e-1 + y * (0.1430488765e-3 + y * (-0.6911147651e-5 + y * (0.7621095161e-6 - y * 0.934935152e-7)));
        double sx, cx;
        sincos(xx, &sx, &cx);
        ans = sqrt(0.636619772 / ax) * (cx * ans1 - z * sx * ans2);
    }
    return ans;
}

// gpuAAK.cu:34-60
__host__ __device__ inline double aak_j1(double x)
{
    double ax, z, xx, y, ans, ans1, ans2;
    if ((ax = fabs(x)) < 8.0) {
        y = x * x;
        ans1 = x * (72362614232.0 + y * (-7895059235.0 + y * (242396853.1 + y * (-2972611.439 + y * (15704.48260 + y * (-30.16036606))))));
        ans2 = 144725228442.0 + y * (2300535178.0 + y * (18583304.74 + y * (99447.43394 + y * (376.9991397 + y * 1.0))));
        ans = ans1 / ans2;
    } else {
        z = 8.0 / ax;
        y = z * z;
        xx = ax - 2.356194491;
        ans1 = 1.0 + y * (0.183105e-2 + y * (-0.3516396496e-4 + y * (0.2457520174e-5 + y * (-0.240337019e-6))));
        ans2 = 0.04687499995 + y * (-0.2002690873e-3 + y * (0.8449199096e-5 + y * (-0.88228987e-6 + y * 0.105787412e-6)));
        double sx, cx;
        sincos(xx, &sx, &cx);
        ans = sqrt(0.636619772 / ax) * (cx * ans1 - z * sx * ans2);
        if (x < 0.0) ans = -ans;
    }
    return ans;
}

// gpuAAK.cu:62-118: forward recurrence above the order, Miller's backward
// recurrence normalised by the sum rule below it
__host__ __device__ inline double aak_jn(int n, double x)
{
    int j, jsum, m;
    double ax, bj, bjm, bjp, sum, tox, ans;
    ax = fabs(x);
    if (n == 0) return aak_j0(ax);
    if (n == 1) return aak_j1(ax);
    if (ax == 0.0) return 0.0;
    if (ax > (double)n) {
        tox = 2.0 / ax;
        bjm = aak_j0(ax);
        bj = aak_j1(ax);
        for (j = 1; j < n; j++) {
            bjp = j * tox * bj - bjm;
            bjm = bj;
            bj = bjp;
        }
        ans = bj;
    } else {
        tox = 2.0 / ax;
        m = 2 * ((n + (int)sqrt(40.0 * n)) / 2);
        jsum = 0;
        bjp = ans = sum = 0.0;
        bj = 1.0;
        for (j = m; j > 0; j--) {
            bjm = j * tox * bj - bjp;
            bjp = bj;
            bj = bjm;
            if (fabs(bj) > 1.0e10) {
                bj *= 1.0e-10;
                bjp *= 1.0e-10;
                ans *= 1.0e-10;
                sum *= 1.0e-10;
            }
            if (jsum) sum += bj;
            jsum = !jsum;
            if (j == n) ans = bjp;
        }
        sum = 2.0 * sum - bj;
        ans /= sum;
    }
    return x < 0.0 && n % 2 == 1 ? -ans : ans;
}

// One step j of Miller's backward recurrence (aak_jn's x <= n branch) for
// the chains LO..4 of aak_jn5: each chain's operations and their order are
// aak_jn's (update, renormalisation, the sum over even terms -- j odd, as m
// is even -- and the capture at j = order).
template <int LO>
__host__ __device__ inline void aak_miller_step(int j, double tox, int nu0, double (&bj)[5],
                                                double (&bjp)[5], double (&sum)[5],
                                                double (&ans)[5])
{
    const double jt = j * tox;
    bool big = false;
#pragma unroll
    for (int i = LO; i < 5; i++) {
        const double bjm = jt * bj[i] - bjp[i];
        bjp[i] = bj[i];
        bj[i] = bjm;
        big = big || fabs(bjm) > 1.0e10;
    }
    if (big) {
#pragma unroll
        for (int i = LO; i < 5; i++) {
            if (fabs(bj[i]) > 1.0e10) {
                bj[i] *= 1.0e-10;
                bjp[i] *= 1.0e-10;
                ans[i] *= 1.0e-10;
                sum[i] *= 1.0e-10;
            }
        }
    }
    if (j & 1) {
#pragma unroll
        for (int i = LO; i < 5; i++) sum[i] += bj[i];
    }
#pragma unroll
    for (int i = LO; i < 5; i++)
        if (j == nu0 + i) ans[i] = bjp[i];
}

// aak_jn(nu0 + i, x), i = 0..4, for x >= 0 (the five Bessel values of one
// harmonic), bit for bit.  Miller's recurrences of the five orders run in
// lockstep (independent chains, instruction-level parallelism; the start
// m(order) grows with the order, so chain i joins at j = m_i); orders 0 and
// 1, x = 0 and x > order take aak_jn's other branches (the lockstep chain of
// such an order runs and is discarded).
__host__ __device__ inline void aak_jn5(int nu0, double x, double (&out)[5])
{
    const double ax = fabs(x);
    const double tox = 2.0 / ax;
    int m[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const int nu = nu0 + i;
        m[i] = nu >= 2 ? 2 * ((nu + (int)sqrt(40.0 * nu)) / 2) : 0;
    }
    double bj[5], bjp[5], sum[5], ans[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        bj[i] = 1.0;
        bjp[i] = ans[i] = sum[i] = 0.0;
    }
    int j = m[4];
    for (; j > m[3]; j--) aak_miller_step<4>(j, tox, nu0, bj, bjp, sum, ans);
    for (; j > m[2]; j--) aak_miller_step<3>(j, tox, nu0, bj, bjp, sum, ans);
    for (; j > m[1]; j--) aak_miller_step<2>(j, tox, nu0, bj, bjp, sum, ans);
    for (; j > m[0]; j--) aak_miller_step<1>(j, tox, nu0, bj, bjp, sum, ans);
    for (; j > 0; j--) aak_miller_step<0>(j, tox, nu0, bj, bjp, sum, ans);
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const int nu = nu0 + i;
        if (nu >= 2 && ax != 0.0 && !(ax > (double)nu)) {
            const double sm = 2.0 * sum[i] - bj[i];
            out[i] = ans[i] / sm;
        } else {
            out[i] = aak_jn(nu, x);
        }
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
        aak_jn5(n - 2, ne, Jv);
        const double J0 = (n == 1) ? -1.0 * aak_j1(ne) : Jv[0];
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
