/*
 * few_oracle_aak.c -- CPU FP64 restatement of FastEMRIWaveforms' Augmented
 * Analytic Kludge summation (the CPU branch of `get_waveform` in
 * src/few/cutils/gpuAAK.cu:1-578): the trajectory's dense-output splines,
 * the source-frame geometry, the rotation to the wave frame and the sum of
 * Peters-Mathews harmonics with the Numerical Recipes Bessel functions.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only the test suite, the smoke
 * check of __graft_entry__ and the cpu_baseline leg of bench.py may load
 * this library, and only as the checker / CPU baseline.  The shipped
 * library (fastemriwaveforms_amd/csrc) never links or calls it.
 *
 * Built with -ffp-contract=off: every expression rounds as written.
 */

#define _GNU_SOURCE 1
#include <math.h>
#include <stdlib.h>

/* The reference calls cos and sin of one argument separately; an optimising
 * compiler merges such pairs into one sincos call, whose last bit can differ
 * from separate calls.  The restatement calls sincos for every such pair so
 * its result does not depend on this file's compiler or flags. */

/* src/few/cutils/global.h:23-27 */
#define ORC_YRSID_SI 31558149.763545595
#define ORC_MTSUN_SI 4.9254909491978065e-06
#define ORC_GPCINSEC 1.02927125054339e+17
#define ORC_AUSEC 499.00478383615643

/* gpuAAK.cu:3-6 */
#define ORC_ACC 40.0
#define ORC_BIGNO 1.0e10
#define ORC_BIGNI 1.0e-10

/* gpuAAK.cu:8-32 -- J0 (rational approximation below 8, asymptotic form above) */
static double orc_bessel_j0(double x)
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

/* gpuAAK.cu:34-60 -- J1 */
static double orc_bessel_j1(double x)
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

/* gpuAAK.cu:62-118 -- J_n: forward recurrence above n, Miller's backward
 * recurrence (normalised by the sum rule) below */
double orc_bessel_jn(int n, double x)
{
    int j, jsum, m;
    double ax, bj, bjm, bjp, sum, tox, ans;
    ax = fabs(x);
    if (n == 0) return orc_bessel_j0(ax);
    if (n == 1) return orc_bessel_j1(ax);
    if (ax == 0.0) return 0.0;
    else if (ax > (double)n) {
        tox = 2.0 / ax;
        bjm = orc_bessel_j0(ax);
        bj = orc_bessel_j1(ax);
        for (j = 1; j < n; j++) {
            bjp = j * tox * bj - bjm;
            bjm = bj;
            bj = bjp;
        }
        ans = bj;
    } else {
        tox = 2.0 / ax;
        m = 2 * ((n + (int)sqrt(ORC_ACC * n)) / 2);
        jsum = 0;
        bjp = ans = sum = 0.0;
        bj = 1.0;
        for (j = m; j > 0; j--) {
            bjm = j * tox * bj - bjp;
            bjp = bj;
            bj = bjm;
            if (fabs(bj) > ORC_BIGNO) {
                bj *= ORC_BIGNI;
                bjp *= ORC_BIGNI;
                ans *= ORC_BIGNI;
                sum *= ORC_BIGNI;
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

double orc_bessel_j1_export(double x) { return orc_bessel_j1(x); }

static double dot3(const double *u, const double *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; }

static void cross3(const double *u, const double *v, double *w)
{
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

static double norm3(const double *u) { return sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]); }

/* gpuAAK.cu:136-176 -- rotation from the source frame to the wave frame */
void orc_aak_rot_coeff(double *rot, double *n, double *L, double *S, double *nxL, double *nxS,
                       double iota, double theta_S, double phi_S, double theta_K, double phi_K,
                       double alpha)
{
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
    cross3(n, L, nxL);
    cross3(n, S, nxS);
    double norm = norm3(nxL) * norm3(nxS);
    double dot, cosrot, sinrot;
    dot = dot3(nxL, nxS);
    if (norm < 1e-6) norm = 1e-6;
    cosrot = dot / norm;
    dot = dot3(L, nxS);
    sinrot = dot;
    dot = dot3(S, nxL);
    sinrot -= dot;
    sinrot /= norm;
    rot[0] = 2. * cosrot * cosrot - 1.;
    rot[1] = cosrot * sinrot;
    rot[2] = -rot[1];
    rot[3] = rot[0];
}

static double orc_dense(const double *r, double s, double s1)
{
    return r[0] + s * (r[1] + s1 * (r[2] + s * (r[3] + s1 * (r[4] + s * (r[5] + s1 * (r[6] + s * r[7]))))));
}

/* the time derivative of the dense-output form times the segment width
 * (gpuAAK.cu:330-332) */
static double orc_dense_dot(const double *r, double s, double s2, double s3, double s4, double s5,
                            double s6)
{
    return r[1] + r[2] * (1. - 2. * s) + r[3] * (2. * s - 3. * s2) + r[4] * (2. * s - 6. * s2 + 4. * s3) +
           r[5] * (3. * s2 - 8. * s3 + 5. * s4) + r[6] * (3. * s2 - 12. * s3 + 15 * s4 - 6. * s5) +
           r[7] * (4. * s3 - 15. * s4 + 18. * s5 - 7. * s6);
}

/* gpuAAK.cu:421-446 -- the Bessel brackets of harmonic n (Peters & Mathews
 * 1963; Barack & Cutler 2004, eq. 9), with J = J_k(n e):
 *   X = J_{n-2} - 2e J_{n-1} + (2/n) J_n + 2e J_{n+1} - J_{n+2},
 *   Y = J_{n-2} - 2 J_n + J_{n+2},   Jn = J_n;
 * the harmonic's a, b, c are -n Amp X cos(n Phi), -n Amp sqrt(1-e^2) Y sin(n Phi)
 * and 2 Amp Jn cos(n Phi).  J_{-1} = -J_1 at n = 1. */
void orc_aak_harmonic(int n, double e, double *X, double *Y, double *Jn)
{
    const double ne = n * e;
    double J0, J1, J2, J3, J4;
    if (n == 1) J0 = -1.0 * orc_bessel_j1(ne);
    else J0 = orc_bessel_jn(n - 2, ne);
    J1 = orc_bessel_jn(n - 1, ne);
    J2 = orc_bessel_jn(n, ne);
    J3 = orc_bessel_jn(n + 1, ne);
    J4 = orc_bessel_jn(n + 2, ne);
    *X = J0 - 2. * e * J1 + 2. / n * J2 + 2. * e * J3 - J4;
    *Y = J0 - 2. * J2 + J4;
    *Jn = J2;
}

/* gpuAAK.cu:181-452 (CPU branch) for samples [start_ind, end_ind) of
 * segment old_ind: overwrites waveform[i] = hI - i hII.  interp_array is
 * (8, init_length - 1, 6) (coefficient, segment, parameter p e Y Phi_phi
 * Phi_theta Phi_r). */
static void orc_aak_segment(double *waveform, const double *interp_array, double M_phys,
                            double S_phys, double mu, double qS, double phiS, double qK,
                            double phiK, double dist, int nmodes, int mich, double delta_t,
                            double start_t, int old_ind, int start_ind, int end_ind,
                            int init_length, double segwidth)
{
    double sc[6][8];
    for (int i = 0; i < 8 * 6; i++) {
        const int coeff_num = i / 6, par_num = i % 6;
        sc[par_num][coeff_num] = interp_array[(coeff_num * (init_length - 1) + old_ind) * 6 + par_num];
    }
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
    const double mu_sec = mu * ORC_MTSUN_SI;
    const double zeta = mu_sec / dist / ORC_GPCINSEC;
    for (int i = start_ind; i < end_ind; i++) {
        double rot[4], n_rot[3], L_rot[3], S_rot[3], nxL_rot[3], nxS_rot[3];
        waveform[2L * i] = 0.0;
        waveform[2L * i + 1] = 0.0;
        const double t = delta_t * i;
        const double s = (t - start_t) / segwidth;
        const double s1 = (1.0 - s);
        const double s2 = s * s, s3 = s * s2, s4 = s * s3, s5 = s * s4, s6 = s * s5;
        const double e = orc_dense(sc[1], s, s1);
        const double Y = orc_dense(sc[2], s, s1);
        const double Phi_phi = orc_dense(sc[3], s, s1);
        const double Phi_theta = orc_dense(sc[4], s, s1);
        const double Phi_r = orc_dense(sc[5], s, s1);
        const double OmegaPhi = orc_dense_dot(sc[3], s, s2, s3, s4, s5, s6) / segwidth;
        const double OmegaTheta = orc_dense_dot(sc[4], s, s2, s3, s4, s5, s6) / segwidth;
        const double OmegaR = orc_dense_dot(sc[5], s, s2, s3, s4, s5, s6) / segwidth;
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
        if (S_phys == 0.0 || lam < 1e-10 || lam > (M_PI - 1e-10)) {
            beta = 0.0;
        } else {
            const double betaup = -Sdotn + coslam * Ldotn;
            const double betadown = sinqS * sindphi * sinlam * cosalp + (cosqK * Sdotn - cosqS) / sinqK * sinlam * sinalp;
            beta = atan2(betaup, betadown);
        }
        const double gam = 2. * (gim + beta);
        double cos2gam, sin2gam;
        sincos(gam, &sin2gam, &cos2gam);
        double orbphs = 0, cosorbphs = 0, sinorbphs = 0, FplusI, FcrosI, FplusII, FcrosII;
        if (mich) {
            orbphs = 2. * M_PI * t / ORC_YRSID_SI;
            double sinorb, cosorb;
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
        const double Amp = pow(fabs(OmegaPhi) * M_phys * ORC_MTSUN_SI, 2. / 3.) * zeta;
        orc_aak_rot_coeff(rot, n_rot, L_rot, S_rot, nxL_rot, nxS_rot, lam, qS, phiS, qK, phiK, alp);
        double hItemp = 0.0, hIItemp = 0.0;
        for (int n = 1; n <= nmodes; n++) {
            double fn, Doppler, nPhi;
            if (mich) {
                fn = n * nu + gimdot / M_PI;
                Doppler = 2. * M_PI * fn * ORC_AUSEC * sinqS * cosorbphs;
                nPhi = n * Phi + Doppler;
            } else {
                nPhi = n * Phi;
            }
            double X, Y, Jn;
            orc_aak_harmonic(n, e, &X, &Y, &Jn);
            double snp, cnp;
            sincos(nPhi, &snp, &cnp);
            const double a = -n * Amp * X * cnp;
            const double b = -n * Amp * sqrt(1 - e * e) * Y * snp;
            const double c = 2. * Amp * Jn * cnp;
            double Aplus = -(1. + Ldotn2) * (a * cos2gam - b * sin2gam) + c * (1 - Ldotn2);
            double Acros = 2. * Ldotn * (b * cos2gam + a * sin2gam);
            const double Aplusold = Aplus, Acrosold = Acros;
            Aplus = Aplusold * rot[0] + Acrosold * rot[1];
            Acros = Aplusold * rot[2] + Acrosold * rot[3];
            double hnI, hnII;
            if (mich) {
                hnI = halfsqrt3 * (FplusI * Aplus + FcrosI * Acros);
                hnII = halfsqrt3 * (FplusII * Aplus + FcrosII * Acros);
            } else {
                hnI = FplusI * Aplus + FcrosI * Acros;
                hnII = FplusII * Aplus + FcrosII * Acros;
            }
            hItemp += hnI;
            hIItemp += hnII;
        }
        waveform[2L * i] = hItemp;
        waveform[2L * i + 1] = -hIItemp;
    }
}

/* gpuAAK.cu:458-483 -- the segment start indices; *length becomes the number
 * of knots used (one past the segment that reaches the end).  The reference
 * leaves *length = init_len + 1 when no knot reaches the end (and then reads
 * past its arrays); here the segments stop at the last knot in that case. */
int orc_aak_start_inds(int *start_inds, const double *t_arr, double delta_t, int length,
                       int new_length)
{
    const double T = (new_length - 1) * delta_t;
    start_inds[0] = 0;
    int i;
    for (i = 1; i < length; i += 1) {
        const double t = t_arr[i];
        if (t < T) {
            start_inds[i] = (int)ceil(t / delta_t);
        } else {
            start_inds[i] = new_length;
            break;
        }
    }
    return i < length ? i + 1 : length;
}

/* gpuAAK.cu:486-577 (CPU branch) -- the whole waveform: waveform (out_len,)
 * complex, each sample of a used segment overwritten.  Returns 0. */
int orc_aak_waveform(double *waveform, const double *interp_array, double M_phys, double S_phys,
                     double mu, double qS, double phiS, double qK, double phiK, double dist,
                     int nmodes, int mich, int init_len, int out_len, double delta_t,
                     const double *h_t)
{
    int *start_inds = (int *)malloc(sizeof(int) * (size_t)(init_len + 1));
    const int npts = orc_aak_start_inds(start_inds, h_t, delta_t, init_len, out_len);
    for (int i = 0; i < npts - 1; i++)
        orc_aak_segment(waveform, interp_array, M_phys, S_phys, mu, qS, phiS, qK, phiK, dist,
                        nmodes, mich, delta_t, h_t[i], i, start_inds[i], start_inds[i + 1],
                        init_len, h_t[i + 1] - h_t[i]);
    free(start_inds);
    return 0;
}
