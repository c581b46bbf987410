/*
 * few_oracle_amp.c -- CPU FP64 restatement of FastEMRIWaveforms' bicubic
 * amplitude interpolation kernel (the CPU branch of `interp2D_wrap`,
 * src/few/cutils/AmpInterp2D.cu:1-208): FITPACK's fpbspl (de Boor-Cox
 * recurrence for the non-zero B-splines) and fpbisp (tensor-product sum).
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only the test suite, the smoke
 * check of __graft_entry__ and the cpu_baseline leg of bench.py may load
 * this library, and only as the checker / CPU baseline.  The shipped
 * library (fastemriwaveforms_amd/csrc) never links or calls it.
 *
 * Built with -ffp-contract=off: every expression rounds as written.
 */

/* AmpInterp2D.cu:7-26 -- the k+1 non-zero B-splines of degree k at
 * t[l] <= x < t[l+1] (l is the index of the knot interval, 0-based). */
static void orc_fpbspl(const double *t, int k, double x, int l, double *h)
{
    int i, j, li, lj;
    double f;
    double hh[5] = {0};
    h[0] = 1;
    for (j = 1; j <= k; j++) {
        for (i = 0; i < k; i++) hh[i] = h[i];
        h[0] = 0;
        for (i = 0; i < j; i++) {
            li = l + i;
            lj = li - j;
            f = hh[i] / (t[li] - t[lj]);
            h[i] += f * (t[li] - x);
            h[i + 1] = f * (x - t[lj]);
        }
    }
}

/* AmpInterp2D.cu:28-84 -- one point of the tensor-product spline; the
 * argument is clamped to [t[k], t[n-k-1]] and the interval found by a
 * linear walk from the first interior knot. */
static double orc_fpbisp(const double *tx, int nx, const double *ty, int ny, const double *c,
                         int kx, int ky, double x, double y)
{
    int i1, j1, kx1, ky1, l, l1, l2, nkx1, nky1, lx, ly;
    double arg, sp, tb, te;
    double wx[6] = {0.}, wy[6] = {0.};
    kx1 = kx + 1;
    nkx1 = nx - kx1;
    tb = tx[kx1 - 1];
    te = tx[nkx1];
    l = kx1;
    arg = x;
    if (arg < tb) arg = tb;
    if (arg > te) arg = te;
    while (!(arg < tx[l] || l == nkx1)) l++;
    orc_fpbspl(tx, kx, arg, l, wx);
    lx = l - kx1;
    ky1 = ky + 1;
    nky1 = ny - ky1;
    tb = ty[ky1 - 1];
    te = ty[nky1];
    l = ky1;
    arg = y;
    if (arg < tb) arg = tb;
    if (arg > te) arg = te;
    while (!(arg < ty[l] || l == nky1)) l++;
    orc_fpbspl(ty, ky, arg, l, wy);
    ly = l - ky1;
    l = lx * nky1;
    l1 = l + ly;
    sp = 0;
    for (i1 = 0; i1 < kx1; i1++) {
        l2 = l1;
        for (j1 = 0; j1 < ky1; j1++) sp += c[l2++] * wx[i1] * wy[j1];
        l1 += nky1;
    }
    return sp;
}

/* AmpInterp2D.cu:86-162 (CPU branch) -- z[c_i * mx + i] for every
 * coefficient set c_i (num_indiv_c of len_indiv_c) and point i.  Returns 0,
 * or 1 when mx != my (the reference throws invalid_argument). */
int orc_interp2d(double *z, const double *tx, int nx, const double *ty, int ny, const double *c,
                 int kx, int ky, const double *x, int mx, const double *y, int my,
                 int num_indiv_c, int len_indiv_c)
{
    if (mx != my) return 1;
    for (int c_i = 0; c_i < num_indiv_c; c_i++)
        for (int i = 0; i < mx; i++)
            z[(long)c_i * mx + i] = orc_fpbisp(tx, nx, ty, ny, c + (long)c_i * len_indiv_c, kx,
                                               ky, x[i], y[i]);
    return 0;
}
