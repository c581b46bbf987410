/*
 * few_rocm.h -- C ABI of the MI355X (gfx950) interpolated mode-summation
 * library (libfew_rocm.so), a drop-in for FastEMRIWaveforms' native
 * interpolation backend.
 *
 * Conventions
 *   - Plain pointers and sizes only.  "Device" pointers are HIP device
 *     memory; the drop-in twins keep the reference's host/device split.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *   - Every entry returns 0 (FEW_OK) or a FEW_ERR_* code; the message of the
 *     calling thread's last failure is few_last_error().  Nothing calls
 *     exit() (the reference aborts via gpuErrchk, src/few/cutils/global.h:49-57).
 *   - complex128 buffers are interleaved {double re, double im}, like the
 *     reference's cmplx (src/few/cutils/global.h:30).
 *
 * Citations are to the FastEMRIWaveforms source tree.
 */
#ifndef FEW_ROCM_H
#define FEW_ROCM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FEW_ROCM_ABI_VERSION 1

#define FEW_PRECISION_FP64 0   /* plain FP64 direct sum (reference arithmetic) */
#define FEW_PRECISION_FACT64 1 /* factorised sum in FP64 on the matrix cores (sum and adjoint) */
#define FEW_PRECISION_MIXED 2  /* factorised MFMA sum: FP64 phases, FP32 MFMA, FP64 output */
#define FEW_PRECISION_MIXED16 3 /* factorised sum on FP16 MFMA with split (hi+lo) operands */

const char *few_last_error(void);
int few_abi_version(void);

/*
 * Replaces interpolate_arrays (src/few/cutils/interpolate.hh:27,
 * src/few/cutils/interpolate.cu:284-325; bound at src/few/cutils/pyinterp.pyx:25-38).
 * Same arguments, all device pointers:
 *   t_arr        (ninterps, length) knot times, one row per array
 *   interp_array (4, ninterps, length): slot 0 = y on entry, slots 1..3
 *                receive c1..c3 (slot 0 is left as y)
 *   B, upper_diag, diag, lower_diag  (ninterps*length) scratch; contents on
 *                return are unspecified (the reference leaves the solve there)
 * Synchronous on `stream`, like the reference.  length >= 3 (the reference
 * reads out of bounds for length < 3).
 */
int few_interpolate_arrays(double *t_arr, double *interp_array, int ninterps, int length,
                           double *B, double *upper_diag, double *diag, double *lower_diag,
                           void *stream);

/*
 * Replaces get_waveform (src/few/cutils/interpolate.hh:29-31,
 * src/few/cutils/interpolate.cu:726-797; bound at src/few/cutils/pyinterp.pyx:41-61).
 * Same arguments and host/device split:
 *   d_waveform          device complex128 (out_len,), ACCUMULATED INTO
 *   interp_array        device (4, init_len, 2*num_teuk_modes) coefficients
 *   phase_interp_t      HOST (init_len,) dense-output segment times
 *   phase_interp_coeffs device (8, init_len-1, 3)
 *   d_m, d_k, d_n       device int32 (num_teuk_modes,)
 *   d_Ylms              device complex128 (2*num_teuk_modes,), Y+ then Y-
 *   h_t                 HOST (init_len,) amplitude knot times
 *   dev                 device ordinal (hipSetDevice; restored on return)
 * Plain FP64 arithmetic with the reference's phase rounding.  Synchronous.
 */
int few_get_waveform(void *d_waveform, double *interp_array, double *phase_interp_t,
                     double *phase_interp_coeffs, int *d_m, int *d_k, int *d_n, int init_len,
                     int out_len, int num_teuk_modes, void *d_Ylms, double delta_t, double *h_t,
                     int dev, void *stream);

/*
 * Asynchronous get_waveform on device-resident time arrays (the caller's
 * CubicSplineInterpolant / InterpolatedModeSum keep t on the device; the
 * reference copies it to the host, src/few/summation/interpolatedmodesum.py:339).
 * `precision` selects the arithmetic (FEW_PRECISION_*); `workspace` is a
 * device buffer of few_waveform_workspace_bytes() bytes, owned by the
 * caller, that must not be shared by calls in flight at the same time.
 * No host synchronisation and no allocation: safe to capture in a hipGraph.
 */
int few_get_waveform_async(void *d_waveform, const double *interp_array,
                           const double *d_phase_interp_t, const double *phase_interp_coeffs,
                           const int *d_m, const int *d_k, const int *d_n, int init_len,
                           int out_len, int num_teuk_modes, const void *d_Ylms, double delta_t,
                           const double *d_h_t, int precision, void *workspace,
                           size_t workspace_bytes, void *stream);

size_t few_waveform_workspace_bytes(int init_len, int out_len, int num_teuk_modes, int precision);

/*
 * Strided batched not-a-knot spline build (the computation of
 * interpolate_arrays, src/few/cutils/interpolate.cu:284-325, on arbitrary
 * layouts).  System s, knot i:
 *   y[s*y_sys + i*y_knot], t[s*t_sys + i*t_knot]   (t_sys = 0: shared knots)
 *   dense == 0: out[c*o_coef + s*o_sys + i*o_knot] = c-th cubic coefficient
 *   dense == 1: out[c*(length-1)*nsys + i*nsys + s] = c-th dense-output
 *               coefficient (8 per interval, the layout get_waveform reads
 *               for phases, src/few/cutils/interpolate.cu:516-545)
 * scratch: 4*length*nsys doubles.  Asynchronous.
 */
int few_spline_build_async(const double *y, long y_sys, long y_knot, const double *t,
                           long t_sys, long t_knot, double *out, long o_coef, long o_sys,
                           long o_knot, int nsys, int length, int dense, double *scratch,
                           void *stream);

/*
 * One launch that prepares everything get_waveform reads from knot values:
 * the 2K amplitude splines of teuk (length, K) complex128 written to
 * interp_out (4, length, 2K) -- exactly what InterpolatedModeSum.sum builds
 * (src/few/summation/interpolatedmodesum.py:326-337, 99-101) -- and the phase
 * splines of phase_knots (3, length) [Phi_phi, Phi_theta, Phi_r] written as
 * dense-output coefficients pcoef_out (8, length-1, 3).  t (length,) is
 * shared; scratch holds 4*length*(2K+3) doubles.  phase_knots may be NULL
 * (amplitudes only; scratch 4*length*2K doubles).  Asynchronous.
 */
int few_prepare_modes_async(const double *teuk, const double *phase_knots, const double *t,
                            int length, int num_modes, double *interp_out, double *pcoef_out,
                            double *scratch, void *stream);

/*
 * Mode plan for the factorised sum (no reference counterpart; the reference
 * re-reads m/k/n per launch, src/few/cutils/interpolate.cu:591-593).  Built
 * on the host from HOST copies of the int32 mode indices and uploaded to the
 * current device once; reused by every call with the same mode set.
 * Returns NULL on error.  info7 = {K, rows, cols, n_lo, row blocks,
 * k-steps, flags}; flags is a bit mask (1: FEW_PRECISION_MIXED supported,
 * 2: FEW_PRECISION_MIXED16 supported, 4: FEW_PRECISION_FACT64 supported,
 * 8: the coefficient tiles of the factorised sums fold column -n onto n,
 * max|n| + 1 tile columns);
 * unsupported sets (too many rows/columns for the MFMA tiles) can still use
 * FEW_PRECISION_FP64.
 */
void *few_mode_plan_create(const int *h_m, const int *h_k, const int *h_n, int num_modes);
int few_mode_plan_info(const void *plan, int *info7);
void few_mode_plan_destroy(void *plan);
size_t few_plan_workspace_bytes(const void *plan, int init_len, int out_len, int precision);

/*
 * get_waveform (src/few/cutils/interpolate.cu:726-797) with a mode plan and a
 * precision choice, asynchronous on device-resident time arrays.
 * FEW_PRECISION_FP64 runs the reference arithmetic; FEW_PRECISION_MIXED and
 * FEW_PRECISION_MIXED16 the factorised MFMA sum (rel-L2 ~1e-7 against the
 * FP64 result); FEW_PRECISION_FACT64 the factorised sum in FP64 on
 * v_mfma_f64_16x16x4f64 (FP64 tiles, phasors and accumulation; it differs from
 * the reference arithmetic only by the rounding of the phase products,
 * rel-L2 ~1e-10; variant 0 = 16 samples per wave, 4 = 32).  `variant`
 * selects the kernel shape (0 = default).  FEW_PRECISION_MIXED: 0 = 32
 * samples per wave, 2 waves per SIMD; 4 = 64 samples per wave, 1 wave per
 * SIMD.  FEW_PRECISION_MIXED16, a bit set: 4 = 4 waves per workgroup (one
 * per SIMD) instead of 8; 8 = no raised wave priority over the MFMA phase;
 * 16 = the row-streamed contraction (same results bit for bit); 32 = twice
 * as many persistent workgroups as fit at once; other bits are refused
 * (FEW_ERR_ARG).  The autotuner (fastemriwaveforms_amd.autotune) times them.
 * `stage` is a bitmask:
 * 1 = prepare (interval planning, coefficient tiles), 2 = sum (the output
 * kernel), 3 = both; stage 2 reuses what the last stage-1 call on the same
 * workspace prepared, and is refused (FEW_ERR_ARG) unless that call had the
 * same plan, precision, variant, init_len and out_len.  Adding 4 to a stage with the sum (6, 7) overwrites
 * instead of accumulating: every output sample is written (h where an
 * interval covers it, 0 after the last one), so the caller need not zero
 * d_waveform first.  Workspace:
 * few_plan_workspace_bytes() bytes, caller-owned, one call in flight at a
 * time.  No host synchronisation, no allocation: capturable in a hipGraph.
 */
int few_get_waveform_plan_async(void *d_waveform, const double *interp_array,
                                const double *d_phase_interp_t, const double *phase_interp_coeffs,
                                const void *plan, const int *d_m, const int *d_k, const int *d_n,
                                int init_len, int out_len, int num_teuk_modes, const void *d_Ylms,
                                double delta_t, const double *d_h_t, int precision, int variant,
                                int stage, void *workspace, size_t workspace_bytes, void *stream);

/*
 * Batched mode sum: B waveforms with their own knots (L_b), modes (K_b, a
 * subset of any mode set), output lengths (N_b) and dt_b, one launch per
 * stage for the whole batch.  No reference counterpart (the reference sums
 * one waveform per call; its only batching is sequential, in Python,
 * src/few/waveform/base.py:256-268).
 *
 * few_batch_create takes HOST arrays: h_L, h_K, h_N, h_dt (B,) and the
 * int32 mode indices of all waveforms concatenated (sum K_b).  Returns NULL
 * on error.  Packed device buffers, per waveform b in order, in doubles:
 *   t (L_b) | teuk complex (L_b, K_b) | phase knots (3, L_b) |
 *   ylm complex (2 K_b): Y+ then Y- | interp (4, L_b, 2K_b) |
 *   pcoef (8, L_b-1, 3) | wave complex (N_b), accumulated into.
 * few_batch_layout fills totals10 = the 7 packed sizes (doubles), then
 * spline systems, intervals and a supported mask (1: FP64, 2: MIXED16), and
 * optionally offsets (B x 7).
 * few_batch_run_async: precision FEW_PRECISION_FP64 (control run, the
 * reference arithmetic) or FEW_PRECISION_MIXED16; stage is a mask: 1 =
 * splines of every waveform, 2 = interval plan (+ coefficient tiles for
 * MIXED16), 4 = the mode sum, 8 (with 4) = the sum overwrites: every output
 * sample is written (0 after a waveform's last interval), so the outputs
 * need no zeroing first.  Stage 1 overwrites the coefficient tiles, so
 * run 2 after it.  phase_knots may be NULL: pcoef then holds the caller's
 * dense-output phase coefficients (the trajectory's DOPR853 output,
 * src/few/trajectory/inspiral.py:143-147), which stage 1 leaves as they are.
 * Workspace: few_batch_workspace_bytes(), caller-owned.
 * No host synchronisation and no allocation: capturable in a hipGraph.
 */
void *few_batch_create(int B, const int *h_L, const int *h_K, const int *h_N, const double *h_dt,
                       const int *h_m, const int *h_k, const int *h_n);
void few_batch_destroy(void *batch);
int few_batch_layout(const void *batch, long *totals10, long *offsets);
size_t few_batch_workspace_bytes(const void *batch, int precision);
int few_batch_run_async(const void *batch, const double *t, const double *teuk,
                        const double *phase_knots, const double *ylm, double *interp, double *pcoef,
                        void *wave, int precision, int variant, int stage, void *workspace,
                        size_t workspace_bytes, void *stream);

/*
 * Reverse mode (no reference counterpart: the reference has no backward
 * pass).  For a real loss L and the upstream gradient g = dL/dh (complex,
 * out_len), these compute the exact derivatives of get_waveform
 * (src/few/cutils/interpolate.cu:726-797 with make_waveform :444-689) and
 * of interpolate_arrays (:284-325) for fixed knot times.
 *
 * few_waveform_adjoint_async: same mode-sum arguments as get_waveform
 * (device pointers, phase_interp_t and h_t included) plus grad_wave;
 * writes g_interp (4, init_len, 2K) = dL/d(interp_array) in its layout
 * (row init_len-1, which the sum never reads, is zero) and g_phase
 * (4, init_len, 3) = dL/dc_p of the cubic of phase q on interval j at
 * [p][j][q] (the dense-output coefficients equal that cubic; see
 * few_prepare_modes_async).  Segmented per-mode reductions over each
 * interval's samples, in a fixed order (deterministic).  precision:
 * FEW_PRECISION_FP64.  Workspace: few_adjoint_workspace_bytes().
 */
size_t few_adjoint_workspace_bytes(int init_len, int out_len, int num_modes, int precision);
int few_waveform_adjoint_async(const void *grad_wave, const double *interp_array,
                               const double *phase_interp_t, const double *phase_interp_coeffs,
                               const int *d_m, const int *d_k, const int *d_n, int init_len,
                               int out_len, int num_teuk_modes, const void *d_Ylms, double delta_t,
                               const double *h_t, double *g_interp, double *g_phase, int precision,
                               void *workspace, size_t workspace_bytes, void *stream);

/*
 * The same adjoint with a mode plan (few_mode_plan_create), so it can run in
 * mixed precision.  precision FEW_PRECISION_FP64 is few_waveform_adjoint_async;
 * FEW_PRECISION_MIXED16 computes the per-interval moments
 * sum_i sigma^s conj(g_i) u_r v_n (s = 0..6, u_r v_n the factorised phasors)
 * on split-FP16 MFMA with FP32 accumulation per 1024-sample chunk, adds the
 * chunks in FP64 and finishes per mode in FP64.  FEW_PRECISION_FACT64
 * computes the same moments with FP64 operands and FP64 accumulation on the
 * matrix cores (v_mfma_f64_16x16x4f64): FP64-accurate gradients, phase
 * gradients included (their moments cancel by up to ~1e4, which the FP32
 * accumulation of MIXED16 cannot resolve: ~1e-4 relative there).  Both
 * factorised paths need phase_interp_t == h_t (one knot array, as
 * InterpolatedModeSum passes it).
 * Workspace: few_plan_adjoint_workspace_bytes().
 */
size_t few_plan_adjoint_workspace_bytes(const void *plan, int init_len, int out_len, int precision);
int few_waveform_adjoint_plan_async(const void *grad_wave, const double *interp_array,
                                    const double *phase_interp_t, const double *phase_interp_coeffs,
                                    const void *plan, const int *d_m, const int *d_k, const int *d_n,
                                    int init_len, int out_len, int num_teuk_modes, const void *d_Ylms,
                                    double delta_t, const double *h_t, double *g_interp,
                                    double *g_phase, int precision, void *workspace,
                                    size_t workspace_bytes, void *stream);

/*
 * Transpose of the spline build (interpolate_arrays) for one group of
 * systems of `length` >= 4 knots sharing t: g_coef[s*g_sys + p*g_p +
 * i*g_knot] = dL/dc_p(i) of system s -> g_y[s*gy_sys + i*gy_knot] = dL/dy_i.
 * The final-derivative row, the coefficient formulas, the tridiagonal solve
 * (transposed, LAPACK dgtsv's pivoting; the elimination of the shared
 * matrix is done once for all systems) and the right-hand side are reversed
 * in turn.  scratch: length * (nsys + 7) doubles.
 */
int few_spline_adjoint_async(const double *t, const double *g_coef, long g_sys, long g_p,
                             long g_knot, int nsys, int length, double *g_y, long gy_sys,
                             long gy_knot, double *scratch, void *stream);

/*
 * Adjoint of few_prepare_modes_async: g_interp (4, length, 2K) and g_phase
 * (4, length, 3) (as written by few_waveform_adjoint_async) -> grad_teuk
 * (length, K) complex (dL/dRe + i dL/dIm) and, if g_phase is not NULL,
 * grad_phase_knots (3, length).  One launch; scratch as in
 * few_prepare_modes_async (4 * length * (2K + 3) doubles).
 */
int few_modes_adjoint_async(const double *g_interp, const double *g_phase, const double *t,
                            int length, int num_modes, double *grad_teuk, double *grad_phase_knots,
                            double *scratch, void *stream);

/*
 * Replaces CubicSplineInterpolant.__call__ (src/few/summation/
 * interpolatedmodesum.py:139-257; the reference evaluates with numpy/cupy
 * gathers after a per-array searchsorted loop, :145-149).  All device
 * pointers: knots t[s*t_sys + j*t_knot] (length >= 2), coefficient p of array
 * s at knot j coef[p*c_coef + s*c_sys + j*c_knot], points tnew[s*tn_sys + q]
 * (tn_sys = 0 shares one row), out[s*npts + q].  deriv_order 0..3 with the
 * reference's expressions (bitwise equal to numpy on the same coefficients);
 * order 0 fills points outside [t0, t_last] with the edge values.  flags
 * (device int[2], cleared by the caller) receive 1 when some point lies left
 * / right of the range: the caller warns (order 0) or raises (derivatives),
 * as the reference does.
 */
int few_spline_eval_async(const double *t, long t_sys, long t_knot, int length, int ninterps,
                          const double *coef, long c_coef, long c_sys, long c_knot,
                          const double *tnew, long tn_sys, int npts, int deriv_order, double *out,
                          int *flags, void *stream);

/*
 * Replaces DirectModeSum.sum (src/few/summation/directmodesum.py:24-126):
 * h[i] = sum_q Y+_q A[i,q] e^{-i phi_q(i)} + sum_{q: m_q > 0} (-1)^l_q Y-_q
 * conj(A[i,q]) e^{+i phi_q(i)}, phi_q = m Phi_phi + k Phi_theta + n Phi_r, on
 * dense per-sample inputs: teuk (num_samples, num_modes) complex128, ylms
 * (2 num_modes) complex128 (Y+ then Y-, as passed to the reference), the three
 * phases (num_samples,), l/m/k/n int32.  Writes out (num_samples) complex128.
 * FP64 throughout.
 */
int few_direct_mode_sum_async(const void *teuk, const void *ylms, const double *phi_phi,
                              const double *phi_theta, const double *phi_r, const int *l,
                              const int *m, const int *k, const int *n, int num_samples,
                              int num_modes, void *out, void *stream);

/* ---- frequency-domain interpolated sum --------------------------------- */

/* Bytes of workspace few_get_waveform_generic_fd_async needs (the per-(mode,
 * segment) bin counts and their prefix): written to *bytes. */
int few_fd_workspace_bytes(int num_teuk_modes, int num_segments, size_t *bytes);

/* Replaces get_waveform_generic_fd (src/few/cutils/interpolate.hh:33-38,
 * bound at src/few/cutils/pyinterp.pyx:64-92; kernel interpolate.cu:1142-1500).
 * Same arguments, all device pointers:
 *   waveform        (data_length,) complex128, accumulated into; with
 *                   separate_modes (num_teuk_modes, data_length), each mode's
 *                   bins overwritten
 *   interp_array    (4, 2K + 3, init_length) spline coefficients of Re A, Im A,
 *                   f_phi, f_theta, f_r (fdinterp.py:508-512 layout)
 *   phase_interp_t  (init_length,)  phase_interp_coeffs (8, 3, init_length - 1)
 *   m, k, n         (K,) int32      old_time_arr (init_length,) knot times
 *   frequencies     (data_length,)  seg_start_inds / seg_end_inds (K, num_segments)
 *                   int32 inclusive frequency-index ranges
 *   ylm_all         (2K,) complex128, Y- already times (-1)^l
 *   zero_index      index of f = 0; the -m term of bin i lands in bin 2 zero - i
 * delta_t is accepted and unused, as in the reference kernel.  Differences
 * from the reference, both on inputs the reference mishandles: bins outside
 * [0, data_length) are skipped before the frequency is read, and a -m bin
 * outside the array is dropped instead of written out of bounds.  Colliding
 * bins are added with FP64 atomics, so their summation order is not fixed
 * (as with the reference's atomic adds).  Stream-ordered. */
int few_get_waveform_generic_fd_async(void *waveform, const double *interp_array,
                                      const double *phase_interp_t,
                                      const double *phase_interp_coeffs, const int *m,
                                      const int *k, const int *n, int num_teuk_modes,
                                      double delta_t, const double *old_time_arr, int init_length,
                                      int data_length, const double *frequencies,
                                      const int *seg_start_inds, const int *seg_end_inds,
                                      int num_segments, const void *ylm_all, int zero_index,
                                      int include_minus_m, int separate_modes, void *workspace,
                                      size_t workspace_bytes, void *stream);

/* The reference's synchronous call shape (interpolate.hh:33-38): library-owned
 * workspace, returns after the stream has finished. */
int few_get_waveform_generic_fd(void *waveform, double *interp_array, double *phase_interp_t,
                                double *phase_interp_coeffs, int *m, int *k, int *n,
                                int num_teuk_modes, double delta_t, double *old_time_arr,
                                int init_length, int data_length, double *frequencies,
                                int *seg_start_inds, int *seg_end_inds, int num_segments,
                                void *ylm_all, int zero_index, int include_minus_m,
                                int separate_modes, void *stream);

/* ---- bicubic amplitude interpolation ----------------------------------- */

/* Bytes of workspace few_interp2d_async needs for mx points (their B-spline
 * intervals and weights). */
size_t few_interp2d_workspace_bytes(int mx);

/* Replaces interp2D_wrap (src/few/cutils/AmpInterp2D.hh, kernel
 * AmpInterp2D.cu:86-206, bound at pyampinterp2D.pyx:9-29).  Same arguments,
 * device pointers: z (num_indiv_c * mx) written, z[c_i * mx + i] the spline
 * c_i (len_indiv_c coefficients, FITPACK layout) at (x[i], y[i]); knots tx
 * (nx), ty (ny), degrees kx, ky.  Arguments outside the box of the knots are
 * clamped to it, as in fpbisp.  Errors (the reference throws or reads out of
 * bounds): mx != my, degrees outside 0..5, fewer than 2 (k + 1) knots, or
 * len_indiv_c below (nx - kx - 1)(ny - ky - 1).  Bit-identical to the
 * reference's arithmetic.  Stream-ordered. */
int few_interp2d_async(double *z, const double *tx, int nx, const double *ty, int ny,
                       const double *c, int kx, int ky, const double *x, int mx, const double *y,
                       int my, int num_indiv_c, int len_indiv_c, void *workspace,
                       size_t workspace_bytes, void *stream);

/* The reference's synchronous call shape: library-owned workspace, returns
 * after the stream has finished. */
int few_interp2D_wrap(double *z, const double *tx, int nx, const double *ty, int ny, double *c,
                      int kx, int ky, const double *x, int mx, const double *y, int my,
                      int num_indiv_c, int len_indiv_c, void *stream);

/* ---- Augmented Analytic Kludge summation -------------------------------- */

/* Bytes of workspace few_aak_waveform_async needs (segment start indices). */
size_t few_aak_workspace_bytes(int init_len);

/* Replaces get_waveform of src/few/cutils/gpuAAK.hh (kernel gpuAAK.cu:181-577,
 * bound as pyWaveform at gpuAAKWrap.pyx:17-33).  Same arguments, all device
 * pointers here (d_h_t included): waveform (out_len,) complex128, each sample
 * of a used knot segment overwritten with hI - i hII; interp_array (8,
 * init_len - 1, 6) dense-output coefficients of (p, e, Y, Phi_phi,
 * Phi_theta, Phi_r); masses in solar masses, dist in Gpc; nmodes harmonics;
 * mich selects the long-wavelength LISA response.  Segments follow
 * gpuAAK.cu:458-483; when no knot reaches the last sample, the segments stop
 * at the last knot (the reference reads past its arrays there).
 * Stream-ordered. */
int few_aak_waveform_async(void *waveform, const double *interp_array, double M_phys,
                           double S_phys, double mu, double qS, double phiS, double qK,
                           double phiK, double dist, int nmodes, int mich, int init_len,
                           int out_len, double delta_t, const double *d_h_t, void *workspace,
                           size_t workspace_bytes, void *stream);

/* The reference's synchronous call shape (gpuAAK.hh: get_waveform), h_t on the
 * host; library-owned workspace; returns after the stream has finished. */
int few_aak_get_waveform(void *waveform, double *interp_array, double M_phys, double S_phys,
                         double mu, double qS, double phiS, double qK, double phiK, double dist,
                         int nmodes, int mich, int init_len, int out_len, double delta_t,
                         double *h_t, void *stream);

/* Frees the library-owned workspaces of the synchronous drop-in entries. */
int few_release_workspace(void);

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): few_get_waveform_plan_async with the same arguments
 * except the workspace, which comes from a library-owned static device
 * array of FEW_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends).
 * A call whose few_plan_workspace_bytes() exceeds that size is refused with
 * FEW_ERR_LIMIT (FP64 needs ~36 B per knot; MIXED16 adds the coefficient
 * tiles: 5.4 MB for 120 knots and the 3843-mode set, 18 MB for 200 knots
 * and n in [-55, 55]; FACT64 its FP64 tiles, 11 MB and 32 MB).  The call allocates nothing: no hipMalloc / hipFree
 * and no allocating library routine.  Every call on one GPU shares that
 * GPU's array, which keeps the last call's contents: a stage-2 call sums
 * with what the last stage-1 call on the same GPU prepared, and calls that
 * overlap on one GPU (other streams or host threads) race.  The first call
 * on a device looks the array's address up once (hipGetSymbolAddress).
 */
#define FEW_NOALLOC_SCRATCH_BYTES (32u << 20)
int few_get_waveform_plan_noalloc(void *d_waveform, const double *interp_array,
                                  const double *d_phase_interp_t,
                                  const double *phase_interp_coeffs, const void *plan,
                                  const int *d_m, const int *d_k, const int *d_n, int init_len,
                                  int out_len, int num_teuk_modes, const void *d_Ylms,
                                  double delta_t, const double *d_h_t, int precision, int variant,
                                  int stage, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FEW_ROCM_H */
