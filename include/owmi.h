/*
 * owmi: MI355X (gfx950) kernels for optweight's conjugate-gradient
 * Wiener-filter path, exported as a plain C ABI.
 *
 * Conventions (all functions):
 *   - return 0 on success, nonzero on error; owmi_last_error() returns a
 *     thread-local message for the last failure. No C++ exception crosses
 *     the ABI.
 *   - every pointer named *dev* or described as "device" is a device
 *     (HBM) pointer owned by the caller; host arrays are marked "host".
 *   - every call is asynchronous on the given hipStream_t (passed as void*;
 *     NULL means the default stream).
 *   - dtype: OWMI_F32 (float / complex float) or OWMI_F64 (double /
 *     complex double). Complex arrays are interleaved (re, im).
 *   - alm arrays use optweight's triangular m-major layout with
 *     mmax = lmax: index(l, m) = m (2 lmax + 1 - m) / 2 + l,
 *     nelem = (lmax^2 + 3 lmax) / 2 + 1, components contiguous blocks of
 *     nelem (optweight src/optweight_alm_c_utils.c:3-10).
 *   - maps are ring-major with per-ring offsets (optweight
 *     map_utils.MapInfo, map_utils.py:19-297).
 *
 * Reference interfaces replaced are cited per function (paths relative to
 * the optweight repository).
 */
#ifndef OWMI_H
#define OWMI_H

#ifdef __cplusplus
extern "C" {
#endif

#define OWMI_F32 0
#define OWMI_F64 1

/* Message of the last failed call on this thread ("" if none). */
const char *owmi_last_error(void);

/* ABI version (increments on incompatible change). */
int owmi_abi_version(void);

/* ------------------------------------------------------------------ */
/* Streaming kernels on alm                                            */
/* ------------------------------------------------------------------ */

/*
 * out[t, i, lm] = sum_j lmat[i, j, l(lm)] in[t, j, lm]   (diag == 0)
 * out[t, i, lm] = lmat[i, l(lm)] in[t, i, lm]             (diag != 0)
 * for t < ntrans, arrays (ntrans, ncomp, nelem). lmat has the alm precision
 * (float for OWMI_F32, double for OWMI_F64), shape (ncomp, ncomp, lmax+1)
 * or (ncomp, lmax+1). out may equal in (in-place).
 * Replaces lmul_dp / lmul_sp / lmul_diag_dp / lmul_diag_sp /
 * lmul_inplace_dp / lmul_inplace_sp / lmul_diag_inplace_dp /
 * lmul_diag_inplace_sp (include/optweight_alm_c_utils.h:17-90, 116-140)
 * and the leading-dimension loop of cython/alm_c_utils.pyx:157-250.
 */
int owmi_lmul(int dtype, const void *lmat, int diag, const void *in,
              void *out, int lmax, int ncomp, long long ntrans,
              void *stream);

/*
 * out[t, c, p] = x[t, c, p] + dcomp[c] mask[c, p] (a[t, c, p] - b[t, c, p])
 * on maps (ntrans, ncomp, npix) of dtype's real type; x, b, dcomp (ncomp) and
 * mask (ncomp, npix) may be NULL (0, 0, 1, 1); any array may alias out. The
 * array operations of a multigrid V-cycle in one pass (multigrid.py:
 * 400-447): the Jacobi update x += D (r - G x), the masked residual
 * M (r - G x) that is restricted, and x += M (interpolated correction).
 */
int owmi_map_fused(int dtype, void *out, const void *x, const void *a,
                   const void *b, const void *dcomp, const void *mask, int ncomp,
                   long long npix, long long ntrans, void *stream);

/*
 * out[t, i, lm] = sum_j La[i, j, l] a[t, j, lm] + sum_j Lb[i, j, l] b[t, j, lm]
 * with full (ncomp, ncomp, lmax+1) matrices of the alm precision, ncomp <= 4;
 * lmat_a NULL takes La as the identity. a and b may alias out. The sum of
 * two EllMatVecAlm products in one pass: the signal term S^-1 p added to
 * the beam-weighted noise term of CGWienerMap.a_matrix (solvers.py:217-242),
 * and H (tau B)^T n + H S^-1 a, the last stages of PseudoInvPreconditioner
 * (preconditioners.py:126-249).
 */
int owmi_lmul_add(int dtype, const void *lmat_a, const void *a,
                  const void *lmat_b, const void *b, void *out, int lmax,
                  int ncomp, long long ntrans, void *stream);

/*
 * owmi_lmul_add with the real-field <b_t, out_t> of every transform t
 * written to dev_dot[t] (float64; as owmi_alm_dot_batch of b with out
 * would give, summed in another fixed order): the last pass of the fused
 * CGWienerMap.a_matrix (solvers.py:217-242) with the CG coefficient
 * <p, A p> of pixell.utils.CG.step folded in (b = p, out = A p), so p and
 * A p are not read again. ntrans <= 1024; b may not alias out. Uses the
 * stream's partial-sum buffer of owmi_alm_dot_batch.
 */
int owmi_lmul_add_dot(int dtype, const void *lmat_a, const void *a,
                      const void *lmat_b, const void *b, void *out, int lmax,
                      int ncomp, long long ntrans, double *dev_dot, void *stream);

/*
 * Copy l <= lmax_out, m <= lmax_out of (ntrans*ncomp) alm rows at lmax into
 * rows at lmax_out. Replaces trunc_alm_dp / trunc_alm_sp
 * (include/optweight_alm_c_utils.h:92-101, 142-146).
 */
int owmi_trunc_alm(int dtype, const void *alm, void *alm_out, int lmax,
                   int lmax_out, long long nrows, void *stream);

/*
 * alm[r, lm(l, m)] += w_ell[l] wlm[r, lm'(l, m)] for l, m <= lmax_w over
 * nrows rows; w_ell has the alm precision. Replaces wlm2alm_dp /
 * wlm2alm_sp (include/optweight_alm_c_utils.h:41-58, 128-133).
 */
int owmi_wlm2alm(int dtype, const void *w_ell, const void *wlm, void *alm,
                 int lmax_w, int lmax_a, long long nrows, void *stream);

/*
 * Real-field inner product over nrows alm rows at lmax (complex, dtype):
 *   *dev_out = 2 Re sum a conj(b) - Re sum_{m=0} a conj(b)
 * accumulated in double; dev_out is a device double. Deterministic
 * (fixed-order two-pass reduction). The partial sums go to a buffer of the
 * library private to the (device, stream) pair, allocated on the first call
 * on that stream (also inside a HIP-graph recording, outside the graph):
 * dots on different streams may overlap; this and every other reduction
 * here (owmi_alm_dot_batch, owmi_fourier_dot) follow the same rule.
 * Replaces alm_utils.contract_almxblm (optweight/alm_utils.py:9-44).
 */
int owmi_alm_dot(int dtype, const void *a, const void *b, int lmax,
                 long long nrows, double *dev_out, void *stream);

/* ------------------------------------------------------------------ */
/* Streaming kernels on maps                                           */
/* ------------------------------------------------------------------ */

/*
 * In-place imap[k, offsets[r] + j * stride[r]] *= weight[r] for j < nphi[r],
 * r < nrow, k < nmaps (maps map_stride elements apart). weight, nphi,
 * offsets, stride are device arrays. For OWMI_F32 the weight is rounded
 * to float before the multiply, as the reference does.
 * Replaces _apply_ringweight_core_sp / _apply_ringweight_core_dp
 * (include/optweight_map_c_utils.h:10-13).
 */
int owmi_apply_ringweight(int dtype, void *imap, const double *weight,
                          const long long *nphi, const long long *offsets,
                          const long long *stride, long long nrow,
                          long long nmaps, long long map_stride,
                          void *stream);

/*
 * Pixel-diagonal matrix-vector product on ntrans maps sharing one matrix,
 * out may equal in:
 *   kind 0: out[t, i, p] = mat[i, p] in[t, i, p]           (mat (ncomp, npix))
 *   kind 1: out[t, i, p] = sum_j mat[i, j, p] in[t, j, p]  (mat (ncomp, ncomp, npix))
 * Replaces mat_utils.matvec (optweight/mat_utils.py:475-532) as used by
 * operators.PixMatVecMap (operators.py:162-204).
 */
int owmi_pix_matvec(int dtype, const void *mat, int kind, const void *in,
                    void *out, int ncomp, long long npix, long long ntrans,
                    void *stream);

/* ------------------------------------------------------------------ */
/* Conjugate-gradient vector updates (scalars stay on the device)      */
/* ------------------------------------------------------------------ */

/*
 * alpha = *dev_num / *dev_den (doubles on device);
 * x += alpha p; r -= alpha q; over n complex elements.
 * The x/r update of pixell.utils.CG.step as called through
 * solvers.CGWienerMap (solvers.py:127-141).
 */
int owmi_cg_xr(int dtype, void *x, void *r, const void *p, const void *q,
               const double *dev_num, const double *dev_den, long long n,
               void *stream);

/* beta = *dev_num / *dev_den; p = z + beta p over n complex elements. */
int owmi_cg_p(int dtype, void *p, const void *z, const double *dev_num,
              const double *dev_den, long long n, void *stream);

/*
 * Mixed-precision x/r update of a reliable-update CG: alpha = *dev_num /
 * *dev_den; x += alpha p accumulated in x's precision (dtype_x), r -= alpha q
 * in dtype_r (p, q and r share dtype_r); n complex elements.
 */
int owmi_cg_xr_mixed(int dtype_x, void *x, int dtype_r, void *r, const void *p,
                     const void *q, const double *dev_num, const double *dev_den,
                     long long n, void *stream);

/*
 * nbatch independent real-field inner products (as owmi_alm_dot), product b
 * over rows [b * nrows, (b + 1) * nrows) of a and b; dev_out[b] (device
 * doubles). 1 <= nbatch <= 1024. The per-right-hand-side scalars of a
 * batch of pixell CG solves (solvers.py:127-141) and of
 * alm_utils.contract_almxblm (alm_utils.py:9-44) applied per RHS.
 */
int owmi_alm_dot_batch(int dtype, const void *a, const void *b, int lmax,
                       long long nrows, long long nbatch, double *dev_out,
                       void *stream);

/*
 * owmi_alm_dot_batch without any allocation: the same products, kernels and
 * summation order (results equal owmi_alm_dot_batch bit for bit), with the
 * partial sums in a fixed 8 KiB device array of the library (one per GPU,
 * held until the process ends) instead of a buffer allocated on a stream's
 * first call. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP-graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Limit: nbatch <= 1024 (the
 * array holds 1024 partial sums); larger values fail with an error.
 * Every call on one GPU shares that array: calls that overlap in time on
 * one GPU (different streams or threads) race and must be serialised by
 * the caller.
 */
int owmi_alm_dot_batch_noalloc(int dtype, const void *a, const void *b, int lmax,
                               long long nrows, long long nbatch, double *dev_out,
                               void *stream);

/*
 * Batched mixed-precision x/r update: for each b < nbatch, alpha_b =
 * dev_num[b] / dev_den[b] (0 when dev_den[b] == 0); x_b += alpha_b p_b in
 * dtype_x, r_b -= alpha_b q_b in dtype_r, each over n complex elements
 * (arrays (nbatch, n)). The x/r update of pixell.utils.CG.step per RHS.
 */
int owmi_cg_xr_batch(int dtype_x, void *x, int dtype_r, void *r,
                     const void *p, const void *q, const double *dev_num,
                     const double *dev_den, long long n, long long nbatch,
                     void *stream);

/*
 * owmi_cg_xr_batch on alm (nbatch, nrows, nelem(lmax)) with the real-field
 * <r_b, r_b> of every new residual written to dev_rr[b] (float64, as
 * owmi_alm_dot_batch of r with itself would give, summed in another fixed
 * order): the x/r update and the residual norm of pixell.utils.CG.step in
 * one pass over r. nbatch <= 1024. Uses the stream's partial-sum buffer of
 * owmi_alm_dot_batch.
 */
int owmi_cg_xr_rr_batch(int dtype_x, void *x, int dtype_r, void *r, const void *p,
                        const void *q, const double *dev_num, const double *dev_den, int lmax,
                        long long nrows, long long nbatch, double *dev_rr, void *stream);

/* Batched p_b = z_b + beta_b p_b, beta_b = dev_num[b] / dev_den[b] (0 when
 * dev_den[b] == 0); arrays (nbatch, n). */
int owmi_cg_p_batch(int dtype, void *p, const void *z, const double *dev_num,
                    const double *dev_den, long long n, long long nbatch,
                    void *stream);

/* y = a x + b y over n complex elements with host scalars a, b. */
int owmi_axpby(int dtype, double a, const void *x, double b, void *y,
               long long n, void *stream);

/* Convert n complex elements between precisions (dst may not alias src). */
int owmi_convert(int dtype_src, const void *src, int dtype_dst, void *dst,
                 long long n, void *stream);

/* ------------------------------------------------------------------ */
/* 2-D Fourier coefficients of flat-sky (CAR) maps                     */
/* ------------------------------------------------------------------ */
/* Coefficient arrays are (nlead, ny, nlx) complex, nlx = nx / 2 + 1 for
 * maps of nx columns. */

/*
 * out[c, p] = sum_j M[c, j, p] in[j, p] (kind 1, M (ncomp, ncomp, npix)
 * real, ncomp <= 4) or d[c * diag_stride + p] in[c, p] (kind 0; diag_stride
 * npix for a per-component diagonal, 0 for one shared by all components),
 * for ntrans consecutive (ncomp, npix) complex arrays; in may alias out.
 * Replaces dft.fmul_2d (optweight/dft.py:332-365).
 */
int owmi_fmul(int dtype, const void *mat, int kind, const void *in, void *out,
              int ncomp, long long npix, long long ntrans, long long diag_stride,
              void *stream);

/* The full-matrix product of owmi_fmul for any ncomp, through a caller
 * buffer tmp of the output's size (dft.py:332-365). */
int owmi_fmul_generic(int dtype, const void *mat, const void *in, void *out,
                      void *tmp, int ncomp, long long npix, long long ntrans,
                      void *stream);

/*
 * Kernel slices of the Fourier-wavelet transform: block j of the packed
 * buffer small (offset nlead * desc[8j], shape (nlead, ny_j = desc[8j+1],
 * nx_j = desc[8j+2])) receives full[b, rows[desc[8j+4] + r], desc[8j+3] + c] times
 * kern[desc[8j+5] + r nx_j + c]; the launch has nblocks = sum_j
 * ceil(ny_j nx_j / 256) thread blocks, scale j's first one at desc[8j+6]
 * (ascending in j). desc, rows and kern are device arrays. The gather half of wavtrans.f2wav
 * (optweight/wavtrans.py:687-712, dft.slice_fmap dft.py:498-533).
 */
int owmi_fwav_gather(int dtype, const void *full, long long nlead, long long ny,
                     long long nlx, int nscale, const long long *desc,
                     const int *rows, const void *kern, long long nblocks,
                     void *small, void *stream);

/*
 * The adjoint: full[b, y, x] = sum over scales j (in table order) whose
 * slice holds (y, x) of small_j[b, invrows[j ny + y], x - x0_j] times the
 * kernel value; invrows[j ny + y] is the block row of full row y or -1.
 * Writes every element of full. The accumulate half of wavtrans.wav2f
 * (wavtrans.py:714-745, dft.add_to_fmap dft.py:535-560).
 */
int owmi_fwav_scatter(int dtype, const void *small, long long nlead,
                      long long ny, long long nlx, int nscale,
                      const long long *desc, const int *invrows,
                      const void *kern, void *full, void *stream);

/*
 * Real-field conventions around library FFT passes, in place. mode 0,
 * after a real-to-complex transform: zero the imaginary parts of the
 * self-conjugate coefficients (ly in {0, ny/2}, lx in {0, nx/2}). mode 1,
 * between the inverse transform along y and the complex-to-real pass
 * along x: zero the imaginary parts of column 0 and, for even nx, of
 * column nlx - 1 of every row. Used by dft.rfft / dft.irfft
 * (dft.py:13-67) for the semantics of a 2-D real FFT.
 */
int owmi_fourier_fix(int dtype, void *data, long long nlead, long long ny,
                     long long nlx, int nx_even, int mode, void *stream);

/*
 * Real-field inner product of half-plane coefficients into *dev_out
 * (float64, device): 2 sum Re(f g*) - sum over column 0 - [nx even] sum
 * over column nlx - 1. nx_parity 1 / 0 states the parity of nx; -1 takes
 * nx as even when f has no imaginary part at (ly = 0, and ly = ny/2 for
 * even ny; lx = nlx - 1). Replaces dft.contract_fxg (dft.py:449-496).
 */
int owmi_fourier_dot(int dtype, const void *f, const void *g, long long nlead,
                     long long ny, long long nlx, int nx_parity, double *dev_out,
                     void *stream);

/* ------------------------------------------------------------------ */
/* Spherical harmonic transforms                                       */
/* ------------------------------------------------------------------ */

typedef struct owmi_sht_plan owmi_sht_plan;

/*
 * Build a transform plan for a ring geometry (host arrays of length nrings:
 * theta (colatitude), weight (per-pixel quadrature weight), nphi, offsets,
 * phi0) and triangular alm with mmax = lmax. Mirror rings
 * (|theta_i + theta_j - pi| < 1e-12) share one Legendre evaluation. The
 * plan holds its tables on the current device and grows a workspace as
 * transforms of more rows are requested.
 * Replaces the geometry arguments optweight passes to
 * ducc0.sht.experimental.{synthesis, adjoint_synthesis}
 * (optweight/sht.py:56-93, 141-174).
 */
int owmi_sht_plan_create(owmi_sht_plan **plan, long long nrings,
                         const double *theta, const double *weight,
                         const long long *nphi, const long long *offsets,
                         const double *phi0, int lmax);

/*
 * Plan over a subset of a map's rings: the arrays describe nrings rings of a
 * larger ring-major map whose components are map_npix (>= 1) pixels apart
 * (offsets index that map). Synthesis writes only these rings' pixels;
 * the adjoint reads only them, which equals the full adjoint when the
 * map is zero on every other ring. Mirror pairs are formed within the
 * subset. Serves masked operators whose input and output vanish outside
 * the given rings (optweight's multigrid G = M Y d_ell Y^T M on
 * the unobserved pixels, multigrid.py:13-471, operators.py:991-1102).
 */
int owmi_sht_plan_create_rings(owmi_sht_plan **plan, long long nrings,
                               const double *theta, const double *weight,
                               const long long *nphi, const long long *offsets,
                               const double *phi0, int lmax,
                               long long map_npix);

int owmi_sht_plan_destroy(owmi_sht_plan *plan);

/* Number of pixels of the plan's geometry. */
long long owmi_sht_plan_npix(const owmi_sht_plan *plan);

/*
 * Launch choices of a plan's Legendre kernels, set by an autotuner. No
 * counterpart in the reference (ducc0 picks its blocking internally); the
 * values only change which kernel variant runs and the order of float
 * sums, never the transform computed. A setting applies to transforms of
 * at least min_batch right-hand sides: the entry with the largest
 * min_batch <= ntrans is used, and a value of 0 keeps the library's rule.
 */
#define OWMI_TUNE_MFMA_MIN 1 /* float32 batches from this size on the matrix cores (>= 1) */
#define OWMI_TUNE_ADJ_CT0 2  /* matrix-core adjoint, spin 0: 2 or 4 column tiles (16 / 32 transforms per block) */
#define OWMI_TUNE_ADJ_CT2 3  /* matrix-core adjoint, spin 2: 2 or 4 column tiles (8 / 16 transforms per block) */
#define OWMI_TUNE_SYNTH_NB 4 /* vector synthesis: at most 1, 2 or 4 transforms per thread */
#define OWMI_TUNE_ADJ_NB 5   /* vector adjoint: at most 1, 2 or 4 transforms per block row */
#define OWMI_TUNE_ADJ_K0 6   /* vector adjoint, spin 0: 1, 2 or 4 ring pairs per lane */
#define OWMI_TUNE_ADJ_K2 7   /* vector adjoint, spin 2: 1 or 2 ring pairs per lane */
#define OWMI_TUNE_MFMA64_MIN 8 /* float64 batches from this size take the Legendre synthesis and adjoint on the float64 matrix cores (>= 1; default 8) */
int owmi_sht_plan_tune(owmi_sht_plan *plan, long long min_batch, int param, int value);

/* Drop every tuned setting of a plan (back to the library's rules). */
int owmi_sht_plan_tune_reset(owmi_sht_plan *plan);

/*
 * Synthesis of ntrans transforms of one spin group (spin 0: 1 component,
 * spin 2: 2 components (E, B) -> (Q, U), HEALPix convention).
 * alm: transform t, component c at alm + (t * alm_tstride + c * nelem);
 * map: transform t, component c at map + (t * map_tstride + c * npix)
 * (strides in elements of the complex / real type). weighted != 0
 * multiplies by the ring weights (W Y).
 * Replaces sht.alm2map (optweight/sht.py:95-178) for one spin group.
 */
int owmi_sht_synthesis(owmi_sht_plan *plan, int dtype, int spin,
                       const void *alm, long long alm_tstride, void *map,
                       long long map_tstride, long long ntrans, int weighted,
                       void *stream);

/*
 * Adjoint synthesis (Y^T), or Y^T W when weighted != 0; same layouts.
 * Replaces sht.map2alm (optweight/sht.py:10-93) for one spin group.
 */
int owmi_sht_adjoint(owmi_sht_plan *plan, int dtype, int spin,
                     const void *map, long long map_tstride, void *alm,
                     long long alm_tstride, long long ntrans, int weighted,
                     void *stream);

/*
 * alm_out = Y^T D Y alm_in for one spin group, D a per-component pixel
 * diagonal (pixdiag: component c of transform t at
 * pixdiag + c * npix + t * pix_tstride; pix_tstride may be 0 to share D).
 * The map never leaves on-chip memory: ring synthesis FFT, the pixel
 * multiply and the ring analysis FFT run in one kernel.
 * Fuses the map-space middle of operators.PixMatVecAlm.call
 * (optweight/operators.py:614-643) and of CGWienerMap.a_matrix
 * (solvers.py:217-242) when mask and inverse noise are pixel-diagonal.
 */
int owmi_sht_pixel_sandwich(owmi_sht_plan *plan, int dtype, int spin,
                            const void *alm_in, long long alm_in_tstride,
                            void *alm_out, long long alm_out_tstride,
                            const void *pixdiag, long long pix_tstride,
                            long long ntrans, void *stream);

/*
 * Spectral form of owmi_sht_pixel_sandwich for a pixel diagonal D shared by
 * every transform: the ring FFT stage becomes a convolution over m with
 * per-ring real multipliers of D (two FFTs of length M per ring pair
 * instead of four Bluestein FFTs); same result up to rounding.
 *
 * owmi_sht_sandwich_length: M (the smallest power of two >= 4 mmax + 1),
 *   or 0 when the plan's ring FFTs do not cover it (rings shorter than
 *   2 mmax + 1); the spectral entry points then fail.
 * owmi_sht_sandwich_multipliers: mult[c][ring][k] (ncomp x nrings x M
 *   values of dtype) from pixdiag (ncomp x npix of dtype), computed in
 *   float64; once per diagonal.
 * owmi_sht_pixel_sandwich_spectral: alm_out = Y^T D Y alm_in for one spin
 *   group with the multipliers of its components (mult points at the
 *   group's first component; component stride nrings * M).
 * Same reference interfaces as owmi_sht_pixel_sandwich
 * (optweight/operators.py:614-643, solvers.py:217-242).
 */
long long owmi_sht_sandwich_length(owmi_sht_plan *plan);
int owmi_sht_sandwich_multipliers(owmi_sht_plan *plan, int dtype,
                                  const void *pixdiag, int ncomp, void *mult,
                                  void *stream);
int owmi_sht_pixel_sandwich_spectral(owmi_sht_plan *plan, int dtype, int spin,
                                     const void *alm_in,
                                     long long alm_in_tstride, void *alm_out,
                                     long long alm_out_tstride,
                                     const void *mult, long long ntrans,
                                     void *stream);

/*
 * The Legendre stage alone, between alm and the plan's ring-coefficient
 * workspace (direction 0: alm -> workspace, the first half of synthesis;
 * direction 1: workspace -> alm, the second half of the adjoint). Used to
 * time and profile the Legendre kernels in isolation; the workspace holds
 * whatever the previous transform on this plan left (zero when new).
 * Building block of sht.alm2map / sht.map2alm (optweight/sht.py:10-178).
 */
int owmi_sht_legendre(owmi_sht_plan *plan, int dtype, int spin, int direction,
                      void *alm, long long alm_tstride, long long ntrans,
                      void *stream);

#ifdef __cplusplus
}
#endif

#endif /* OWMI_H */
