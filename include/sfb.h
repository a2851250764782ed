/*
 * sfb.h -- C ABI of libsfb_mi355x.so, the MI355X-native catalogue -> SFB
 * coefficient transform (cat2amln) and the host setup it needs.
 *
 * Conventions
 *  - Every pointer argument is a plain host pointer unless the function name
 *    ends in _device; those take device pointers and an optional hipStream_t
 *    passed as void* (NULL = the plan's own stream).
 *  - Arrays follow the reference's Julia memory layout (column-major):
 *      rtp       3 x N, (r, theta, phi) per galaxy, i.e. r0 th0 ph0 r1 ...
 *      win dense npix x (lmax+1) x nmax   (win_rhat_ln, src/windows.jl:244-257)
 *      win_b     (lmax+1) x nmax          (separable w_ln, src/windows.jl:260-270)
 *      knl       nmax x (lmax+1), NaN where there is no mode
 *      anlm      getnlmsize complex numbers, interleaved (re, im), ordered
 *                n slowest, then l, then m = 0..l (src/modes.jl:80-82,223-232)
 *  - Basis-function spline tables are concatenated per (n,l) pair in the same
 *    n-major, l-minor order ("map order"); knot_off has nmaps+1 entries.
 *  - Pixel indices are 0-based HEALPix ring-scheme indices.
 *  - Every int-returning function returns SFB_OK or an SFB_ERR_* code; the text
 *    of the last error on the calling thread is returned by sfb_last_error().
 *  - A plan may be used by one host thread at a time; different plans may run
 *    concurrently.  Calls are synchronous at return unless named _device.
 *  - Fault injection for error-path tests: the environment variable SFB_FAULT=alloc
 *    makes every device allocation fail (SFB_ERR_DEVICE), SFB_FAULT=maxevals gives the
 *    k_nl root finder no evaluations (SFB_ERR_SETUP, as the reference's doublesecant_method
 *    returning NaN after maxevals).  Read at each use; unset, nothing changes.
 */
#ifndef SFB_MI355X_H
#define SFB_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFB_OK 0
#define SFB_ERR_ARG 1        /* bad argument (null pointer, size, kind) */
#define SFB_ERR_LMAX 2       /* lmax > 4 nside: src/healpix_helpers.jl:51-53 error("exiting") */
#define SFB_ERR_NONFINITE 3  /* non-finite a_nlm: src/cat2anlm.jl:308 @assert */
#define SFB_ERR_SETUP 4      /* root finding or normalisation failed: SphericalBesselGNLs.jl:126-127,:416 */
#define SFB_ERR_DEVICE 5     /* HIP runtime error (includes device out-of-memory) */
#define SFB_ERR_NOMEM 6      /* host allocation failed */

#define SFB_PREC_FP64 0      /* plain path: reference method, FP64 throughout */
#define SFB_PREC_MIXED 1     /* FP32 spline values, delta maps and phi-transform; the first Legendre
                                analysis on FP64 MFMA over FP32 ring coefficients, the Jacobi passes
                                on FP32 MFMA (FP32 iterate and aliasing buffers); FP64 setup, segment
                                search, pixel sums, lambda tables, GEMM epilogues, aliasing sums and
                                the l = 0 maps' monopoles.  Every (n,l) block within 1e-6 of FP64. */
#define SFB_PREC_BF16X3 2    /* SFB_PREC_MIXED with the Jacobi GEMMs on bf16 MFMA: each FP32 operand split
                                into hi + mid + lo bf16 terms, six products, FP32 accumulation */
#define SFB_PREC_BF16X2 3    /* the same with hi + lo terms, three products (~2^-16 per product; the
                                Jacobi steps only correct the FP64-MFMA first analysis, so every
                                (n,l) block still stays within 1e-6 of FP64) */

#define SFB_WIN_DENSE 0      /* win_a = W[npix, lmax+1, nmax] */
#define SFB_WIN_SEPARABLE 1  /* win_a = mask[npix], win_b = w_ln[lmax+1, nmax] */

typedef struct sfb_modes sfb_modes;
typedef struct sfb_plan sfb_plan;

/* Text of the last error raised on this thread ("" if none). */
const char* sfb_last_error(void);

/* Library version string. */
const char* sfb_version(void);

/* ---- host setup: AnlmModes ---------------------------------------------- */

/* AnlmModes(kmax, rmin, rmax; nside, boundary=potential): src/modes.jl:122-158,
 * roots src/SphericalBesselGNLs.jl:309-336, c/d :404-421, splines :558-579.
 * nside <= 0 selects estimate_nside(lmax) (src/modes.jl:68). */
int sfb_modes_create_kmax(sfb_modes** out, double kmax, double rmin, double rmax, int64_t nside);

/* AnlmModes(nmax, lmax, rmin, rmax; nside): src/modes.jl:161-175,
 * roots src/SphericalBesselGNLs.jl:294-306. */
int sfb_modes_create_nl(sfb_modes** out, int64_t nmax, int64_t lmax, double rmin, double rmax,
                        int64_t nside);

/* Sizes: nmax, lmax, nside, number of (n,l) maps, getnlmsize (src/modes.jl:183-189),
 * total spline knots. */
int sfb_modes_sizes(const sfb_modes* m, int64_t* nmax, int64_t* lmax, int64_t* nside,
                    int64_t* nmaps, int64_t* nlmsize, int64_t* nknots);

/* Copies out lmax_n[nmax], nmax_l[lmax+1], knl[nmax*(lmax+1)], cnl, dnl (same shape),
 * knot_off[nmaps+1], knot_x/y/ypp[nknots].  Any output pointer may be NULL. */
int sfb_modes_tables(const sfb_modes* m, int64_t* lmax_n, int64_t* nmax_l, double* knl, double* cnl,
                     double* dnl, int64_t* knot_off, double* knot_x, double* knot_y, double* knot_ypp);

void sfb_modes_destroy(sfb_modes* m);

/* Spline1D{3} build (src/Splines.jl:206-271): fills ypp[n] from knots x[n], y[n]. */
int sfb_spline_build(int64_t n, const double* x, const double* y, double* ypp);

/* Spline1D{3} evaluation (src/Splines.jl:61-69) and derivative (:367-377) at r[nr]. */
int sfb_spline_eval(int64_t n, const double* x, const double* y, const double* ypp, int64_t nr,
                    const double* r, double* out, int32_t derivative);

/* ---- the transform -------------------------------------------------------- */

/* Plan for cat2amln on one device.  The spline tables are taken as given
 * (from the reference's amodes.basisfunctions.gnl[n,l+1].x/.y/.ypp, or from
 * sfb_modes_tables), so the knots are bit-identical to the caller's.
 * niter is the map2alm Jacobi iteration count (src/healpix_helpers.jl:50: 3).
 * Fails with SFB_ERR_LMAX when lmax > 4 nside. */
int sfb_plan_create(sfb_plan** out, int64_t nside, int64_t lmax, int64_t nmax, const int64_t* lmax_n,
                    const int64_t* knot_off, const double* knot_x, const double* knot_y,
                    const double* knot_ypp, int32_t niter, int32_t precision, int32_t device);

void sfb_plan_destroy(sfb_plan* p);

/* Sizes the plan works with: number of maps, getnlmsize, npix. */
int sfb_plan_sizes(const sfb_plan* p, int64_t* nmaps, int64_t* nlmsize, int64_t* npix);

/* Switches the Legendre-step precision of an existing plan (SFB_PREC_*). */
int sfb_plan_set_precision(sfb_plan* p, int32_t precision);

/* cat2amln(rθϕ, amodes, nbar, win_rhat_ln, weight): src/cat2anlm.jl:257-310.
 * w == NULL means unit weights (the reference default ones(N)).  N may be 0
 * (field2anlm_v1, src/cat2anlm.jl:315-321).  anlm_out holds 2*nlmsize doubles. */
int sfb_cat2amln(sfb_plan* p, int64_t N, const double* rtp, const double* w, double nbar,
                 int32_t win_kind, const double* win_a, const double* win_b, double* anlm_out);

/* Same transform with every array already in device memory; enqueued on
 * stream (NULL = the plan's stream) and synchronised before returning, so the
 * non-finite check can be reported. */
int sfb_cat2amln_device(sfb_plan* p, int64_t N, const double* d_rtp, const double* d_w, double nbar,
                        int32_t win_kind, const double* d_win_a, const double* d_win_b,
                        double* d_anlm_out, void* stream);

/* ---- batched mocks ----------------------------------------------------------- */

/* cat2amln of B catalogues through one plan (typically mock catalogues of one survey
 * geometry): catalogue b has N_b[b] galaxies, stored after catalogue b-1 in rtp (3 x sum N)
 * and w (sum N, or NULL for unit weights), and its own nbar_b[b]; the window is shared.
 * anlm_out holds B x 2*nlmsize doubles, catalogue-major.  Equivalent to B calls of
 * sfb_cat2amln (the reference loops cat2amln over mocks); N_b and nbar_b are host
 * arrays in both variants.  An error names the failing catalogue. */
int sfb_cat2amln_batch(sfb_plan* p, int32_t B, const int64_t* N_b, const double* rtp, const double* w,
                       const double* nbar_b, int32_t win_kind, const double* win_a, const double* win_b,
                       double* anlm_out);

/* Same with rtp, w, the window and the output in device memory. */
int sfb_cat2amln_batch_device(sfb_plan* p, int32_t B, const int64_t* N_b, const double* d_rtp, const double* d_w,
                              const double* nbar_b, int32_t win_kind, const double* d_win_a,
                              const double* d_win_b, double* d_anlm_out, void* stream);

/* Batched mocks as one captured HIP graph (no reference counterpart): with enable = 1 the
 * batch entry points capture the whole batch into a graph and launch it, and replay that
 * graph when a later batch call repeats the same sizes, n-bar values, pointers, window kind
 * and precision (no plan buffer reallocated since).  Before a capture the largest mock runs
 * once eagerly so that no buffer grows under capture.  enable = 0 enqueues the mocks
 * directly (the default); enable < 0 leaves the setting.  *capture_ms (if not NULL)
 * receives the wall time of the last capture and instantiation. */
int sfb_plan_set_batch_graph(sfb_plan* p, int32_t enable, double* capture_ms);

/* ---- catalogue preparation and the C_l(n, n') consumer ------------------------ */

/* xyz2rtp(xyz) (src/SphericalFourierBesselDecompositions.jl:517-532): 3 x N Cartesian
 * positions -> 3 x N (r, theta, phi), phi = atan(y/r, x/r) in (-pi, pi]. */
int sfb_xyz2rtp(int32_t device, int64_t N, const double* xyz, double* rtp);

/* winweights2galweights(weights, wmodes, rθϕ) (src/cat2anlm.jl:218-238): w_out[i] =
 * weights[bin, pix] for the nr x npix (column-major) weights, bin from the bin edges
 * rbounds[nr+1] (get_rbounds(wmodes), r == rmax -> the last bin).  A radius outside the
 * bins is an error (the reference's BoundsError). */
int sfb_winweights2galweights(int32_t device, int64_t nside, int64_t nr, const double* rbounds,
                              const double* weights, int64_t N, const double* rtp, double* w_out);

/* apply_window(rθϕ, win, rmin, rmax, r, Δr) (src/windows.jl:169-210): keeps galaxy i when
 * rmin <= r <= rmax and u[i] <= win[idx_r, pix] / maximum(win), idx_r = ceil((r - rmin)/Δr)
 * (0 -> 1, > nr -> rejected).  u[N] are the caller's uniform draws (the reference draws
 * them from its own RNG).  rtp_out receives the kept galaxies in input order, *n_out
 * their count. */
int sfb_apply_window(int32_t device, int64_t nside, int64_t nr, double rmin, double rmax, double dr,
                     const double* win, int64_t N, const double* rtp, const double* u, double* rtp_out,
                     int64_t* n_out);

/* apply_window (as sfb_apply_window) with every array in device memory and no allocation:
 * no counterpart in the reference; it is for callers that cannot allocate device memory
 * (code under HIP-graph capture, tight serving loops).  Its scratch is one static device
 * array of about 16.1 MiB per GPU, held until the process ends, so N is at most 4194304
 * (2^22; SFB_ERR_ARG beyond).  The window maximum is taken on the device.  Enqueued on
 * stream (NULL = the default stream) and not synchronised: *d_n_out (device memory) and
 * d_rtp_out are valid once the stream has reached this point.  Calls that overlap on one
 * GPU share the scratch and race. */
int sfb_apply_window_noalloc(int32_t device, int64_t nside, int64_t nr, double rmin, double rmax, double dr,
                             const double* d_win, int64_t N, const double* d_rtp, const double* d_u,
                             double* d_rtp_out, int64_t* d_n_out, void* stream);

/* The HEALPix pixel window w_l, l = 0 .. lmax, that pixwin(cmodes)
 * (src/SphericalFourierBesselDecompositions.jl:158-167) reads from HEALPix's tables, computed
 * from its definition: w_l^2 = (1/Npix) sum_p (4 pi/(2l+1)) sum_m |<Y_lm>_p|^2, the pixel
 * means taken over the centres of the 4^depth sub-pixels at nside 2^depth and extrapolated
 * from depths depth-1 and depth (depth 0: 5 up to nside 32, 4 up to 256, 3 above; the
 * work grows as nside^2 16^depth lmax).  lmax <= min(4 nside, 3000), nside <= 4096.
 * Runs on `device`; synchronous. */
int sfb_pixwin(int64_t nside, int64_t lmax, int32_t depth, int32_t device, double* w_out);

/* amln2clnn(anlm1, anlm2, cmodes) (src/SphericalFourierBesselDecompositions.jl:130-150)
 * for the (l, n1, n2) triples lnn[3 x nlnn] of a ClnnModes over the plan's modes:
 * C = (Re(a1_l0 a2_l0) + 2 sum_{m>=1} Re(a1_lm conj(a2_lm))) / (2l + 1). */
int sfb_amln2clnn(sfb_plan* p, int64_t nlnn, const int64_t* lnn, const double* anlm1, const double* anlm2,
                  double* clnn);

/* The same for a ClnnModes over two AnlmModes (src/modes.jl:396-399, ClnnModes{false}):
 * anlm1 in the order of plan pa's modes, anlm2 in that of pb's, n1 a mode of pa and n2 of
 * pb (src/SphericalFourierBesselDecompositions.jl:133-134).  Both plans on one device. */
int sfb_amln2clnn_cross(sfb_plan* pa, sfb_plan* pb, int64_t nlnn, const int64_t* lnn, const double* anlm1,
                        const double* anlm2, double* clnn);

/* ---- adjoint ----------------------------------------------------------------- */

/* Vector-Jacobian product of cat2amln at (rθϕ, w, nbar): given dL_danlm (2*nlmsize
 * doubles, (dL/dRe a, dL/dIm a) interleaved in a_nlm order), returns dL/dr_i and
 * dL/dw_i (N each, input order) and dL/dnbar.  dL/dθ and dL/dφ are zero (the
 * pixel assignment is piecewise constant).  The window cotangent is not returned:
 * a is affine in the window with dL/dW[p, l, n] = -dbar_{nl}(p), the synthesised
 * cotangent map this function forms internally, but no entry point hands it out.
 * A non-finite angle is an error (SFB_ERR_ARG), as in the forward call.  With a plan of
 * SFB_PREC_MIXED the adjoint runs in FP32 (Jacobi recursion, synthesis, delta-bar maps,
 * per-galaxy products) with FP64 sums over the maps; otherwise in FP64.
 * g'(r) is the spline derivative of src/Splines.jl:367-377
 * (derivative(gnl, n, l, r), src/GNLs.jl:110).  No counterpart in the reference,
 * which has no derivative of cat2amln; it backs the ChainRulesCore.rrule of the
 * Julia shim.  Any of dL_dr, dL_dw, dL_dnbar may be NULL. */
int sfb_cat2amln_vjp(sfb_plan* p, int64_t N, const double* rtp, const double* w, double nbar,
                     const double* dL_danlm, double* dL_dr, double* dL_dw, double* dL_dnbar);

/* Same with device arrays (d_dr or d_dw may be NULL), enqueued on stream
 * (NULL = the plan's stream) and synchronised before returning. */
int sfb_cat2amln_vjp_device(sfb_plan* p, int64_t N, const double* d_rtp, const double* d_w, double nbar,
                            const double* d_dL_danlm, double* d_dr, double* d_dw, void* stream);

/* ---- window and field transforms (the path's input producers) ----------------- */

/* win_rhat_ln(win, wmodes, amodes) (src/windows.jl:244-270) on the device.
 * r[nr] are the window's radial bin centres and dr their width (wmodes.r, wmodes.Δr).
 * Dense: win_a is the nr x npix window (column-major, r fastest) and out is
 * W[npix, lmax+1, nmax] = dr * sum_k win[k, p] r_k^2 g_nl(r_k), NaN where l > lmax_n[n]
 * (one MFMA GEMM per n).  Separable (SeparableArray(phi, mask), :260-270): win_a = phi[nr]
 * and out = w_ln[lmax+1, nmax]; the mask passes through unchanged. */
int sfb_win_rhat_ln(sfb_plan* p, int64_t nr, const double* r, double dr, int32_t win_kind, const double* win_a,
                    double* out);
int sfb_win_rhat_ln_device(sfb_plan* p, int64_t nr, const double* d_r, double dr, int32_t win_kind,
                           const double* d_win_a, double* d_out, void* stream);

/* field2anlm(f_xyz, wmodes, amodes) (src/cat2anlm.jl:323-358, the default field2anlm_v2):
 * f is nr x npix (column-major); every shell goes through the plan's map2alm, then
 * f_nlm += dr r^2 g_nl(r) a_lm.  out holds 2*nlmsize doubles. */
int sfb_field2anlm(sfb_plan* p, int64_t nr, const double* r, double dr, const double* f, double* out);
int sfb_field2anlm_device(sfb_plan* p, int64_t nr, const double* d_r, double dr, const double* d_f, double* d_out,
                          void* stream);

/* anlm2field(f_nlm, wmodes, amodes) (src/cat2anlm.jl:383-422): per shell
 * a_lm = sum_n g_nl(r) f_nlm, then alm2map; f_out is nr x npix (column-major). */
int sfb_anlm2field(sfb_plan* p, int64_t nr, const double* r, const double* f_nlm, double* f_out);
int sfb_anlm2field_device(sfb_plan* p, int64_t nr, const double* d_r, const double* d_f_nlm, double* d_f_out,
                          void* stream);

/* ---- per-phase timing ------------------------------------------------------ */

#define SFB_PHASE_CATALOG 0    /* ang2pix, the two stable sorts, gathers, pixel segments */
#define SFB_PHASE_SCATTER 1    /* spline evaluation + ordered per-pixel sums -> delta maps */
#define SFB_PHASE_PHI 2        /* phi-transform of the rings with 4 nside pixels (one FFT kernel) */
#define SFB_PHASE_ANALYSIS 3   /* Legendre analysis GEMMs (1 + niter launches) */
#define SFB_PHASE_SYNTHESIS 4  /* Legendre synthesis GEMMs (niter launches) */
#define SFB_PHASE_RESIDUAL 5   /* ring-Fourier residual with aliasing (niter launches) */
#define SFB_PHASE_PACK 6       /* a_nlm packing and the non-finite check */
#define SFB_PHASE_SEGMENT 7    /* per knot group and galaxy: spline segment and t (shared by the group's maps) */
#define SFB_PHASE_MONOPOLE 8   /* mixed precision: FP64 monopole correction of the l = 0 maps */
#define SFB_PHASE_ADJOINT 9    /* sfb_cat2amln_vjp(_device) after its catalogue stage */
#define SFB_PHASE_PHI_POLAR 10 /* phi-transform of the polar rings (one DFT kernel) */
#define SFB_NPHASE 11

/* Turns per-phase HIP-event timing of sfb_cat2amln(_device) on or off and
 * clears the totals.  No reference counterpart (the reference prints @time). */
int sfb_plan_profile(sfb_plan* p, int32_t enable);

/* Accumulated milliseconds and launch counts per phase (SFB_NPHASE entries each)
 * and the number of profiled calls since sfb_plan_profile. */
int sfb_plan_phase_times(const sfb_plan* p, double* ms, int64_t* launches, int64_t* calls);

/* The phi-transform tile shapes the plan's tile search chose at creation (no reference
 * counterpart; SFB_TUNE=0 in the environment skips the search and keeps the defaults):
 * tiles[0..5] = maps per FFT block FP64, FFT threads per block FP64, the same two for FP32,
 * maps per polar block at the largest polar ring FP64, FP32 (fixed by the ring sizes, not
 * searched); *tune_ms = the search's wall time. */
int sfb_plan_tiles(const sfb_plan* p, int32_t* tiles, double* tune_ms);

/* ---- pieces of the path, exposed for parity checks ------------------------ */

/* ang2pixRing (Healpix.jl, called at src/cat2anlm.jl:263) on the GPU, 0-based. */
int sfb_ang2pix_ring(int32_t device, int64_t nside, int64_t N, const double* theta, const double* phi,
                     int32_t* pix_out);

/* The delta map of map q (the (n,l) pair in map order) as the reference forms it
 * before map2alm (src/cat2anlm.jl:284-291), length npix: FP64 and bit-exact on an
 * SFB_PREC_FP64 plan; on the other precisions the FP32 map their scatter writes,
 * widened to double. */
int sfb_delta_map(sfb_plan* p, int64_t N, const double* rtp, const double* w, double nbar,
                  int32_t win_kind, const double* win_a, const double* win_b, int64_t q,
                  double* map_out);

/* map2alm with niter refinements (Healpix.jl map2alm!, src/healpix_helpers.jl:50-56)
 * of nmap FP64 maps [nmap][npix] on the GPU with the plan's nside/lmax and
 * precision; alm_out is [nmap][nalm] interleaved complex in the Healpix m-major
 * index m(2 lmax+1-m)/2 + l.  nmap must not exceed the plan's map count. */
int sfb_map2alm(sfb_plan* p, int64_t nmap, const double* maps, double* alm_out);

/* The batched Legendre GEMM kernel on one matrix: C[n][i] = sum_k A[k][i] B[k][n],
 * A is K x rows, B is K x N, C is N x rows, all row-major.  No reference
 * counterpart: it checks the MFMA kernel that carries the map2alm Legendre sums. */
int sfb_gemm_f64(int32_t device, int64_t rows, int64_t K, int64_t N, const double* A, const double* B, double* C);

/* The same GEMM through the FP32-operand MFMA kernel (operands rounded to FP32, FP32
 * accumulation on the MFMA, FP64 epilogue), operands stored as FP64. */
int sfb_gemm_f32(int32_t device, int64_t rows, int64_t K, int64_t N, const double* A, const double* B, double* C);

/* The kernels of the SFB_PREC_MIXED Jacobi steps and syntheses: A, B (and Cin) rounded to
 * FP32 and stored as FP32, 8-deep k-tiles, C stored as FP32 and returned widened.  With
 * 0 < K1 < K, rows k >= K1 of A and B go through the split-K operands (the merged
 * (I - C0) + residual step); Cin (N x rows, or NULL) is added: C = A^T B + Cin.  No
 * reference counterpart: a kernel check. */
int sfb_gemm_f32_stored(int32_t device, int64_t rows, int64_t K, int64_t K1, int64_t N, const double* A,
                        const double* B, const double* Cin, double* C);

/* The same GEMM with split-BF16 operands in `terms` bf16 terms (2: hi + lo, three products;
 * 3: hi + mid + lo, six products), the kernel of the SFB_PREC_BF16X2 / X3 Jacobi steps. */
int sfb_gemm_bf16(int32_t device, int64_t rows, int64_t K, int64_t N, const double* A, const double* B, double* C,
                  int32_t terms);

/* Launch shapes of the ring transforms for a geometry, computed on the host (no device):
 * out[0] polar ring pairs on the forward radix-4 kernel (nside - 1), out[1] its LDS bytes
 * per block (FP64), out[2] its fewest maps per block; out[3] polar ring pairs on the
 * inverse radix-4 kernel, out[4] ring pairs on the inverse direct sum, out[5] the inverse
 * radix-4 LDS bytes, out[6] its fewest maps per block, out[7] maps per inverse-FFT block.
 * SFB_ERR_ARG when the forward blocks do not fit in lds_bytes.  No reference counterpart:
 * it lets a host-only test check the shapes for every nside. */
int sfb_ring_kernel_shapes(int64_t nside, int64_t lmax, int64_t lds_bytes, int64_t* out);

/* The per-ring phi-transform of nmap maps [nmap][npix]: x_out[((m*2+par)*K + k)*2*nmap + 2q+c],
 * K = 2 nside ring pairs, par 0 = F_north + F_south, 1 = F_north - F_south,
 * F_m = sum_j map_j e^{-i m phi_j} (the ring DFT inside libsharp's map2alm). */
int sfb_phi_transform(sfb_plan* p, int64_t nmap, const double* maps, double* x_out);

#ifdef __cplusplus
}
#endif

#endif /* SFB_MI355X_H */
