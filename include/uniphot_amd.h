/*
 * uniphot_amd -- C ABI of the MI355X (gfx950) forced-photometry fit path.
 *
 * Drop-in boundary for the hot path of uniphot's ``tract7dt.fit`` island
 * architecture.  All array arguments are device pointers (row-major,
 * C-contiguous) owned by the caller; the library owns nothing between calls.
 * Every entry point is stream-ordered on ``stream`` (a hipStream_t; NULL =
 * the default stream) and performs no host synchronisation unless stated.
 *
 * Status codes: UPH_OK, UPH_ERR_INVALID_ARGUMENT, UPH_ERR_OOM (the caller may
 * halve the island batch and retry, as the reference halves a chunk on
 * RESOURCE_EXHAUSTED, tract7dt/fit/grid_gpu/island_solver.py:3134-3230) and
 * UPH_ERR_INTERNAL.  ``uph_last_error()`` returns the calling thread's last
 * error message.
 */
#ifndef UNIPHOT_AMD_H
#define UNIPHOT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UPH_OK 0
#define UPH_ERR_INVALID_ARGUMENT 1
#define UPH_ERR_OOM 2
#define UPH_ERR_INTERNAL 3

#define UPH_PRECISION_FP32 0 /* pixel arithmetic in float32 (production) */
#define UPH_PRECISION_FP64 1 /* pixel arithmetic in float64 (GPU oracle)  */

/* Library version (major*10000 + minor*100 + patch). */
int uph_version(void);

/* Last error message of the calling thread ("" when none). */
const char* uph_last_error(void);

/*
 * Legacy LM constants of one island class.  Mirrors the knobs the reference
 * bakes into its compiled step program (make_class_kernels,
 * tract7dt/fit/grid_gpu/island_solver.py:1874-1922; values
 * island_plan.py:55-74 and island_solver.py:146-152).
 */
typedef struct uph_lm_options {
  float reg;            /* ridge on real member diagonals (1e-9)          */
  float pos_step_max;   /* per-step clip on x, y (0.5 px)                  */
  float shape_step_max; /* per-step clip on logre, ee1, ee2 (0.15)         */
  float un_step_max;    /* per-step clip on u_n (0.5)                      */
  float ftol_rel;       /* prong 1: dchi2 < ftol_rel * max(chi2, 1) (1e-4) */
  float step_tol;       /* prong 2: position step < step_tol px (0.1)      */
  float shape_step_tol; /* prong 2: shape steps < shape_step_tol (0.01)    */
  int max_steps;        /* LM steps per island (8 main + 22 re-queue = 30) */
  int resume;           /* 0: fresh state (lam 1e-3, chi2 at p0, counters 0).
                           1: resume each island from p0 and the LM state the
                           caller wrote into the outputs chi2, lam, consec,
                           nrej and nst (run_class_chunk's init_state,
                           island_solver.py:2902-2930); conv and stall start
                           false                                               */
} uph_lm_options;

/*
 * One island class (static shape) and its chunk of islands.  Replaces the
 * per-class compiled program family of ``_make_bucket_kernels`` plus
 * ``run_class_chunk`` (island_solver.py:221-1848, :2793-3131) for the
 * legacy step protocol.  Layouts follow ``class_host_arrays``
 * (island_plan.py:682-699) and the padded field stacks
 * (island_stage.py:1711-1751).
 */
typedef struct uph_island_class {
  int n_isl;      /* islands in this call                                  */
  int N;          /* nsrc class (member slots per island), <= 64           */
  int B;          /* window side (bbox class), multiple of 16              */
  int T;          /* stamp tier (odd)                                      */
  int n_terms;    /* PSF polynomial terms (1, 3, 6, 10, 15 or 21)          */
  int n_bands;    /* bands                                                 */
  int n_live_max; /* max live parameters per island (2N star, 6N mixed)   */
  int Hp, Wp;     /* padded field-stack height / width                     */
  int mixed;      /* 0: star-only class, 1: class with extended members    */
} uph_island_class;

typedef struct uph_island_inputs {
  const float* dstack;    /* (n_bands, Hp, Wp) data - frozen background    */
  const float* wstack;    /* (n_bands, Hp, Wp) inverse variance (0 = bad)  */
  const float* psf_coefs; /* (n_bands, n_terms, T, T) planar PSF coefs     */
  const float* psf_geo;   /* (n_bands, 4) x_origin, x_scale, y_origin, y_scale */
  const double* sersic;   /* Sersic mixture table block (float64): n_knots, n_min,
                             n_max, K, knots, amps (n_knots, K), log_vars
                             (n_knots, K), core (n_knots); NULL for star-only */
  const int32_t* ox;      /* (n_isl) window origin x in the padded stack   */
  const int32_t* oy;      /* (n_isl)                                        */
  const float* x0b;       /* (n_isl) window origin x in image coordinates  */
  const float* y0b;       /* (n_isl)                                        */
  const int32_t* stype;   /* (n_isl, N) 0 STAR, 1 EXP, 2 DEV, 3 SERSIC     */
  const float* n_frozen;  /* (n_isl, N) Sersic index when not free          */
  const uint8_t* free_n;  /* (n_isl, N)                                     */
  const uint8_t* mmask;   /* (n_isl, N) real member slots                   */
  const uint8_t* keep;    /* (n_isl, N, 6) fitted parameters                */
  const float* p0;        /* (n_isl, 6N) seeds x, y, logre, ee1, ee2, u_n   */
  const float* lower;     /* (n_isl, 6N) box bounds                         */
  const float* upper;     /* (n_isl, 6N)                                    */
  const int32_t* order;   /* (n_isl) processing order (work queue), or NULL */
} uph_island_inputs;

typedef struct uph_island_outputs {
  float* p;          /* (n_isl, 6N) fitted parameters                       */
  float* fluxes;     /* (n_isl, n_bands, N) VarPro member fluxes            */
  float* ferr;       /* (n_isl, n_bands, N) sqrt(diag(A^-1))                */
  double* chi2;      /* (n_isl) float64: the accept test compares float64 */
  double* lam;       /* (n_isl) final LM damping, float64                   */
  uint8_t* conv;     /* (n_isl)                                             */
  uint8_t* stall;    /* (n_isl) converged by 4 consecutive rejects          */
  int32_t* nst;      /* (n_isl) LM steps taken                              */
  int32_t* consec;   /* (n_isl)                                             */
  int32_t* nrej;     /* (n_isl)                                             */
  int32_t* n_good;   /* (n_isl) good pixels summed over bands               */
} uph_island_outputs;

/* Bytes of device workspace one persistent workgroup needs for a class. */
size_t uph_island_scratch_per_block(int N, int B, int T, int n_live_max, int precision);

/*
 * Fit every island of the chunk: fresh LM state (lam 1e-3, chi2 at p0), up
 * to max_steps legacy steps with device-side accept/reject and convergence,
 * then final fluxes.  One persistent kernel; islands are pulled from a work
 * queue in ``order``.  ``scratch`` holds n_blocks workgroups' workspace
 * (n_blocks = scratch_bytes / uph_island_scratch_per_block(...)); ``counter``
 * is one device int32 the library resets.  Replaces
 * ``run_class_chunk`` (island_solver.py:2793) with fresh state.
 */
int uph_island_fit(const uph_island_class* cls, const uph_island_inputs* in,
                   const uph_island_outputs* out, const uph_lm_options* opt, int precision,
                   void* scratch, size_t scratch_bytes, int32_t* counter, void* stream);

/*
 * Heterogeneous island batch: islands of different classes (N, B, stamp
 * tier) in one persistent launch, pulled from the work queue in ``order``
 * (longest first is the caller's choice).  Member arrays of
 * uph_island_inputs / outputs are concatenated over islands: island i owns
 * member slots [mem_off[i], mem_off[i] + N_i) (stype, n_frozen, free_n,
 * mmask; keep, p0, lower, upper and p at 6x that) and flux elements
 * [flux_off[i], flux_off[i] + n_bands * N_i).  Per-island scalars (ox, oy,
 * x0b, y0b, order and the scalar outputs) are indexed by island.
 * ``psf_coefs`` of the inputs is ignored: tier t uses psf_coefs[t].
 */
typedef struct uph_island_batch {
  int n_isl, n_bands, n_terms, Hp, Wp;
  int max_B;                   /* largest window side in the batch                */
  int mixed;                   /* any extended member                             */
  int n_tiers;                 /* 1..3                                            */
  int tier_T[3];               /* stamp size of each tier                         */
  const float* psf_coefs[3];   /* (n_bands, n_terms, T_t, T_t) per tier           */
  const int32_t* shape;        /* (n_isl, 4): N, B, tier index, n_live_max        */
  const int64_t* mem_off;      /* (n_isl)                                          */
  const int64_t* flux_off;     /* (n_isl)                                          */
} uph_island_batch;

/* ``scratch_per_block`` = max over the batch's shapes of
   uph_island_scratch_per_block (a multiple of 256). */
int uph_island_fit_batch(const uph_island_batch* batch, const uph_island_inputs* in,
                         const uph_island_outputs* out, const uph_lm_options* opt, int precision,
                         void* scratch, size_t scratch_bytes, size_t scratch_per_block,
                         int32_t* counter, void* stream);

/*
 * Band-split form of uph_island_fit_batch, for islands whose cost would make
 * one workgroup the critical path of a batch: every evaluation pass (chi2
 * at p0, H and g at p, chi2 at the trial, final fluxes) runs as
 * (island, band) items spread over all compute units, and a per-island
 * launch sums the band blocks in band order and takes the legacy LM step.
 * Same inputs, outputs and results (bit for bit) as uph_island_fit_batch;
 * the LM loop is 4 * max_steps + 3 launches on ``stream`` with no host
 * synchronisation.  Replaces run_class_chunk's per-step program
 * (island_solver.py:2890-2977) with band-parallel evaluation.
 *   item_isl, item_band (n_items = n_isl * n_bands): every (island, band)
 *     pair once, in the order the work queue should take them;
 *   hg_off (n_isl): first double of island i's n_bands blocks of
 *     (n_live_max^2 + n_live_max + 2) doubles each, inside hg (hg_doubles);
 *   p_try: laid out like outputs->p;
 *   counters: uph_island_split_counters(max_steps) int32 (reset here; the
 *     last one is set to 1 if an island's or a galaxy's blocks did not fit).
 * With the member pre-render, scratch_per_block must also cover
 * uph_prerender_scratch_bytes(T) (float32) or uph_prerender64_scratch_bytes(T)
 * (float64) of every stamp tier.
 */
typedef struct uph_island_split {
  int n_items;
  const int32_t* item_isl;
  const int32_t* item_band;
  const int64_t* hg_off;
  double* hg;
  size_t hg_doubles;
  float* p_try;
  int32_t* counters;
  void* const* eval_events; /* NULL, or 6 * (2 * max_steps + 2 + post)
                               hipEvent_t: per evaluation pass (chi2 at p0,
                               then H/g and trial per step, then final
                               fluxes, then the post-fit pass) one at its
                               start, one after the pre-render taps/spectra
                               launch, one after the extended-member planes,
                               one after the star planes, one after the
                               large-island Gram launch (before the
                               (island, band) evaluation) and one at its end */
  /* member pre-render (n_pre = 0 renders members in place): before each
     pass every real member of the active islands is rendered and shifted
     by (member, band) items into its planes, and the evaluation only
     places their window part.  With precision UPH_PRECISION_FP64, pbuf and
     fspec hold doubles (pbuf_elems, fspec_elems, pre_off and pre_spec count
     doubles), an extended member's spectra block is
     uph_prerender64_spectra_doubles(T) long, gdt and gtap may be NULL, and
     post, gram_bf16 and big_n must be 0 */
  int n_pre;
  int n_pre_ext;              /* members [0, n_pre_ext) are extended, the
                                 rest stars                                    */
  const int32_t* pre_isl;     /* (n_pre) island                                */
  const int32_t* pre_slot;    /* (n_pre) member slot within the island         */
  const int64_t* pre_off;     /* (n_pre) first float of its n_bands x np planes
                                 of min(T, B)^2 floats in pbuf, np = 7 for an
                                 extended member (model and 6 tangents), 3 for
                                 a star (model, d/dx, d/dy); a plane holds the
                                 part of the stamp inside the island window,
                                 stored from that part's corner                */
  const int64_t* pre_spec;    /* (n_pre) first float of its 3 x (pad/2+1) x pad
                                 Fourier spectra in fspec; -1 for stars        */
  const int32_t* pre_of_slot; /* (total member slots) pre-render index, -1
                                 for padding slots                             */
  float* pbuf;
  size_t pbuf_elems;
  float* fspec;
  size_t fspec_elems;
  double* gdt;                /* (n_pre, 9) scratch                            */
  float* gtap;                /* (n_pre, 28) scratch: Lanczos taps per pass    */
  /* post-fit pass (post = 1; needs the pre-render): after the final fluxes,
     one more evaluation at the converged parameters and a per-island
     covariance launch (island_solver.py:1484-1715 final_errors,
     final_flux_marg, source_metrics without priors):
       post_off (n_isl): first double of the island's block of
         nLc * (n_bands * N + nLc) doubles in post_buf (post_doubles);
       post_var (member slots, 7): marginal variances of x, y, logre, ee1,
         ee2, u and cov(ee1, ee2) (1 and 0 where a parameter is not live);
       post_dvar: laid out like outputs->fluxes, the flux variance increment
         of marginalising over extended members' parameters (0 for stars);
       post_metrics (member slots, n_bands, 6): per band sum m, sum (f m)^2,
         sum (model - f m) f m, sum m over zero-weight pixels,
         sum m w (d - model)^2 and f, m the unit model on the window;
       post_kappa (n_isl): 1-norm condition number of the band-summed
         Gauss-Newton matrix scaled to a unit diagonal (1 for an island
         with nothing to fit) -- how strongly an LM step amplifies float32
         model error; a caller refits islands above its threshold on the
         float64 kernels.                                                  */
  int post;
  const int64_t* post_off;
  double* post_buf;
  size_t post_doubles;
  float* post_var;
  float* post_dvar;
  float* post_metrics;
  /* gram_bf16 = 1 (needs the pre-render): the island Gram matrices on
     bfloat16 matrix cores (operands rounded to nearest even, float32
     accumulation flushed to float64 every 256 products); the float64 flux
     solve then takes three refinement steps on float64 residual
     projections instead of one.  gram_unsettled (n_isl int32, zeroed by
     the caller) is set to 1 for an island where the last step still moved
     a flux by more than 1e-6 of itself (Gram error times condition number
     near 1); the caller refits those islands with gram_bf16 = 0. */
  int gram_bf16;
  int32_t* gram_unsettled;
  /* Gram of the large islands on many workgroups (float32 Gram only;
     big_n = 0 turns it off): islands with big_slot[isl] = s >= 0 (s <
     big_n) get their (island, band) Gram computed before each evaluation,
     one 16 x 16 column tile pair per wave, into big_part + ((s * n_bands +
     band) * big_ntp + pair) * 256 doubles (big_ntp >= the island's tile
     pairs t (t + 1) / 2, t = ceil((1 + N + nLc) / 16)); big_isl (big_n)
     names the island of each slot.  The evaluation then reads the Gram
     instead of forming it in one workgroup; it is the same Gram, summed in
     the same order. */
  int big_n;
  int big_ntp;
  const int32_t* big_slot;
  const int32_t* big_isl;
  double* big_part;
  float* post_kappa;          /* (n_isl), see post above                       */
  /* optional (NULL: not written): with post, the band sums of the
     Gauss-Newton H (nL x nL over the island's live parameters, members in
     slot order, parameters x, y, logre, ee1, ee2, u_n within a member) and
     g (nL) at the converged parameters, from post_hg + post_hg_off[i] --
     what the host's stall probes (island_solver.py:1335-1465) need        */
  double* post_hg;
  const int64_t* post_hg_off;
} uph_island_split;

size_t uph_island_split_counters(int max_steps);
/* Per-workgroup scratch of the member pre-render for stamp tier T (0 when
   the pre-render does not support T), and the spectra floats of one
   extended member. */
size_t uph_prerender_scratch_bytes(int T);
size_t uph_prerender_spectra_floats(int T);
/* The same for the float64 member pre-render: per-workgroup scratch bytes
   for stamp tier T, and the doubles of one extended member's spectra block
   (F, F1, Fn on the (pad, pad/2+1) grid, then core, dcore, the nine
   ellipse-term derivatives and dn/du).  Sizing helpers of this build's
   pre-render launches; the reference renders members inside its per-island
   programs (island_solver.py:441-562) and has no counterpart. */
size_t uph_prerender64_scratch_bytes(int T);
size_t uph_prerender64_spectra_doubles(int T);
int uph_island_fit_split(const uph_island_batch* batch, const uph_island_inputs* in,
                         const uph_island_outputs* out, const uph_lm_options* opt, int precision,
                         const uph_island_split* split, void* scratch, size_t scratch_bytes,
                         size_t scratch_per_block, void* stream);

/*
 * Batched unit-flux stamp renderer: PSF polynomial at (x, y), galaxy
 * mixture convolution for extended sources, unit-sum renormalisation when
 * ``renorm`` (all-star grids only), Lanczos-3 shift by the remainder of
 * (pos - origin).  Replaces render_source_stamps / render_star_stamps /
 * forced_phot_sparse.render_stamps (source_model.py:331-390,
 * forced_phot_sparse.py:126-220).
 *   src_band (n) : band index of each stamp; x, y (n) full-image position;
 *   gx0, gy0 (n) : origin the sub-pixel remainder is taken against;
 *   stype (n) or NULL (all stars); logre, ee1, ee2, sersic_n (n) :
 *   morphology of extended sources; out (n, T, T).
 * Extended sources, and stamps too large for one workgroup's LDS (T > 111),
 * need ``scratch`` of uph_render_scratch_bytes(T, k) bytes (k persistent
 * workgroups, k >= 1).
 */
size_t uph_render_scratch_bytes(int T, int n_blocks);
int uph_render_stamps(int n, int T, int n_terms, const float* psf_coefs, const float* psf_geo,
                      const double* sersic, const int32_t* src_band, const float* x,
                      const float* y, const float* gx0, const float* gy0,
                      const int32_t* stype, const float* logre, const float* ee1,
                      const float* ee2, const float* sersic_n, int renorm, float* out,
                      void* scratch, size_t scratch_bytes, void* stream);

/*
 * Direct render backend (replaces render_source_stamp_at with
 * render_backend="direct", source_model.py:218-256, :298-303): STAR = PSF
 * polynomial at (x, y); EXP (stype 1) = the pixel-centre 6-Gaussian
 * exponential profile (unit sum) convolved directly with the PSF stamp
 * ('same' mode) and renormalised to unit box sum; then the Lanczos-3 shift
 * by the remainder of (pos - origin).  Arguments as uph_render_stamps
 * (stype NULL = all stars); DEV/SERSIC have no direct form and must not be
 * passed.  Stamps above 63 px need scratch of
 * uph_render_direct_scratch_bytes(T, k) bytes (k >= 1 workgroups).
 */
size_t uph_render_direct_scratch_bytes(int T, int n_blocks);
int uph_render_stamps_direct(int n, int T, int n_terms, const float* psf_coefs,
                             const float* psf_geo, const int32_t* src_band, const float* x,
                             const float* y, const float* gx0, const float* gy0,
                             const int32_t* stype, const float* logre, const float* ee1,
                             const float* ee2, float* out, void* scratch, size_t scratch_bytes,
                             void* stream);

/*
 * Batched per grid-band linear solve (member fluxes + polynomial background
 * in one normal-equation system, float64 solve and inverse diagonal),
 * replacing solve_band_sparse / launch_band_sparse
 * (forced_phot_sparse.py:583-709, solver_inner :342-554).  One problem per
 * (grid, band).  Problem q owns sources [src_off[q], src_off[q+1]) of the
 * flattened per-problem source arrays and the grid window
 * win[q] = (band, x0, y0, h, w) of the (n_bands, H, W) image arrays, whose
 * pixel (0, 0) sits at full-image pixel (img_x0, img_y0).
 *   data, invvar (n_bands, H, W) f32; good (n_bands, H, W) u8;
 *   background basis X^m Y^n of order bg_order in X = (x - bg_xo) / bg_xs,
 *   Y = (y - bg_yo) / bg_ys (full-image pixel coordinates, computed in
 *   float64 and rounded to float32 as bg_polynomial.make_bg_basis does);
 *   stamps (total_src, T, T) f32 (from uph_render_stamps);
 *   ix_top, iy_top (total_src) i32 window-local stamp top-left;
 *   nbr_idx (total_src, k_max) i32 problem-local neighbour (-1 = none);
 *   outputs fluxes, flux_var (total_src) and bg, bg_var (n_prob, M) f32.
 * ``scratch`` must hold uph_band_scratch_bytes(n_prob, max_src, bg_order).
 */
/*
 * Wing images of large extended sources for the linear solve's wing
 * peeling (solver_grid.py:1578-1626 _compute_wing_images; the final solve
 * fits data - wing, which by linearity is the solve of data minus the
 * solve of the wing image).  Problem q writes its window win[q] = (band,
 * x0, y0, h, w) of out (n_bands, H, W) f32: per pixel, the sum over its
 * sources s in [src_off[q], src_off[q+1]) of flux[s] * (big[s] - small[s]),
 * big (n, Tb, Tb) unit stamps at window-local top-left (ix[s], iy[s]) and
 * small (n, Ts, Ts) unit stamps centred on the same pixel (Tb, Ts odd).
 * Sources are summed in order; the windows of one call must be disjoint.
 */
int uph_wing_images(int n_prob, int Tb, int Ts, const int64_t* src_off, const int32_t* win,
                    const float* big, const float* small, const int32_t* ix, const int32_t* iy,
                    const float* flux, int H, int W, float* out, void* stream);

/*
 * Weighted stamp sums into disjoint windows of an image stack: the scene
 * subtraction's field-wide STAR subtract at S1 flux (replaces the per-source
 * scatter `dmb[bi, ...] -= _f * _st[...]`, island_stage.py:2040-2117) and the
 * island member add-back (replaces addback_v, island_solver.py:1795-1818,
 * applied at :2779-2789).  Window q = win[q] = (x0, y0, h, w) of band
 * band0 + bl (bl < nbc) of out (n_bands, H, W) f32 gets
 *   out += scale * sum_{k in [off[q], off[q+1])} flux[s*nbc + bl] * stamp(bl, s)
 * with s = list[k], stamp(bl, s) the T x T block s of band chunk bl of
 * stamps (nbc, n_st, T, T) f32 placed at top-left (ix[s], iy[s]) in out
 * coordinates (zero outside it).  Windows lie inside out and are disjoint;
 * max_pix is the largest h * w.  Entries are summed in list order.
 */
int uph_stamp_sum(int n_win, int max_pix, int T, int n_st, int nbc, int band0, int n_bands,
                  const int64_t* off, const int32_t* win, const int32_t* list, const int32_t* ix,
                  const int32_t* iy, const float* stamps, const float* flux, float scale, int H,
                  int W, float* out, void* stream);

/*
 * Per-grid background planes subtracted from the padded field stack
 * (replaces the host loop `dmb[bi, core] -= basis @ bg_coefs[bi]`,
 * island_stage.py:1731-1751): grid q's core window win[q] = (x0, y0, h, w)
 * in image coordinates gets, in every band, out[band][pad + y][pad + x] -=
 * sum_m bg[(q * n_bands + band) * M + m] * X^i Y^j, M = (o + 1)(o + 2) / 2
 * terms of order o = bg_order in X = (x - bg_xo) / bg_xs, Y = (y - bg_yo) /
 * bg_ys (float64, rounded to float32).  out (n_bands, Hp, Wp) f32; the
 * windows (with pad) lie inside it and are disjoint; max_pix >= h * w.
 */
int uph_subtract_bg_planes(int n_grid, int max_pix, int n_bands, int bg_order, const int32_t* win,
                           const float* bg, int pad, double bg_xo, double bg_xs, double bg_yo,
                           double bg_ys, int Hp, int Wp, float* out, void* stream);

/*
 * Island window gather (replaces _gather_jit / gather, island_solver.py:
 * 1820-1830): window q = win[q] = (sx, sy, dx, dy, side) copies the
 * side x side block at (sx, sy) of every band of src_a (n_bands, Hs, Ws) to
 * (dx, dy) of dst_a (n_bands, Hd, Wd), and likewise src_b -> dst_b when both
 * are given.  Blocks lie inside both stacks; max_side is the largest side.
 */
int uph_gather_windows(int n_win, int max_side, int n_bands, const int32_t* win, int Hs, int Ws,
                       const float* src_a, const float* src_b, int Hd, int Wd, float* dst_a,
                       float* dst_b, void* stream);

/*
 * Blank-response (empty-aperture) correction of a pass-2 grid-band solve
 * (solver_grid.py:2867-2943): for problem q with window win[q] = (band, x0,
 * y0, h, w) of the (n_bands, H, W) images, the model of its sources
 * [src_off[q], src_off[q+1]) -- flux[s] times the unit stamp s (T x T at
 * window-local top-left (ix_top[s], iy_top[s])) plus the background bg[q]
 * (order bg_order, basis as in uph_band_solve) -- the weighted residual
 * w (d - wing - model) on good pixels (wing (n_bands, H, W) where
 * use_wing[q], else none), resp = corr(w r, phi) / corr(w, phi^2) ('same'
 * correlation with the template phi[q] (T x T, clipped >= 0, unit sum);
 * NaN where the denominator is 0 or the pixel is not good) and
 * b_hat[q] = nanmedian(resp) (NaN if no value).  The caller subtracts
 * b_hat from the STAR fluxes.  work: uph_blank_work_floats(n_prob,
 * max_pix) floats, max_pix >= every h * w.
 */
size_t uph_blank_work_floats(int n_prob, int max_pix);
int uph_blank_response(int n_prob, int T, int bg_order, const int64_t* src_off,
                       const int32_t* win, int H, int W, const float* data, const float* invvar,
                       const uint8_t* good, const float* wing, const uint8_t* use_wing,
                       const float* stamps, const int32_t* ix_top, const int32_t* iy_top,
                       const float* flux, const float* bg, double bg_xo, double bg_xs,
                       double bg_yo, double bg_ys, const float* phi, float* work,
                       size_t work_floats, int max_pix, double* b_hat, void* stream);

/*
 * Confusion-noise flux error of the post-fit error model (replaces
 * _confusion_band_numeric / compute_grid_confusion, tract7dt/fit/grid_gpu/
 * error_model.py:268-325, :663-930).  Problem q = one (grid, band) window
 * win[q] = (band, x0, y0, h, w) of the (n_bands, H, W) images with sources
 * [src_off[q], src_off[q+1]) (n_src in all): the model is flux x unit stamp
 * (T x T at ix_top, iy_top) plus the background bg[q] (basis as in
 * uph_band_solve), wing images subtracted from the data where use_wing[q];
 * residual r = d - model on good pixels (0 elsewhere).  With phi[q] the
 * unit-sum band PSF at the window centre (sumsq[q] = sum(phi^2) in float32):
 * MF = corr(r, phi) / corr(good, phi^2) and Vstat = corr(var, phi^2) /
 * corr(good, phi^2)^2 ('same' correlations; var = 1/invvar on good pixels
 * with invvar > 0) where corr(good, phi^2) > 0.6 sumsq (valid).  For source s
 * at window pixel (xi[s], yi[s]): the float64 means of MF, MF^2 and Vstat over
 * the valid pixels of the square annulus radii[q] = (r_in, r_out) (clipped to
 * the window; NaN below 50 pixels), excess = max(0, <MF^2> - <MF>^2 -
 * <Vstat>), and sigma[s] = sqrt(R excess), R = 2 g^2 / ((g + sp2) sp2),
 * g = sp2[q] + sgal2[s]; NaN where no_fit[s].  work: uph_confusion_work_floats
 * (n_prob, max_pix) floats, max_pix >= every h * w; T up to 63 (LDS).
 */
size_t uph_confusion_work_floats(int n_prob, int max_pix);
int uph_confusion(int n_prob, int64_t n_src, int T, int bg_order, const int64_t* src_off,
                  const int32_t* win, int H, int W, const float* data, const float* invvar,
                  const uint8_t* good, const float* wing, const uint8_t* use_wing,
                  const float* stamps, const int32_t* ix_top, const int32_t* iy_top,
                  const float* flux, const float* bg, double bg_xo, double bg_xs, double bg_yo,
                  double bg_ys, const float* phi, const float* sumsq, const int32_t* radii,
                  const double* sp2, const int32_t* xi, const int32_t* yi, const double* sgal2,
                  const uint8_t* no_fit, float* work, size_t work_floats, int max_pix,
                  float* sigma, void* stream);

size_t uph_band_scratch_bytes(int n_prob, int max_src, int bg_order);
int uph_band_solve(int n_prob, int T, int bg_order, int k_max, const int64_t* src_off,
                   const int32_t* win, int H, int W, int img_x0, int img_y0, const float* data,
                   const float* invvar, const uint8_t* good, double bg_xo, double bg_xs, double bg_yo, double bg_ys,
                   const float* stamps, const int32_t* ix_top, const int32_t* iy_top,
                   const int32_t* nbr_idx, float reg, float* fluxes, float* flux_var, float* bg,
                   float* bg_var, void* scratch, size_t scratch_bytes, void* stream);

/*
 * Allocation-free form of uph_band_solve.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Same inputs, outputs and results as
 * uph_band_solve; the workspace is a static device array of
 * UPH_BAND_NOALLOC_BYTES per GPU (held from library load until the process
 * ends), so the call makes no allocation.  ``max_src`` is the largest
 * problem's source count; batches whose uph_band_scratch_bytes(n_prob,
 * max_src, bg_order) exceed UPH_BAND_NOALLOC_BYTES are rejected with
 * UPH_ERR_INVALID_ARGUMENT (split them).  Every call on one GPU uses the
 * same array: calls that overlap on one GPU race.
 */
#define UPH_BAND_NOALLOC_BYTES (64u << 20)
int uph_band_solve_noalloc(int n_prob, int T, int bg_order, int k_max, int max_src,
                           const int64_t* src_off, const int32_t* win, int H, int W, int img_x0,
                           int img_y0, const float* data, const float* invvar,
                           const uint8_t* good, double bg_xo, double bg_xs, double bg_yo,
                           double bg_ys, const float* stamps, const int32_t* ix_top,
                           const int32_t* iy_top, const int32_t* nbr_idx, float reg,
                           float* fluxes, float* flux_var, float* bg, float* bg_var,
                           void* stream);

/*
 * PSF diagnostics of the band flags at fitted positions (replaces
 * psf_lobe_zone and psf_uncovered_fraction, tract7dt/fit/grid_gpu/
 * band_flags.py:115-196, both of which evaluate evaluate_psf_at_batch_np,
 * psf_polynomial.py:229-270).  For source s and band b: the band's PSF
 * model stamp at (x[s], y[s]) (size[b] x size[b] native pixels, centred in
 * the planar coefs (n_bands, n_terms, T, T) f32; geo (n_bands, 4) f64 =
 * x_origin, x_scale, y_origin, y_scale; basis in f64 rounded to f32),
 * peak[b*n+s] = its maximum and vmin[b*n+s] its minimum (NaN for a
 * non-finite position); where has_nodata[b] is set, uncovered[b*n+s] = the
 * fraction of the stamp's positive part on pixels of nodata (n_bands, H, W)
 * u8 (the stamp centred on the pixel nearest (x, y), round half to even;
 * pixels beyond the image count as no data): 0 for a non-finite position
 * or a stamp touching no such pixel, 1 for a position off the image or a
 * stamp with no positive part; 0 where has_nodata[b] is clear.  nodata and
 * has_nodata are both null or both given.
 */
int uph_band_psf_flags(int n, int n_bands, int n_terms, int T, const int32_t* size,
                       const float* coefs, const double* geo, const double* x, const double* y,
                       int H, int W, const uint8_t* nodata, const uint8_t* has_nodata,
                       float* peak, float* vmin, double* uncovered, void* stream);

/*
 * Correlated-noise variance factor F_corr of the post-fit error model
 * (replaces the numerics of measure_correlated_noise_factor and
 * measure_correlated_noise_factor_residual, tract7dt/fit/grid_gpu/
 * error_model.py:933-1113, :1116-1302; the host combines the passes below
 * as those functions do).  Maps: float64 images packed back to back, map k
 * of map_hw[k] = (h, w) pixels starting at map_off[k]; max_pix >= every
 * h * w.
 *
 * uph_fcorr_load: v[k] = float64(src[src_off[k] + y * src_stride[k] + x]),
 * g0 = finite & good (good_u8 != 0 at the same index, or good_f32 > 0.5;
 * exactly one of the two is given); n_good[k] = count of g0.
 */
int uph_fcorr_load(int n_maps, int64_t max_pix, const int64_t* map_off, const int32_t* map_hw,
                   const float* src, const int64_t* src_off, const int32_t* src_stride,
                   const uint8_t* good_u8, const float* good_f32, double* v, uint8_t* g0,
                   int64_t* n_good, void* stream);
/*
 * uph_fcorr_select: out[q] = the q_rank[q]-th smallest (0-based) of v over
 * g0 of map q_map[q], or of |v - q_center[q]| where q_center[q] is not NaN
 * (exact: radix select on order-preserving keys).  The medians of np.median
 * are the mean of ranks (n-1)/2 and n/2.  work: uph_fcorr_select_work_words
 * (n_q) 32-bit words.
 */
size_t uph_fcorr_select_work_words(int n_q);
int uph_fcorr_select(int n_q, int64_t max_pix, const int32_t* q_map, const int64_t* q_rank,
                     const double* q_center, const int64_t* map_off, const int32_t* map_hw,
                     const double* v, const uint8_t* g0, uint32_t* work, size_t work_words,
                     double* out, void* stream);
/*
 * uph_fcorr_clean: struct = !g0 | (symmetric ? |v - med| > clip mad :
 * v > med + clip mad | v < med - low mad) with med[k], mad[k];
 * clean = no struct pixel within |dy| + |dx| <= 2 (scipy binary_dilation by
 * the 4-neighbour cross, iterations = 2, border not struct), inverted.
 */
int uph_fcorr_clean(int n_maps, int64_t max_pix, const int64_t* map_off, const int32_t* map_hw,
                    const double* v, const uint8_t* g0, const double* med, const double* mad,
                    double clip, double low, int symmetric, uint8_t* clean, void* stream);
/*
 * uph_fcorr_tile_stats: tile t = the ts x ts block at tile_yx[t] = (y0, x0)
 * of map tile_map[t] (clipped at the map edge): stats[t] = (count, sum, sum
 * of squared deviations from the tile mean) of v over mask.
 */
int uph_fcorr_tile_stats(int n_tiles, const int32_t* tile_map, const int32_t* tile_yx, int ts,
                         const int64_t* map_off, const int32_t* map_hw, const double* v,
                         const uint8_t* mask, double* stats, void* stream);
/*
 * uph_fcorr_mask: with tile t = tile_base[k] + (y / ts) * ceil(w / ts) +
 * x / ts: m = mask & keep[t]; v = m ? v - tile_mean[t] : 0; mask = m.
 */
int uph_fcorr_mask(int n_maps, int64_t max_pix, const int64_t* map_off, const int32_t* map_hw,
                   const int32_t* tile_base, int ts, const uint8_t* keep, const double* tile_mean,
                   double* v, uint8_t* mask, void* stream);
/*
 * uph_fcorr_autocov: circular autocovariance of (x, m) over the half window
 * of lags: A(d) = sum_i x_i x_{i+d}, K(d) = sum_i m_i m_{i+d} (indices mod
 * h, w) for lag (0, dx), dx = 0..L at index dx, and (dy, dx), dy = 1..L,
 * dx = -L..L at index (L + 1) + (dy - 1)(2L + 1) + dx + L; out[k] =
 * (A[nlag], K[nlag]) with nlag = uph_fcorr_nlag(L); L <= 56, L < h / 2 and
 * w / 2.  bpm workgroups per map each write partial sums (partial: n_maps *
 * bpm * 2 * nlag doubles), reduced in a fixed order.
 */
int uph_fcorr_nlag(int L);
int uph_fcorr_autocov(int n_maps, int bpm, int L, const int64_t* map_off, const int32_t* map_hw,
                      const double* x, const uint8_t* m, double* partial, size_t partial_doubles,
                      double* out, void* stream);

/*
 * PSF model build from training stars (replaces the numerics of
 * _extract_centered_stamps_gw, _annulus_median, _gw_centroid_vec,
 * fourier_shift, _fit_polynomial_one_iter, _per_star_chi2 and the pixelwise
 * median of _replace_model_wings_with_median, tract7dt/fit/grid_gpu/
 * psf_from_stars.py:1478-2147).  image (H, W) f32, bad (H, W) u8.
 *
 * uph_psf_cutouts: star k at the pixel (ox[k], oy[k]): status[k] = 1 if the
 * (2 half + 1)^2 box leaves the image, 2 if it holds a bad pixel, else 0 with
 * bkg[k] = the sigma-clipped annulus median (float32 values, finite and not
 * bad, with rin^2 <= r^2 <= rout^2; median, 1.4826 MAD, keep med +- sigma MAD,
 * the kept median; the median where the MAD is 0 or fewer than minpix are
 * kept; NaN below minpix; NaN without do_bkg) and cut[k] = box - float32(bkg)
 * (the box where bkg is NaN).  half <= 30, (2 ceil(rout) + 1)^2 <= 4096, minpix >= 1.
 */
int uph_psf_cutouts(int n, int H, int W, const float* image, const uint8_t* bad,
                    const int32_t* ox, const int32_t* oy, int half, double rin, double rout,
                    double sigma, int minpix, int do_bkg, uint8_t* status, double* bkg, float* cut,
                    void* stream);
/*
 * uph_psf_centroids: Gaussian-weighted centroid of max(cut[k], 0) (weight
 * FWHM fwhm) from the box centre, iters iterations; hist[(k * iters + i) * 3]
 * = (x, y, largest coordinate move) after iteration i (the centre kept
 * where the weighted sum is 0).
 */
int uph_psf_centroids(int n, int half, double fwhm, int iters, const float* cut, double* hist,
                      void* stream);
/*
 * uph_psf_stamps: star k centred at (fx[k], fy[k]): the box at the rounded
 * centre (round half to even; status 1 off the image, 2 on a bad pixel, 3 for
 * a non-finite centre), minus its background (bkg[k] where the rounded centre
 * is (ox[k], oy[k]), else the annulus median there), moved by minus the
 * fractional part with the Fourier shift of the box zero-padded by pad
 * (2 half + 1 + 2 pad odd), the central psf x psf crop divided by its sum
 * into stamps[k] (status 4 where the sum is not positive and finite).
 */
int uph_psf_stamps(int n, int H, int W, const float* image, const uint8_t* bad, const int32_t* ox,
                   const int32_t* oy, const double* fx, const double* fy, const double* bkg,
                   int half, int psf, int pad, double rin, double rout, double sigma, int minpix,
                   int do_bkg, uint8_t* status, float* stamps, void* stream);
/*
 * uph_psf_polyfit: coefs[p * nt + t] = float32(sum_s pinv[t * ns + s] *
 * stamps[sel[s] * P + p]) (float64 sums; pinv the pseudo-inverse of the
 * selected stars' design matrix).  uph_psf_chi2: chi2[s] = sum_p
 * (stamps[s * P + p] - sum_t A[s * nt + t] coefs[p * nt + t])^2.  nt <= 21.
 * uph_psf_pixel_median: med[p] = median of stamps[sel[s] * P + p] over the
 * ns <= 4096 selected stamps, in float64.
 */
int uph_psf_polyfit(int ns, int nt, int P, const int32_t* sel, const double* pinv,
                    const float* stamps, float* coefs, void* stream);
int uph_psf_chi2(int n, int nt, int P, const double* A, const float* stamps, const float* coefs,
                 double* chi2, void* stream);
int uph_psf_pixel_median(int ns, int P, const int32_t* sel, const float* stamps, double* med,
                         void* stream);

#ifdef __cplusplus
}
#endif

#endif /* UNIPHOT_AMD_H */
