// Island VarPro-LM argument block and per-block scratch layout, shared by
// the kernel and the C-ABI launcher.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace uph {

constexpr int kMaxMembers = 64;     // nsrc class cap (joint fits use <= 32)
constexpr int kMaxLive = 6 * kMaxMembers;

struct IslandArgs {
  int n_isl, N, B, T, nt, nb, nL_max, Hp, Wp;
  int max_steps;
  int resume;  // LM state from the outputs (see uph_lm_options)
  float reg, pos_step_max, shape_step_max, un_step_max, ftol_rel, step_tol, shape_step_tol;
  const float* dstack;
  const float* wstack;
  const float* psf_coefs;  // (nb, nt, T, T)
  const float* psf_geo;    // (nb, 4): x_origin, x_scale, y_origin, y_scale
  const double* sersic;    // Sersic table block (see uph_galaxy.hpp), may be null
  const int* ox;
  const int* oy;
  const float* x0b;
  const float* y0b;
  const int* stype;
  const float* n_frozen;
  const uint8_t* free_n;
  const uint8_t* mmask;
  const uint8_t* keep;
  const float* p0;
  const float* lower;
  const float* upper;
  const int* order;
  float* p_out;
  float* flux_out;
  float* ferr_out;
  double* chi2_out;  // float64 LM state (accept test and damping)
  double* lam_out;
  uint8_t* conv_out;
  uint8_t* stall_out;
  int* nst_out;
  int* consec_out;
  int* nrej_out;
  int* ngood_out;
  char* scratch;
  size_t scratch_per_block;
  int* counter;
  // heterogeneous batches (NULL shape = every island has the class shape)
  const int32_t* shape;    // (n_isl, 4): N, B, tier index, n_live_max
  const int64_t* mem_off;  // (n_isl) first member slot of the island
  const int64_t* flux_off; // (n_isl) first element of its (nb, N) flux block
  const float* tier_coefs[3];
  int tier_T[3];
  // band-split path: (island, band) evaluation items and per-band H, g, chi2
  int mode;                // kSplit* phase of the launch
  int n_items;
  const int32_t* item_isl;
  const int32_t* item_band;
  const int64_t* hg_off;   // (n_isl) first double of island i's band blocks
  double* hg;
  size_t hg_doubles;
  float* p_try;            // trial parameters, laid out like p_out
  int* err;                // set to 1 when an island's band blocks overflow hg
};

// Member pre-render (band-split path, float32): before each evaluation pass
// every real member of the active islands is rendered by (member, band)
// items -- stars from the PSF polynomial, extended members through the
// Fourier mixture -- and shifted to its sub-pixel position, into
// n_bands x np planes of min(T, B)^2 floats per member (the model, then the
// tangents of x, y and, for an extended member (np = 7), logre, ee1, ee2, n;
// np = 3 for a star); a plane holds the part of the stamp that lands in the
// island window (pre_side), which the evaluation kernel places.
// Only the pre-render launches see these fields; every other kernel takes
// the base IslandArgs.
struct IslandArgsPre : IslandArgs {
  int n_pre;
  int n_pre_ext;               // members [0, n_pre_ext) are extended, the rest stars
  const int32_t* pre_isl;      // (n_pre) island of the member
  const int32_t* pre_slot;     // (n_pre) member slot within the island
  const int64_t* pre_off;      // (n_pre) first float of the member's planes in pbuf
  const int64_t* pre_spec;     // (n_pre) first float of its Fourier spectra in fspec, -1 for stars
  const int32_t* pre_of_slot;  // (total member slots) pre-render index, -1 for padding
  float* pbuf;                 // shifted planes
  size_t pbuf_elems;
  float* fspec;                // per extended member: F + core, F1, Fn + dcore on the (H, pad) grid
  size_t fspec_elems;
  double* pbuf64;              // float64 islands: shifted planes (pre_off in doubles)
  double* fspec64;             // float64 islands: F, F1, Fn on the (pad, H) grid, then core,
                               // dcore, 9 ellipse-term derivatives, dn/du (pre_spec in doubles)
  double* gdt;                 // (n_pre, 9) ellipse-term derivatives of extended members
  float* gtap;                 // (n_pre, 28) Lanczos taps and tap derivatives, x then y
  size_t star_stride;          // per-workgroup scratch bytes of the star launch
  int* pre_counter;            // work queue of the planes kernel (float64: the next counter
                               // is the queue of the larger-tier planes kernel)
  // post-fit pass at the converged parameters (post != 0):
  int post;
  const int64_t* post_off;     // (n_isl) first double of the island's block in post_buf:
                               // D (n_bands, nLc, N) then Sigma (nLc, nLc)
  double* post_buf;
  size_t post_doubles;
  float* post_var;             // (member slots, 7) var x, y, logre, ee1, ee2, u, cov(ee1, ee2)
  float* post_dvar;            // flux layout: shape-marginalisation variance increment
  float* post_metrics;         // (member slots, n_bands, 6) per-band metric sums
  float* post_kappa;           // (n_isl) 1-norm condition number of the diagonally scaled H
  double* post_hg;             // optional: band-summed H (nL x nL) then g (nL) per island
  const int64_t* post_hg_off;  // (n_isl) first double of the island's block in post_hg
  int gram_bf16;               // island Gram on bfloat16 matrix cores (float32 accumulate)
  int* gram_unsettled;         // (n_isl) set to 1 where a bfloat16 refinement did not settle
  // large islands: Gram tile pairs computed by a separate many-workgroup launch
  int big_n;                   // islands with a slot (0: off)
  int big_ntp;                 // tile-pair capacity per (slot, band)
  const int32_t* big_slot;     // (n_isl) slot or -1
  const int32_t* big_isl;      // (big_n) island of each slot
  double* big_part;            // (big_n, n_bands, big_ntp, 256) tile pairs
};

// Band-split phases.  Evaluation launches run one (island, band) item per
// workgroup visit; step launches run one workgroup per island.
constexpr int kSplitEvalInit = 0;   // chi2 per band at p0
constexpr int kSplitEvalGrad = 1;   // H, g, chi2 per band at p
constexpr int kSplitEvalTry = 2;    // chi2 per band at p_try
constexpr int kSplitEvalFinal = 3;  // fluxes, flux errors, good-pixel count at p
constexpr int kSplitStepInit = 4;   // LM state from the initial chi2
constexpr int kSplitStepSolve = 5;  // damped solve -> p_try
constexpr int kSplitStepAccept = 6; // accept / reject, convergence
constexpr int kSplitEvalPost = 7;   // H, g, flux sensitivities and metric sums at p (converged)

// Passes that need the tangent planes.
__host__ __device__ inline bool split_grads(int mode) {
  return mode == kSplitEvalGrad || mode == kSplitEvalPost;
}

// Doubles of one island's post-fit block: the per-band flux sensitivities
// D (n_bands, nLc, N) then the parameter covariance (nLc, nLc).
__host__ __device__ inline size_t post_block(int nb, int nLc, int N) {
  return (size_t)nLc * ((size_t)nb * N + nLc);
}

// Planes per band of a pre-rendered member.
__host__ __device__ inline int pre_planes(bool star) { return star ? 3 : 7; }

// Side of one stored plane: the part of a T x T stamp inside a B x B window
// is at most min(T, B) on a side.  Pixel (r, c) of the stamp's window part
// [sx0, sx1) x [sy0, sy1) is stored at (r - sy0) * side + (c - sx0).
__host__ __device__ inline int pre_side(int T, int B) { return T < B ? T : B; }

// Doubles of one band block of an island with n_live_max = nLc:
// H (nL x nL, row-major with the actual nL), g (nL), chi2, one spare.
__host__ __device__ inline size_t split_band_block(int nLc) {
  return (size_t)nLc * nLc + nLc + 2;
}

__host__ __device__ inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Per-block scratch layout (bytes).  Planes are plane-major B*B arrays:
// [w, d, R, models (N), tangents (nL_max)].
struct IslandScratch {
  size_t planes, render, gram, part, small, total;
  int cmax;  // Gram columns: d + models + tangents
};

__host__ __device__ inline IslandScratch island_scratch_layout(int N, int B, int T, int nL,
                                                               size_t real_bytes) {
  IslandScratch s;
  size_t off = 0;
  s.planes = off;
  off = align256(off + (size_t)(3 + N + nL) * B * B * real_bytes);
  s.render = off;
  off = align256(off + (size_t)6 * T * T * real_bytes);
  s.cmax = 1 + N + nL;
  s.gram = off;
  off = align256(off + (size_t)s.cmax * s.cmax * 8);
  const int nt = (s.cmax + 15) / 16;
  s.part = off;  // matrix-core Gram partial tiles: tile pairs (+ pixel splits)
  off = align256(off + (size_t)(nt * (nt + 1) / 2 + 4) * 256 * 8);
  s.small = off;
  const int L = nL > N ? nL : N;
  // A_raw, LU, b, f, c, U, Gt, s, D, AD, H, g, Hd, delta, inverse-column ws, pivots
  size_t nd = (size_t)2 * N * N + 3 * N + (size_t)N * nL + (size_t)nL * nL + nL +
              (size_t)2 * nL * N + (size_t)2 * nL * nL + 2 * nL + (size_t)256 * L + 8;
  off = align256(off + nd * 8 + (size_t)(N + nL + 8) * 4);
  s.total = off;
  return s;
}

}  // namespace uph
