// Member pre-render of the band-split island fit (float32): entry points of
// prerender.hip used by the island launcher.
#pragma once
#include <hip/hip_runtime.h>

#include "uph_island.hpp"

namespace uph {

// Whether the pre-render serves stamp tier T (its line buffer must fit LDS).
bool prerender_supports(int T);

// Per-workgroup scratch bytes the planes kernel needs for stamp tier T.
size_t prerender_scratch(int T);

// Floats of one extended member's Fourier spectra for tier T.
size_t prerender_spectra_floats(int T);

// Render and shift the planes of every pre-rendered member for one
// evaluation pass (`mode` is a kSplitEval* phase): the taps/spectra launch
// (one workgroup per member), the extended-member (member, band) launch on
// its work queue `counter`, then the star (member, band) launch; `ev_spec`
// and `ev_ext`, when given, are recorded after the first and second.
void prerender_pass(const IslandArgsPre& a, int mode, int* counter, int nblk, hipStream_t s,
                    hipEvent_t ev_spec, hipEvent_t ev_ext);

// Unit stamps of the extended sources among n (source, band) items for the
// linear solve through the planes pipeline (stars are skipped).  Returns
// false, launching nothing, when stamp side T is not served (its line
// buffer does not fit LDS) or the scratch is too small for one workgroup;
// the caller then renders them another way.  n_cu sizes the grid.
bool launch_galaxy_stamps(int n, int T, int nt, const float* coefs, const float* geo,
                          const double* sersic, const int32_t* src_band, const float* x,
                          const float* y, const float* gx0, const float* gy0,
                          const int32_t* stype, const float* logre, const float* ee1,
                          const float* ee2, const float* sersic_n, float* out, char* scratch,
                          size_t scratch_bytes, int n_cu, hipStream_t s);

// Float64 member pre-render (prerender64.hip): per-workgroup scratch bytes
// for stamp tier T, doubles of one extended member's spectra block, and
// the spectra + planes launches of one evaluation pass (pbuf64, fspec64);
// `ev_spec` and `ev_lds`, when given, are recorded after the spectra launch
// and after the stars / tier-55 planes launch (the larger-tier launches
// follow it).
size_t prerender64_scratch(int T);
size_t prerender64_spectra_doubles(int T);
void prerender64_pass(const IslandArgsPre& a, int mode, int* counter, int nblk, hipStream_t s,
                      hipEvent_t ev_spec, hipEvent_t ev_lds);

}  // namespace uph
