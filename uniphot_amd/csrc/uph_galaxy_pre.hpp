// Galaxy pre-render of the band-split island fit (float32): entry points of
// galaxy.hip used by the island launcher.
#pragma once
#include <hip/hip_runtime.h>

#include "uph_island.hpp"

namespace uph {

// Largest stamp tier the pre-render handles (its line buffer must fit LDS).
bool galaxy_prerender_supports(int T);

// Per-workgroup scratch bytes the planes kernel needs for stamp tier T.
size_t galaxy_prerender_scratch_bytes(int T);

// Floats of one galaxy's Fourier spectra (F + core, F1, Fn + dcore) for tier T.
size_t galaxy_spectra_floats(int T);

// Render the raw planes of every extended member for one evaluation pass
// (`mode` is a kSplitEval* phase): spectra launch, then the (galaxy, band)
// planes launch on its work queue `counter`; `ev_mid`, when given, is
// recorded between the two.
void galaxy_prerender(const IslandArgsPre& a, int mode, int* counter, int nblk, hipStream_t s,
                      hipEvent_t ev_mid);

}  // namespace uph
