// Extended-source (EXP/DEV/SERSIC) member rendering for the island kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "uph_island.hpp"

namespace uph {

// Extra per-workgroup workspace of the galaxy renderer (bytes).
__host__ __device__ inline size_t galaxy_scratch_bytes(int T, size_t real_bytes) {
  (void)T;
  (void)real_bytes;
  return 0;
}

// Galaxy members are rejected by the launcher until the FFT renderer lands;
// this body is never reached for star-only classes.
template <typename Real>
__device__ void render_galaxy(const IslandArgs& a, int b, const float* p6, int stype, float nfz,
                              int free_n, float x0f, float y0f, Real* m_out, Real* tan_base,
                              const int* live_j, int l0, int l1, Real* ws, Real* taps, Real* bas) {
  (void)a; (void)b; (void)p6; (void)stype; (void)nfz; (void)free_n; (void)x0f; (void)y0f;
  (void)m_out; (void)tan_base; (void)live_j; (void)l0; (void)l1; (void)ws; (void)taps; (void)bas;
}

}  // namespace uph
