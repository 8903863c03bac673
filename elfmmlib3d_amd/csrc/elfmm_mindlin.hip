// Mindlin harmonic-image far field (lfmm3dmindlintriatarg,
// src/lfmm3dtriamindlin_iter_1.f:41) on gfx950.
#include <hip/hip_runtime.h>

#include "elfmm_half.h"

namespace elfmm {

int mindlin_far_dev(int, double, double, int, const double *, const double *, const double *, int, const double *,
                    int, const double *, int, const double *, int, double *, int, double *, hipStream_t)
{
    return 16;   // not available yet
}

}  // namespace elfmm
